// conv_driver.cpp -- host driver of the 3D FFT convolution.
//
// Pass sequence for one convolution (reference: the body of
// convolution3DfftCUDAInPlace, src/convolution3Dfft.cu:399-578, which pads
// the kernel with fftShiftKernel, runs cuFFT R2C on image and kernel,
// modulateAndNormalize_kernel and cuFFT C2R):
//   K1  kernel rows   x-R2C of the ky*kz compact rows, taps placed at
//                     (x - cx) mod nx                 -> Kx  [kz][ky][nxp]
//   K2  kernel planes y-FFT of Kx columns, taps placed at (y - cy) mod ny
//                                                     -> Kxy [kz][ny][nxp]
//   P1  image x-R2C                                   -> S   [nz][ny][nxp]
//   P2  y-FFT of S in place
//   P3  z pass: FFTz(S) * FFTz(Kxy placed at (z - cz) mod nz) / N, IFFTz
//   P4  inverse y-FFT of S in place
//   P5  x-C2R of S into the output volume
// The kernel is never materialised at image size.
#include "conv_driver.hpp"
#include "../../include/fourierconvolution_amd.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <tuple>

#include "fft_static.hpp"
#include "kernels_long.hpp"
#include "static_kernels.hpp"

namespace fcamd {

// ---------------------------------------------------------------------------
// plans and twiddle tables
// ---------------------------------------------------------------------------
int radix_plan_into(int L, int* rad, int cap) {
  int k = 0, n = L;
  auto put = [&](int r) {
    if (k < cap) rad[k] = r;
    ++k;
  };
  while (n % 8 == 0) { put(8); n /= 8; }
  if (n % 4 == 0) { put(4); n /= 4; }
  if (n % 2 == 0) { put(2); n /= 2; }
  for (int p : {3, 5, 7})
    while (n % p == 0) { put(p); n /= p; }
  for (int p = 11; (long)p * p <= n; p += 2)
    while (n % p == 0) { put(p); n /= p; }
  if (n > 1) put(n);
  return k;
}

std::vector<int> radix_plan(int L) {
  int rad[64];
  const int k = radix_plan_into(L, rad, 64);
  return std::vector<int>(rad, rad + (k < 64 ? k : 64));
}

int stage_table_segments(int L, TwSegment* seg, int* total) {
  const StaticPlanData p = static_plan(L);
  *total = 0;
  if (p.n <= 0) return 0;
  int n = 0;
  for (int rev = 0; rev < 2; ++rev) {
    int ns = 1;
    for (int i = 0; i < p.n; ++i) {
      const int r = rev ? p.r[p.n - 1 - i] : p.r[i];
      if (n < kMaxTwSegments) seg[n] = TwSegment{*total, ns, r, even_up(r)};
      ++n;
      *total += ns * even_up(r);
      ns *= r;
    }
  }
  return n;
}

namespace {

std::mutex g_tw_mu;
std::map<std::pair<int, int>, float2*> g_tw;  // (device, L) -> table

float2* twiddle_table(int dev, int L) {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(dev, L);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  std::vector<float2> h(L);
  for (int j = 0; j < L; ++j) {
    // exact integer j < L keeps the angle in [0, 2pi); FP64 then one rounding
    const double a = -2.0 * M_PI * (double)j / (double)L;
    h[j] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2* d = nullptr;
  FC_HIP_CHECK(hipMalloc(&d, sizeof(float2) * (size_t)L));
  FC_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(float2) * (size_t)L, hipMemcpyHostToDevice));
  g_tw[key] = d;
  return d;
}

// Per-stage tables of the compile-time plan of L, forward then reversed
// stage order (layout in fft_static.hpp); empty when L has no static plan.
std::vector<float2> static_stage_table(int L) {
  TwSegment seg[kMaxTwSegments];
  int total = 0;
  const int n = stage_table_segments(L, seg, &total);
  FC_REQUIRE(n <= kMaxTwSegments, kInternal, "too many twiddle stages");
  std::vector<float2> out((size_t)total);
  for (int i = 0; i < n; ++i) {
    const TwSegment& g = seg[i];
    const long span = (long)g.ns * g.r;
    for (long k = 0; k < g.ns; ++k)
      for (int j = 0; j < g.rp; ++j) {
        float2& e = out[(size_t)g.offset + (size_t)k * g.rp + j];
        if (j >= g.r) {
          e = make_float2(1.f, 0.f);
          continue;
        }
        const long m = ((long)j * k) % span;
        const double a = -2.0 * M_PI * (double)m / (double)span;
        e = make_float2((float)std::cos(a), (float)std::sin(a));
      }
  }
  return out;
}

std::map<std::pair<int, int>, float2*> g_stw;  // (device, L) -> stage tables

float2* stage_table(int dev, int L) {
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto key = std::make_pair(dev, L);
  auto it = g_stw.find(key);
  if (it != g_stw.end()) return it->second;
  std::vector<float2> h = static_stage_table(L);
  float2* d = nullptr;
  if (!h.empty()) {
    FC_HIP_CHECK(hipMalloc(&d, sizeof(float2) * h.size()));
    FC_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
  }
  g_stw[key] = d;
  return d;
}

int current_device() {
  int dev = 0;
  FC_HIP_CHECK(hipGetDevice(&dev));
  return dev;
}

std::once_flag g_attr_once;

}  // namespace

namespace {
thread_local const FftDescSet* t_desc_override = nullptr;
}  // namespace

const float2* long_twiddle_table(int L, hipStream_t st) {
  const int dev = current_device();
  bool have = false;
  {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    have = g_tw.count(std::make_pair(dev, L)) != 0;
  }
  if (!have && st != nullptr) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FC_HIP_CHECK(hipStreamIsCapturing(st, &cs));
    FC_REQUIRE(cs == hipStreamCaptureStatusNone, kInvalidArgument,
               "transform length " + std::to_string(L) +
                   " is used for the first time inside a stream capture; run one call of "
                   "this shape before capturing");
  }
  return twiddle_table(dev, L);
}

void set_fft_desc_override(const FftDescSet* set) { t_desc_override = set; }

namespace {

// Bluestein's chirp-z replaces plans with a prime factor above this (each
// output of a direct p-point stage costs p complex multiply-adds)
constexpr int kBluesteinPrime = 31;
// ... when the convolution length fits the smallest tile of the passes
// (the runtime-length z pass: three buffers in 160 KiB)
constexpr int kBluesteinMaxM = 6826;

int largest_prime_factor(int n) {
  int best = 1;
  for (int p = 2; (long)p * p <= n; ++p)
    while (n % p == 0) {
      best = p;
      n /= p;
    }
  return n > 1 ? n : best;
}

int smooth_at_least(int n) {  // smallest 2^a 3^b 5^c 7^d >= n
  for (int m = n;; ++m) {
    int r = m;
    for (int f : {2, 3, 5, 7})
      while (r % f == 0) r /= f;
    if (r == 1) return m;
  }
}

// FP64 forward DFT of a 7-smooth length, recursive decimation in time (host)
void host_dft64(std::vector<std::complex<double>>& a) {
  const size_t n = a.size();
  if (n <= 1) return;
  size_t p = 2;
  while (n % p) ++p;
  const size_t m = n / p;
  std::vector<std::vector<std::complex<double>>> sub(p, std::vector<std::complex<double>>(m));
  for (size_t j = 0; j < m; ++j)
    for (size_t r = 0; r < p; ++r) sub[r][j] = a[j * p + r];
  for (auto& v : sub) host_dft64(v);
  for (size_t k = 0; k < n; ++k) {
    std::complex<double> acc = 0.0;
    for (size_t r = 0; r < p; ++r) {
      const double ang = -2.0 * M_PI * (double)((r * k) % n) / (double)n;
      acc += sub[r][k % m] * std::complex<double>(std::cos(ang), std::sin(ang));
    }
    a[k] = acc;
  }
}

// chirp[n] = exp(-i pi n^2 / L) (n < L) and bfft = DFT_M(b) / M with
// b[m] = conj(chirp[|m|]) for |m| < L placed at m mod M, zero elsewhere
void bluestein_tables(int L, int M, std::vector<float2>* chirp, std::vector<float2>* bfft) {
  std::vector<std::complex<double>> w(L), b(M, 0.0);
  for (int n = 0; n < L; ++n) {
    const long long q = ((long long)n * n) % (2LL * L);  // exact angle
    const double ang = -M_PI * (double)q / (double)L;
    w[n] = std::complex<double>(std::cos(ang), std::sin(ang));
  }
  for (int n = 0; n < L; ++n) {
    b[n] = std::conj(w[n]);
    if (n > 0) b[M - n] = std::conj(w[n]);
  }
  host_dft64(b);
  chirp->resize(L);
  bfft->resize(M);
  for (int n = 0; n < L; ++n) (*chirp)[n] = make_float2((float)w[n].real(), (float)w[n].imag());
  for (int k = 0; k < M; ++k)
    (*bfft)[k] = make_float2((float)(b[k].real() / M), (float)(b[k].imag() / M));
}

float2* upload_table(const std::vector<float2>& h) {
  float2* d = nullptr;
  FC_HIP_CHECK(hipMalloc(&d, sizeof(float2) * h.size()));
  FC_HIP_CHECK(hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
  return d;
}

std::map<std::pair<int, int>, FftDesc> g_desc;  // (device, L) -> descriptor

struct ChirpKey {
  int dev, L, M, M1;
  bool operator<(const ChirpKey& o) const {
    return std::tie(dev, L, M, M1) < std::tie(o.dev, o.L, o.M, o.M1);
  }
};
std::map<ChirpKey, std::pair<const float2*, const float2*>> g_chirp;

}  // namespace

int direct_stage_prime(int L) {
  if (L < 2) return 0;
  const int p = largest_prime_factor(L);
  if (p <= kBluesteinPrime || smooth_at_least(2 * L - 1) <= kBluesteinMaxM) return 0;
  return p;
}

void long_chirp_tables(int L, int M, int M1, hipStream_t st, const float2** chirp,
                       const float2** bfft) {
  FC_REQUIRE(M >= 2 * L - 1 && M1 >= 0 && (M1 == 0 || M % M1 == 0), kInternal,
             "bad chirp-z length");
  const ChirpKey key{current_device(), L, M, M1};
  {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_chirp.find(key);
    if (it != g_chirp.end()) {
      *chirp = it->second.first;
      *bfft = it->second.second;
      return;
    }
  }
  if (st != nullptr) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FC_HIP_CHECK(hipStreamIsCapturing(st, &cs));
    FC_REQUIRE(cs == hipStreamCaptureStatusNone, kInvalidArgument,
               "transform length " + std::to_string(L) +
                   " is used for the first time inside a stream capture; run one call of "
                   "this shape before capturing");
  }
  std::vector<float2> c, bf;
  bluestein_tables(L, M, &c, &bf);
  // the forward four-step M = M1 * M2 leaves frequency k1 + M1 k2 at
  // position k1 M2 + k2 (M1 = 0: one tile, natural order)
  std::vector<float2> perm(bf);
  if (M1 > 0) {
    const int M2 = M / M1;
    for (int k1 = 0; k1 < M1; ++k1)
      for (int k2 = 0; k2 < M2; ++k2)
        perm[(size_t)k1 * M2 + k2] = bf[(size_t)k1 + (size_t)M1 * k2];
  }
  float2* dc = upload_table(c);
  float2* db = upload_table(perm);
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto ins = g_chirp.emplace(key, std::make_pair((const float2*)dc, (const float2*)db));
  if (!ins.second) {  // another thread made them first: keep one
    (void)hipFree(dc);
    (void)hipFree(db);
  }
  *chirp = ins.first->second.first;
  *bfft = ins.first->second.second;
}

FftDesc fft_desc(int L, hipStream_t st) {
  if (const FftDescSet* o = t_desc_override) {
    for (int i = 0; i < o->n; ++i)
      if (o->d[i].L == L) return o->d[i];
    throw Error{kInternal, "transform length without prepared tables"};
  }
  const int dev = current_device();
  const auto key = std::make_pair(dev, L);
  {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    auto it = g_desc.find(key);
    if (it != g_desc.end()) return it->second;
  }
  if (st != nullptr) {
    // creating the tables allocates and copies synchronously, which a
    // stream capture cannot record: refuse instead of breaking the capture
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FC_HIP_CHECK(hipStreamIsCapturing(st, &cs));
    FC_REQUIRE(cs == hipStreamCaptureStatusNone, kInvalidArgument,
               "transform length " + std::to_string(L) +
                   " is used for the first time inside a stream capture; run one call of "
                   "this shape (or create a plan) before capturing");
  }
  FftDesc d{};
  d.L = L;
  int M = 0;
  if (largest_prime_factor(L) > kBluesteinPrime) {
    const int m = smooth_at_least(2 * L - 1);
    if (m <= kBluesteinMaxM) M = m;
  }
  const int plan_len = M > 0 ? M : L;
  d.nst = radix_plan_into(plan_len, d.rad, kMaxStages);
  FC_REQUIRE(d.nst <= kMaxStages, kUnsupportedSize, "FFT plan too deep");
  d.tw = twiddle_table(dev, plan_len);
  if (M > 0) {
    std::vector<float2> chirp, bfft;
    bluestein_tables(L, M, &chirp, &bfft);
    d.M = M;
    d.chirp = upload_table(chirp);
    d.bfft = upload_table(bfft);
  } else {
    d.stw = stage_table(dev, L);
  }
  std::lock_guard<std::mutex> lk(g_tw_mu);
  auto ins = g_desc.emplace(key, d);
  if (!ins.second) {  // another thread made it first: keep one
    if (d.chirp) (void)hipFree(const_cast<float2*>(d.chirp));
    if (d.bfft) (void)hipFree(const_cast<float2*>(d.bfft));
  }
  return ins.first->second;
}

// ---------------------------------------------------------------------------
// generic tile sizes (bytes of LDS per workgroup)
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kLdsMax = 160 * 1024;
constexpr size_t kLdsTarget = 64 * 1024;

int x_pairs_per_block(int L) {
  const size_t per = 2 * sizeof(float2) * (size_t)(L + 1);
  int p = (int)(32768 / per);
  if (p < 1) p = 1;
  if (p > 8) p = 8;
  return p;
}

int col_tile(int L, int nbuf) {
  const size_t per = (size_t)nbuf * sizeof(float2) * (size_t)L;
  int tx = 16;
  while (tx > 1 && per * tx > kLdsTarget) tx /= 2;
  return tx;
}
}  // namespace

void validate_shape(const ConvShape& s) {
  FC_REQUIRE(s.nx >= 1 && s.ny >= 1 && s.nz >= 1, kInvalidArgument,
             "image dimensions must be >= 1");
  FC_REQUIRE(s.kx >= 1 && s.ky >= 1 && s.kz >= 1, kInvalidArgument,
             "kernel dimensions must be >= 1");
  FC_REQUIRE(s.kx <= s.nx && s.ky <= s.ny && s.kz <= s.nz, kInvalidArgument,
             "kernel extent exceeds image extent on some axis");
  FC_REQUIRE(s.cx >= 0 && s.cx < s.kx && s.cy >= 0 && s.cy < s.ky && s.cz >= 0 && s.cz < s.kz,
             kInvalidArgument, "kernel centre outside the kernel");
  // LDS capacity of the runtime-length kernels
  FC_REQUIRE(2 * sizeof(float2) * (size_t)(s.nx + 1) <= kLdsMax, kUnsupportedSize,
             "x extent too large for a single-tile FFT");
  FC_REQUIRE(2 * sizeof(float2) * (size_t)s.ny <= kLdsMax, kUnsupportedSize,
             "y extent too large for a single-tile FFT");
  FC_REQUIRE(3 * sizeof(float2) * (size_t)s.nz <= kLdsMax, kUnsupportedSize,
             "z extent too large for a single-tile FFT");
  FC_REQUIRE(s.ny <= 65536 && s.nz <= 65536, kUnsupportedSize, "y/z extent exceeds grid limits");
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

WorkspaceLayout workspace_layout(const ConvShape& s) {
  const size_t nxp = (size_t)padded_bins(s.nx);
  WorkspaceLayout w;
  w.spectrum = align256(sizeof(float2) * nxp * s.ny * s.nz);
  w.krows = align256(sizeof(float2) * nxp * s.ky * s.kz);
  w.kplanes = align256(sizeof(float2) * nxp * s.ny * s.kz);
  return w;
}

// ---------------------------------------------------------------------------
// per-pass event timing
// ---------------------------------------------------------------------------
namespace {
std::mutex g_pt_mu;
bool g_pt_on = false;
struct PassEvents {
  int pass;
  hipEvent_t a, b;
};
std::vector<PassEvents> g_pt_events;

class PassScope {
 public:
  PassScope(int pass, hipStream_t st) : st_(st) {
    std::lock_guard<std::mutex> lk(g_pt_mu);
    if (!g_pt_on) return;
    rec_.pass = pass;
    FC_HIP_CHECK(hipEventCreate(&rec_.a));
    FC_HIP_CHECK(hipEventCreate(&rec_.b));
    FC_HIP_CHECK(hipEventRecord(rec_.a, st_));
    active_ = true;
  }
  ~PassScope() {
    if (!active_) return;
    (void)hipEventRecord(rec_.b, st_);
    std::lock_guard<std::mutex> lk(g_pt_mu);
    g_pt_events.push_back(rec_);
  }

 private:
  hipStream_t st_;
  PassEvents rec_{};
  bool active_ = false;
};
}  // namespace

const char* pass_name(int p) {
  static const char* names[kNumPasses] = {"kernel_rows_xfwd", "kernel_planes_yfwd", "image_xfwd",
                                          "image_yfwd",       "zfused",             "image_yinv",
                                          "image_xinv"};
  return (p >= 0 && p < kNumPasses) ? names[p] : "?";
}

void pass_timing_enable(bool on) {
  std::lock_guard<std::mutex> lk(g_pt_mu);
  g_pt_on = on;
}

bool pass_timing_enabled() {
  std::lock_guard<std::mutex> lk(g_pt_mu);
  return g_pt_on;
}

void pass_timing_collect(double* ms, int* count) {
  std::vector<PassEvents> ev;
  {
    std::lock_guard<std::mutex> lk(g_pt_mu);
    ev.swap(g_pt_events);
  }
  for (int p = 0; p < kNumPasses; ++p) {
    ms[p] = 0.0;
    count[p] = 0;
  }
  for (auto& e : ev) {
    float t = 0.f;
    FC_HIP_CHECK(hipEventSynchronize(e.b));
    FC_HIP_CHECK(hipEventElapsedTime(&t, e.a, e.b));
    ms[e.pass] += t;
    count[e.pass] += 1;
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
}

void driver_init() {
  std::call_once(g_attr_once, [] {
    generic_kernels_allow_large_lds();
    static_kernels_allow_large_lds();
    long_kernels_allow_large_lds();
  });
}

namespace {

void ensure_kernel_attributes() { driver_init(); }

ColArgs spectrum_column_args(const ConvShape& s, const FftDesc& fy, float2* S) {
  const long nxp = padded_bins(s.nx);
  ColArgs ya{};
  ya.fft = fy;
  ya.tx = col_tile(fy.buf(), 2);
  ya.nbins = half_bins(s.nx);
  ya.src = S;
  ya.dst = S;
  ya.src_outer = ya.dst_outer = (long)s.ny * nxp;
  ya.src_elem = ya.dst_elem = nxp;
  ya.nvalid = s.ny;
  ya.center = 0;
  return ya;
}

}  // namespace

void pass_x_forward(const ConvShape& s, const float* src, float2* S, const float* kernel,
                    float2* Kx, hipStream_t st, const VPad* vp) {
  ensure_kernel_attributes();
  const int nb = half_bins(s.nx);
  const long nxp = padded_bins(s.nx);
  const FftDesc fx = fft_desc(s.nx, st);
  const long nrows = (long)s.ny * s.nz;
  // x-R2C of the image rows (and of the kernel rows, in the same launch
  // when a compile-time kernel serves nx)
  {
    PassScope ps(kPassXForward, st);
    if (!static_xfwd(s, fx, src, S, kernel, Kx, st, vp)) {
      XPassArgs a{};
      a.fft = fx;
      a.nrows = nrows;
      a.nbins = nb;
      a.pairs_per_block = x_pairs_per_block(fx.buf());
      a.src = src;
      a.dst = S;
      a.src_row = s.nx;
      a.dst_row = nxp;
      if (vp) a.vp = *vp;
      if (src) launch_gk_xfwd(a, (nrows + 1) / 2, st);
      if (kernel) {
        PassScope pk(kPassKernelRows, st);
        XPassArgs k = a;
        k.vp.on = 0;
        k.nrows = (long)s.ky * s.kz;
        k.src = kernel;
        k.dst = Kx;
        k.src_row = s.kx;
        k.kshift = 1;
        k.klen = s.kx;
        k.kcenter = s.cx;
        launch_gk_xfwd(k, (k.nrows + 1) / 2, st);
      }
    }
  }
}

void pass_y_forward(const ConvShape& s, float2* S, const float2* Kx, float2* Kxy, hipStream_t st) {
  ensure_kernel_attributes();
  const int nb = half_bins(s.nx);
  const long nxp = padded_bins(s.nx);
  const FftDesc fy = fft_desc(s.ny, st);
  const bool kernel = Kx != nullptr;
  // y-FFT of the spectrum (and kernel rows -> kernel planes)
  {
    PassScope ps(kPassYForward, st);
    if (!static_ycol(s, fy, S, -1, Kx, Kxy, st)) {
      if (S && s.ny > 1) launch_gk_col(spectrum_column_args(s, fy, S), -1, s.nz, st);
      if (kernel) {
        PassScope pk(kPassKernelPlanes, st);
        ColArgs a{};
        a.fft = fy;
        a.tx = col_tile(fy.buf(), 2);
        a.nbins = nb;
        a.src = Kx;
        a.dst = Kxy;
        a.src_outer = (long)s.ky * nxp;
        a.src_elem = nxp;
        a.dst_outer = (long)s.ny * nxp;
        a.dst_elem = nxp;
        a.nvalid = s.ky;
        a.center = s.cy;
        launch_gk_col(a, -1, s.kz, st);
      }
    }
  }
}

void forward_xy(const ConvShape& s, const float* src, float2* S, const float* kernel, float2* Kx,
                float2* Kxy, hipStream_t st, const VPad* vp) {
  pass_x_forward(s, src, S, kernel, Kx, st, vp);
  pass_y_forward(s, S, kernel ? Kx : nullptr, Kxy, st);
}

void z_pass(const ConvShape& s, float2* S, const float2* operand, bool dense, bool conj,
            hipStream_t st, int out_planes, int out_shift) {
  ensure_kernel_attributes();
  const long nxp = padded_bins(s.nx);
  PassScope ps(kPassZFused, st);
  ZArgs a{};
  a.fft = fft_desc(s.nz, st);
  a.tx = col_tile(a.fft.buf(), 3);
  a.nbins = half_bins(s.nx);
  a.s = S;
  a.s_row = nxp;
  a.s_plane = (long)s.ny * nxp;
  a.k = operand;
  a.k_row = nxp;
  a.k_plane = (long)s.ny * nxp;
  a.nvalid = dense ? s.nz : s.kz;
  a.center = dense ? 0 : s.cz;
  a.conj = conj ? 1 : 0;
  a.scale = (float)(1.0 / (double)s.voxels());
  // 0 keeps the plain in-place store; a shift always takes the rotated one
  a.out_planes = (out_planes < s.nz || out_shift != 0) ? std::min(out_planes, s.nz) : 0;
  a.out_shift = out_shift;
  FC_REQUIRE(a.out_shift >= 0 && a.out_shift < s.nz, kInternal, "z-pass output shift");
  if (!static_zfused(s, a, st)) launch_gk_zfused(a, s.ny, st);
}

void pass_y_inverse(const ConvShape& s, float2* S, hipStream_t st) {
  ensure_kernel_attributes();
  const FftDesc fy = fft_desc(s.ny, st);
  if (s.ny > 1) {
    PassScope ps(kPassYInverse, st);
    if (!static_ycol(s, fy, S, +1, nullptr, nullptr, st))
      launch_gk_col(spectrum_column_args(s, fy, S), +1, s.nz, st);
  }
}

void inverse_yx(const ConvShape& s, float2* S, float* dst, hipStream_t st, const VPad* vp) {
  pass_y_inverse(s, S, st);
  pass_x_inverse(s, S, dst, st, vp);
}

void pass_x_inverse(const ConvShape& s, const float2* S, float* dst, hipStream_t st,
                    const VPad* vp) {
  ensure_kernel_attributes();
  const FftDesc fx = fft_desc(s.nx, st);
  PassScope ps(kPassXInverse, st);
  if (!static_xinv(s, fx, S, dst, st, vp)) {
    const long nrows = (long)s.ny * s.nz;
    XPassArgs a{};
    a.fft = fx;
    a.nrows = nrows;
    a.nbins = half_bins(s.nx);
    a.pairs_per_block = x_pairs_per_block(fx.buf());
    a.csrc = S;
    a.rdst = dst;
    a.src_row = padded_bins(s.nx);
    a.dst_row = s.nx;
    if (vp) a.vp = *vp;
    launch_gk_xinv(a, (nrows + 1) / 2, st);
  }
}

namespace {
PlaneArgs plane_args(const ConvShape& s, int dir, hipStream_t st) {
  PlaneArgs a{};
  a.fx = fft_desc(s.nx, st);
  a.fy = fft_desc(s.ny, st);
  a.nz = s.nz;
  a.nxp = padded_bins(s.nx);
  if (plane_split(s, dir)) {
    const FftDesc h = fft_desc(plane_half_length(s, dir), st);
    FC_REQUIRE(h.stw != nullptr, kInternal, "half-length plan without a stage table");
    (dir < 0 ? a.hx_stw : a.hy_stw) = h.stw;
  }
  return a;
}
}  // namespace

int table_lengths(const ConvShape& s, int* L) {
  int n = 0;
  auto add = [&](int len) {
    for (int i = 0; i < n; ++i)
      if (L[i] == len) return;
    L[n++] = len;
  };
  add(s.nx);
  add(s.ny);
  add(s.nz);
  for (int dir : {-1, +1})
    if (plane_split(s, dir)) add(plane_half_length(s, dir));
  return n;
}

void plane_forward(const ConvShape& s, const float* src, float2* S, const float* kernel,
                   float2* Kxy, hipStream_t st) {
  ensure_kernel_attributes();
  PlaneArgs a = plane_args(s, -1, st);
  if (!src) a.nz = 0;  // the kernel's planes only
  a.src = src;
  a.spec = S;
  a.kernel = kernel;
  a.kplanes = Kxy;
  a.kx = s.kx;
  a.ky = s.ky;
  a.kz = s.kz;
  a.cx = s.cx;
  a.cy = s.cy;
  PassScope ps(kPassXForward, st);
  plane_launch(s, a, -1, st);
}

void plane_inverse(const ConvShape& s, const float2* S, float* dst, hipStream_t st) {
  ensure_kernel_attributes();
  PlaneArgs a = plane_args(s, +1, st);
  a.spec = const_cast<float2*>(S);
  a.dst = dst;
  PassScope ps(kPassXInverse, st);
  plane_launch(s, a, +1, st);
}

int fast_length(int n) {
  FC_REQUIRE(n >= 1, kInvalidArgument, "transform length must be positive");
  int smooth = n;  // smallest 7-smooth length >= n
  for (;; ++smooth) {
    int r = smooth;
    for (int f : {2, 3, 5, 7})
      while (r % f == 0) r /= f;
    if (r == 1) break;
  }
  // a 7-smooth compile-time length unless it is more than 25% longer
  auto smooth7 = [](int m) {
    for (int f : {2, 3, 5, 7})
      while (m % f == 0) m /= f;
    return m == 1;
  };
  int best = 0;
  int lens[128];
  const int cnt = fcamd_static_lengths(lens, 128);
  for (int i = 0; i < cnt && i < 128; ++i)
    if (lens[i] >= n && smooth7(lens[i]) && (best == 0 || lens[i] < best)) best = lens[i];
  return best && 4L * best <= 5L * smooth ? best : smooth;
}

PaddedProblem padded_problem(const ConvShape& s, int boundary) {
  FC_REQUIRE(boundary >= kMapPeriodic && boundary <= kMapMirror, kInvalidArgument,
             "unknown boundary mode");
  FC_REQUIRE(s.nx >= 1 && s.ny >= 1 && s.nz >= 1 && s.kx >= 1 && s.ky >= 1 && s.kz >= 1,
             kInvalidArgument, "dimensions must be positive");
  PaddedProblem p{};
  p.padded = s;
  VPad& v = p.vp;
  v.on = 1;
  const int n[3] = {s.nx, s.ny, s.nz}, k[3] = {s.kx, s.ky, s.kz}, c[3] = {s.cx, s.cy, s.cz};
  int P[3];
  for (int d = 0; d < 3; ++d) {
    v.mode[d] = boundary;
    v.n[d] = n[d];
    v.m[d] = n[d];
    v.off[d] = 0;
    // image at lo = k - 1 - c: output i reads padded i - (k - 1 - c) .. i + c,
    // all inside [0, n + k - 1) for i in [lo, lo + n), so nothing wraps
    if (k[d] > 1 && boundary != kMapPeriodic) {
      FC_REQUIRE((long)n[d] + k[d] - 1 <= (1L << 30), kUnsupportedSize, "padded extent too large");
      P[d] = fast_length(n[d] + k[d] - 1);
      v.lo[d] = k[d] - 1 - c[d];
    } else {
      P[d] = n[d];
      v.lo[d] = 0;
    }
  }
  p.padded.nx = P[0];
  p.padded.ny = P[1];
  p.padded.nz = P[2];
  for (int d = 0; d < 3; ++d) v.hi[d] = P[d];
  v.py = P[1];
  validate_shape(p.padded);
  return p;
}

void enqueue_convolution_padded(const PaddedProblem& p, float* d_out, const float* d_image,
                                const float* d_kernel, Op op, void* ws, hipStream_t st) {
  const ConvShape& s = p.padded;
  const WorkspaceLayout wl = workspace_layout(s);
  char* base = static_cast<char*>(ws);
  float2* S = reinterpret_cast<float2*>(base);
  float2* Kx = reinterpret_cast<float2*>(base + wl.spectrum);
  float2* Kxy = reinterpret_cast<float2*>(base + wl.spectrum + wl.krows);
  forward_xy(s, d_image, S, d_kernel, Kx, Kxy, st, &p.vp);
  z_pass(s, S, Kxy, false, op == Op::kCorrelate, st);
  inverse_yx(s, S, d_out, st, &p.vp);
  FC_HIP_CHECK(hipGetLastError());
}

void enqueue_convolution(const ConvShape& s, float* d_out, const float* d_image,
                         const float* d_kernel, Op op, void* ws, hipStream_t st) {
  const WorkspaceLayout wl = workspace_layout(s);
  char* base = static_cast<char*>(ws);
  float2* S = reinterpret_cast<float2*>(base);
  float2* Kx = reinterpret_cast<float2*>(base + wl.spectrum);
  float2* Kxy = reinterpret_cast<float2*>(base + wl.spectrum + wl.krows);
  if (plane_path(s, -1))
    plane_forward(s, d_image, S, d_kernel, Kxy, st);
  else
    forward_xy(s, d_image, S, d_kernel, Kx, Kxy, st);
  z_pass(s, S, Kxy, false, op == Op::kCorrelate, st);
  if (plane_path(s, +1))
    plane_inverse(s, S, d_out, st);
  else
    inverse_yx(s, S, d_out, st);
  FC_HIP_CHECK(hipGetLastError());
}

GradWorkspaceLayout grad_workspace_layout(const ConvShape& s) {
  const size_t nxp = (size_t)padded_bins(s.nx);
  GradWorkspaceLayout w;
  w.spectrum = align256(sizeof(float2) * nxp * s.ny * s.nz);
  w.corr = align256(sizeof(float) * (size_t)s.nx * s.ny * s.kz);  // the kz support planes
  return w;
}

void enqueue_grad_kernel(const ConvShape& s, float* d_dkernel, const float* d_grad_out,
                         const float* d_image, void* ws, hipStream_t st) {
  const GradWorkspaceLayout wl = grad_workspace_layout(s);
  char* base = static_cast<char*>(ws);
  float2* G = reinterpret_cast<float2*>(base);
  float2* I = reinterpret_cast<float2*>(base + wl.spectrum);
  float* U = reinterpret_cast<float*>(base + 2 * wl.spectrum);
  // u = IFFT(FFT(g) * conj(FFT(im))) / N, the circular cross-correlation
  // u[d] = sum_i g[i] im[i - d]; dk[j] = u[(j - c) mod n].  The image goes
  // through the x and y passes as their second operand (the slot the
  // kernel's rows and planes use, here with the image's extents and no
  // shift), so both volumes share each launch.
  ConvShape both = s;
  both.kx = s.nx;
  both.ky = s.ny;
  both.kz = s.nz;
  both.cx = both.cy = both.cz = 0;
  forward_xy(both, d_grad_out, G, d_image, I, I, st);
  // The gather reads u only on the z-planes (j - cz) mod nz, j < kz.  The z
  // pass writes just those, rotated by cz so plane (j - cz) mod nz lands on
  // plane j: the support is the first kz planes, inverted (y and x act
  // within a plane) as a volume of its own; U holds kz planes.
  z_pass(s, G, I, true, true, st, s.kz, s.cz);
  ConvShape sub = s;
  sub.nz = s.kz;
  const FftDesc fx = fft_desc(s.nx, st);
  if (fx.M == 0 && gather_support_c2r_lds(s.nx) <= 65536) {
    // the gather reads kx of each row's nx outputs on ky of its ny rows:
    // the y inverse of the support planes, then each support row's C2R at
    // the support positions only (no x inverse pass, no real planes)
    // (the y pass over the kz support planes takes its own tile: the tuned
    // one fits full volumes.  At 256 x 256 x 128 with 21^3 the 16-bin tile
    // leaves 170 workgroups for 256 CUs, the 8-bin tile twice as many:
    // PSF gradient 89.1 -> 86.5 us)
    const FftDesc fy = fft_desc(s.ny, st);
    if (s.ny > 1) {
      ensure_kernel_attributes();
      PassScope ps(kPassYInverse, st);
      if (!static_ycol_fill(sub, fy, G, +1, cu_count(), st))
        launch_gk_col(spectrum_column_args(sub, fy, G), +1, sub.nz, st);
    }
    launch_gather_support_c2r(s, G, padded_bins(s.nx), fx.tw, d_dkernel, st);
  } else {
    // Bluestein x (its table is not the length-nx one the direct C2R reads)
    // or a row and table past 64 KiB of LDS
    inverse_yx(sub, G, U, st);
    ConvShape g = sub;
    g.cz = 0;
    launch_gather_kernel_support(g, U, d_dkernel, st);
  }
  FC_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// pools and device guard
// ---------------------------------------------------------------------------
DevicePool& device_pool(int dev) {
  static std::mutex mu;
  static std::map<int, std::unique_ptr<DevicePool>> pools;
  std::lock_guard<std::mutex> lk(mu);
  auto& p = pools[dev];
  if (!p) p.reset(new DevicePool());
  return *p;
}

void* pool_reserve(void** p, size_t* cap, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (*p && *cap >= bytes) return *p;
  if (*p) {
    FC_HIP_CHECK(hipDeviceSynchronize());
    FC_HIP_CHECK(hipFree(*p));
    *p = nullptr;
    *cap = 0;
  }
  FC_HIP_CHECK(hipMalloc(p, bytes));
  *cap = bytes;
  return *p;
}

DeviceGuard::DeviceGuard(int dev) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    throw Error{kNoDevice, "no HIP device available"};
  FC_REQUIRE(dev >= 0 && dev < n, kInvalidArgument, "device index out of range");
  FC_HIP_CHECK(hipGetDevice(&prev_));
  if (prev_ != dev) FC_HIP_CHECK(hipSetDevice(dev));
}

DeviceGuard::~DeviceGuard() {
  int cur = -1;
  if (prev_ >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev_) (void)hipSetDevice(prev_);
}

}  // namespace fcamd
