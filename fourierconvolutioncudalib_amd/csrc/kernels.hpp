// kernels.hpp -- argument blocks and launchers of the pass kernels.
#pragma once

#include "fc_internal.hpp"

namespace fcamd {

// x pass (rows along the contiguous axis).  Forward reads real rows (`src`)
// and writes half-spectrum rows (`dst`); inverse reads `csrc`, writes `rdst`.
// Virtual padding of the x passes.  The transform runs on a padded volume
// (rows of length L = padded nx, py rows per padded plane); padded element p
// of axis d reads source element vpad_map(p - lo[d] + off[d], n[d],
// mode[d]) (zero where the map is undefined), and the inverse writes only
// output elements p - lo[d] in [0, m[d]) of a dense m[0] x m[1] x m[2]
// output.  Modes follow FCAMD_BOUNDARY_* of the public header.
constexpr int kMapPeriodic = 0;
constexpr int kMapZero = 1;
constexpr int kMapNearest = 2;
constexpr int kMapReflect = 3;  // half-sample symmetric: d c b a | a b c d | d c b a
constexpr int kMapMirror = 4;   // whole-sample symmetric: d c b | a b c d | c b a

struct VPad {
  int on;      // 0: dense rows, no virtual padding
  int mode[3];  // x, y, z
  int n[3];     // source extents
  int lo[3];    // padded position of element 0
  int off[3];   // source index offset (slab tiling)
  int m[3];     // output extents of the cropped inverse
  int hi[3];    // padded positions >= hi read zero (a slab's window)
  int py;       // padded rows per plane
};

__host__ __device__ inline int vpad_map(int i, int n, int mode, bool* ok) {
  *ok = true;
  switch (mode) {
    case kMapZero:
      *ok = i >= 0 && i < n;
      return *ok ? i : 0;
    case kMapNearest:
      return i < 0 ? 0 : (i >= n ? n - 1 : i);
    case kMapReflect: {
      const int per = 2 * n;
      int r = i % per;
      if (r < 0) r += per;
      return r < n ? r : per - 1 - r;
    }
    case kMapMirror: {
      if (n == 1) return 0;
      const int per = 2 * n - 2;
      int r = i % per;
      if (r < 0) r += per;
      return r < n ? r : per - r;
    }
    default: {
      int r = i % n;
      if (r < 0) r += n;
      return r;
    }
  }
}

// source index of padded position p of axis d; *ok false where it reads zero
__host__ __device__ inline int vpad_src(const VPad& v, int d, int p, bool* ok) {
  if (p >= v.hi[d]) {
    *ok = false;
    return 0;
  }
  return vpad_map(p - v.lo[d] + v.off[d], v.n[d], v.mode[d], ok);
}

// source row of padded row r (row index over py x padded nz); false when the
// row reads zero
__host__ __device__ inline bool vpad_row(const VPad& v, long r, long* src_row) {
  const int yp = (int)(r % v.py), zp = (int)(r / v.py);
  bool oky, okz;
  const int ys = vpad_src(v, 1, yp, &oky);
  const int zs = vpad_src(v, 2, zp, &okz);
  *src_row = (long)zs * v.n[1] + ys;
  return oky && okz;
}

// output row of padded row r under the crop; false when cropped away
__host__ __device__ inline bool vpad_out_row(const VPad& v, long r, long* out_row) {
  const int yo = (int)(r % v.py) - v.lo[1], zo = (int)(r / v.py) - v.lo[2];
  *out_row = (long)zo * v.m[1] + yo;
  return yo >= 0 && yo < v.m[1] && zo >= 0 && zo < v.m[2];
}

struct XPassArgs {
  FftDesc fft;          // length nx
  long nrows;           // rows to transform
  int nbins;            // nx/2 + 1
  int pairs_per_block;  // row pairs per workgroup
  const float* src;     // forward input
  float2* dst;          // forward output
  const float2* csrc;   // inverse input
  float* rdst;          // inverse output
  long src_row;         // element stride between input rows
  long dst_row;         // element stride between output rows
  int kshift;           // forward only: rows are compact kernel rows
  int klen;             // compact row length (kx)
  int kcenter;          // kx / 2
  VPad vp;              // image rows: virtual padding / cropped inverse
};

// column pass along y or z over a dense half spectrum (or from the compact
// kernel rows into the kernel planes).
struct ColArgs {
  FftDesc fft;   // column length
  int tx;        // bins per tile
  int nbins;     // valid bins per row (nx/2 + 1)
  const float2* src;
  float2* dst;
  long src_outer, src_elem;  // strides: tile row (grid.y) and column position
  long dst_outer, dst_elem;
  int nvalid;  // source positions present (column length, or ky)
  int center;  // circular placement offset (0, or ky / 2)
  long nouter;  // tile rows (grid rows): nz for the spectrum, kz for kernel planes
};

// fused z pass: s (in place) <- IFFTz( FFTz(s) * op(FFTz(k)) * scale )
struct ZArgs {
  FftDesc fft;  // length nz
  int tx;
  int nbins;
  float2* s;
  long s_row, s_plane;  // strides of y and z in s
  const float2* k;
  long k_row, k_plane;  // strides of y and z in k
  int nvalid;           // k planes present (kz, or nz for a dense operand)
  int center;           // kz / 2, or 0
  int conj;             // use conj(FFTz(k)) (adjoint / correlation)
  float scale;          // 1 / (nx*ny*nz)
  long ny;              // tile rows
  // 0: the result overwrites the column in place.  > 0: only planes
  // [0, out_planes) of the result are written, result position p going to
  // plane (p + out_shift) mod nz (the PSF gradient keeps just the planes its
  // gather reads, in one contiguous run)
  int out_planes;
  int out_shift;
};

// Whole-plane x and y transforms (plane_kernels.hpp): one workgroup per
// z-plane of the image, then one per plane of the compact kernel.
struct PlaneArgs {
  FftDesc fx, fy;
  int nz;               // image planes (the first nz workgroups)
  long nxp;             // spectrum row pitch (bins)
  const float* src;     // forward: image planes [nz][ny][nx]
  float2* spec;         // spectrum planes [nz][ny][nxp]
  const float* kernel;  // forward: compact kernel [kz][ky][kx], or null
  float2* kplanes;      // kernel planes [kz][ny][nxp]
  int kx, ky, kz, cx, cy;
  float* dst;  // inverse: real planes [nz][ny][nx]
  // split kernels: stage tables of the half-length plans (nx/2 forward,
  // ny/2 inverse)
  const float2* hx_stw;
  const float2* hy_stw;
};

// Column tiles of TX consecutive bins over `nouter` rows.  When the bins end
// in a single trailing bin (even nx: bin nx/2 after full tiles), that bin is
// not given a nearly empty tile per row; instead "Nyquist tiles" gather it
// from TX consecutive rows, lane c taking row j*TX + c.
__host__ __device__ inline long column_tiles(int nbins, int TX, long nouter) {
  const int nfull = nbins / TX, rem = nbins - nfull * TX;
  if (rem == 1 && nfull > 0) return (long)nfull * nouter + (nouter + TX - 1) / TX;
  return (long)((nbins + TX - 1) / TX) * nouter;
}

// Tile indices, counts and rows stay below 2^31 (a grid dimension), so the
// decode divides in 32 bits: it runs before a workgroup's first load, and a
// 64-bit division is some hundred instructions.
__host__ __device__ inline void column_tile_decode(int nbins, int TX, long nouter, long tile, int c,
                                                   long* o, int* x, bool* ok) {
  const unsigned utx = (unsigned)TX, t32 = (unsigned)tile, rows = (unsigned)nouter;
  const unsigned nfull = (unsigned)nbins / utx, rem = (unsigned)nbins - nfull * utx;
  if (rem == 1 && nfull > 0) {
    // the Nyquist tiles (strided single-bin gathers, slower than a full
    // tile) take the lowest indices so they run under the bulk of the grid
    // instead of trailing it
    const unsigned nnyq = (rows + utx - 1) / utx;
    if (t32 < nnyq) {
      const unsigned r = t32 * utx + (unsigned)c;
      *o = (long)r;
      *x = nbins - 1;
      *ok = r < rows;
    } else {
      const unsigned f = t32 - nnyq, q = f / nfull;
      *o = (long)q;
      *x = (int)((f - q * nfull) * utx) + c;
      *ok = true;
    }
    return;
  }
  const unsigned ntx = ((unsigned)nbins + utx - 1) / utx, q = t32 / ntx;
  *o = (long)q;
  *x = (int)((t32 - q * ntx) * utx) + c;
  *ok = *x < nbins;
}

// A pass launch that also carries the compact kernel's rows / planes as
// extra workgroups (they share the transform length with the image).
struct XPass2 {
  XPassArgs m;      // image rows
  XPassArgs k;      // kernel rows (kshift)
  long m_blocks;    // workgroups of the image part
  long k_blocks;    // workgroups of the kernel part (they come first)
};
struct Col2 {
  ColArgs m;        // image columns
  ColArgs k;        // kernel planes
  long m_outer;     // image tile rows (grid rows of the image part)
  long k_outer;     // kernel tile rows (0: no kernel part)
  // filled by the launcher for the chosen tile width:
  int ntx;          // bin tiles per row
  long m_tiles;     // tiles of the image part (after the kernel's)
  long ntiles;      // total tiles
};

// fused multiview Richardson-Lucy passes (rl_kernels.hpp)
constexpr int kMaxViews = 8;

struct RLZArgs {
  FftDesc fft;  // length nz
  int nbins;
  long s_row, s_plane;  // strides of y and z (all spectra share the layout)
  int nviews;
  float scale;                       // 1 / N
  long ny;                           // tile rows
  const float2* src;                 // fan-out input (FFTxy of psi)
  float2* dst;                       // sum output
  float2* view[kMaxViews];           // per-view spectra (fan-out out / sum in)
  const float2* kplanes[kMaxViews];  // per-view kernel planes Kxy
  int kz[kMaxViews];
  int cz[kMaxViews];
};

struct RLXArgs {
  FftDesc fft;  // length nx
  long nrows;
  int nbins;
  long srow;  // spectrum row stride (nxp)
  long rrow;  // real row stride (nx)
  float eps;
  float inv_views;
  float2* spec[kMaxViews];       // per-view spectra (xratio) / summed spectrum (xupdate: [0])
  const float* phi[kMaxViews];   // per-view data (xratio)
  float* psi;                    // estimate (xupdate, read and written)
};

void launch_gk_xfwd(const XPassArgs& a, long npairs, hipStream_t st);
void launch_gk_xinv(const XPassArgs& a, long npairs, hipStream_t st);
void launch_gk_col(const ColArgs& a, int dir, long nouter, hipStream_t st);
void launch_gk_zfused(const ZArgs& a, long ny, hipStream_t st);
void generic_kernels_allow_large_lds();

}  // namespace fcamd
