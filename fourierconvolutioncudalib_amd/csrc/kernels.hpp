// kernels.hpp -- argument blocks and launchers of the pass kernels.
#pragma once

#include "fc_internal.hpp"

namespace fcamd {

// x pass (rows along the contiguous axis).  Forward reads real rows (`src`)
// and writes half-spectrum rows (`dst`); inverse reads `csrc`, writes `rdst`.
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

__host__ __device__ inline void column_tile_decode(int nbins, int TX, long nouter, long tile, int c,
                                                   long* o, int* x, bool* ok) {
  const int nfull = nbins / TX, rem = nbins - nfull * TX;
  if (rem == 1 && nfull > 0) {
    const long full = (long)nfull * nouter;
    if (tile < full) {
      *o = tile / nfull;
      *x = (int)(tile - *o * nfull) * TX + c;
      *ok = true;
    } else {
      *o = (tile - full) * TX + c;
      *x = nbins - 1;
      *ok = *o < nouter;
    }
    return;
  }
  const int ntx = (nbins + TX - 1) / TX;
  *o = tile / ntx;
  *x = (int)(tile - *o * ntx) * TX + c;
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
