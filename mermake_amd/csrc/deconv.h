// Device-side descriptors shared by the deconvolution, maxima and fit kernels.
#pragma once
#include "fft_rader.h"

// Pointer into global memory.  Pointers read from descriptors are generic, and
// loads through them compile to flat instructions; a cast to this type makes
// them global loads.
template <class X>
using gptr = X __attribute__((address_space(1)))*;

// In-plane layout of the tile spectra (T rows x, H half-spectrum columns ky):
// natural, x * H + ky (blk = 0), or column-blocked, (ky / blk) * T * blk +
// x * blk + ky % blk, where the x transforms of one block of blk columns are a
// single contiguous run.  `plane` is the element stride of one (tile, z) plane.
struct SpecLayout {
  int T, H, blk;
  long long plane;
  __host__ __device__ __forceinline__ long long row(int x) const {
    return blk ? (long long)x * blk : (long long)x * H;
  }
};

// One overlapped tile of a z-stack (reference: deconvolver.py:186-206 for the
// window, :118-148 for the placement and reflections).  Filled on the host.
// Stack pixel types: uint16, float32, or the microscope's uint8 values whose
// squares are the pixels (read_im's astype(uint16) ** 2, io.py:140-142; exact in
// float32, at most 255^2), squared on load.
enum { MM_PIX_U16 = 0, MM_PIX_F32 = 1, MM_PIX_U8SQ = 2 };

struct mm_tile_dev {
  const void* img;          // stack base (see img_type), element (0,0,0)
  const float* flat;        // flat-field base (2D, element (0,0)) or nullptr
  long long img_sz;         // stack stride between z planes (elements)
  long long img_sx;         // stack stride between x rows (elements); y stride is 1
  long long flat_sx;        // flat stride between x rows (elements)
  int X0, Y0;               // window origin in the stack
  int xd, yd;               // data extents of the window
  int xs, ys;               // where the data starts in the padded tile (overlap or 0)
  int xfirst, yfirst;       // tile origin is 0 along x / y ('in' reflection)
  int xe, ye;               // extents of the yielded view (clipped to T)
  int psf;                  // index of the Wiener spectrum
  int ox, oy;               // tile origin in the stack (for spot offsets)
  int img_type;              // MM_PIX_U16, MM_PIX_F32 or MM_PIX_U8SQ
  int pad0, pad1;
};

// Closed form of one axis' reflection (deconvolver.py:131-144): padded index j
// -> data index, or -1 for a position that stays zero.
__device__ __forceinline__ int pad_map(int j, int s, int d, int first, int T) {
  const int rel = j - s;
  if (rel >= 0 && rel < d) return rel;
  if (first) {                      // reflect(ov, 'in'): j < s takes 2s - j if < s + d
    if (j < s) {
      const int r2 = s - j;
      return r2 < d ? r2 : -1;
    }
    return -1;
  }
  if (d < T && j > d - 1) {         // reflect(d-1, 'out'): j > d-1 takes 2(d-1) - j if >= 0
    const int src = 2 * (d - 1) - j;
    return src >= 0 ? src : -1;
  }
  return -1;
}

__device__ __forceinline__ float pix_at(const void* p, int type, long long o) {
  if (type == MM_PIX_F32) return ((const float*)p)[o];
  if (type == MM_PIX_U8SQ) {
    const unsigned u = ((const uint8_t*)p)[o];
    return (float)(u * u);
  }
  return (float)((const uint16_t*)p)[o];
}

// Value of the flat-divided padded tile at data position (z, r, c) of the window.
__device__ __forceinline__ float tile_src(const mm_tile_dev& t, int z, int r, int c) {
  const long long o = (long long)z * t.img_sz + (long long)(t.X0 + r) * t.img_sx + (t.Y0 + c);
  float v = pix_at(t.img, t.img_type, o);
  if (t.flat) v = v / t.flat[(long long)(t.X0 + r) * t.flat_sx + (t.Y0 + c)];
  return v;
}

// `raw` view value at (z, x, y) of tile_pad[zpad:-zpad, :xend, :yend] after the
// flat re-multiply of deconvolver.py:166-168: (v/f)*f on the data mask,
// flat-divided reflected values in the pads, 0 where nothing was reflected.
__device__ __forceinline__ float tile_raw(const mm_tile_dev& t, int T, int z, int x, int y) {
  const int r = pad_map(x, t.xs, t.xd, t.xfirst, T);
  const int c = pad_map(y, t.ys, t.yd, t.yfirst, T);
  if (r < 0 || c < 0) return 0.0f;
  const long long o = (long long)z * t.img_sz + (long long)(t.X0 + r) * t.img_sx + (t.Y0 + c);
  float v = pix_at(t.img, t.img_type, o);
  if (t.flat) {
    const float f = t.flat[(long long)(t.X0 + r) * t.flat_sx + (t.Y0 + c)];
    v = v / f;
    const bool inmask = (x - t.xs) >= 0 && (x - t.xs) < t.xd && (y - t.ys) >= 0 && (y - t.ys) < t.yd;
    if (inmask) v = v * f;
  }
  return v;
}

// Plan of one Deconvolver geometry (host object behind the opaque mm_deconv*).
struct mm_deconv {
  int sz, ts, ov, zpad, T, Zp, H;
  int* zmap_dev;
  FftPlanDev<float2> pT, pZ;
  FftPlanDev<double2> pT64, pZ64;
  int gT_r2c, gT_x, gZ, gT64_r2c, gT64_x, gZ64;
  int zconv;                // z pass as a folded z-kernel matrix (sz <= 48) instead of z FFTs
  int zG;                   // columns per block of the z-kernel pass
  int* zsrc_dev;            // per data plane, the padded planes that copy it (4 slots, -1 padded)
  int zsrc_h[4 * 48];       // host copy (kernel argument of the z-kernel pass)
  int pfM, pfG;             // prime-first T passes: T = p * pfM, pfG transforms per block (pfM = 0: off)
  PfShape pfS;
  int rd, rdP;              // the prime rdP (113 or 239) runs by Rader's algorithm (tables rdF / rdI)
  int zfft;                 // z pass by length-80 DFTs (zpad = sz - 1, sz <= 40): operator planes, not w
  int zdual;                // z-kernel plan without zfft: W holds w (Zp planes) then W itself (Zp planes), so
                            // that batches with one PSF per tile run the z FFT form
  int xy_ok, z_ok;          // masks (1 << form) of the MM_XY_* / MM_Z_* forms the geometry admits
  int xy_form, z_form;      // the schedule in use (mm_deconv_set_schedule)
  int pfM0, pfG0, rdG;      // prime-first radix and group width; Rader group width
  int blk;                  // spectra column-blocked by blk columns (0: natural layout)
  int Hs;                   // columns of one spectrum row in memory (H rounded up to blk)
  RdTab rdF, rdI;
  int prof;                 // record events around the five passes of mm_deconv_tiles
  hipEvent_t ev[6];
};

// mm_deconv_tiles_window that also writes the maxima threshold bits (!(v <= thr))
// of every computed voxel into `marks`, the masks of the maxima workspace (rows
// (tile * sz + z) * T + x, (T + 63) / 64 words each), when the plan's C2R is
// the Rader form (*marked = 1); other forms leave `marks` alone (*marked = 0).
int deconv_tiles_marked(mm_deconv* d, const struct mm_tile_desc* tiles_dev, int ntiles, const void* W, int n_psf,
                        int tiles_per_block, void* spec_ws, float* res, int lo, int hi, unsigned long long* marks,
                        float thr, int* marked, void* stream);

