/*
 * C ABI of libmermake_hip.so: the MI355X (gfx950) implementation of MERMAKE's
 * image hot path (deconvolution -> blur subtraction -> local maxima -> fit,
 * plus the FFT correlation used by the drift).
 *
 * Conventions
 *  - Every pointer argument named *_dev, img, res, raw, out..., ws, W is a
 *    device pointer owned by the caller (the library never frees it).
 *  - `stream` is a hipStream_t (0 = the null stream); all work is enqueued on
 *    it and the calls return without synchronising.
 *  - Every function returns 0 on success, a hipError_t value for a HIP failure,
 *    or an MM_E* code (>= 1000); mm_last_error() describes the last failure of
 *    the calling thread.  No C++ exception crosses this boundary.
 *  - Volumes are C-ordered (z, x, y) with unit stride along y; strides are in
 *    elements.
 */
#ifndef MERMAKE_HIP_H
#define MERMAKE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_EINVAL 1000
#define MM_ECAP 1001
#define MM_ENOMEM 1002

int mm_version(void);
const char* mm_last_error(void);

/* ------------------------------------------------------------------ blur
 * Replaces the CUDA kernel box_1d (mermake/blur.cu:1-52) launched by
 * blur.box_1d (mermake/blur.py:157-188): mean over the window
 * [i-delta, i+delta] clipped to the axis, summed sequentially from the low to
 * the high index in float32, then sum/num.  `in` and `out` must not overlap.
 */
int mm_box_1d(const float* in, float* out, int size_z, int size_x, int size_y, int delta, int axis, void* stream);
/* out = res - box_1d(buf, axis): the subtraction of deconvolver.py:160-163
 * fused into the second blur pass.  `out` may alias `res`, not `buf`. */
int mm_box_1d_sub(const float* res, const float* buf, float* out, int size_z, int size_x, int size_y, int delta,
                  int axis, void* stream);

/* out = res - box_1d(box_1d(res, axis 1), axis 2) over nplanes (T, T) planes:
 * the blur subtraction of deconvolver.py:160-163 for a batch of tiles.
 * Bit-identical to the two reference passes.  `out` must not overlap `res`;
 * delta 15 and 25 run as one fused kernel, other widths need `tmp`
 * (nplanes*T*T floats). */
int mm_blur_sub_xy(const float* res, float* out, float* tmp, int nplanes, int T, int delta, void* stream);
/* mm_blur_sub_xy with outputs only at rows and columns [lo, hi) of every plane
 * (0 <= lo < hi <= T); the box windows still clip at the plane edges and read
 * every input they cover, so those outputs are bit-identical, and `out` is left
 * untouched elsewhere.  No counterpart in the reference: for the hyb caller,
 * whose keep filter (mermake/__main__.py:256-261) discards every spot fitted
 * outside the overlap margins, so the margins' blur is never read.  Widths
 * without the fused kernel compute whole planes. */
int mm_blur_sub_xy_window(const float* res, float* out, float* tmp, int nplanes, int T, int delta, int lo, int hi,
                          void* stream);

/* ------------------------------------------------------------------ fill
 * Replaces the CUDA kernel `mirror` (mermake/fill.py:6-52) launched by
 * fill.reflect (mermake/fill.py:54-111), in place on `arr` (float32, up to 8
 * dims, element strides).  mode 0 = 'out', 1 = 'in'.
 */
int mm_reflect(float* arr, int ndim, const long long* shape, const long long* strides, int axis, int i, int mode,
               void* stream);

/* ------------------------------------------------------------- deconvolution
 * One Deconvolver geometry (mermake/deconvolver.py:47-96): stack depth sz,
 * tile_size, overlap, zpad.  Padded tile T = tile_size + 2*overlap,
 * Zp = sz + 2*zpad, half spectrum H = T/2 + 1.
 */
typedef struct mm_deconv mm_deconv;

/* Tile descriptor, one per overlapped tile (deconvolver.py:186-206 window,
 * :118-148 placement/reflection).  Layout is part of the ABI (104 bytes). */
typedef struct mm_tile_desc {
  const void* img;   /* stack base (see img_type), element (0,0,0) */
  const float* flat; /* 2D flat field base, or NULL */
  long long img_sz;  /* stack stride between z planes */
  long long img_sx;  /* stack stride between x rows */
  long long flat_sx; /* flat stride between x rows */
  int X0, Y0;        /* window origin max(x-ov,0), max(y-ov,0) */
  int xd, yd;        /* window extents */
  int xs, ys;        /* data start in the padded tile: ov if the tile origin is 0 else 0 */
  int xfirst, yfirst;/* tile origin is 0 ('in' reflection) */
  int xe, ye;        /* extents of the yielded views, clipped to T */
  int psf;           /* Wiener spectrum index */
  int ox, oy;        /* tile origin (x, y) */
  int img_type;  /* 0: uint16, 1: float32, 2: uint8 whose squares are the pixels (read_im, io.py:140-142) */
  int pad0, pad1;
} mm_tile_desc;

/* replaces Deconvolver.__init__'s buffers and spectra set-up (deconvolver.py:48-96) */
int mm_deconv_create(int sz, int tile_size, int overlap, int zpad, mm_deconv** out);
int mm_deconv_destroy(mm_deconv* d);
int mm_deconv_dims(const mm_deconv* d, int* T, int* Zp, int* H);
/* Schedule of mm_deconv_tiles: the form of the T-length passes along y and x
 * and of the z pass.  No counterpart in the reference, whose FFTs are cuFFT's
 * plans (deconvolver.py:130-158); the project's autotuner times the forms a
 * geometry admits and selects one.
 *   MM_XY_STOCKHAM     LDS-staged Stockham radix passes (any T)
 *   MM_XY_PRIME_FIRST  T = p * M: a matrix-product stage of the prime p, then radix M
 *   MM_XY_RADER        T = 113 * M or 239 * 2: the prime by Rader's algorithm in
 *                      Good-Thomas form, spectra column-blocked
 *   MM_Z_FFT           z FFTs of length Zp with the Wiener multiply
 *   MM_Z_MATRIX        the folded z-kernel matrix (sz <= 48)
 *   MM_Z_DFT80         length-80 DFTs (zpad = sz - 1, sz <= 40)
 * mm_deconv_schedule reports the forms in use and the masks (1 << form) of
 * those the geometry admits; a plan starts with the fastest admitted form of
 * each.  mm_deconv_set_schedule selects forms (MM_EINVAL if not admitted);
 * the spectrum layout, mm_deconv_spec_bytes and the operator planes follow
 * it, so W is built again with mm_deconv_build_wiener after a change. */
#define MM_XY_STOCKHAM 0
#define MM_XY_PRIME_FIRST 1
#define MM_XY_RADER 2
#define MM_Z_FFT 0
#define MM_Z_MATRIX 1
#define MM_Z_DFT80 2
int mm_deconv_schedule(const mm_deconv* d, int* xy, int* z, int* xy_mask, int* z_mask);
int mm_deconv_set_schedule(mm_deconv* d, int xy, int z);
/* In-plane layout of the half spectra (T rows x, H columns ky) in W and in the
 * spectrum workspace: blk = 0 is natural (x * H + ky); blk > 0 is
 * column-blocked, (ky / blk) * T * blk + x * blk + ky % blk with Hs = H rounded
 * up to blk columns (zero padding), which makes every x transform group one
 * contiguous run.  A plane holds T * Hs complex64; W holds wplanes of them.
 * No counterpart in the reference (its spectra are cupy arrays). */
int mm_deconv_spec_layout(const mm_deconv* d, int* blk, int* Hs, int* wplanes);
size_t mm_deconv_wiener_ws_bytes(const mm_deconv* d);
/* W = conj(H)/(H conj(H) + beta |L|^2) / (Zp*T*T) as Zp planes of complex64
 * half spectra in the mm_deconv_spec_layout layout, from one centred PSF
 * (sz, tile_size, tile_size) float32 (deconvolver.py:60-74; the PSF FFT runs
 * in float64).  What the planes hold follows the plan's z form
 * (mm_deconv_set_schedule): MM_Z_FFT, W itself (Zp planes); MM_Z_MATRIX, the
 * z-domain kernels w = ifft_z(W) and then W (2 Zp planes; W for batches with
 * one PSF per tile); MM_Z_DFT80 (zpad = sz - 1, sz <= 40), the 240 planes of
 * the length-80 z-pass operator built from w (A = DFT80 of w on |m| < sz, G =
 * DFT80 of w[n] + w[n - 2(sz - 1)] on n < 2 sz - 1, both / 80; w[t] and
 * w[t - sz + 1] for t < 40). */
int mm_deconv_build_wiener(mm_deconv* d, const float* psf_centered, double beta, void* W_out, void* ws,
                           size_t ws_bytes, void* stream);
/* mm_deconv_build_wiener with the blur subtraction of deconvolver.py:160-163
 * (blur.py box_1d of half-width box_half along x, then along y, subtracted)
 * folded into W: W(kz, kx, ky) * (1 - b(kx) b(ky)), b(k) = the transfer
 * function of the (2 box_half + 1)-tap mean on the T-periodic tile, in
 * float64.  A tile deconvolved with it holds res - box_y(box_x(res)) at
 * every voxel whose box stays inside the tile (rows and columns [box_half,
 * T - box_half)): exactly there in real arithmetic, and within float32
 * rounding of the reference's sequential sums.  box_half = 0 is
 * mm_deconv_build_wiener.  0 <= 2 box_half + 1 <= T. */
int mm_deconv_build_wiener_box(mm_deconv* d, const float* psf_centered, double beta, int box_half, void* W_out,
                               void* ws, size_t ws_bytes, void* stream);
/* bytes of the per-tile half-spectrum workspace (sz planes of T*Hs complex64) */
size_t mm_deconv_spec_bytes(const mm_deconv* d);
/* Replaces the FFT part of Deconvolver.tile_wise (deconvolver.py:116-157) for
 * a batch of tiles: res[t] (sz, T, T) float32 = ifftn(fftn(pad_t) * W)[zpad:-zpad].real.
 * tiles_dev: device array of mm_tile_desc.  W: n_psf outputs of mm_deconv_build_wiener.
 * tiles_per_block: tiles sharing one load of W (use ntiles when all tiles use
 * the same PSF, 1 otherwise).  spec_ws: ntiles * mm_deconv_spec_bytes. */
int mm_deconv_tiles(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W, int n_psf,
                    int tiles_per_block, void* spec_ws, float* res, void* stream);
/* mm_deconv_tiles with res computed only at rows and columns [lo, hi) of every
 * plane (0 <= lo < hi <= T; the rest of res is left undefined), bit-identical
 * there.  No counterpart in the reference: for the hyb caller, whose blur and
 * maxima read only the margins its keep filter can use (see
 * mm_blur_sub_xy_window).  Tiles T = 113*M skip the other rows in the final
 * inverse y transform and the other columns' stores; other sizes compute
 * whole planes. */
int mm_deconv_tiles_window(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W, int n_psf,
                           int tiles_per_block, void* spec_ws, float* res, int lo, int hi, void* stream);
/* Timing of mm_deconv_tiles with HIP events recorded on its stream around
 * its five passes (no counterpart in the reference; used by the benchmark):
 * ms5 = r2c, x forward, z Wiener, x inverse, c2r of the last profiled call. */
int mm_deconv_set_profiling(mm_deconv* d, int on);
int mm_deconv_pass_times(mm_deconv* d, float* ms5);
/* The `raw` views of tile_wise (deconvolver.py:165-170), (sz, T, T) per tile,
 * zero outside [0,xe) x [0,ye). */
int mm_deconv_tile_raw(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, float* raw, void* stream);
/* Deconvolver.apply's stitch (deconvolver.py:179-183) into out (sz, SX, SY). */
int mm_deconv_stitch(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const float* res, float* out,
                     long long o_sz, long long o_sx, int SX, int SY, void* stream);
/* mm_deconv_tiles followed by mm_deconv_stitch without the result tiles: the
 * inverse y transform stores the owned rows and columns [ov, xe - ov) x
 * [ov, ye - ov) of every tile straight into out (sz, SX, SY) (Deconvolver.apply,
 * deconvolver.py:116-183), bit for bit what the two calls store.  Only for
 * plans whose y passes are the Rader form (MM_EINVAL otherwise, then use the
 * two calls). */
int mm_deconv_tiles_stitch(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W, int n_psf,
                           int tiles_per_block, void* spec_ws, float* out, long long o_sz, long long o_sx, int SX,
                           int SY, void* stream);

/* DAPI normalisation (mermake/__main__.py:221-222): a /= ||a||_2/sqrt(n) in place;
 * *rms_out (device float) receives the divisor.  ws: mm_rms_ws_bytes(). */
size_t mm_rms_ws_bytes(void);
int mm_rms_normalize(float* a, long long n, void* ws, float* rms_out, void* stream);

/* *out (device float) = the median of the n float32 values at a (device): the
 * middle value, or for an even n the float32 mean of the two middle values,
 * NaN when any value is NaN (the cp.median of load_flats, mermake/io.py:52).
 * Exact, by radix selection.  ws: mm_median_ws_bytes() bytes. */
size_t mm_median_ws_bytes(void);
int mm_median(const float* a, long long n, float* out, void* ws, size_t ws_bytes, void* stream);

/* In-place 3D C2C FFT of a contiguous complex64 (n0,n1,n2) volume over the axes
 * in axes_mask (bit i = axis i); inverse scales by 1/n per axis.  Replaces the
 * cupy.fft.fftn/ifftn calls of utils.fftconvolve (mermake/utils.py:10-36). */
int mm_fft3d_c2c(void* data, int n0, int n1, int n2, int axes_mask, int inverse, void* stream);

/* Forward mm_fft3d_c2c over all three axes of a grid whose values are zero outside
 * [0,m0) x [0,m1) x [0,n2): the lines that are zero before their pass are skipped.
 * For the zero-padded rasters of the drift correlation, whose cupy.fft.fftn in
 * utils.fftconvolve (mermake/utils.py:10-36, called at mermake/align.py:104-110)
 * transforms the whole padded grid; the result is the same. */
int mm_fft3d_c2c_box(void* data, int n0, int n1, int n2, int m0, int m1, void* stream);

/* ------------------------------------------------------------------ maxima
 * Replaces local_maxima_count + local_maxima (mermake/maxima.cu:1-145) as
 * driven by find_local_maxima (mermake/maxima.py:30-77): voxels with
 * sign*v > threshold and sign*v >= every neighbour of the delta sphere
 * (single reflection at the borders), written as int32 (z, x, y) triplets in
 * raster order; *count (device) receives the number found (entries beyond
 * cap are dropped).  ws: mm_maxima_ws_bytes(D, H, W, 1) bytes.
 */
/* workspace bytes for nvol volumes of at most D x H x W voxels (row bit masks and row counts) */
size_t mm_maxima_ws_bytes(int D, int H, int W, int nvol);
int mm_local_maxima(const float* img, int D, int H, int W, long long sZ, long long sX, float sign, float threshold,
                    int delta, int* coords, int cap, int* count, void* ws, size_t ws_bytes, void* stream);
/* mm_local_maxima of sign +1 and of sign -1 in one pass over the volume: the
 * two find_local_maxima calls of the DAPI block on +buffer and -buffer
 * (mermake/__main__.py:224-225).  coords: (2, cap, 3) int32, the + maxima then
 * the - maxima; counts: 2 device ints.  ws: mm_maxima_ws_bytes(D, H, W, 2). */
int mm_local_maxima_pm(const float* img, int D, int H, int W, long long sZ, long long sX, float threshold, int delta,
                       int* coords, int cap, int* counts, void* ws, size_t ws_bytes, void* stream);
/* mm_local_maxima without a workspace argument and without any allocation: the
 * row masks and counts live in a library-owned device array of
 * MM_MAXIMA_SCRATCH_BYTES bytes (one per GPU, held until the process ends).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Volumes whose mm_maxima_ws_bytes(D, H, W, 1) exceeds the array
 * (about 40 x 2800 x 3300 voxels) fail with MM_ECAP.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (other streams,
 * other threads) race; calls in stream order do not. */
#define MM_MAXIMA_SCRATCH_BYTES (48u << 20)
int mm_local_maxima_noalloc(const float* img, int D, int H, int W, long long sZ, long long sX, float sign,
                            float threshold, int delta, int* coords, int cap, int* count, void* stream);
/* Replaces delta_fit_cross_corr (mermake/maxima.cu:148-341): out8 rows
 * [zc, xc, yc, background, a, habs, hn, h] for the *count candidates.
 * raw: volume with strides rZ, rX whose pixels are raw_type 0 uint16,
 * 1 float32 or 2 uint8 values standing for their squares (as img_type). */
int mm_delta_fit(const float* img, int D, int H, int W, long long sZ, long long sX, float sign, const void* raw,
                 int raw_type, long long rZ, long long rX, const int* coords, const int* count, int cap, int delta_fit,
                 float sigmaZ, float sigmaXY, float* out8, void* stream);

/* Batched forms over the tile views res[t][:, :xe, :ye] of mm_deconv_tiles:
 * the hyb loop of mermake/__main__.py:254-265.  coords: ntiles*cap*3,
 * counts: ntiles, out8: ntiles*cap*8, ws: mm_maxima_ws_bytes(sz, T, T, ntiles). */
int mm_tiles_maxima(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const float* res, float threshold,
                    int delta, int* coords, int cap, int* counts, void* ws, size_t ws_bytes, void* stream);
/* mm_tiles_maxima with candidates searched at x, y in [lo, hi) only (every z);
 * the neighbour tests read res at [lo - delta, hi + delta).  No counterpart in
 * the reference: for the hyb caller, whose keep filter discards candidates
 * whose fitted x, y leave [ov, extent - ov) (a fit moves a candidate by at most
 * delta_fit), so candidates outside [ov - delta_fit - 1, T - ov + delta_fit + 1)
 * are never kept. */
int mm_tiles_maxima_window(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const float* res,
                           float threshold, int delta, int lo, int hi, int* coords, int cap, int* counts, void* ws,
                           size_t ws_bytes, void* stream);
/* mm_blur_sub_xy_window(res -> out, rows and columns [blo, bhi), blur_delta 15
 * or 25) followed by mm_tiles_maxima_window(out, [lo, hi) inside [blo, bhi)),
 * with the same outputs as those two calls.  The blur kernel also records the
 * threshold test of every output it writes, as bits in the maxima workspace,
 * so the maxima pass reads only candidates and their neighbours instead of the
 * whole window.  The neighbour tests read out at [lo - delta, hi + delta),
 * which must lie inside [blo, bhi) except where blo = 0 or bhi = T (MM_EINVAL
 * otherwise).  Replaces the pair deconvolver.py:160-163 (blur) +
 * maxima.py:17-98 (find_local_maxima) of the hyb caller (__main__.py:254). */
int mm_tiles_blur_maxima_window(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const float* res, float* out,
                                int blur_delta, int blo, int bhi, float threshold, int delta, int lo, int hi,
                                int* coords, int cap, int* counts, void* ws, size_t ws_bytes, void* stream);
/* mm_deconv_tiles_window(res, rows and columns [dlo, dhi)) followed by
 * mm_tiles_maxima_window(res, [lo, hi) inside [dlo, dhi)), with the same
 * outputs as those two calls.  With W from mm_deconv_build_wiener_box, res is
 * the deconvolved tile minus its blur, so this is the hyb caller's
 * tile_wise + find_local_maxima (deconvolver.py:116-170, maxima.py:17-98,
 * __main__.py:254) with no blur pass.  When the plan's y passes are the Rader
 * form, the inverse y transform records the threshold test of every voxel it
 * stores as bits in the maxima workspace, so the maxima pass reads only
 * candidates and their neighbours; other forms run the whole maxima pass.
 * The neighbour tests read out at [lo - delta, hi + delta), which must lie
 * inside [dlo, dhi) except where dlo = 0 or dhi = T (MM_EINVAL otherwise). */
int mm_tiles_deconv_maxima_window(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W, int n_psf,
                                  int tiles_per_block, void* spec_ws, float* res, int dlo, int dhi, float threshold,
                                  int delta, int lo, int hi, int* coords, int cap, int* counts, void* ws,
                                  size_t ws_bytes, void* stream);
int mm_tiles_fit(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const float* res, const int* coords,
                 const int* counts, int cap, int delta_fit, float sigmaZ, float sigmaXY, float* out8, void* stream);
/* keep-filter on the fitted x,y (ov <= . < view extent - ov), offset to stack
 * coordinates, stacked in tile order into Xh (out_cap rows); *total (device)
 * = rows kept; *max_count (device, may be NULL) = the largest of the ntiles
 * candidate counts (above cap, candidates were dropped).  kept_ws: ntiles
 * ints. (mermake/__main__.py:256-265) */
int mm_tiles_keep(const mm_tile_desc* tiles_dev, int ntiles, int overlap, const float* fit, const int* counts,
                  int cap, int* kept_ws, float* Xh, int out_cap, int* total, int* max_count, void* stream);

/* ------------------------------------------------------------------ drift
 * Building blocks of Aligner (mermake/align.py:37-118) and utils.fftconvolve
 * (mermake/utils.py:10-36).  Points: n rows of >= 3 float32 (row stride ld).
 */
/* Xi (n*3 int32) = round_half_even(X[:, :3] / resc); bounds = min xyz, max xyz */
int mm_points_round_bounds(const float* X, int n, int ld, float resc, int* Xi, int* bounds, void* stream);
/* over the points inside [lo, hi] (host int[3] each): out4 = max xyz, count */
int mm_points_kept_max(const int* Xi, int n, const int* lo, const int* hi, int* out4, void* stream);
/* writes 1+0i at every kept point of a zeroed complex64 (L0,L1,L2) grid; flip
 * mirrors each axis within the image size R (im[::-1,::-1,::-1]) */
int mm_points_raster(const int* Xi, int n, const int* lo, const int* hi, void* buf, int L0, int L1, int L2, int flip,
                     int R0, int R1, int R2, void* stream);
/* a *= b elementwise, complex64 */
int mm_cmul(void* a, const void* b, long long n, void* stream);
/* first raster index (over [0,F0)x[0,F1)x[0,F2)) of the maximum of rint(Re(buf)) */
size_t mm_argmax_ws_bytes(void);
int mm_argmax_rounded(const void* buf, int L0, int L1, int L2, int F0, int F1, int F2, long long* out_i, float* out_v,
                      void* ws, void* stream);
/* one refinement round of Aligner.get_Xtzxy (align.py:66-76): nearest reference
 * point of X - t (float64), keep distances < th, t = mean of the kept
 * differences when any; *npts = kept count.  ds_ws: n doubles, idx_ws: n ints. */
int mm_nn_round(const float* X, int n, int ldx, const float* R, int m, int ldr, double* t, double th, double* ds_ws,
                int* idx_ws, int* npts, void* stream);

#ifdef __cplusplus
}
#endif
#endif
