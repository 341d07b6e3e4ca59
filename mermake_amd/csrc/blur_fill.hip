// Box blur (reference mermake/blur.cu:1-52) and mirror reflection
// (reference mermake/fill.py:6-52) for gfx950.
//
// The blur keeps the reference arithmetic exactly: each output sums its
// clipped window sequentially from the lowest to the highest index in float32
// and divides once by the element count, so results are bit-identical to the
// reference kernel for the same input.
#include "common.h"
#include "mermake_hip.h"

template <bool SUB>
__global__ __launch_bounds__(256) void k_box_1d(const float* __restrict__ in, const float* __restrict__ res,
                                                float* out, int sz, int sx, int sy, int delta, int axis) {
  const long long total = (long long)sz * sx * sy;
  const long long plane = (long long)sx * sy;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(idx / plane);
    const int rem = (int)(idx - (long long)z * plane);
    const int x = rem / sy, y = rem - (rem / sy) * sy;
    float sum = 0.0f;
    int num = 0;
    if (axis == 0) {
      const int s = max(0, z - delta), e = min(sz - 1, z + delta);
      for (int i = s; i <= e; ++i) { sum += in[(long long)i * plane + (long long)x * sy + y]; ++num; }
    } else if (axis == 1) {
      const int s = max(0, x - delta), e = min(sx - 1, x + delta);
      const float* p = in + (long long)z * plane + y;
      for (int i = s; i <= e; ++i) { sum += p[(long long)i * sy]; ++num; }
    } else {
      const int s = max(0, y - delta), e = min(sy - 1, y + delta);
      const float* p = in + (long long)z * plane + (long long)x * sy;
      for (int i = s; i <= e; ++i) { sum += p[i]; ++num; }
    }
    const float m = num > 0 ? sum / (float)num : 0.0f;
    if (SUB) out[idx] = res[idx] - m;
    else out[idx] = m;
  }
}

static int box_launch(const float* in, const float* res, float* out, int sz, int sx, int sy, int delta, int axis,
                      bool sub, void* stream) {
  if (!in || !out || sz < 0 || sx < 0 || sy < 0 || delta < 0 || axis < 0 || axis > 2) {
    mm_set_error("mm_box_1d: bad argument");
    return MM_EINVAL;
  }
  const long long total = (long long)sz * sx * sy;
  if (total == 0) return MM_OK;
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (sub)
    hipLaunchKernelGGL(k_box_1d<true>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, in, res, out, sz,
                       sx, sy, delta, axis);
  else
    hipLaunchKernelGGL(k_box_1d<false>, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, in, nullptr, out,
                       sz, sx, sy, delta, axis);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_box_1d(const float* in, float* out, int size_z, int size_x, int size_y, int delta, int axis,
                         void* stream) {
  return box_launch(in, nullptr, out, size_z, size_x, size_y, delta, axis, false, stream);
}

extern "C" int mm_box_1d_sub(const float* res, const float* buf, float* out, int size_z, int size_x, int size_y,
                             int delta, int axis, void* stream) {
  if (!res) { mm_set_error("mm_box_1d_sub: null res"); return MM_EINVAL; }
  return box_launch(buf, res, out, size_z, size_x, size_y, delta, axis, true, stream);
}

// ------------------------------------------------------- fused blur-subtract
// out = res - box_y(box_x(res)) on (nplanes, T, T) float32 planes, for a window
// half-width D known at compile time (deconvolver.py:160-163 with
// blur_radius = 2D or 2D+1).  Every output keeps the arithmetic of
// blur.cu:20-51 -- the window summed sequentially from its lowest to its
// highest index in float32, starting from +0, then one correctly rounded
// division by the element count -- so results equal the two-pass reference
// form bit for bit.  Two facts make the kernel branch-free:
//   * windows clipped at the plane border are summed over zero padding: the
//     sequential sum meets the +0 terms before (or after) the in-plane ones,
//     and +0 + x = x, x + 0 = x exactly, so the sum is unchanged;
//   * the division uses q = s*y, r = fma(-q, n, s), q' = fma(r, y, q) with
//     y = RN(1/n), which equals RN(s/n) whenever |s| >= 2^-100 (checked
//     exhaustively over whole binades for n = 1..64); smaller |s| takes the
//     IEEE division.
// A block owns BX = 16 consecutive rows of one plane across all T columns.
// x pass: a thread takes a column, reads its BX + 2D rows from global memory
// into registers and writes the BX box_x values into an LDS row buffer with D
// zero columns on each side; y pass: a thread takes 9 consecutive outputs of
// one row from that buffer and writes res - box_y.
__device__ __forceinline__ float div_count(float s, float n, float y) {
  if (__builtin_expect(fabsf(s) < 0x1p-100f, 0)) return s / n;
  const float q = s * y;
  return fmaf(fmaf(-q, n, s), y, q);
}

// div_count of K sums at once: the rare tiny sums take the IEEE division in one
// branch for the whole group instead of one branch per sum.
template <int K>
__device__ __forceinline__ void div_counts(const float (&s)[K], const float (&n)[K], const float (&y)[K],
                                           float (&q)[K]) {
  bool tiny = false;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float t = s[k] * y[k];
    q[k] = fmaf(fmaf(-t, n[k], s[k]), y[k], t);
    tiny |= fabsf(s[k]) < 0x1p-100f;
  }
  if (__builtin_expect(tiny, 0)) {
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (fabsf(s[k]) < 0x1p-100f) q[k] = s[k] / n[k];
  }
}

#define BLUR_BX 16
#define BLUR_YB 9
// Outputs are computed for rows and columns in [lo, hi) of each plane (the
// whole plane: lo = 0, hi = T); the windows still read every input they cover.
__host__ __device__ __forceinline__ int blur_pitch(int T, int D, int lo, int hi) {
  const int nyb = (hi - lo + BLUR_YB - 1) / BLUR_YB;
  const int need = lo + nyb * BLUR_YB + 2 * D;  // every y chunk's window stays inside its row
  return (need > T + 2 * D ? need : T + 2 * D) | 1;
}

// x pass of column c: the BX box_x values of the patch rows into tb.  IN: every
// row of the strip and every window lies inside the plane (counts 2D + 1).
template <int D, bool IN>
__device__ __forceinline__ void blur_x_col(const float* __restrict__ rp, const float* __restrict__ strip, int T,
                                           int x0, int c, int pitch, const float* __restrict__ rown,
                                           const float* __restrict__ rowy, float* __restrict__ tb) {
  constexpr int BX = BLUR_BX, NR = BX + 2 * D;
  float v[NR];
  if (IN) {  // unconditional loads from one scalar strip base with 32-bit lane offsets
#pragma unroll
    for (int i = 0; i < NR; ++i) v[i] = strip[(unsigned)(i * T + c)];
  } else {  // strips crossing the plane edge: clamped rows, zeros outside
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const int gx = x0 - D + i;
      const float x = rp[(unsigned)(min(max(gx, 0), T - 1) * T + c)];
      v[i] = (gx >= 0 && gx < T) ? x : 0.0f;
    }
  }
  float sum[BX], n[BX], y[BX], q[BX];
#pragma unroll
  for (int xo = 0; xo < BX; ++xo) {
    sum[xo] = 0.0f;
#pragma unroll
    for (int w = 0; w <= 2 * D; ++w) sum[xo] += v[xo + w];
    n[xo] = IN ? (float)(2 * D + 1) : rown[xo];
    y[xo] = IN ? 1.0f / (float)(2 * D + 1) : rowy[xo];
  }
  div_counts<BX>(sum, n, y, q);
#pragma unroll
  for (int xo = 0; xo < BX; ++xo)
    if (IN || x0 + xo < T) tb[xo * pitch + D + c] = q[xo];
}

// y pass of one item: outputs y0 .. y0 + YB - 1 (< hi) of row gx.  IN: every
// window inside the row (counts 2D + 1).  MARK: returns bit yo set for each
// output o with !(sign o <= thr), the maxima threshold test.
template <int D, bool IN, bool MARK>
__device__ __forceinline__ unsigned blur_y_item(const float* __restrict__ trow, const float* __restrict__ rrow,
                                                float* __restrict__ orow, int T, int y0, int hi,
                                                const float* __restrict__ rcp, float sign, float thr) {
  constexpr int YB = BLUR_YB, W = YB + 2 * D;
  float u[W];
#pragma unroll
  for (int w = 0; w < W; ++w) u[w] = trow[w];
  float rv[YB];  // the item's res values, in flight during the sums (columns clamped into the window)
#pragma unroll
  for (int yo = 0; yo < YB; ++yo) rv[yo] = rrow[min(y0 + yo, hi - 1)];
  float sum[YB], n[YB], y[YB], q[YB];
#pragma unroll
  for (int yo = 0; yo < YB; ++yo) {
    sum[yo] = 0.0f;
#pragma unroll
    for (int w = 0; w <= 2 * D; ++w) sum[yo] += u[yo + w];
    if (IN) {
      n[yo] = (float)(2 * D + 1);
      y[yo] = 1.0f / (float)(2 * D + 1);
    } else {
      const int gy = y0 + yo;
      const int nn = min(T - 1, gy + D) - max(0, gy - D) + 1;
      n[yo] = (float)nn;
      y[yo] = rcp[min(nn, 63)];
    }
  }
  div_counts<YB>(sum, n, y, q);
  unsigned bits = 0;
#pragma unroll
  for (int yo = 0; yo < YB; ++yo)
    if (y0 + yo < hi) {
      const float o = rv[yo] - q[yo];
      orow[y0 + yo] = o;
      if (MARK && !(sign * o <= thr)) bits |= 1u << yo;
    }
  return bits;
}

// MARK: also the maxima threshold bits of every computed output, as ww 64-bit
// words per row of the (nplanes, T) rows (word k bit b: column 64 k + b; zero
// outside [lo, hi)), for rows [lo, hi).
template <int D, bool MARK>
__global__ __launch_bounds__(256) void k_blur_sub_xy(const float* __restrict__ res, float* __restrict__ out, int T,
                                                     int nplanes, int npatch, int lo, int hi, float sign, float thr,
                                                     unsigned long long* __restrict__ marks, int ww) {
  constexpr int BX = BLUR_BX, YB = BLUR_YB;
  const int nyb = (hi - lo + YB - 1) / YB;
  const int pitch = blur_pitch(T, D, lo, hi);
  extern __shared__ __align__(16) float smem_f[];
  unsigned long long* aw = (unsigned long long*)smem_f;  // [BX][ww] threshold words (MARK)
  float* const sf = smem_f + (MARK ? 2 * BX * ww : 0);
  float* rcp = sf;                 // [64]: RN(1/n)
  float* rown = sf + 64;           // [BX]: window counts of the patch rows
  float* rowy = sf + 64 + BX;      // [BX]: their RN(1/n)
  float* tb = sf + 64 + 2 * BX;    // [BX][pitch]: box_x of the patch rows, D zero columns each side
  // neighbouring patches of a plane run on one XCD and share its L2 for the halo rows
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int plane = bid / npatch, patch = bid - (bid / npatch) * npatch;
  if (plane >= nplanes) return;
  const int x0 = lo + patch * BX;
  const float* rp = res + (long long)plane * T * T;
  if (threadIdx.x < 64) rcp[threadIdx.x] = 1.0f / (float)(threadIdx.x ? threadIdx.x : 1);
  if (threadIdx.x < BX) {
    const int gx = x0 + threadIdx.x;
    const int n = min(T - 1, gx + D) - max(0, gx - D) + 1;
    rown[threadIdx.x] = (float)n;
    rowy[threadIdx.x] = 1.0f / (float)n;
  }
  const int nz = pitch - T;  // zero columns per row: [0, D) and [D + T, pitch)
  for (int i = threadIdx.x; i < BX * nz; i += blockDim.x) {
    const int r = i / nz, c = i - r * nz;
    tb[r * pitch + (c < D ? c : T + c)] = 0.0f;
  }
  if (MARK)
    for (int i = threadIdx.x; i < BX * ww; i += blockDim.x) aw[i] = 0ull;
  __syncthreads();
  // x pass over the columns the y windows of [lo, hi) read
  const int chi = min(T, hi + D);
  const bool rows_in = x0 - D >= 0 && x0 + BX + D <= T;  // uniform
  const float* strip = rp + (long long)(rows_in ? x0 - D : 0) * T;
  if (rows_in) {
    for (int c = max(0, lo - D) + threadIdx.x; c < chi; c += blockDim.x)
      blur_x_col<D, true>(rp, strip, T, x0, c, pitch, rown, rowy, tb);
  } else {
    for (int c = max(0, lo - D) + threadIdx.x; c < chi; c += blockDim.x)
      blur_x_col<D, false>(rp, strip, T, x0, c, pitch, rown, rowy, tb);
  }
  __syncthreads();
  // y pass
  for (int it = threadIdx.x; it < BX * nyb; it += blockDim.x) {
    const int xo = it / nyb, yb = it - (it / nyb) * nyb;
    const int gx = x0 + xo;
    if (gx >= hi) continue;
    const int y0 = lo + yb * YB;
    const float* trow = tb + xo * pitch + y0;  // trow[w] = box_x at column y0 - D + w
    const float* rrow = rp + (long long)gx * T;
    float* orow = out + (long long)plane * T * T + (long long)gx * T;
    const unsigned bits = y0 - D >= 0 && y0 + YB - 1 + D <= T - 1
                              ? blur_y_item<D, true, MARK>(trow, rrow, orow, T, y0, hi, rcp, sign, thr)
                              : blur_y_item<D, false, MARK>(trow, rrow, orow, T, y0, hi, rcp, sign, thr);
    if (MARK && bits) {  // rare: LDS OR into the row's word(s); bits past hi are zero
      const int w = y0 >> 6, sh = y0 & 63;
      atomicOr(aw + xo * ww + w, (unsigned long long)bits << sh);
      const unsigned long long up = sh ? (unsigned long long)bits >> (64 - sh) : 0ull;
      if (up) atomicOr(aw + xo * ww + w + 1, up);
    }
  }
  if (MARK) {
    __syncthreads();
    for (int i = threadIdx.x; i < BX * ww; i += blockDim.x) {
      const int xo = i / ww;
      if (x0 + xo < hi) marks[((long long)plane * T + x0 + xo) * ww + (i - xo * ww)] = aw[i];
    }
  }
}

template <int D>
static int launch_blur_sub(const float* res, float* out, int nplanes, int T, int lo, int hi, hipStream_t st,
                           unsigned long long* marks = nullptr, float sign = 1.0f, float thr = 0.0f) {
  if (2 * D + 1 > 63) { mm_set_error("mm_blur_sub_xy: window too wide"); return MM_EINVAL; }
  const int ww = (T + 63) / 64;
  const size_t lds = sizeof(float) * (64 + 2 * BLUR_BX + (size_t)BLUR_BX * blur_pitch(T, D, lo, hi)) +
                     (marks ? sizeof(unsigned long long) * BLUR_BX * ww : 0);
  if (lds > 163840) { mm_set_error("mm_blur_sub_xy: plane of %d columns too wide", T); return MM_EINVAL; }
  const int npatch = (hi - lo + BLUR_BX - 1) / BLUR_BX;
  const long long nblk = (long long)npatch * nplanes;
  if (nblk > 0x7fffffffLL) { mm_set_error("mm_blur_sub_xy: too many planes"); return MM_EINVAL; }
  if (marks)
    hipLaunchKernelGGL((k_blur_sub_xy<D, true>), dim3((unsigned)nblk), dim3(256), lds, st, res, out, T, nplanes,
                       npatch, lo, hi, sign, thr, marks, ww);
  else
    hipLaunchKernelGGL((k_blur_sub_xy<D, false>), dim3((unsigned)nblk), dim3(256), lds, st, res, out, T, nplanes,
                       npatch, lo, hi, 0.0f, 0.0f, nullptr, 0);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

// The windowed fused blur that also writes the maxima threshold bits (see
// k_blur_sub_xy): the first stage of mm_tiles_blur_maxima_window.
int mm_blur_sub_mark(const float* res, float* out, int nplanes, int T, int delta, int lo, int hi, float sign,
                     float thr, unsigned long long* marks, hipStream_t st) {
  if (delta == 15) return launch_blur_sub<15>(res, out, nplanes, T, lo, hi, st, marks, sign, thr);
  if (delta == 25) return launch_blur_sub<25>(res, out, nplanes, T, lo, hi, st, marks, sign, thr);
  mm_set_error("mm_tiles_blur_maxima_window: no fused blur for delta %d (15 and 25 only)", delta);
  return MM_EINVAL;
}

// out = res - box_y(box_x(res)) over nplanes (T, T) planes; `out` must not overlap
// `res`.  delta 15 and 25 (blur_radius 30/31 and 50/51) run fused; other widths
// run the two single-axis passes through `tmp` (nplanes*T*T floats, unused and
// may be NULL for the fused widths).
extern "C" int mm_blur_sub_xy(const float* res, float* out, float* tmp, int nplanes, int T, int delta, void* stream) {
  if (!res || !out || nplanes < 0 || T < 1 || delta < 0) { mm_set_error("mm_blur_sub_xy: bad argument"); return MM_EINVAL; }
  if (nplanes == 0) return MM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (delta == 15) return launch_blur_sub<15>(res, out, nplanes, T, 0, T, st);
  if (delta == 25) return launch_blur_sub<25>(res, out, nplanes, T, 0, T, st);
  if (!tmp) { mm_set_error("mm_blur_sub_xy: delta %d needs a temporary buffer", delta); return MM_EINVAL; }
  const int rc = box_launch(res, nullptr, tmp, nplanes, T, T, delta, 1, false, stream);
  if (rc) return rc;
  return box_launch(tmp, res, out, nplanes, T, T, delta, 2, true, stream);
}

// mm_blur_sub_xy restricted to rows and columns [lo, hi) of every plane: out
// holds res - box_y(box_x(res)) there (the box windows still clip at the plane
// edges 0 and T) and is left untouched elsewhere.  Widths without a fused
// kernel compute the whole plane.
extern "C" int mm_blur_sub_xy_window(const float* res, float* out, float* tmp, int nplanes, int T, int delta, int lo,
                                     int hi, void* stream) {
  if (!res || !out || nplanes < 0 || T < 1 || delta < 0 || lo < 0 || hi > T || lo >= hi) {
    mm_set_error("mm_blur_sub_xy_window: bad argument (lo %d hi %d T %d)", lo, hi, T);
    return MM_EINVAL;
  }
  if (nplanes == 0) return MM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (delta == 15) return launch_blur_sub<15>(res, out, nplanes, T, lo, hi, st);
  if (delta == 25) return launch_blur_sub<25>(res, out, nplanes, T, lo, hi, st);
  return mm_blur_sub_xy(res, out, tmp, nplanes, T, delta, stream);
}

// ------------------------------------------------------------------ reflect
struct NdDesc {
  int ndim;
  long long shape[8];
  long long stride[8];
};

// fill.py:21-40: mode 0 ('out') writes j > i from 2i-j >= 0, mode 1 ('in')
// writes j < i from 2i-j < n.  Each mode reads one side of i and writes the
// other, so the in-place update has no race.
__global__ __launch_bounds__(256) void k_reflect(float* a, NdDesc d, int axis, int i, int mode, long long total) {
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    long long tmp = idx, off = 0, j = 0;
    for (int k = d.ndim - 1; k >= 0; --k) {
      const long long c = tmp % d.shape[k];
      tmp /= d.shape[k];
      off += c * d.stride[k];
      if (k == axis) j = c;
    }
    const long long n = d.shape[axis];
    const long long src = 2LL * i - j;
    if (mode == 0) {
      if (j <= i || src < 0) continue;
    } else {
      if (j >= i || src >= n) continue;
    }
    a[off] = a[off + (src - j) * d.stride[axis]];
  }
}

extern "C" int mm_reflect(float* arr, int ndim, const long long* shape, const long long* strides, int axis, int i,
                          int mode, void* stream) {
  if (!arr || ndim < 1 || ndim > 8 || axis < 0 || axis >= ndim || (mode != 0 && mode != 1)) {
    mm_set_error("mm_reflect: bad argument");
    return MM_EINVAL;
  }
  NdDesc d;
  d.ndim = ndim;
  long long total = 1;
  for (int k = 0; k < 8; ++k) {
    d.shape[k] = k < ndim ? shape[k] : 1;
    d.stride[k] = k < ndim ? strides[k] : 0;
  }
  for (int k = 0; k < ndim; ++k) total *= shape[k];
  if (i < 0 || i >= shape[axis]) { mm_set_error("mm_reflect: index out of range"); return MM_EINVAL; }
  if (total == 0) return MM_OK;
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_reflect, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, arr, d, axis, i, mode,
                     total);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

// ------------------------------------------------------------ RMS normalise
// DAPI path of mermake/__main__.py:221-222: buffer /= ||buffer||_2 / sqrt(N),
// with the norm accumulated in float64 and the divisor rounded to float32.
#define RMS_BLOCKS 2048
// Element forms (any alignment).
__global__ __launch_bounds__(256) void k_sumsq1(const float* __restrict__ a, long long n, double* __restrict__ part) {
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double v = a[i];
    s += v * v;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_div_by1(float* __restrict__ a, long long n, const float* __restrict__ d) {
  const float q = *d;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    a[i] = a[i] / q;
}

// Partial sums of squares (float64) per block: float4 loads when the array is
// 16-byte aligned (n4 of them), then the ntail elements after them.
__global__ __launch_bounds__(256) void k_sumsq(const float* __restrict__ a, long long n4, int ntail,
                                               double* __restrict__ part) {
  double s = 0.0;
  const float4* a4 = (const float4*)a;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
    const float4 v = a4[i];
    const double x = v.x, y = v.y, z = v.z, w = v.w;
    s += x * x;
    s += y * y;
    s += z * z;
    s += w * w;
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < ntail; i += blockDim.x) {
      const double v = a[4 * n4 + i];
      s += v * v;
    }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void k_rms_final(const double* __restrict__ part, int np, long long n,
                                                   float* __restrict__ rms) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < np; k += blockDim.x) s += part[k];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *rms = (float)(sqrt(sh[0]) / sqrt((double)n));
}

// a[i] / rms for the n4 float4 of the aligned part, then the ntail elements after them.
__global__ __launch_bounds__(256) void k_div_by(float* __restrict__ a, long long n4, int ntail,
                                                const float* __restrict__ d) {
  const float q = *d;
  float4* a4 = (float4*)a;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x) {
    float4 v = a4[i];
    v.x = v.x / q;
    v.y = v.y / q;
    v.z = v.z / q;
    v.w = v.w / q;
    a4[i] = v;
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < ntail; i += blockDim.x) a[4 * n4 + i] = a[4 * n4 + i] / q;
}

extern "C" size_t mm_rms_ws_bytes(void) { return RMS_BLOCKS * sizeof(double) + 256; }

extern "C" int mm_rms_normalize(float* a, long long n, void* ws, float* rms_out, void* stream) {
  if (!a || !ws || !rms_out || n < 1) { mm_set_error("mm_rms_normalize: bad argument"); return MM_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  double* part = (double*)ws;
  // float4 loads for a 16-byte aligned array (the last n % 4 elements by block 0); an unaligned
  // array of more than a few elements takes the element kernels
  const bool al = ((uintptr_t)a & 15) == 0;
  const long long n4 = al ? n / 4 : 0;
  const long long rest = n - 4 * n4;
  if (rest > 1024) {
    hipLaunchKernelGGL(k_sumsq1, dim3(RMS_BLOCKS), dim3(256), 0, st, a, n, part);
    MM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rms_final, dim3(1), dim3(256), 0, st, part, RMS_BLOCKS, n, rms_out);
    MM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_div_by1, dim3(4096), dim3(256), 0, st, a, n, rms_out);
    MM_LAUNCH_CHECK();
    return MM_OK;
  }
  hipLaunchKernelGGL(k_sumsq, dim3(RMS_BLOCKS), dim3(256), 0, st, a, n4, (int)rest, part);
  MM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_rms_final, dim3(1), dim3(256), 0, st, part, RMS_BLOCKS, n, rms_out);
  MM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_div_by, dim3(4096), dim3(256), 0, st, a, n4, (int)rest, rms_out);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

// ------------------------------------------------------------------ median (load_flats)
// Exact median of a float32 array by radix selection on the order-preserving
// 32-bit keys of the values (sign-flipped for negatives, so -0 < +0): four
// passes of 8-bit digit histograms, most significant digit first, each over the
// elements whose higher digits equal the prefix found so far; after each pass one
// thread walks the histogram to the digit holding the k-th element.  Any NaN
// makes the result NaN (numpy.median).  Workspace: 256 counters, the state
// (prefix, k) and a NaN count.
struct KthState {
  unsigned prefix, mask;
  unsigned long long k;
  unsigned nan;
  unsigned pad;
};

__device__ __forceinline__ unsigned f32_key(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_f32(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void k_kth_init(unsigned* hist, KthState* st, unsigned long long k) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0u;
  if (threadIdx.x == 0) *st = KthState{0u, 0u, k, 0u, 0u};
}

__global__ __launch_bounds__(256) void k_kth_hist(const float* __restrict__ a, long long n, int shift,
                                                  unsigned* __restrict__ hist, KthState* __restrict__ st) {
  __shared__ unsigned h[256];
  __shared__ unsigned nn;
  h[threadIdx.x] = 0u;
  if (threadIdx.x == 0) nn = 0u;
  __syncthreads();
  const unsigned prefix = st->prefix, mask = st->mask;
  unsigned nan = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float v = a[i];
    if (v != v) {
      ++nan;
      continue;
    }
    const unsigned key = f32_key(v);
    if ((key & mask) == prefix) atomicAdd(&h[(key >> shift) & 255u], 1u);
  }
  if (shift == 24 && nan) atomicAdd(&nn, nan);
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
  if (shift == 24 && threadIdx.x == 0 && nn) atomicAdd(&st->nan, nn);
}

// Digit of the k-th key; the histogram is cleared for the next pass.  On the last
// pass the selected value goes to out.
__global__ void k_kth_select(unsigned* hist, KthState* st, int shift, float* out) {
  if (threadIdx.x != 0) return;
  unsigned long long k = st->k, run = 0;
  unsigned d = 255;
  for (unsigned i = 0; i < 256; ++i) {
    const unsigned long long c = hist[i];
    if (k < run + c) {
      d = i;
      break;
    }
    run += c;
  }
  st->k = k - run;
  st->prefix |= d << shift;
  st->mask |= 255u << shift;
  for (int i = 0; i < 256; ++i) hist[i] = 0u;
  if (shift == 0 && out) *out = st->nan ? __uint_as_float(0x7fc00000u) : key_f32(st->prefix);
}

__global__ void k_median_pair(const float* lo, const float* hi, float* out) {
  if (threadIdx.x == 0) *out = (*lo + *hi) / 2.0f;
}

extern "C" size_t mm_median_ws_bytes(void) { return 256 * sizeof(unsigned) + 2 * sizeof(KthState) + 2 * sizeof(float) + 64; }

static int kth_smallest(const float* a, long long n, unsigned long long k, unsigned* hist, KthState* st, float* out,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_kth_init, dim3(1), dim3(256), 0, s, hist, st, k);
  MM_LAUNCH_CHECK();
  long long blocks = (n + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_kth_hist, dim3((unsigned)blocks), dim3(256), 0, s, a, n, shift, hist, st);
    MM_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_kth_select, dim3(1), dim3(64), 0, s, hist, st, shift, out);
    MM_LAUNCH_CHECK();
  }
  return MM_OK;
}

extern "C" int mm_median(const float* a, long long n, float* out, void* ws, size_t ws_bytes, void* stream) {
  if (!a || !out || !ws || n < 1 || ws_bytes < mm_median_ws_bytes()) {
    mm_set_error("mm_median: bad argument");
    return MM_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  unsigned* hist = (unsigned*)ws;
  KthState* st = (KthState*)(hist + 256);
  float* tmp = (float*)(st + 2);
  int rc;
  if (n % 2) return kth_smallest(a, n, (unsigned long long)(n / 2), hist, st, out, s);
  if ((rc = kth_smallest(a, n, (unsigned long long)(n / 2 - 1), hist, st, tmp, s))) return rc;
  if ((rc = kth_smallest(a, n, (unsigned long long)(n / 2), hist, st, tmp + 1, s))) return rc;
  hipLaunchKernelGGL(k_median_pair, dim3(1), dim3(64), 0, s, tmp, tmp + 1, out);
  MM_LAUNCH_CHECK();
  return MM_OK;
}
