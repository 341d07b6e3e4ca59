// Drift between acquisitions from point clouds (reference mermake/align.py:12-177,
// mermake/utils.py:10-36) on gfx950.
//
//   k_round_bounds  round(X/resc) half-to-even -> int32, min/max per axis
//   k_kept_max      max of the points inside the trim box (image size)
//   k_raster        binary image written into a zero-padded complex64 FFT grid
//                   (optionally flipped, for the correlation kernel)
//   k_cmul          spectrum product
//   k_argmax_*      first index of the maximum of rint(Re(corr)) over the
//                   linear ('full') correlation region: correlations of binary
//                   images are integers, so this is exact and deterministic
//   k_nn            brute-force nearest reference point (lowest index on ties)
//   k_nn_mean       deterministic mean of the matched differences (float64)
#include "common.h"
#include "mermake_hip.h"

__global__ __launch_bounds__(1024) void k_round_bounds(const float* __restrict__ X, int n, int ld, float resc,
                                                       int* __restrict__ Xi, int* __restrict__ bounds) {
  int mn[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
  int mx[3] = {-0x7fffffff - 1, -0x7fffffff - 1, -0x7fffffff - 1};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    for (int k = 0; k < 3; ++k) {
      const int r = (int)rintf(X[(long long)i * ld + k] / resc);  // float32 divide, half-to-even
      Xi[i * 3 + k] = r;
      mn[k] = min(mn[k], r);
      mx[k] = max(mx[k], r);
    }
  }
  __shared__ int s[6][1024];
  for (int k = 0; k < 3; ++k) { s[k][threadIdx.x] = mn[k]; s[3 + k][threadIdx.x] = mx[k]; }
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int k = 0; k < 3; ++k) {
        s[k][threadIdx.x] = min(s[k][threadIdx.x], s[k][threadIdx.x + st]);
        s[3 + k][threadIdx.x] = max(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) bounds[threadIdx.x] = s[threadIdx.x][0];
}

struct Box {
  int lo[3], hi[3];
};

__device__ __forceinline__ bool in_box(const int* p, const Box& b) {
  return p[0] >= b.lo[0] && p[0] <= b.hi[0] && p[1] >= b.lo[1] && p[1] <= b.hi[1] && p[2] >= b.lo[2] &&
         p[2] <= b.hi[2];
}

__global__ __launch_bounds__(1024) void k_kept_max(const int* __restrict__ Xi, int n, Box b, int* __restrict__ out) {
  int mx[3] = {-1, -1, -1};
  int cnt = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (in_box(Xi + i * 3, b)) {
      ++cnt;
      for (int k = 0; k < 3; ++k) mx[k] = max(mx[k], Xi[i * 3 + k]);
    }
  }
  __shared__ int s[4][1024];
  for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = mx[k];
  s[3][threadIdx.x] = cnt;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = max(s[k][threadIdx.x], s[k][threadIdx.x + st]);
      s[3][threadIdx.x] += s[3][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_raster(const int* __restrict__ Xi, int n, Box b, float2* __restrict__ buf, int L1, int L2, int flip,
                         int R0, int R1, int R2) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int* p = Xi + i * 3;
    if (!in_box(p, b)) continue;
    // image origin is 0 (align.py:55); negative coordinates wrap like the
    // reference's NumPy/CuPy fancy index imf[tuple(X.T)] = 1 (align.py:58)
    int z = p[0] < 0 ? p[0] + R0 : p[0];
    int x = p[1] < 0 ? p[1] + R1 : p[1];
    int y = p[2] < 0 ? p[2] + R2 : p[2];
    if (flip) { z = R0 - 1 - z; x = R1 - 1 - x; y = R2 - 1 - y; }
    buf[((long long)z * L1 + x) * L2 + y] = make_float2(1.0f, 0.0f);
  }
}

__global__ void k_cmul(float2* __restrict__ a, const float2* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 x = a[i], y = b[i];
    a[i] = make_float2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
  }
}

// argmax over the region [0,F0) x [0,F1) x [0,F2) of the (L0,L1,L2) grid.
__device__ __forceinline__ void better(float& bv, long long& bi, float v, long long i) {
  if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
}

__global__ __launch_bounds__(256) void k_argmax_part(const float2* __restrict__ buf, int L1, int L2, int F0, int F1,
                                                     int F2, float* __restrict__ pv, long long* __restrict__ pi) {
  const long long n = (long long)F0 * F1 * F2;
  float bv = -3.0e38f;
  long long bi = 0x7fffffffffffffffLL;
  for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n; f += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(f / ((long long)F1 * F2));
    const int rem = (int)(f - (long long)z * F1 * F2);
    const int x = rem / F2, y = rem - (rem / F2) * F2;
    const float v = rintf(buf[((long long)z * L1 + x) * L2 + y].x);
    better(bv, bi, v, f);  // f is the raster index of the 'full' volume
  }
  __shared__ float sv[256];
  __shared__ long long si[256];
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      float v = sv[threadIdx.x];
      long long i = si[threadIdx.x];
      better(v, i, sv[threadIdx.x + st], si[threadIdx.x + st]);
      sv[threadIdx.x] = v;
      si[threadIdx.x] = i;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { pv[blockIdx.x] = sv[0]; pi[blockIdx.x] = si[0]; }
}

__global__ __launch_bounds__(256) void k_argmax_final(const float* __restrict__ pv, const long long* __restrict__ pi,
                                                      int np, long long* __restrict__ out_i, float* __restrict__ out_v) {
  float bv = -3.0e38f;
  long long bi = 0x7fffffffffffffffLL;
  for (int k = threadIdx.x; k < np; k += blockDim.x) better(bv, bi, pv[k], pi[k]);
  __shared__ float sv[256];
  __shared__ long long si[256];
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      float v = sv[threadIdx.x];
      long long i = si[threadIdx.x];
      better(v, i, sv[threadIdx.x + st], si[threadIdx.x + st]);
      sv[threadIdx.x] = v;
      si[threadIdx.x] = i;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { *out_i = si[0]; *out_v = sv[0]; }
}

// nearest reference point of X[i] - t (align.py:67-68); float64 distances
// 16 lanes per query point: lane s of a query scans the reference points j = s (mod 16) of each
// 256-point LDS tile, and the 16 partial minima are merged by (distance, then lowest index), so the
// result is that of one ascending scan with a strict comparison, while a launch has n / 16 blocks of
// work instead of n / 256 (a few thousand queries would otherwise leave most CUs idle).
#define NN_LANES 16
__global__ __launch_bounds__(256) void k_nn(const float* __restrict__ X, int n, int ldx, const float* __restrict__ R,
                                            int m, int ldr, const double* __restrict__ t, double* __restrict__ ds,
                                            int* __restrict__ idx) {
  __shared__ double sr[256][3];
  const int s = threadIdx.x % NN_LANES;
  const int i = blockIdx.x * (blockDim.x / NN_LANES) + threadIdx.x / NN_LANES;
  double q0 = 0, q1 = 0, q2 = 0;
  if (i < n) {
    q0 = (double)X[(long long)i * ldx + 0] - t[0];
    q1 = (double)X[(long long)i * ldx + 1] - t[1];
    q2 = (double)X[(long long)i * ldx + 2] - t[2];
  }
  double best = 1.0e300;
  int bi = 0x7fffffff;
  for (int j0 = 0; j0 < m; j0 += 256) {
    const int j = j0 + threadIdx.x;
    if (j < m) {
      sr[threadIdx.x][0] = R[(long long)j * ldr + 0];
      sr[threadIdx.x][1] = R[(long long)j * ldr + 1];
      sr[threadIdx.x][2] = R[(long long)j * ldr + 2];
    }
    __syncthreads();
    const int jm = min(256, m - j0);
    for (int k = s; k < jm; k += NN_LANES) {
      const double a = q0 - sr[k][0], b = q1 - sr[k][1], c = q2 - sr[k][2];
      const double d = a * a + b * b + c * c;
      if (d < best) { best = d; bi = j0 + k; }
    }
    __syncthreads();
  }
  // merge the 16 partial minima of the query: the smaller distance, the lower index on equal distances
#pragma unroll
  for (int off = NN_LANES / 2; off > 0; off >>= 1) {
    const double ob = __shfl_xor(best, off, NN_LANES);
    const int oi = __shfl_xor(bi, off, NN_LANES);
    if (ob < best || (ob == best && oi < bi)) { best = ob; bi = oi; }
  }
  if (i < n && s == 0) { ds[i] = sqrt(best); idx[i] = bi == 0x7fffffff ? 0 : bi; }
}

// One refinement round of align.py:66-76: keep pairs with ds < th; if any,
// t = mean(X[keep] - R[idx[keep]]); npts = number kept.  Single block.
__global__ __launch_bounds__(256) void k_nn_mean(const float* __restrict__ X, int n, int ldx,
                                                 const float* __restrict__ R, int ldr, const double* __restrict__ ds,
                                                 const int* __restrict__ idx, double th, double* __restrict__ t,
                                                 int* __restrict__ npts) {
  double s0 = 0, s1 = 0, s2 = 0;
  int c = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (ds[i] < th) {
      const int j = idx[i];
      s0 += (double)X[(long long)i * ldx + 0] - (double)R[(long long)j * ldr + 0];
      s1 += (double)X[(long long)i * ldx + 1] - (double)R[(long long)j * ldr + 1];
      s2 += (double)X[(long long)i * ldx + 2] - (double)R[(long long)j * ldr + 2];
      ++c;
    }
  }
  __shared__ double sh[3][256];
  __shared__ int sc[256];
  sh[0][threadIdx.x] = s0; sh[1][threadIdx.x] = s1; sh[2][threadIdx.x] = s2; sc[threadIdx.x] = c;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + st];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + st];
      sh[2][threadIdx.x] += sh[2][threadIdx.x + st];
      sc[threadIdx.x] += sc[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int k = sc[0];
    if (k > 0) { t[0] = sh[0][0] / k; t[1] = sh[1][0] / k; t[2] = sh[2][0] / k; }
    *npts = k;
  }
}

// ------------------------------------------------------------------- host
extern "C" int mm_points_round_bounds(const float* X, int n, int ld, float resc, int* Xi, int* bounds, void* stream) {
  if (!X || !Xi || !bounds || n < 1) { mm_set_error("mm_points_round_bounds: bad argument"); return MM_EINVAL; }
  hipLaunchKernelGGL(k_round_bounds, dim3(1), dim3(1024), 0, (hipStream_t)stream, X, n, ld, resc, Xi, bounds);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_points_kept_max(const int* Xi, int n, const int* lo, const int* hi, int* out4, void* stream) {
  if (!Xi || !lo || !hi || !out4) { mm_set_error("mm_points_kept_max: bad argument"); return MM_EINVAL; }
  Box b;
  for (int k = 0; k < 3; ++k) { b.lo[k] = lo[k]; b.hi[k] = hi[k]; }
  hipLaunchKernelGGL(k_kept_max, dim3(1), dim3(1024), 0, (hipStream_t)stream, Xi, n, b, out4);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_points_raster(const int* Xi, int n, const int* lo, const int* hi, void* buf, int L0, int L1, int L2,
                                int flip, int R0, int R1, int R2, void* stream) {
  if (!Xi || !buf || R0 > L0 || R1 > L1 || R2 > L2) { mm_set_error("mm_points_raster: bad argument"); return MM_EINVAL; }
  Box b;
  for (int k = 0; k < 3; ++k) { b.lo[k] = lo[k]; b.hi[k] = hi[k]; }
  // kept points lie in [lo, R) per axis (R = kept max + 1); wrapping keeps them in
  // [0, R) when lo >= -R, the range a NumPy negative index accepts
  const int R[3] = {R0, R1, R2};
  for (int k = 0; k < 3; ++k)
    if (R[k] < 1 || b.lo[k] < -R[k]) { mm_set_error("mm_points_raster: coordinate out of the image"); return MM_EINVAL; }
  hipLaunchKernelGGL(k_raster, dim3(64), dim3(256), 0, (hipStream_t)stream, Xi, n, b, (float2*)buf, L1, L2, flip, R0,
                     R1, R2);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_cmul(void* a, const void* b, long long n, void* stream) {
  if (!a || !b) { mm_set_error("mm_cmul: null"); return MM_EINVAL; }
  hipLaunchKernelGGL(k_cmul, dim3(2048), dim3(256), 0, (hipStream_t)stream, (float2*)a, (const float2*)b, n);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" size_t mm_argmax_ws_bytes(void) { return 1024 * (sizeof(float) + sizeof(long long)); }

extern "C" int mm_argmax_rounded(const void* buf, int L0, int L1, int L2, int F0, int F1, int F2, long long* out_i,
                                 float* out_v, void* ws, void* stream) {
  if (!buf || !out_i || !out_v || !ws || F0 > L0 || F1 > L1 || F2 > L2) {
    mm_set_error("mm_argmax_rounded: bad argument");
    return MM_EINVAL;
  }
  float* pv = (float*)ws;
  long long* pi = (long long*)((char*)ws + 1024 * sizeof(float));
  hipLaunchKernelGGL(k_argmax_part, dim3(1024), dim3(256), 0, (hipStream_t)stream, (const float2*)buf, L1, L2, F0, F1,
                     F2, pv, pi);
  MM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_argmax_final, dim3(1), dim3(256), 0, (hipStream_t)stream, pv, pi, 1024, out_i, out_v);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_nn_round(const float* X, int n, int ldx, const float* R, int m, int ldr, double* t, double th,
                           double* ds_ws, int* idx_ws, int* npts, void* stream) {
  if (!X || !R || !t || !ds_ws || !idx_ws || !npts || m < 1) { mm_set_error("mm_nn_round: bad argument"); return MM_EINVAL; }
  if (n < 1) {
    MM_CHECK(hipMemsetAsync(npts, 0, sizeof(int), (hipStream_t)stream));
    return MM_OK;
  }
  hipLaunchKernelGGL(k_nn, dim3((n + 256 / NN_LANES - 1) / (256 / NN_LANES)), dim3(256), 0, (hipStream_t)stream, X, n,
                     ldx, R, m, ldr, t, ds_ws, idx_ws);
  MM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_nn_mean, dim3(1), dim3(256), 0, (hipStream_t)stream, X, n, ldx, R, ldr, ds_ws, idx_ws, th, t,
                     npts);
  MM_LAUNCH_CHECK();
  return MM_OK;
}
