// Shared helpers for the MERMAKE image-path kernels (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MM_WAVE 64

// error reporting of the C-ABI layer (defined in capi.hip)
void mm_set_error(const char* fmt, ...);

#define MM_CHECK(expr)                                                            \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      mm_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),         \
                   __FILE__, __LINE__);                                           \
      return (int)e_;                                                             \
    }                                                                             \
  } while (0)

#define MM_LAUNCH_CHECK()                                                         \
  do {                                                                            \
    hipError_t e_ = hipGetLastError();                                            \
    if (e_ != hipSuccess) {                                                       \
      mm_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),     \
                   __FILE__, __LINE__);                                           \
      return (int)e_;                                                             \
    }                                                                             \
  } while (0)

enum { MM_OK = 0, MM_EINVAL = 1000, MM_ECAP = 1001, MM_ENOMEM = 1002 };

// fused windowed blur subtraction (delta 15 or 25) that also writes the maxima
// threshold bits, ceil(T / 64) words per row of the nplanes x T rows (blur_fill.hip)
int mm_blur_sub_mark(const float* res, float* out, int nplanes, int T, int delta, int lo, int hi, float sign,
                     float thr, unsigned long long* marks, hipStream_t st);

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap so that consecutive logical block ids run on the same XCD: neighbours
// then share L2 lines (adjacent column blocks of a strided pass).
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nblocks) {
  const unsigned nx = 8;
  if (nblocks < nx * 2) return bid;
  unsigned per = nblocks / nx;           // full rounds
  unsigned full = per * nx;
  if (bid >= full) return bid;           // tail keeps identity
  return (bid % nx) * per + bid / nx;
}
