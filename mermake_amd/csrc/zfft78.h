// Length-78 z transforms in registers for the z FFT form of the DAPI tile
// (sz 40, zpad 19: Zp = 78 padded planes), one column per thread.
//
// 78 = 2 * 3 * 13 with coprime factors, so the DFT is a 2 x 3 x 13
// multidimensional DFT without twiddles (Good's prime-factor mapping):
//   input   x[n],  n = (39 n1 + 26 n2 + 6 n3) mod 78
//   output  X[k],  k = k1 (mod 2) = k2 (mod 3) = k3 (mod 13)
//   X[k] = sum x[n] w2^(n1 k1) w3^(n2 k2) w13^(n3 k3).
// The values live in a[pos], pos = i1 * 39 + i2 * 13 + i3 (i = n before the
// transforms, k after them).  The inverse takes the spectrum in that same
// arrangement and returns x in the input one, so forward -> multiply ->
// inverse needs no reordering.  The DFT-13 pairs x[j] with x[13 - j]
// (6 cosine and 6 sine sums per output pair); DFT-3 and DFT-2 are direct.
#pragma once
#include "common.h"

// A scheduling fence between the sub-transforms: the compiler would otherwise
// interleave them for instruction-level parallelism and hold all their
// temporaries at once, past the register file.  (Nothing in the host pass,
// which only the CPU check of the index maps compiles.)
#ifdef __HIP_DEVICE_COMPILE__
#define ZF78_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define ZF78_FENCE()
#endif

namespace zf78 {

__host__ __device__ constexpr int n_of(int pos) {  // input index held at pos
  return (39 * (pos / 39) + 26 * ((pos / 13) % 3) + 6 * (pos % 13)) % 78;
}
// (closed forms, loop-free, so that an unrolled loop index folds to a constant
// register index)
__host__ __device__ constexpr int k_of(int pos) {  // output index held at pos: CRT with 39 = 1 (2),
  return (39 * (pos / 39) + 52 * ((pos / 13) % 3) + 66 * (pos % 13)) % 78;  // 52 = 1 (3), 66 = 1 (13)
}
__host__ __device__ constexpr int pos_of_n(int n) {  // position of input index n: n1 = n mod 2,
  return 39 * (n % 2) + 13 * ((2 * (n % 3)) % 3) + (11 * (n % 13)) % 13;  // 2 n2 = n (3), 6 n3 = n (13)
}

// cos and sin of 2 pi m / 13, m = 1..6
__host__ __device__ constexpr float c13(int m) {
  return m == 1 ? 0.8854560256532099f : m == 2 ? 0.5680647467311559f : m == 3 ? 0.120536680255323f
       : m == 4 ? -0.35460488704253545f : m == 5 ? -0.7485107481711012f : -0.970941817426052f;
}
__host__ __device__ constexpr float s13(int m) {
  return m == 1 ? 0.4647231720437685f : m == 2 ? 0.8229838658936564f : m == 3 ? 0.992708874098054f
       : m == 4 ? 0.9350162426854148f : m == 5 ? 0.6631226582407952f : 0.23931566428755768f;
}

// DFT-13 of a[b], a[b + 1], ..., a[b + 12] in place (INV: positive exponent).
template <bool INV, int B>
__host__ __device__ __forceinline__ void dft13(float2* a) {
  float2 p[6], q[6];
#pragma unroll
  for (int j = 1; j <= 6; ++j) {
    p[j - 1] = make_float2(a[B + j].x + a[B + 13 - j].x, a[B + j].y + a[B + 13 - j].y);
    q[j - 1] = make_float2(a[B + j].x - a[B + 13 - j].x, a[B + j].y - a[B + 13 - j].y);
  }
  const float2 x0 = a[B];
  float2 s = x0;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    s.x += p[j].x;
    s.y += p[j].y;
  }
#pragma unroll
  for (int k = 1; k <= 6; ++k) {
    float2 A = x0, Bs = make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 1; j <= 6; ++j) {
      const int m = (j * k) % 13;
      const float cm = m <= 6 ? c13(m) : c13(13 - m);
      const float sm = m <= 6 ? s13(m) : -s13(13 - m);
      A.x = fmaf(p[j - 1].x, cm, A.x);
      A.y = fmaf(p[j - 1].y, cm, A.y);
      Bs.x = fmaf(q[j - 1].x, sm, Bs.x);
      Bs.y = fmaf(q[j - 1].y, sm, Bs.y);
    }
    // forward: X[k] = A - i Bs, X[13 - k] = A + i Bs; the inverse swaps them
    const float2 lo = make_float2(A.x + Bs.y, A.y - Bs.x), hi = make_float2(A.x - Bs.y, A.y + Bs.x);
    a[B + k] = INV ? hi : lo;
    a[B + 13 - k] = INV ? lo : hi;
  }
  a[B] = s;
}

// DFT-3 of a[b], a[b + 13], a[b + 26].
template <bool INV, int B>
__host__ __device__ __forceinline__ void dft3(float2* a) {
  const float h = 0.8660254037844386f;  // sqrt(3) / 2
  const float2 x0 = a[B], x1 = a[B + 13], x2 = a[B + 26];
  const float2 t = make_float2(x0.x - 0.5f * (x1.x + x2.x), x0.y - 0.5f * (x1.y + x2.y));
  const float2 u = make_float2(h * (x1.x - x2.x), h * (x1.y - x2.y));
  a[B] = make_float2(x0.x + x1.x + x2.x, x0.y + x1.y + x2.y);
  const float2 lo = make_float2(t.x + u.y, t.y - u.x), hi = make_float2(t.x - u.y, t.y + u.x);
  a[B + 13] = INV ? hi : lo;
  a[B + 26] = INV ? lo : hi;
}

template <bool INV, int I>
struct Dft3Cols {  // DFT-3 over i2 for every (i1, i3), i3 = I .. 12
  __host__ __device__ static __forceinline__ void run(float2* a) {
    dft3<INV, I>(a);
    dft3<INV, 39 + I>(a);
    ZF78_FENCE();
    Dft3Cols<INV, I + 1>::run(a);
  }
};
template <bool INV>
struct Dft3Cols<INV, 13> {
  __host__ __device__ static __forceinline__ void run(float2*) {}
};

// The 78-point DFT of a (arrangement above), in place.
template <bool INV>
__host__ __device__ __forceinline__ void fft78(float2* a) {
  dft13<INV, 0>(a);
  ZF78_FENCE();
  dft13<INV, 13>(a);
  ZF78_FENCE();
  dft13<INV, 26>(a);
  ZF78_FENCE();
  dft13<INV, 39>(a);
  ZF78_FENCE();
  dft13<INV, 52>(a);
  ZF78_FENCE();
  dft13<INV, 65>(a);
  ZF78_FENCE();
  Dft3Cols<INV, 0>::run(a);
#pragma unroll
  for (int i = 0; i < 39; ++i) {  // DFT-2 over i1
    const float2 x0 = a[i], x1 = a[39 + i];
    a[i] = make_float2(x0.x + x1.x, x0.y + x1.y);
    a[39 + i] = make_float2(x0.x - x1.x, x0.y - x1.y);
  }
}

}  // namespace zf78
