// Prime-first FFT with the prime radix p run by Rader's algorithm, for
// T = p * M: p = 113 (678 = 113 * 6, the full-FOV tile) and p = 239 (478 =
// 239 * 2, the tile of tile_size 300 and overlap 89).
//
// For a butterfly j of the prime-first order (stride 1, no twiddles) the DFT
//   X_k = sum_{r=0}^{p-1} x_r w^{r k},  w = exp(-/+ 2 pi i / p)
// is, with a primitive root gamma, r = gamma^-a and k = gamma^b,
//   X_{gamma^b} = x_0 + (f (*) g)_b,  f_a = x_{gamma^-a},  g_n = w^{gamma^n},
//   X_0 = x_0 + sum_a f_a,
// a cyclic convolution of length L = p - 1 = A * B with A and B coprime
// (112 = 16 * 7, 238 = 17 * 14), so the FFT-L runs in the Good-Thomas
// (prime-factor) form without twiddles: with n = B n1 + A n2 and
// k = cA k1 + cB k2 (mod L; cA = 1 mod A, 0 mod B and cB the other way round),
//   WL^{n k} = WA^{n1 k1} WB^{n2 k2}.
// Forward: DFT-A over n1, then DFT-B over n2; the product with
// G = FFT(g) / L (precomputed in float64, stored in the (k1, k2) order); the
// inverse: DFT-B over k2, then DFT-A over k1, whose outputs (m1, m2) are
// C_b at b = B m1 + A m2.  The index maps live in the tables: slot_in places
// f_a at its (n1, n2) slot, perm_out maps (m1, m2) straight to gamma^b.  The
// two radix-B passes act on the same B elements, so they run back to back in
// registers.  About 2.6 times fewer operations than the direct symmetric DFT
// for p = 113.
//
// LDS per transform: M * S complex (slot L holds x_0, slot L + 1 holds X_0)
// in one buffer worked in place; the last stage leaves the transforms X of
// the M sub-sequences at stride p, in natural order for the radix-M final
// stage of fft_pf.h (row passes) or in Rader's output order for rd_final
// (strided pass: no scattered LDS stores).
#pragma once
#include "fft_pf.h"

#define RD_NI 16  // fetch items (single elements) per thread: G * M * p <= RD_NI * blockDim.x

// Device tables (float64-accurate, float32 storage): ghat forward / inverse
// [L] each (already divided by L) at k1 + A k2; perm_out[m2 + B m1] =
// gamma^b mod p (b = B m1 + A m2 mod L); slot_in[p]: the slot n2 + B n1 of
// element r = gamma^-a (a = B n1 + A n2 mod L), L for r = 0.
struct RdTab {
  const float2* ghat;  // of the transform direction
  const int* perm_out;
  const int* slot_in;
};

// One data buffer per block, the stages in place (the third stage reads
// everything before a barrier, then writes).  Needs G M B <= blockDim.x.

// LDS bytes: twiddles of N, ghat, perm_out, slot_in, pos, the data buffer.
template <int P>
__host__ __device__ __forceinline__ size_t rd_lds_bytes(int N, int M, int G, size_t extra);

struct RdLds {
  float2* tw;
  float2* ghat;
  int* perm_out;
  int* slot_in;
  int* pos;  // [p]: slot of X_k in Rader's output order (0 for k = 0, 1 + s for k = perm_out[s])
  float2* bufA;
  float2* bufB;  // the same buffer (the stages run in place)
  unsigned char* extra;
};

// cos / sin (2 pi m / 17), m = 1..8
__device__ __constant__ static const double kCos17[8] = {
    0.9324722294043558, 0.7390089172206591, 0.4457383557765383, 0.09226835946330202,
    -0.2736629900720829, -0.6026346363792563, -0.850217135729614, -0.9829730996839018};
__device__ __constant__ static const double kSin17[8] = {
    0.3612416661871529, 0.6736956436465572, 0.8951632913550623, 0.9957341762950345,
    0.961825643172819, 0.7980172272802396, 0.5264321628773561, 0.18374951781657037};

// ------------------------------------------------------------ radix 16
// w16^m (forward: exp(-2 pi i m / 16)); m is a compile-time constant after unrolling.
template <bool INV>
__device__ __forceinline__ float2 w16mul(float2 a, int m) {
  const float c1 = 0.9238795325112867f, s1 = 0.3826834323650898f, h = 0.7071067811865476f;
  float c, s;  // w = (c, -s) forward
  switch (m & 15) {
    case 0: return a;
    case 4: return vmi<INV>(a);
    case 1: c = c1; s = s1; break;
    case 2: c = h; s = h; break;
    case 3: c = s1; s = c1; break;
    case 6: c = -h; s = h; break;
    case 9: c = -c1; s = -s1; break;
    default: c = 1.0f; s = 0.0f; break;  // not reached: bfly16 uses m in {0, 1, 2, 3, 4, 6, 9}
  }
  if (INV) s = -s;
  return make_float2(a.x * c + a.y * s, a.y * c - a.x * s);
}

// DFT-16 as 4 x 4 with twiddles w16^{n2 k1}; natural order in and out.
template <bool INV>
__device__ __forceinline__ void bfly16(float2* v) {
  float2 y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
    bfly4<INV>(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = w16mul<INV>(t[k1], n2 * k1);
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    bfly4<INV>(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

// ------------------------------------------------------------ radix 14
// DFT-14 in the Good-Thomas form 2 x 7: n = 7 a + 2 b, k = 7 k1 + 8 k2 (mod 14),
// W14^{n k} = W2^{a k1} W7^{b k2}; natural order in and out.
template <bool INV>
__device__ __forceinline__ void bfly14(float2* v) {
  float2 u[2][7];
#pragma unroll
  for (int b = 0; b < 7; ++b) {
    float2 t[2] = {v[(2 * b) % 14], v[(7 + 2 * b) % 14]};
    bfly2<INV>(t);
    u[0][b] = t[0];
    u[1][b] = t[1];
  }
  bfly_odd<7, INV>(u[0], kCos7, kSin7);
  bfly_odd<7, INV>(u[1], kCos7, kSin7);
#pragma unroll
  for (int k1 = 0; k1 < 2; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 7; ++k2) v[(7 * k1 + 8 * k2) % 14] = u[k1][k2];
}

// ------------------------------------------------------------ primes
// L = p - 1 = A * B (coprime); S: LDS slots per transform (L for f, then x_0
// and X_0, padded so that the transforms of a group start on different banks:
// 122 measured 0.5 ms faster than 114 per 40 x 2800^2 stack for p = 113).
template <int P> struct RdSpec;
template <> struct RdSpec<113> {
  static constexpr int L = 112, A = 16, B = 7, S = 122;
  template <bool INV> __device__ static __forceinline__ void bflyA(float2* v) { bfly16<INV>(v); }
  template <bool INV> __device__ static __forceinline__ void bflyB(float2* v) { bfly_odd<7, INV>(v, kCos7, kSin7); }
};
template <> struct RdSpec<239> {
  static constexpr int L = 238, A = 17, B = 14, S = 244;
  template <bool INV> __device__ static __forceinline__ void bflyA(float2* v) { bfly_odd<17, INV>(v, kCos17, kSin17); }
  template <bool INV> __device__ static __forceinline__ void bflyB(float2* v) { bfly14<INV>(v); }
};

template <int P>
__host__ __device__ __forceinline__ size_t rd_lds_bytes(int N, int M, int G, size_t extra) {
  using R = RdSpec<P>;
  return mm_align16((size_t)N * 8) + (size_t)R::L * 8 + mm_align16((size_t)(R::L + 2 * P) * 4) +
         (size_t)G * M * R::S * 8 + extra;
}

template <int P>
__device__ __forceinline__ RdLds rd_carve(unsigned char* smem, int N, int M, int G, const float2* __restrict__ tw,
                                          const RdTab& T) {
  using R = RdSpec<P>;
  RdLds L;
  size_t off = 0;
  L.tw = (float2*)smem;
  off += mm_align16((size_t)N * 8);
  L.ghat = (float2*)(smem + off);
  off += R::L * 8;
  L.perm_out = (int*)(smem + off);
  L.slot_in = L.perm_out + R::L;
  L.pos = L.slot_in + P;
  off += mm_align16((size_t)(R::L + 2 * P) * 4);
  L.bufA = (float2*)(smem + off);
  L.bufB = L.bufA;
  off += (size_t)G * M * R::S * 8;
  L.extra = smem + off;
  for (int i = threadIdx.x; i < N; i += blockDim.x) L.tw[i] = tw[i];
  for (int i = threadIdx.x; i < R::L; i += blockDim.x) {
    L.ghat[i] = T.ghat[i];
    L.perm_out[i] = T.perm_out[i];
    L.pos[T.perm_out[i]] = i + 1;  // perm_out is a permutation of 1 .. p - 1
  }
  if (threadIdx.x == 0) L.pos[0] = 0;
  for (int i = threadIdx.x; i < P; i += blockDim.x) L.slot_in[i] = T.slot_in[i];
  return L;
}

// ------------------------------------------------------------ fetch / put
// Item (g, j, r), r < p: element n = j + r M of transform g goes to slot
// slot_in[r] of DFT (g, j).  nn = n | g << 16; sl = LDS slot.
struct RdItems {
  unsigned nn[RD_NI], sl[RD_NI];
  int nval;
};

// GFAST: g fastest (strided pass); otherwise j, then r, then g (rows).
template <int P, int M, bool GFAST>
__device__ __forceinline__ void rd_items(int G, const int* __restrict__ slot_in, RdItems& I) {
  const int tot = G * M * P;
  const FastDiv fG(G), fp(P);
  I.nval = 0;
#pragma unroll
  for (int u = 0; u < RD_NI; ++u) {
    const int t = threadIdx.x + u * blockDim.x;
    I.nn[u] = 0;
    I.sl[u] = 0;
    if (t < tot) {
      int g, j, r;
      if (GFAST) {  // t = (r M + j) G + g
        int q;
        fG.divmod(t, q, g);
        r = q / M;
        j = q - r * M;
      } else {      // t = (g p + r) M + j
        const int q = t / M;
        j = t - q * M;
        fp.divmod(q, g, r);
      }
      I.nn[u] = (unsigned)(j + r * M) | ((unsigned)g << 16);
      I.sl[u] = (unsigned)((g * M + j) * RdSpec<P>::S + slot_in[r]);
      I.nval = u + 1;
    }
  }
}

// ld(u, g, n) issues the unconditional load of element n of transform g into
// the kernel's raw registers of item u.
template <class Ld>
__device__ __forceinline__ void rd_fetch(const RdItems& I, Ld ld) {
#pragma unroll
  for (int u = 0; u < RD_NI; ++u)
    if (u < I.nval) ld(u, (int)(I.nn[u] >> 16), (int)(I.nn[u] & 0xffffu));
}

// get(u, g, n) returns the element of item u; written to its slot of bufA.
template <class Get>
__device__ __forceinline__ void rd_put(float2* __restrict__ bufA, const RdItems& I, Get get) {
#pragma unroll
  for (int u = 0; u < RD_NI; ++u)
    if (u < I.nval) bufA[I.sl[u]] = get(u, (int)(I.nn[u] >> 16), (int)(I.nn[u] & 0xffffu));
}

// ------------------------------------------------------------ Rader stages
// Per-thread constants of the middle stage.  When A divides blockDim.x its
// butterfly q = it % A equals threadIdx.x % A in every round, so its
// G_{q + A r} (r < B) can stay in registers (REGC; p = 113 only).
struct RdMid {
  float2 gh[7];
};

__device__ __forceinline__ RdMid rd_mid(const RdLds& L) {
  const int q = threadIdx.x & 15;
  RdMid c;
#pragma unroll
  for (int r = 0; r < 7; ++r) c.gh[r] = L.ghat[q + 16 * r];
  return c;
}

// After rd_put and a barrier: bufA holds f (slots 0..L-1) and x_0 (slot L) of
// the G*M DFTs.  Ends (after its last barrier) with Y in bufB, at d = g M + j
// (N = p M):
//   BORD false: Y[d p + k] = X_k (natural order),
//   BORD true:  Y[d p] = X_0, Y[d p + 1 + s] = X_{perm_out[s]}, s = m2 + B m1
//               (rd_final reads it).
// Every stage works in place: stage 1 item q = n2 < B reads and writes slots
// q + B n1 (its output (n2, k1) at k1 * B + n2, so that both its stores and
// the middle stage's loads run across the banks), stage 2 item q = k1 < A
// slots q B + r; stage 3's outputs overlap other transforms' slots, so it reads
// everything, waits for the block, then writes (G M B <= blockDim.x).
template <int P, int M, bool BORD, bool REGC>
__device__ __forceinline__ void rd_stages(const RdLds& L, int G, const RdMid& mid) {
  using R = RdSpec<P>;
  constexpr int A = R::A, B = R::B, S = R::S, LL = R::L;
  static_assert(!REGC || (P == 113 && 256 % A == 0), "middle-stage constants in registers: A | blockDim.x");
  const int ndft = G * M;
  // forward DFT-A over n1: butterfly q = n2 < B reads slots q + B n1
  for (int it = threadIdx.x; it < ndft * B; it += blockDim.x) {
    const int d = it / B, q = it - (it / B) * B;
    float2* io = L.bufA + d * S + q;
    float2 v[A];
#pragma unroll
    for (int r = 0; r < A; ++r) v[r] = io[B * r];
    R::template bflyA<false>(v);
#pragma unroll
    for (int r = 0; r < A; ++r) io[B * r] = v[r];
  }
  __syncthreads();
  // DFT-B over n2 -> F(k1 = q, k2), times ghat, inverse DFT-B over k2
  for (int it = threadIdx.x; it < ndft * A; it += blockDim.x) {
    const int d = it / A, q = it - (it / A) * A;
    float2* io = L.bufA + d * S + q * B;
    float2 v[B];
#pragma unroll
    for (int r = 0; r < B; ++r) v[r] = io[r];
    R::template bflyB<false>(v);
    if (q == 0) L.bufA[d * S + LL + 1] = vadd(L.bufA[d * S + LL], v[0]);  // X_0 = x_0 + F_0
#pragma unroll
    for (int r = 0; r < B; ++r) {
      float2 gh;
      if constexpr (REGC) gh = mid.gh[r];
      else gh = L.ghat[q + A * r];
      v[r] = vmul(v[r], gh);
    }
    R::template bflyB<true>(v);
#pragma unroll
    for (int r = 0; r < B; ++r) io[r] = v[r];
  }
  __syncthreads();
  // inverse DFT-A over k1: butterfly q = m2 < B -> C_b, b = B m1 + A m2;
  // X_{gamma^b} = x_0 + C_b
  const int it = threadIdx.x;
  const bool act = it < ndft * B;
  const int d = it / B, q = it - (it / B) * B;
  float2 v[A], x0, X0;
  if (act) {
    const float2* in = L.bufA + d * S + q;
#pragma unroll
    for (int r = 0; r < A; ++r) v[r] = in[B * r];
    x0 = L.bufA[d * S + LL];
    X0 = L.bufA[d * S + LL + 1];
    R::template bflyA<true>(v);
  }
  __syncthreads();
  if (act) {
    float2* Y = L.bufB + d * P;
    if (BORD) {
#pragma unroll
      for (int r = 0; r < A; ++r) Y[1 + q + B * r] = vadd(x0, v[r]);
    } else {
#pragma unroll
      for (int r = 0; r < A; ++r) Y[L.perm_out[q + B * r]] = vadd(x0, v[r]);
    }
    if (q == 0) Y[0] = X0;
  }
  __syncthreads();
}

// Radix-M final stage over Y in the order BORD of rd_stages, consecutive
// lanes on consecutive transforms (strided pass), items in natural output order:
// item k (transform g) reads slot pos[k] of the M sub-transforms, so that
// consecutive lanes store consecutive output rows; st(g, k, v) stores
// v[r] = Z[k + r p].
template <int P, int M, bool INV, class St>
__device__ __forceinline__ void rd_final(const float2* __restrict__ Y, const float2* __restrict__ tw,
                                         const int* __restrict__ pos, int G, St st) {
  const int N = P * M;
  const int tot = G * P;
  const FastDiv fG(G);
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int g, k;
    fG.divmod(t, k, g);
    const float2* y = Y + (size_t)g * N + pos[k];
    float2 v[M];
    v[0] = y[0];
#pragma unroll
    for (int r = 1; r < M; ++r) v[r] = vmul(y[r * P], twl<INV>(tw, k * r));
    pf_bfly<M, INV>(v);
    st(g, k, v);
  }
}
