// Prime-first FFT with the prime radix p = 113 run by Rader's algorithm, for
// T = 113 * M (678 = 113 * 6, the full-FOV tile).
//
// For a butterfly j of the prime-first order (stride 1, no twiddles) the DFT
//   X_k = sum_{r=0}^{p-1} x_r w^{r k},  w = exp(-/+ 2 pi i / p)
// is, with a primitive root gamma (3 for 113), r = gamma^-a and k = gamma^b,
//   X_{gamma^b} = x_0 + (f (*) g)_b,  f_a = x_{gamma^-a},  g_n = w^{gamma^n},
//   X_0 = x_0 + sum_a f_a,
// a cyclic convolution of length L = p - 1 = 112 = 16 * 7.  16 and 7 are
// coprime, so the FFT-112 runs in the Good-Thomas (prime-factor) form without
// twiddles: with n = 7 n1 + 16 n2 and k = 49 k1 + 64 k2 (mod 112),
//   W112^{n k} = W16^{n1 k1} W7^{n2 k2}.
// Forward: DFT-16 over n1, then DFT-7 over n2; the product with
// G = FFT(g) / L (precomputed in float64, stored in the (k1, k2) order); the
// inverse: DFT-7 over k2, then DFT-16 over k1, whose outputs (m1, m2) are
// C_b at b = 7 m1 + 16 m2.  The index maps live in the tables: slot_in places
// f_a at its (n1, n2) slot, perm_out maps (m1, m2) straight to gamma^b.  The
// two radix-7 passes act on the same 7 elements, so they run back to back in
// registers.  About 2.6 times fewer operations than the direct symmetric DFT.
//
// LDS per transform: M * RD_S complex (slot 112 holds x_0, slot 113 holds
// X_0) in one buffer worked in place (or two ping-ponged, RD_NBUF = 2); the
// last stage leaves the transforms X of the M sub-sequences at stride p, in
// natural order for the radix-M final stage of fft_pf.h (row passes) or in
// Rader's output order for rd_final (strided pass: no scattered LDS stores).
#pragma once
#include "fft_pf.h"

#define RD_P 113
#define RD_L 112
// LDS slots per transform: 112 for f, then x_0 and X_0; the stride is padded
// from 114 to 122 so that the transforms of one group start on different LDS
// banks (about 0.5 ms less per 40 x 2800^2 stack across the four Rader passes)
#define RD_S 122
#define RD_NI 16  // fetch items (single elements) per thread: G * M * p <= RD_NI * blockDim.x

// Device tables (float64-accurate, float32 storage): ghat forward / inverse
// [112] each (already divided by 112) at k1 + 16 k2; perm_out[m2 + 7 m1] =
// gamma^b mod p (b = 7 m1 + 16 m2 mod 112); slot_in[113]: the slot
// n2 + 7 n1 of element r = gamma^-a (n1 = 7 a mod 16, n2 = 4 a mod 7), 112
// for r = 0.
struct RdTab {
  const float2* ghat;  // of the transform direction
  const int* perm_out;
  const int* slot_in;
};

// One data buffer per block, the stages in place (the third stage reads
// everything before a barrier, then writes): half the LDS of ping-pong buffers,
// so 3 (x passes) to 4 (row passes) blocks share a CU.  Needs G M 7 <= blockDim.x.
#ifndef RD_NBUF
#define RD_NBUF 1
#endif

// LDS bytes: twiddles of N, ghat, perm_out, slot_in, pos, RD_NBUF data buffers.
__host__ __device__ __forceinline__ size_t rd_lds_bytes(int N, int M, int G, size_t extra) {
  return mm_align16((size_t)N * 8) + (size_t)RD_L * 8 + mm_align16((size_t)(RD_L + 2 * RD_P) * 4) +
         (size_t)RD_NBUF * G * M * RD_S * 8 + extra;
}

struct RdLds {
  float2* tw;
  float2* ghat;
  int* perm_out;
  int* slot_in;
  int* pos;  // [p]: slot of X_k in Rader's output order (0 for k = 0, 1 + s for k = perm_out[s])
  float2* bufA;
  float2* bufB;
  unsigned char* extra;
};

__device__ __forceinline__ RdLds rd_carve(unsigned char* smem, int N, int M, int G, const float2* __restrict__ tw,
                                          const RdTab& T) {
  RdLds L;
  size_t off = 0;
  L.tw = (float2*)smem;
  off += mm_align16((size_t)N * 8);
  L.ghat = (float2*)(smem + off);
  off += RD_L * 8;
  L.perm_out = (int*)(smem + off);
  L.slot_in = L.perm_out + RD_L;
  L.pos = L.slot_in + RD_P;
  off += mm_align16((size_t)(RD_L + 2 * RD_P) * 4);
  L.bufA = (float2*)(smem + off);
  off += (size_t)G * M * RD_S * 8;
  if (RD_NBUF == 2) {
    L.bufB = (float2*)(smem + off);
    off += (size_t)G * M * RD_S * 8;
  } else {
    L.bufB = L.bufA;
  }
  L.extra = smem + off;
  for (int i = threadIdx.x; i < N; i += blockDim.x) L.tw[i] = tw[i];
  for (int i = threadIdx.x; i < RD_L; i += blockDim.x) {
    L.ghat[i] = T.ghat[i];
    L.perm_out[i] = T.perm_out[i];
    L.pos[T.perm_out[i]] = i + 1;  // perm_out is a permutation of 1 .. p - 1
  }
  if (threadIdx.x == 0) L.pos[0] = 0;
  for (int i = threadIdx.x; i < RD_P; i += blockDim.x) L.slot_in[i] = T.slot_in[i];
  return L;
}

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

// ------------------------------------------------------------ fetch / put
// Item (g, j, r), r < p: element n = j + r M of transform g goes to slot
// slot_in[r] of DFT (g, j).  nn = n | g << 16; sl = LDS slot.
struct RdItems {
  unsigned nn[RD_NI], sl[RD_NI];
  int nval;
};

// GFAST: g fastest (strided pass); otherwise j, then r, then g (rows).
template <int M, bool GFAST>
__device__ __forceinline__ void rd_items(int G, const int* __restrict__ slot_in, RdItems& I) {
  const int tot = G * M * RD_P;
  const FastDiv fG(G), fp(RD_P);
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
      I.sl[u] = (unsigned)((g * M + j) * RD_S + slot_in[r]);
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
// Per-thread constants of the middle stage.  Its butterfly q = it & 15 equals
// threadIdx.x & 15 in every round (blockDim.x is a multiple of 16), so its
// G_{q + 16 r} (r = 0..6) can stay in registers (REGC; kernels short of
// registers read them from LDS).
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

// After rd_put and a barrier: bufA holds f (slots 0..111) and x_0 (slot 112) of
// the G*M DFTs.  Ends (after its last barrier) with Y in bufB, at d = g M + j
// (N = p M):
//   BORD false: Y[d p + k] = X_k (natural order),
//   BORD true:  Y[d p] = X_0, Y[d p + 1 + s] = X_{perm_out[s]}, s = m2 + 7 m1
//               (rd_final reads it).
// The first stage writes its output (n2, k1) at k1 * 7 + n2, so that both its
// stores and the middle stage's loads run across the banks.
template <int M, bool BORD, bool REGC>
__device__ __forceinline__ void rd_stages(const RdLds& L, int G, const RdMid& mid) {
  const int ndft = G * M;
  // forward DFT-16 over n1: butterfly q = n2 < 7 reads slots q + 7 n1
  for (int it = threadIdx.x; it < ndft * 7; it += blockDim.x) {
    const int d = it / 7, q = it - (it / 7) * 7;
    const float2* in = L.bufA + d * RD_S + q;
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = in[7 * r];
    bfly16<false>(v);
    float2* out = L.bufB + d * RD_S + q;
#pragma unroll
    for (int r = 0; r < 16; ++r) out[7 * r] = v[r];
  }
  __syncthreads();
  // DFT-7 over n2 -> F(k1 = q, k2), times ghat, inverse DFT-7 over k2
  for (int it = threadIdx.x; it < ndft * 16; it += blockDim.x) {
    const int d = it >> 4, q = it & 15;
    const float2* in = L.bufB + d * RD_S + q * 7;
    float2 v[7];
    v[0] = in[0];
#pragma unroll
    for (int r = 1; r < 7; ++r) v[r] = in[r];
    bfly_odd<7, false>(v, kCos7, kSin7);
    if (q == 0) L.bufA[d * RD_S + 113] = vadd(L.bufA[d * RD_S + 112], v[0]);  // X_0 = x_0 + F_0
#pragma unroll
    for (int r = 0; r < 7; ++r) v[r] = vmul(v[r], REGC ? mid.gh[r] : L.ghat[q + 16 * r]);
    bfly_odd<7, true>(v, kCos7, kSin7);
    float2* out = L.bufA + d * RD_S + q * 7;
#pragma unroll
    for (int r = 0; r < 7; ++r) out[r] = v[r];
  }
  __syncthreads();
  // inverse DFT-16 over k1: butterfly q = m2 < 7 -> C_b, b = 7 m1 + 16 m2;
  // X_{gamma^b} = x_0 + C_b
  if (RD_NBUF == 1) {  // Y overlaps other transforms' slots: every read, a barrier, then the writes
    const int it = threadIdx.x;   // ndft * 7 <= blockDim.x
    const bool act = it < ndft * 7;
    const int d = it / 7, q = it - (it / 7) * 7;
    float2 v[16], x0, X0;
    if (act) {
      const float2* in = L.bufA + d * RD_S + q;
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = in[7 * r];
      x0 = L.bufA[d * RD_S + 112];
      X0 = L.bufA[d * RD_S + 113];
      bfly16<true>(v);
    }
    __syncthreads();
    if (act) {
      float2* Y = L.bufB + d * RD_P;
      if (BORD) {
#pragma unroll
        for (int r = 0; r < 16; ++r) Y[1 + q + 7 * r] = vadd(x0, v[r]);
      } else {
#pragma unroll
        for (int r = 0; r < 16; ++r) Y[L.perm_out[q + 7 * r]] = vadd(x0, v[r]);
      }
      if (q == 0) Y[0] = X0;
    }
    __syncthreads();
    return;
  }
  for (int it = threadIdx.x; it < ndft * 7; it += blockDim.x) {
    const int d = it / 7, q = it - (it / 7) * 7;
    const float2* in = L.bufA + d * RD_S + q;
    float2 v[16];
    v[0] = in[0];
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = in[7 * r];
    bfly16<true>(v);
    const float2 x0 = L.bufA[d * RD_S + 112];
    float2* Y = L.bufB + d * RD_P;
    if (BORD) {
#pragma unroll
      for (int r = 0; r < 16; ++r) Y[1 + q + 7 * r] = vadd(x0, v[r]);
    } else {
#pragma unroll
      for (int r = 0; r < 16; ++r) Y[L.perm_out[q + 7 * r]] = vadd(x0, v[r]);
    }
    if (q == 0) Y[0] = L.bufA[d * RD_S + 113];
  }
  __syncthreads();
}

// Radix-M final stage over Y in the order BORD of rd_stages, consecutive
// lanes on consecutive transforms (strided pass), items in natural output order:
// item k (transform g) reads slot pos[k] of the M sub-transforms, so that
// consecutive lanes store consecutive output rows; st(g, k, v) stores
// v[r] = Z[k + r p].
template <int M, bool INV, class St>
__device__ __forceinline__ void rd_final(const float2* __restrict__ Y, const float2* __restrict__ tw,
                                         const int* __restrict__ pos, int G, St st) {
  const int N = RD_P * M;
  const int tot = G * RD_P;
  const FastDiv fG(G);
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int g, k;
    fG.divmod(t, k, g);
    const float2* y = Y + (size_t)g * N + pos[k];
    float2 v[M];
    v[0] = y[0];
#pragma unroll
    for (int r = 1; r < M; ++r) v[r] = vmul(y[r * RD_P], twl<INV>(tw, k * r));
    pf_bfly<M, INV>(v);
    st(g, k, v);
  }
}
