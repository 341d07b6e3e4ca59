// Prime-first FFT of length N = p * M (p an odd prime handled as a matrix
// product, M in {2, 3, 4, 6}) for the tile lengths whose largest factor is a
// prime, e.g. 678 = 113 * 6 and 478 = 239 * 2.
//
// Stockham order with the prime radix first (stride Ns = 1, so no twiddles):
//   1. pre-step (fused into the global load): for butterfly j < M and
//      r = 1..h (h = (p-1)/2), s_r = x[j + r M] + x[j + (p-r) M] and
//      d_r = x[j + r M] - x[j + (p-r) M] are written to LDS as one float4;
//      slot 0 holds x[j].
//   2. matrix stage: a work item = (transform, butterfly, tile of 8 outputs k)
//      accumulates A_k = x_j + sum_r cos(2 pi r k / p) s_r and
//      B_k = sum_r sin(2 pi r k / p) d_r with the cos / sin rows read from LDS
//      (32 FMAs per float4 data load and 4 float4 coefficient loads); the DC
//      term x_j + sum_r s_r is accumulated in the same loop.  The outputs
//      X_k = A_k -/+ i B_k, X_{p-k} = A_k +/- i B_k are held in registers
//      until every item has read its data, then written over the same LDS as
//      Y[j p + k].
//   3. final stage: radix M at stride p with twiddles w_N^{j r}: butterfly j < p
//      reads Y[j + r p], r < M, and its outputs are the results n = j + r p,
//      handed to a store functor (global memory, or the R2C post-step, which
//      pairs butterflies j and p - j so that X[n] and X[N - n] are in one
//      thread).
// One LDS buffer of M * (h + 1) float4 per transform; the work-item count of
// the matrix stage, G * M * ceil(h / 8), must not exceed the block size, and
// the pre-step items G * M * h must not exceed PF_NI per thread.
// The kernels built on it are persistent: a block loops over groups of G
// transforms and prefetches the next group into registers (pf_fetch) before
// the matrix stage of the current one.
#pragma once
#include "fft_engine.h"

struct PfShape {
  int N, p, h, hp, nkt, sb;  // sb: float4 slots per butterfly (h + 1)
  __host__ __device__ PfShape() {}
  __host__ __device__ PfShape(int N_, int p_) : N(N_), p(p_) {
    h = (p - 1) >> 1;
    hp = (h + 7) & ~7;
    nkt = (h + 7) >> 3;
    sb = h + 1;
  }
};

// LDS bytes for G transforms: twiddles (N float2), cos/sin tables (2 h hp
// floats), data (G M sb float4), then `extra`.
__host__ __device__ __forceinline__ size_t pf_lds_bytes(const PfShape& S, int M, int G, size_t extra) {
  return mm_align16((size_t)S.N * 8) + mm_align16((size_t)2 * S.h * S.hp * 4) + (size_t)G * M * S.sb * 16 + extra;
}

struct PfLds {
  float2* tw;
  float* coef;
  float4* sd;
  unsigned char* extra;
};

__device__ __forceinline__ PfLds pf_carve(unsigned char* smem, const PfShape& S, int M, int G,
                                          const float2* __restrict__ tw, const float* __restrict__ coef) {
  PfLds L;
  size_t off = 0;
  L.tw = (float2*)smem;
  off += mm_align16((size_t)S.N * 8);
  L.coef = (float*)(smem + off);
  off += mm_align16((size_t)2 * S.h * S.hp * 4);
  L.sd = (float4*)(smem + off);
  off += (size_t)G * M * S.sb * 16;
  L.extra = smem + off;
  for (int i = threadIdx.x; i < S.N; i += blockDim.x) L.tw[i] = tw[i];
  const int cl = 2 * S.h * S.hp;
  const float4* c4 = (const float4*)coef;
  float4* l4 = (float4*)L.coef;
  for (int i = threadIdx.x; i < cl / 4; i += blockDim.x) l4[i] = c4[i];
  return L;
}

// ------------------------------------------------------------ radix-M final
template <bool INV>
__device__ __forceinline__ void bfly6(float2* v) {
  // DFT-6 as two DFT-3 (even / odd samples) and a radix-2 step with w6^k
  float2 e[3] = {v[0], v[2], v[4]};
  float2 o[3] = {v[1], v[3], v[5]};
  bfly3<INV>(e);
  bfly3<INV>(o);
  const float c = 0.5f, s = 0.8660254037844386f;
  // forward w6 = exp(-i pi / 3) = (1/2, -s); w6^2 = (-1/2, -s)
  float2 o1, o2;
  if (INV) {
    o1 = make_float2(c * o[1].x - s * o[1].y, c * o[1].y + s * o[1].x);
    o2 = make_float2(-c * o[2].x - s * o[2].y, -c * o[2].y + s * o[2].x);
  } else {
    o1 = make_float2(c * o[1].x + s * o[1].y, c * o[1].y - s * o[1].x);
    o2 = make_float2(-c * o[2].x + s * o[2].y, -c * o[2].y - s * o[2].x);
  }
  v[0] = vadd(e[0], o[0]);
  v[3] = vsub(e[0], o[0]);
  v[1] = vadd(e[1], o1);
  v[4] = vsub(e[1], o1);
  v[2] = vadd(e[2], o2);
  v[5] = vsub(e[2], o2);
}

template <int M, bool INV>
__device__ __forceinline__ void pf_bfly(float2* v) {
  if constexpr (M == 2) bfly2<INV>(v);
  else if constexpr (M == 3) bfly3<INV>(v);
  else if constexpr (M == 4) bfly4<INV>(v);
  else if constexpr (M == 6) bfly6<INV>(v);
}

// ------------------------------------------------------------ stage 1
// Split in two so that the global loads of the next group of transforms can be
// in flight during the matrix and final stages of the current one:
//   pf_fetch: each thread loads its items (g, j, r), r = 1..h, into registers:
//             a = x_g[j + r M], b = x_g[j + (p - r) M]; threads t < G M also
//             load x_g[j] of butterfly t = g M + j;
//   pf_put:   writes (a + b, a - b) to the LDS slot of the item.
// Item order: GFAST -> (g fastest, then j, then r): consecutive lanes read
// consecutive transforms (columns of a strided pass).  Otherwise (j, r, g):
// consecutive lanes read consecutive elements of one row.
#define PF_NI 8  // items per thread: G * M * h <= PF_NI * blockDim.x

// Per-thread item table, the same for every group (decoded once per kernel):
// nn = n1 | n2 << 16 with n1 = j + r M, n2 = j + (p - r) M; sg = LDS slot | g << 16.
struct PfItems {
  unsigned nn[PF_NI], sg[PF_NI];
  int nval;  // items of this thread (a prefix of u = 0..PF_NI-1)
};

template <int M, bool GFAST>
__device__ __forceinline__ void pf_items(const PfShape& S, int G, PfItems& I) {
  const int tot = G * M * S.h;
  const FastDiv fG(G), fh(S.h);
  I.nval = 0;
#pragma unroll
  for (int u = 0; u < PF_NI; ++u) {
    const int t = threadIdx.x + u * blockDim.x;
    I.nn[u] = 0;
    I.sg[u] = 0;
    if (t < tot) {
      int g, j, r;
      if (GFAST) {  // t = ((r - 1) M + j) G + g
        int q;
        fG.divmod(t, q, g);
        r = q / M;
        j = q - r * M;
      } else {      // t = (g h + r - 1) M + j
        const int q = t / M;
        j = t - q * M;
        fh.divmod(q, g, r);
      }
      ++r;
      I.nn[u] = (unsigned)(j + r * M) | ((unsigned)(j + (S.p - r) * M) << 16);
      I.sg[u] = (unsigned)((g * M + j) * S.sb + r) | ((unsigned)g << 16);
      I.nval = u + 1;
    }
  }
}

// Issues the loads of every item of this thread: ld(u, g, n1, n2) for item u
// (the kernel keeps the raw values in registers of its own, indexed by the
// compile-time u), and ld0(g, j) on threads t = g M + j < G M for x_g[j].
// Loads must be unconditional (clamped addresses, validity kept aside) so
// that none of them waits for another.
template <int M, class Ld, class Ld0>
__device__ __forceinline__ void pf_fetch(int G, const PfItems& I, Ld ld, Ld0 ld0) {
#pragma unroll
  for (int u = 0; u < PF_NI; ++u)
    if (u < I.nval) ld(u, (int)(I.sg[u] >> 16), (int)(I.nn[u] & 0xffffu), (int)(I.nn[u] >> 16));
  if ((int)threadIdx.x < G * M) {
    const int g = threadIdx.x / M;
    ld0(g, (int)threadIdx.x - g * M);
  }
}

// get(u, g, n1, n2, a, b) turns the raw registers of item u into a = x_g[n1],
// b = x_g[n2]; get0(g, j) returns x_g[j].  Writes (a + b, a - b) per item.
template <int M, class Get, class Get0>
__device__ __forceinline__ void pf_put(const PfShape& S, int G, float4* __restrict__ sd, const PfItems& I, Get get,
                                       Get0 get0) {
#pragma unroll
  for (int u = 0; u < PF_NI; ++u) {
    if (u < I.nval) {
      float2 a, b;
      get(u, (int)(I.sg[u] >> 16), (int)(I.nn[u] & 0xffffu), (int)(I.nn[u] >> 16), a, b);
      sd[I.sg[u] & 0xffffu] = make_float4(a.x + b.x, a.y + b.y, a.x - b.x, a.y - b.y);
    }
  }
  if ((int)threadIdx.x < G * M) {
    const int g = threadIdx.x / M;
    const float2 v0 = get0(g, (int)threadIdx.x - g * M);
    sd[threadIdx.x * S.sb] = make_float4(v0.x, v0.y, 0.f, 0.f);
  }
}

// ------------------------------------------------------------ stage 2
// Requires G * M * nkt <= blockDim.x.  Ends with Y (float2, G x N) over the
// same LDS, followed by a barrier.
template <int M, bool INV>
__device__ __forceinline__ void pf_matrix(const PfShape& S, float4* __restrict__ sd, const float* __restrict__ coef,
                                          int G) {
  const int items = G * M * S.nkt;
  const int t = threadIdx.x;
  const bool act = t < items;
  float2 A[8], B[8], dc;
  int bf = 0, kt = 0;
  if (act) {
    const int q = t / S.nkt;  // nkt is block uniform
    kt = t - q * S.nkt;
    bf = q;                   // g * M + j
    const float4* row = sd + (size_t)bf * S.sb;
    const float* cm = coef + kt * 8;
    const float* sm = coef + S.h * S.hp + kt * 8;
    const float4 v0 = row[0];
    dc = make_float2(v0.x, v0.y);
#pragma unroll
    for (int i = 0; i < 8; ++i) { A[i] = dc; B[i] = make_float2(0.f, 0.f); }
    for (int r = 1; r <= S.h; ++r) {
      const float4 v = row[r];
      const float4 c0 = *(const float4*)(cm + (r - 1) * S.hp);
      const float4 c1 = *(const float4*)(cm + (r - 1) * S.hp + 4);
      const float4 s0 = *(const float4*)(sm + (r - 1) * S.hp);
      const float4 s1 = *(const float4*)(sm + (r - 1) * S.hp + 4);
      const float c[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
      const float s[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        A[i].x += v.x * c[i];
        A[i].y += v.y * c[i];
        B[i].x += v.z * s[i];
        B[i].y += v.w * s[i];
      }
      dc.x += v.x;
      dc.y += v.y;
    }
  }
  __syncthreads();
  if (act) {
    const int g = bf / M, j = bf - (bf / M) * M;
    float2* Y = (float2*)sd + (size_t)g * S.N + j * S.p;
    if (kt == 0) Y[0] = dc;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = kt * 8 + 1 + i;
      if (k <= S.h) {
        const float2 iB = vmi<INV>(B[i]);
        Y[k] = vadd(A[i], iB);
        Y[S.p - k] = vsub(A[i], iB);
      }
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------ stage 3
// st(g, j, v) stores the M results v[r] = X_g[j + r p] of butterfly j, which
// reads Y[g N + j + r p].
// GFAST: consecutive lanes take consecutive transforms (the stores of a
// strided pass then cover G adjacent columns of one row).
template <int M, bool INV, bool GFAST = false, class St>
__device__ __forceinline__ void pf_final(const PfShape& S, const float2* __restrict__ Y, const float2* __restrict__ tw,
                                         int G, St st) {
  const int tot = G * S.p;
  const FastDiv fp(S.p), fG(G);
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int g, j;
    if (GFAST) fG.divmod(t, j, g);
    else fp.divmod(t, g, j);
    const float2* y = Y + (size_t)g * S.N + j;
    float2 v[M];
    v[0] = y[0];
#pragma unroll
    for (int r = 1; r < M; ++r) v[r] = vmul(y[r * S.p], twl<INV>(tw, j * r));
    pf_bfly<M, INV>(v);
    st(g, j, v);
  }
}

// Final stage for the packed R2C of two real rows: butterflies j and p - j in
// one thread, so that Z[n] and Z[N - n] meet.  rows(g, pa, pb) gives the
// output rows of transform g (real-part row, imaginary-part row; nullptr for
// an absent row); the half-spectrum bin k <= N/2 goes to row[coff(k)].
struct PfContig {
  __device__ __forceinline__ int operator()(int k) const { return k; }
};

template <int M, class Rows, class Coff = PfContig>
__device__ __forceinline__ void pf_final_r2c(const PfShape& S, const float2* __restrict__ Y,
                                             const float2* __restrict__ tw, int G, Rows rows, Coff coff = Coff()) {
  const int np = (S.p + 1) >> 1;
  const int tot = G * np;
  const int half = S.N >> 1;
  const FastDiv fnp(np);
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int g, j1;
    fnp.divmod(t, g, j1);
    const int j2 = j1 == 0 ? 0 : S.p - j1;
    const float2* y = Y + (size_t)g * S.N;
    float2 u[M], w[M];
    u[0] = y[j1];
    w[0] = y[j2];
#pragma unroll
    for (int r = 1; r < M; ++r) {
      u[r] = vmul(y[j1 + r * S.p], twl<false>(tw, j1 * r));
      w[r] = vmul(y[j2 + r * S.p], twl<false>(tw, j2 * r));
    }
    pf_bfly<M, false>(u);
    pf_bfly<M, false>(w);
    float2 *pa, *pb;
    rows(g, pa, pb);
    // u[r] = Z[j1 + r p]; partner N - (j1 + r p) = j2 + (M - 1 - r) p for j1 > 0,
    // (M - r) p for j1 = 0 (then w == u).
#pragma unroll
    for (int r = 0; r < M; ++r) {
      const int n = j1 + r * S.p;
      const float2 Z = u[r];
      const float2 Zm = j1 == 0 ? u[(M - r) % M] : w[M - 1 - r];
      if (n <= half) {
        const int o = coff(n);
        if (pa) pa[o] = make_float2((Z.x + Zm.x) * 0.5f, (Z.y - Zm.y) * 0.5f);
        if (pb) pb[o] = make_float2((Z.y + Zm.y) * 0.5f, (Zm.x - Z.x) * 0.5f);
      }
      if (j1 > 0) {
        const int n2 = j2 + r * S.p;
        const float2 Z2 = w[r], Z2m = u[M - 1 - r];
        if (n2 <= half) {
          const int o = coff(n2);
          if (pa) pa[o] = make_float2((Z2.x + Z2m.x) * 0.5f, (Z2.y - Z2m.y) * 0.5f);
          if (pb) pb[o] = make_float2((Z2.y + Z2m.y) * 0.5f, (Z2m.x - Z2.x) * 0.5f);
        }
      }
    }
  }
}
