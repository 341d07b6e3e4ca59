// Tiled Wiener deconvolution on gfx950: the FFT passes of the tile pipeline and
// the construction of the Wiener spectra.
//
// Reference behaviour (mermake/deconvolver.py:98-170, one tile at a time):
//   pad (Zp,T,T) <- reflected, flat-divided tile; fftn; *= W; ifftn; crop z; .real
// Here the same result is produced for a batch of tiles with five passes over a
// half spectrum S (per tile: sz x T x H complex, H = T/2+1):
//   1. k_r2c_rows      padded rows (z, x) -> R2C along y, two real rows per
//                      complex transform; the reflection padding is evaluated
//                      in the loader from closed-form index maps, so only the
//                      sz data planes are transformed (the z pads are plane copies);
//   2. k_fft_c2c       C2C along x (in place, strided);
//   3. k_tile_zconv    per (kx,ky) column, the z pass (reflection padding to Zp
//                      planes, FFT, * W, inverse FFT, crop to sz planes) is one
//                      linear map of the sz data planes: y[t] = sum_s K[t][s] x[s]
//                      with K[t][s] = sum over the padded copies p of plane s of
//                      w[(zpad + t - p) mod Zp], w = inverse z-DFT of W (computed
//                      once, in float64, when the spectra are built).  K lives in
//                      registers and, with one PSF, is reused for every tile.
//                      (k_tile_zwiener, the FFT form of the same pass, serves
//                      stacks deeper than 48 planes);
//   4. k_fft_c2c       inverse C2C along x;
//   5. k_tile_c2r      C2R along y (two rows per complex transform) -> res (sz,T,T).
// The 1/(Zp*T*T) normalisation of ifftn is folded into W.
#include "deconv.h"
#include "zfft80.h"
#include "zfft78.h"
#include "fft_pf.h"
#include "fft_plan.h"
#include "mermake_hip.h"

#include <type_traits>
#include <vector>

#define MM_UNR 8  // global loads issued back to back before the LDS stores

static_assert(sizeof(mm_tile_dev) == sizeof(mm_tile_desc) && sizeof(mm_tile_desc) == 104,
              "tile descriptor layout");

// --------------------------------------------------------------- generic C2C
template <bool INV, typename C>
__global__ __launch_bounds__(256) void k_fft_c2c(C* data, long long estride, long long cstride, int ncols,
                                                 long long bstride, int gfast, FftPlanDev<C> P, int G,
                                                 double scale) {
  typedef decltype(data[0].x) S;
  extern __shared__ __align__(16) unsigned char smem[];
  const FftLds<C> L = fft_lds_carve(smem, P, G);
  C* a = L.a;
  const int N = P.N;
  const int ncb = (ncols + G - 1) / G;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = bid % ncb;
  const int batch = bid / ncb;
  C* base = data + (long long)batch * bstride + (long long)cb * G * cstride;
  const int gmax = min(G, ncols - cb * G);
  const int tot = G * N;
  const FastDiv fd(gfast ? G : N);
  for (int t0 = threadIdx.x; t0 < tot; t0 += MM_UNR * blockDim.x) {  // MM_UNR loads in flight per thread
    C v[MM_UNR];
#pragma unroll
    for (int u = 0; u < MM_UNR; ++u) {
      const int t = t0 + u * blockDim.x;
      int q, r, g, n;
      fd.divmod(t, q, r);
      if (gfast) { g = r; n = q; } else { n = r; g = q; }
      v[u].x = 0; v[u].y = 0;
      if (t < tot && g < gmax) v[u] = base[g * cstride + n * estride];
    }
#pragma unroll
    for (int u = 0; u < MM_UNR; ++u) {
      const int t = t0 + u * blockDim.x;
      int q, r, g, n;
      fd.divmod(t, q, r);
      if (gfast) { g = r; n = q; } else { n = r; g = q; }
      if (t < tot) a[g * N + n] = v[u];
    }
  }
  __syncthreads();
  C* r = fft_lds<INV>(L.a, L.b, L.tw, P, G, L.coef);
  const S sc = (S)scale;
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int q, rr, g, n;
    fd.divmod(t, q, rr);
    if (gfast) { g = rr; n = q; } else { n = rr; g = q; }
    if (g < gmax) base[g * cstride + n * estride] = vscale(r[g * N + n], sc);
  }
}

// ---------------------------------------------------------------- R2C rows
template <typename C>
__global__ __launch_bounds__(256) void k_r2c_rows_plain(const float* vol, int nrows, C* out, FftPlanDev<C> P,
                                                        int G) {
  typedef decltype(out[0].x) S;
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = P.N, H = N / 2 + 1;
  const FftLds<C> L = fft_lds_carve(smem, P, G);
  const int q0 = blockIdx.x * 2 * G;
  S* af = (S*)L.a;
  for (int t = threadIdx.x; t < 2 * G * N; t += blockDim.x) {
    const int row = t / N, y = t - row * N, q = q0 + row;
    S v = 0;
    if (q < nrows) v = (S)vol[(long long)q * N + y];
    af[((row >> 1) * N + y) * 2 + (row & 1)] = v;
  }
  __syncthreads();
  C* r = fft_lds<false>(L.a, L.b, L.tw, P, G, L.coef);
  for (int t = threadIdx.x; t < G * H; t += blockDim.x) {
    const int g = t / H, k = t - g * H;
    const C Z = r[g * N + k], Zm = r[g * N + (k == 0 ? 0 : N - k)];
    C A, B;
    A.x = (Z.x + Zm.x) * (S)0.5; A.y = (Z.y - Zm.y) * (S)0.5;
    B.x = (Z.y + Zm.y) * (S)0.5; B.y = (Zm.x - Z.x) * (S)0.5;
    const int q = q0 + 2 * g;
    if (q < nrows) out[(long long)q * H + k] = A;
    if (q + 1 < nrows) out[(long long)(q + 1) * H + k] = B;
  }
}

__global__ __launch_bounds__(256) void k_tile_r2c(const mm_tile_dev* __restrict__ tiles, int T, int nrows,
                                                  float2* S, long long s_tstride, FftPlanDev<float2> P, int G) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = P.N, H = N / 2 + 1;
  const FftLds<float2> L = fft_lds_carve(smem, P, G);
  const int tile = blockIdx.y;
  const int q0 = blockIdx.x * 2 * G;
  const mm_tile_dev td = tiles[tile];
  float* af = (float*)L.a;
  int* ymap = (int*)L.extra;           // column -> data column (or -1), once per block
  for (int y = threadIdx.x; y < N; y += blockDim.x) ymap[y] = pad_map(y, td.ys, td.yd, td.yfirst, T);
  __syncthreads();
  const FastDiv fN(N), fT(T);
  const int tot = 2 * G * N;
  for (int t0 = threadIdx.x; t0 < tot; t0 += MM_UNR * blockDim.x) {  // issue all loads, then divide
    float v[MM_UNR], f[MM_UNR];
#pragma unroll
    for (int u = 0; u < MM_UNR; ++u) {
      const int t = t0 + u * blockDim.x;
      int row, y, z, p;
      fN.divmod(t, row, y);
      const int q = q0 + row;
      v[u] = 0.0f;
      f[u] = 1.0f;
      if (t < tot && q < nrows) {
        fT.divmod(q, z, p);
        const int rr = pad_map(p, td.xs, td.xd, td.xfirst, T);
        const int cc = ymap[y];
        if (rr >= 0 && cc >= 0) {
          const long long o = (long long)z * td.img_sz + (long long)(td.X0 + rr) * td.img_sx + (td.Y0 + cc);
          v[u] = pix_at(td.img, td.img_type, o);
          if (td.flat) f[u] = td.flat[(long long)(td.X0 + rr) * td.flat_sx + (td.Y0 + cc)];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < MM_UNR; ++u) {
      const int t = t0 + u * blockDim.x;
      int row, y;
      fN.divmod(t, row, y);
      if (t < tot) af[((row >> 1) * N + y) * 2 + (row & 1)] = td.flat ? v[u] / f[u] : v[u];
    }
  }
  __syncthreads();
  float2* r = fft_lds<false>(L.a, L.b, L.tw, P, G, L.coef);
  float2* out = S + (long long)tile * s_tstride;
  for (int t = threadIdx.x; t < G * H; t += blockDim.x) {
    const int g = t / H, k = t - g * H;
    const float2 Z = r[g * N + k], Zm = r[g * N + (k == 0 ? 0 : N - k)];
    const float2 A = make_float2((Z.x + Zm.x) * 0.5f, (Z.y - Zm.y) * 0.5f);
    const float2 B = make_float2((Z.y + Zm.y) * 0.5f, (Zm.x - Z.x) * 0.5f);
    const int q = q0 + 2 * g;
    if (q < nrows) out[(long long)q * H + k] = A;
    if (q + 1 < nrows) out[(long long)(q + 1) * H + k] = B;
  }
}

// ------------------------------------------------------------------ z pass
// Columns c = kx*H + ky of the (sz, T*H) half spectrum.  Work item = G adjacent
// columns x a group of `tpb` tiles; the blocks are persistent over the items
// (column groups fastest), so the twiddle tables are set up once per block.
// W layout: (n_psf, planes, TH) with the Zp planes of W from w_off.
__global__ __launch_bounds__(256) void k_tile_zwiener(float2* S, long long s_tstride, int TH, int sz, int zpad,
                                                      const int* __restrict__ zmap,
                                                      const float2* __restrict__ W, long long w_stride,
                                                      const mm_tile_dev* __restrict__ tiles, int ntiles, int tpb,
                                                      FftPlanDev<float2> P, int G) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int Zp = P.N;
  const FftLds<float2> L = fft_lds_carve(smem, P, G);
  float2* a = L.a;
  float2* b = L.b;
  float2* wl = (float2*)L.extra;
  int* zm = (int*)(wl + (size_t)G * Zp);
  const int ncb = (TH + G - 1) / G;
  const int ngroups = (ntiles + tpb - 1) / tpb;
  const long long nitems = (long long)ncb * ngroups;
  for (int i = threadIdx.x; i < Zp; i += blockDim.x) zm[i] = zmap[i];
  for (long long item = xcd_remap(blockIdx.x, gridDim.x); item < nitems; item += gridDim.x) {
    const int cb = (int)(item % ncb);
    const int tg = (int)(item / ncb);
    const int c0 = cb * G;
    const int gmax = min(G, TH - c0);
    int cached = -1;
    const int t_end = min(ntiles, (tg + 1) * tpb);
    for (int ti = tg * tpb; ti < t_end; ++ti) {
      const int psf = tiles[ti].psf;
      __syncthreads();
      if (psf != cached) {
        const float2* Wp = W + (long long)psf * w_stride + c0;
        for (int t = threadIdx.x; t < G * Zp; t += blockDim.x) {
          const int g = t % G, p = t / G;
          wl[g * Zp + p] = g < gmax ? Wp[(long long)p * TH + g] : make_float2(0.f, 0.f);
        }
        cached = psf;
      }
      float2* St = S + (long long)ti * s_tstride + c0;
      for (int t = threadIdx.x; t < G * Zp; t += blockDim.x) {
        const int g = t % G, p = t / G;
        const int zs = zm[p];
        a[g * Zp + p] = (g < gmax && zs >= 0) ? St[(long long)zs * TH + g] : make_float2(0.f, 0.f);
      }
      __syncthreads();
      float2* r = fft_lds<false>(a, b, L.tw, P, G, L.coef);
      for (int t = threadIdx.x; t < G * Zp; t += blockDim.x) r[t] = cmul(r[t], wl[t]);
      __syncthreads();
      float2* o = fft_lds<true>(r, r == a ? b : a, L.tw, P, G, L.coef);
      for (int t = threadIdx.x; t < G * sz; t += blockDim.x) {
        const int g = t % G, p = t / G;
        if (g < gmax) St[(long long)p * TH + g] = o[g * Zp + p + zpad];
      }
    }
    __syncthreads();  // the item's last reads of o before the next item's stores into a
  }
}

// Padded plane p -> data plane of the z map (host_zmap), or -1 for a zero plane:
// data at [zpad, zpad + sz), the lower pad an 'in' reflection about zpad (a
// reflection into the still-empty upper pad reads zero), the upper pad an 'out'
// reflection about zpad + sz - 1 of what is there.
__host__ __device__ constexpr int zmap_src(int sz, int zpad, int p) {
  const int Zp = sz + 2 * zpad, top = zpad + sz - 1;
  if (p >= zpad && p <= top) return p - zpad;
  if (p < zpad) {
    const int s = 2 * zpad - p;
    return s >= Zp || s > top ? -1 : s - zpad;
  }
  const int s = 2 * top - p;
  return s < 0 ? -1 : zmap_src(sz, zpad, s);
}

// z pass of the z FFT form for a tile of SZ planes padded by ZPAD to 78 (the DAPI
// setting sz 40, zpad 19), with the length-78 transforms in registers: one thread
// per column of one tile.  The padding planes are copies of data planes (the z map,
// at compile time); W is read per (tile, column) in natural plane order, and the
// tiles of one block of columns follow each other on one XCD, so a W that they
// share comes from L2.
template <int SZ, int ZPAD>
__global__ __launch_bounds__(256, 2) void k_tile_zfft78(float2* S, long long s_tstride, int TH,
                                                       const float2* __restrict__ W, long long w_stride,
                                                       const mm_tile_dev* __restrict__ tiles, int ntiles) {
  static_assert(SZ + 2 * ZPAD == 78, "length-78 z pass");
  const int ncb = (TH + 255) / 256;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int j = (int)(bid % (unsigned)ntiles), cb = (int)(bid / (unsigned)ntiles);
  const int c = cb * 256 + threadIdx.x;
  if (cb >= ncb || c >= TH) return;
  float2* col = S + (long long)j * s_tstride + c;
  float2 x[SZ];
  {
    const float2* p = col;
#pragma unroll
    for (int s = 0; s < SZ; ++s) {
      x[s] = *p;
      p += TH;
      asm volatile("" : "+v"(p));
    }
  }
  float2 a[78];
#pragma unroll
  for (int pos = 0; pos < 78; ++pos) {
    const int src = zmap_src(SZ, ZPAD, zf78::n_of(pos));
    a[pos] = src >= 0 ? x[src] : make_float2(0.f, 0.f);
  }
  zf78::fft78<false>(a);
  const float2* Wp = W + (long long)tiles[j].psf * w_stride + c;
  // W 13 values at a time (a row of i3), so that its loads do not all wait in registers at once
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    float2 w[13];
#pragma unroll
    for (int i = 0; i < 13; ++i) w[i] = Wp[(long long)zf78::k_of(13 * r + i) * TH];
#pragma unroll
    for (int i = 0; i < 13; ++i) a[13 * r + i] = cmul(a[13 * r + i], w[i]);
    asm volatile("" : "+v"(Wp));
  }
  zf78::fft78<true>(a);
#pragma unroll
  for (int t = 0; t < SZ; ++t) {
    *col = a[zf78::pos_of_n(ZPAD + t)];
    col += TH;
    asm volatile("" : "+v"(col));
  }
}

// ------------------------------------------------------------------ C2R rows
__global__ __launch_bounds__(256) void k_tile_c2r(const float2* S, long long s_tstride, float* res,
                                                  long long r_tstride, int nrows, FftPlanDev<float2> P, int G) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = P.N, H = N / 2 + 1;
  const FftLds<float2> L = fft_lds_carve(smem, P, G);
  float2* a = L.a;
  const int tile = blockIdx.y;
  const int q0 = blockIdx.x * 2 * G;
  const float2* in = S + (long long)tile * s_tstride;
  const int tot = G * N;
  for (int t0 = threadIdx.x; t0 < tot; t0 += MM_UNR * blockDim.x) {
    float2 A[MM_UNR], B[MM_UNR];
#pragma unroll
    for (int u = 0; u < MM_UNR; ++u) {
      const int t = t0 + u * blockDim.x;
      const int g = t / N, k = t - g * N;
      const int q = q0 + 2 * g;
      const int kk = k >= H ? N - k : k;
      A[u] = make_float2(0.f, 0.f);
      B[u] = make_float2(0.f, 0.f);
      if (t < tot && q < nrows) A[u] = in[(long long)q * H + kk];
      if (t < tot && q + 1 < nrows) B[u] = in[(long long)(q + 1) * H + kk];
    }
#pragma unroll
    for (int u = 0; u < MM_UNR; ++u) {
      const int t = t0 + u * blockDim.x;
      const int g = t / N, k = t - g * N;
      float2 a0 = A[u], b0 = B[u];
      if (k >= H) { a0.y = -a0.y; b0.y = -b0.y; }
      if (k == 0 || 2 * k == N) { a0.y = 0.f; b0.y = 0.f; }  // real part semantics of ifftn(...).real
      if (t < tot) a[g * N + k] = make_float2(a0.x - b0.y, a0.y + b0.x);
    }
  }
  __syncthreads();
  float2* r = fft_lds<true>(L.a, L.b, L.tw, P, G, L.coef);
  float* out = res + (long long)tile * r_tstride;
  for (int t = threadIdx.x; t < 2 * G * N; t += blockDim.x) {
    const int row = t / N, y = t - row * N, q = q0 + row;
    if (q < nrows) {
      const float2 v = r[(row >> 1) * N + y];
      out[(long long)q * N + y] = (row & 1) ? v.y : v.x;
    }
  }
}

// ------------------------------------------------ prime-first tile passes
// The same three passes (R2C rows, C2R rows, C2C along x) for T = p * M with a
// matrix-product prime p: load fused with the prime pre-step, one LDS buffer,
// radix-M final stage straight to global memory.  Persistent blocks: each
// loops over groups of G transforms (group index stepped by the grid size,
// XCD-remapped so that neighbouring groups share an L2) and prefetches the
// next group's raw inputs into registers during the current group's matrix
// stage; the loads are unconditional (clamped addresses) so they are all in
// flight at once.
template <bool INV, int M>
__global__ __launch_bounds__(256) void k_pf_c2c(float2* data, long long estride, long long cstride, int ncols,
                                                long long bstride, int nbatch, PfShape S,
                                                const float2* __restrict__ tw, const float* __restrict__ coef, int G) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PfLds L = pf_carve(smem, S, M, G, tw, coef);
  const int ncb = (ncols + G - 1) / G;
  const long long ngrp = (long long)ncb * nbatch;
  const int es = (int)estride;  // element offsets within one batch fit in 32 bits
  const int cs = (int)cstride;
  auto setup = [&](long long gi, float2*& base, int& gmax) {
    const long long batch = gi / ncb;
    const int cb = (int)(gi - batch * ncb);
    base = data + batch * bstride + (long long)cb * G * cstride;
    gmax = min(G, ncols - cb * G);
  };
  PfItems I;
  pf_items<M, true>(S, G, I);
  float2 A[PF_NI], B[PF_NI], V0;
  auto fetch = [&](const float2* base, int gmax) {  // absent columns read the last present one
    pf_fetch<M>(
        G, I,
        [&](int u, int g, int n1, int n2) {
          const float2* col = base + min(g, gmax - 1) * cs;
          A[u] = col[n1 * es];
          B[u] = col[n2 * es];
        },
        [&](int g, int j) { V0 = base[min(g, gmax - 1) * cs + j * es]; });
  };
  long long grp = xcd_remap(blockIdx.x, gridDim.x);
  float2* base = nullptr;
  int gmax = 0;
  if (grp < ngrp) {
    setup(grp, base, gmax);
    fetch(base, gmax);
  }
  for (; grp < ngrp; grp += gridDim.x) {
    pf_put<M>(
        S, G, L.sd, I, [&](int u, int, int, int, float2& a, float2& b) { a = A[u]; b = B[u]; },
        [&](int, int) { return V0; });
    __syncthreads();
    float2* nbase = nullptr;
    int ngmax = 0;
    if (grp + gridDim.x < ngrp) {
      setup(grp + gridDim.x, nbase, ngmax);
      fetch(nbase, ngmax);
    }
    pf_matrix<M, INV>(S, L.sd, L.coef, G);
    pf_final<M, INV, true>(S, (const float2*)L.sd, L.tw, G, [&](int g, int j, const float2* v) {
      if (g >= gmax) return;
      float2* col = base + (g * cs + j * es);
#pragma unroll
      for (int r = 0; r < M; ++r) col[r * S.p * es] = v[r];
    });
    __syncthreads();
    base = nbase;
    gmax = ngmax;
  }
}

// v / f with a reciprocal and one correction step (within 1 ulp of the IEEE
// quotient, equal to it for nearly all operands).
__device__ __forceinline__ float div_nr(float v, float f) {
  const float r = __builtin_amdgcn_rcpf(f);
  const float q = v * r;
  return fmaf(fmaf(-q, f, v), r, q);
}

// R2C row tables of one group: ymap (T ints: padded column -> data column or
// -1) and, per padded row, the stack offset of its data row (-1: zero row)
// and the flat-field offset.
struct R2cTab {
  int* ymap;
  long long* roff;
  long long* foff;
};

// Raw prefetch registers of the R2C pass: pixel bits (uint16 zero-extended or
// float32 bits) and flat values of the 4 samples of each item (rows 2g, 2g+1
// at columns n1, n2), with their validity bits (4 per item).
struct R2cRaw {
  unsigned P[PF_NI][4];
  float F[PF_NI][4];
  unsigned msk;
  unsigned P0[2];
  float F0[2];
  unsigned m0;
};

// Raw bits of a stack element: float32 bits, or the integer pixel value (u8 squared).
template <int PIX>
__device__ __forceinline__ unsigned pix_bits(const void* img, long long o) {
  if (PIX == MM_PIX_F32) return __float_as_uint(((const float*)img)[o]);
  if (PIX == MM_PIX_U8SQ) {
    const unsigned u = ((const uint8_t*)img)[o];
    return u * u;
  }
  return (unsigned)((const uint16_t*)img)[o];
}

template <int M, int PIX, bool FLAT>
__device__ __forceinline__ void r2c_fetch(int G, const PfItems& I, R2cRaw& R, const mm_tile_dev& td,
                                          const R2cTab& tb) {
  const float* fl = td.flat;
  auto raw = [&](long long o) -> unsigned { return pix_bits<PIX>(td.img, o); };
  R.msk = 0;
  R.m0 = 0;
  pf_fetch<M>(
      G, I,
      [&](int u, int g, int n1, int n2) {
        const long long o0 = tb.roff[2 * g], o1 = tb.roff[2 * g + 1];
        const int c1 = tb.ymap[n1], c2 = tb.ymap[n2];
        const long long oo[4] = {o0 + c1, o1 + c1, o0 + c2, o1 + c2};
        const bool ok[4] = {o0 >= 0 && c1 >= 0, o1 >= 0 && c1 >= 0, o0 >= 0 && c2 >= 0, o1 >= 0 && c2 >= 0};
        unsigned m = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          R.P[u][e] = raw(ok[e] ? oo[e] : 0);
          m |= ok[e] ? 1u << e : 0u;
        }
        if (FLAT) {
          const long long f0 = tb.foff[2 * g], f1 = tb.foff[2 * g + 1];
          const long long ff[4] = {f0 + c1, f1 + c1, f0 + c2, f1 + c2};
#pragma unroll
          for (int e = 0; e < 4; ++e) R.F[u][e] = fl[ok[e] ? ff[e] : 0];
        }
        R.msk |= m << (4 * u);
      },
      [&](int g, int j) {
        const int c = tb.ymap[j];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const long long o = tb.roff[2 * g + e];
          const bool ok = o >= 0 && c >= 0;
          R.P0[e] = raw(ok ? o + c : 0);
          if (FLAT) R.F0[e] = fl[ok ? tb.foff[2 * g + e] + c : 0];
          R.m0 |= ok ? 1u << e : 0u;
        }
      });
}

template <bool F32, bool FLAT>
__device__ __forceinline__ float r2c_val(unsigned bits, float f, bool ok) {
  float x = F32 ? __uint_as_float(bits) : (float)bits;
  if (FLAT) x = div_nr(x, f);
  return ok ? x : 0.0f;
}

// The same with rf = RN(1 / flat) (computed once for all planes): x * rf, within
// 1.5 ulp of the IEEE quotient.
template <bool F32, bool FLAT>
__device__ __forceinline__ float r2c_val_rcp(unsigned bits, float rf, bool ok) {
  float x = F32 ? __uint_as_float(bits) : (float)bits;
  if (FLAT) x = x * rf;
  return ok ? x : 0.0f;
}

template <int M, bool F32, bool FLAT>
__device__ __forceinline__ void r2c_put(const PfShape& S, int G, float4* sd, const PfItems& I, const R2cRaw& R) {
  pf_put<M>(
      S, G, sd, I,
      [&](int u, int, int, int, float2& a, float2& b) {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = r2c_val<F32, FLAT>(R.P[u][e], R.F[u][e], (R.msk >> (4 * u + e)) & 1u);
        a = make_float2(v[0], v[1]);
        b = make_float2(v[2], v[3]);
      },
      [&](int, int) {
        return make_float2(r2c_val<F32, FLAT>(R.P0[0], R.F0[0], R.m0 & 1u),
                           r2c_val<F32, FLAT>(R.P0[1], R.F0[1], (R.m0 >> 1) & 1u));
      });
}

// Two padded rows (z, x) per transform: re = row 2g, im = row 2g + 1 of the
// flat-divided, reflection-padded tile; half spectra written to S.  A group is
// 2G consecutive rows of one tile.
template <int M>
__global__ __launch_bounds__(256) void k_pf_r2c(const mm_tile_dev* __restrict__ tiles, int ntiles, int T, int nrows,
                                                float2* Sout, long long s_tstride, PfShape S,
                                                const float2* __restrict__ tw, const float* __restrict__ coef, int G) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PfLds L = pf_carve(smem, S, M, G, tw, coef);
  R2cTab tab0, tab1;  // double-buffered: the next group's tables are built during the current group
  {
    unsigned char* e = L.extra;
    R2cTab* tt[2] = {&tab0, &tab1};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      tt[k]->ymap = (int*)e;
      e += mm_align16((size_t)T * 4);
      tt[k]->roff = (long long*)e;
      tt[k]->foff = tt[k]->roff + 2 * G;
      e += (size_t)4 * G * 8;
    }
  }
  const int nrb = (nrows + 2 * G - 1) / (2 * G);
  const long long ngrp = (long long)nrb * ntiles;
  auto tables = [&](long long gi, const R2cTab& tb) {
    const int tile = (int)(gi / nrb);
    const int q0 = (int)(gi - (long long)tile * nrb) * 2 * G;
    const mm_tile_dev& td = tiles[tile];
    for (int y = threadIdx.x; y < T; y += blockDim.x) tb.ymap[y] = pad_map(y, td.ys, td.yd, td.yfirst, T);
    for (int i = threadIdx.x; i < 2 * G; i += blockDim.x) {
      const int q = q0 + i;
      long long ro = -1, fo = 0;
      if (q < nrows) {
        const int z = q / T, x = q - (q / T) * T;
        const int rr = pad_map(x, td.xs, td.xd, td.xfirst, T);
        if (rr >= 0) {
          ro = (long long)z * td.img_sz + (long long)(td.X0 + rr) * td.img_sx + td.Y0;
          fo = (long long)(td.X0 + rr) * td.flat_sx + td.Y0;
        }
      }
      tb.roff[i] = ro;
      tb.foff[i] = fo;
    }
  };
  PfItems I;
  pf_items<M, false>(S, G, I);
  R2cRaw R;
  // pixel type and flat presence are uniform over a group (one tile): one code path each
  auto fetch = [&](long long gi, const R2cTab& tb) {
    const mm_tile_dev& td = tiles[(int)(gi / nrb)];
    if (td.img_type == MM_PIX_F32) {
      if (td.flat) r2c_fetch<M, MM_PIX_F32, true>(G, I, R, td, tb);
      else r2c_fetch<M, MM_PIX_F32, false>(G, I, R, td, tb);
    } else if (td.img_type == MM_PIX_U8SQ) {
      if (td.flat) r2c_fetch<M, MM_PIX_U8SQ, true>(G, I, R, td, tb);
      else r2c_fetch<M, MM_PIX_U8SQ, false>(G, I, R, td, tb);
    } else {
      if (td.flat) r2c_fetch<M, MM_PIX_U16, true>(G, I, R, td, tb);
      else r2c_fetch<M, MM_PIX_U16, false>(G, I, R, td, tb);
    }
  };
  auto put = [&](long long gi) {
    const mm_tile_dev& td = tiles[(int)(gi / nrb)];
    if (td.img_type == MM_PIX_F32) {
      if (td.flat) r2c_put<M, true, true>(S, G, L.sd, I, R);
      else r2c_put<M, true, false>(S, G, L.sd, I, R);
    } else {
      if (td.flat) r2c_put<M, false, true>(S, G, L.sd, I, R);
      else r2c_put<M, false, false>(S, G, L.sd, I, R);
    }
  };
  const int H = S.N / 2 + 1;
  long long grp = xcd_remap(blockIdx.x, gridDim.x);
  int set = 0;
  if (grp < ngrp) {
    tables(grp, tab0);
    __syncthreads();
    fetch(grp, tab0);
  }
  for (; grp < ngrp; grp += gridDim.x) {
    const long long nxt = grp + gridDim.x;
    const R2cTab& tn = set ? tab0 : tab1;
    if (nxt < ngrp) tables(nxt, tn);
    put(grp);
    __syncthreads();
    if (nxt < ngrp) fetch(nxt, tn);
    pf_matrix<M, false>(S, L.sd, L.coef, G);
    const int tile = (int)(grp / nrb);
    const int q0 = (int)(grp - (long long)tile * nrb) * 2 * G;
    float2* out = Sout + (long long)tile * s_tstride;
    pf_final_r2c<M>(S, (const float2*)L.sd, L.tw, G, [&](int g, float2*& pa, float2*& pb) {
      const int q = q0 + 2 * g;
      pa = q < nrows ? out + (long long)q * H : nullptr;
      pb = q + 1 < nrows ? out + (long long)(q + 1) * H : nullptr;
    });
    __syncthreads();
    set ^= 1;
  }
}

// Two half-spectrum rows per transform -> two real rows of res.  A group is
// 2G consecutive rows of one tile.
template <int M>
__global__ __launch_bounds__(256) void k_pf_c2r(const float2* Sin, long long s_tstride, float* res,
                                                long long r_tstride, int ntiles, int nrows, PfShape S,
                                                const float2* __restrict__ tw, const float* __restrict__ coef, int G) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PfLds L = pf_carve(smem, S, M, G, tw, coef);
  const int N = S.N, H = N / 2 + 1;
  const int nrb = (nrows + 2 * G - 1) / (2 * G);
  const long long ngrp = (long long)nrb * ntiles;
  PfItems I;
  pf_items<M, false>(S, G, I);
  // raw half-spectrum values: [u][0] rows (2g, 2g+1) at bin kk(n1), [u][1] at kk(n2)
  float2 RA[PF_NI][2], RB[PF_NI][2], R0[2];
  auto kk = [&](int n) { return n >= H ? N - n : n; };
  // Z[n] of the packed transform from the two rows' bins (absent rows read as 0)
  auto spec = [&](float2 a, float2 b, int n, bool va, bool vb) -> float2 {
    if (!va) a = make_float2(0.f, 0.f);
    if (!vb) b = make_float2(0.f, 0.f);
    if (n >= H) { a.y = -a.y; b.y = -b.y; }
    if (n == 0 || 2 * n == N) { a.y = 0.f; b.y = 0.f; }  // real part semantics of ifftn(...).real
    return make_float2(a.x - b.y, a.y + b.x);
  };
  auto fetch = [&](long long gi) {
    const int tile = (int)(gi / nrb);
    const int q0 = (int)(gi - (long long)tile * nrb) * 2 * G;
    const float2* in = Sin + (long long)tile * s_tstride;
    auto rows = [&](int g, const float2*& ra, const float2*& rb) {  // absent rows read the last row
      const int q = q0 + 2 * g;
      ra = in + (long long)min(q, nrows - 1) * H;
      rb = in + (long long)min(q + 1, nrows - 1) * H;
    };
    pf_fetch<M>(
        G, I,
        [&](int u, int g, int n1, int n2) {
          const float2 *ra, *rb;
          rows(g, ra, rb);
          const int k1 = kk(n1), k2 = kk(n2);
          RA[u][0] = ra[k1];
          RB[u][0] = rb[k1];
          RA[u][1] = ra[k2];
          RB[u][1] = rb[k2];
        },
        [&](int g, int j) {
          const float2 *ra, *rb;
          rows(g, ra, rb);
          R0[0] = ra[j];
          R0[1] = rb[j];
        });
  };
  long long grp = xcd_remap(blockIdx.x, gridDim.x);
  if (grp < ngrp) fetch(grp);
  for (; grp < ngrp; grp += gridDim.x) {
    const int tile = (int)(grp / nrb);
    const int q0 = (int)(grp - (long long)tile * nrb) * 2 * G;
    pf_put<M>(
        S, G, L.sd, I,
        [&](int u, int g, int n1, int n2, float2& a, float2& b) {
          const int q = q0 + 2 * g;
          a = spec(RA[u][0], RB[u][0], n1, q < nrows, q + 1 < nrows);
          b = spec(RA[u][1], RB[u][1], n2, q < nrows, q + 1 < nrows);
        },
        [&](int g, int j) {
          const int q = q0 + 2 * g;
          return spec(R0[0], R0[1], j, q < nrows, q + 1 < nrows);
        });
    __syncthreads();
    if (grp + gridDim.x < ngrp) fetch(grp + gridDim.x);
    pf_matrix<M, true>(S, L.sd, L.coef, G);
    float* out = res + (long long)tile * r_tstride;
    pf_final<M, true>(S, (const float2*)L.sd, L.tw, G, [&](int g, int j, const float2* v) {
      const int q = q0 + 2 * g;
      float* oa = out + (long long)q * N + j;
#pragma unroll
      for (int r = 0; r < M; ++r) {
        if (q < nrows) oa[r * S.p] = v[r].x;
        if (q + 1 < nrows) oa[N + r * S.p] = v[r].y;
      }
    });
    __syncthreads();
  }
}

// ------------------------------------------------ Rader (p = 113, 239) tile passes
// The prime-first passes with the prime stage of fft_rader.h (T = p * M).
// Same persistent structure: the next group's raw elements are prefetched
// into registers during the current group's Rader and final stages.  Only
// outputs k in [klo, khi) are stored (the rows a windowed C2R reads).
template <int PR, bool INV, int M>
__global__ __launch_bounds__(256, 2) void k_rd_c2c(float2* data, long long estride, long long cstride, int ncols,
                                                long long bstride, int nbatch, long long gstride, RdTab tab,
                                                const float2* __restrict__ tw, int G, int klo, int khi) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = PR * M;
  const RdLds L = rd_carve<PR>(smem, N, M, G, tw, tab);
  __syncthreads();  // slot_in is read by rd_items
  const int ncb = (ncols + G - 1) / G;
  const long long ngrp = (long long)ncb * nbatch;
  const int es = (int)estride;
  const int cs = (int)cstride;
  auto setup = [&](long long gi, float2*& base, int& gmax) {
    const long long batch = gi / ncb;
    const int cb = (int)(gi - batch * ncb);
    base = data + batch * bstride + cb * gstride;
    gmax = min(G, ncols - cb * G);
  };
  RdItems I;
  rd_items<PR, M, true>(G, L.slot_in, I);
  const RdMid mid = PR == 113 ? rd_mid(L) : RdMid{};
  float2 A[RD_NI];
  auto fetch = [&](const float2* base, int gmax) {  // absent columns read the last present one
    rd_fetch(I, [&](int u, int g, int n) { A[u] = base[min(g, gmax - 1) * cs + n * es]; });
  };
  long long grp = xcd_remap(blockIdx.x, gridDim.x);
  float2* base = nullptr;
  int gmax = 0;
  if (grp < ngrp) {
    setup(grp, base, gmax);
    fetch(base, gmax);
  }
  for (; grp < ngrp; grp += gridDim.x) {
    rd_put(L.bufA, I, [&](int u, int, int) { return A[u]; });
    __syncthreads();
    float2* nbase = nullptr;
    int ngmax = 0;
    if (grp + gridDim.x < ngrp) {
      setup(grp + gridDim.x, nbase, ngmax);
      fetch(nbase, ngmax);
    }
    rd_stages<PR, M, true, PR == 113>(L, G, mid);
    rd_final<PR, M, INV>((const float2*)L.bufB, L.tw, L.pos, G, [&](int g, int j, const float2* v) {
      if (g >= gmax) return;
      float2* col = base + (g * cs + j * es);
#pragma unroll
      for (int r = 0; r < M; ++r)
        if (j + r * PR >= klo && j + r * PR < khi) col[r * PR * es] = v[r];
    });
    base = nbase;
    gmax = ngmax;
    __syncthreads();  // the final stage's reads before the next group's stores
  }
}

// Row passes (R2C, C2R) by columns: thread tid loads columns n = tid + k * 256
// (k < NC) of each of the 2G rows of its group, so every load address is a
// uniform row pointer plus a per-lane column offset fixed for the kernel.
// Groups of G transforms (at most 36 row samples per thread; the 2G rows'
// validity bits fit one 32-bit mask).  For M = 6 small groups (6 rows in the
// R2C, 4 in the C2R: 25 KB and 19 KB of LDS) let 4 blocks share a CU; they
// measured 5 to 10% faster than groups of 12 rows.
// For p = 239 (M = 2) groups of 4 transforms (8 rows) in both passes, on 128-thread
// blocks (G M B = 112 items in the in-place stages): the C2R with 3 transforms on
// 128 threads took 1.93 ms per 100-tile T 478 stack, with 4 1.49 ms (5: 1.54, on
// 256 threads; 6: 1.88); R2C groups of 6 on 256 threads were slower (2.52 ms).
template <int PR>
constexpr int rd_row_g(int M, bool c2r) {
  return PR == 239 ? 4 : (M == 2 ? 16 : (M == 6 ? (c2r ? 2 : 3) : 9));
}
template <int PR>
constexpr int rd_row_nc(int M, int bs = 256) { return (PR * M + bs - 1) / bs; }

template <int NC>
struct RdCols {
  int n[NC];   // column
  int cs[NC];  // LDS slot of the column within one transform's M * S block, -1 for n >= N
};

template <int PR, int M, int NC, int BS = 256>
__device__ __forceinline__ RdCols<NC> rd_cols(const int* __restrict__ slot_in) {
  RdCols<NC> c;
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int n = threadIdx.x + k * BS;
    c.n[k] = n;
    c.cs[k] = n < PR * M ? (n % M) * RdSpec<PR>::S + slot_in[n / M] : -1;
  }
  return c;
}

// Raw samples of the R2C group: pixel bits and flat reciprocals of row i (2G rows
// of the padded tile from x0), column k; row i exists and maps to data <=> bit i
// of rowok (uniform), column k maps to data <=> bit k of colok.
template <int NR, int NC>
struct RdR2cRaw {
  unsigned P[NR][NC];
  float F[NR][NC];
  unsigned rowok, colok;
  static_assert(NR <= 32 && NC <= 32, "validity masks are 32-bit");
};

// What a work item (tile, band of NR rows) fixes for all its planes: the data
// column of each of the thread's columns (absent columns read data column 0),
// and each row's offset in plane 0 of the stack (absent rows read the stack's
// first row).
template <int NR, int NC>
struct RdR2cItem {
  unsigned c[NC];
  long long ro[NR];
};

__device__ __forceinline__ long long rd_readlane64(long long v, int lane) {
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)(unsigned long long)v, lane);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)((unsigned long long)v >> 32), lane);
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// Set up rows x0 .. x0 + NR - 1 of tile td (those < T exist): the column and row
// maps, and the reciprocals of the flat values of every sample (the flat is the
// same in every plane).
template <int PR, int M, int NR, int NC, bool FLAT>
__device__ __forceinline__ void rd_r2c_item(RdR2cItem<NR, NC>& It, RdR2cRaw<NR, NC>& R, const mm_tile_dev& td,
                                            const RdCols<NC>& C, int T, int x0) {
  R.colok = 0;
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int cc = C.n[k] < PR * M ? pad_map(C.n[k], td.ys, td.yd, td.yfirst, T) : -1;
    R.colok |= (cc >= 0 ? 1u : 0u) << k;
    It.c[k] = cc < 0 ? 0u : (unsigned)cc;
  }
  // Row i of the group is worked out by lane i (vector work in parallel instead
  // of NR rounds of scalar 64-bit arithmetic), then broadcast with readlane.
  const int lane = threadIdx.x & 63;
  long long ro_l = 0, fo_l = 0;
  bool ok_l = false;
  if (lane < NR) {
    const int x = x0 + lane;
    const int rr = x < T ? pad_map(x, td.xs, td.xd, td.xfirst, T) : -1;
    ok_l = rr >= 0;
    if (ok_l) {
      ro_l = (long long)(td.X0 + rr) * td.img_sx + td.Y0;
      fo_l = (long long)(td.X0 + rr) * td.flat_sx + td.Y0;
    }
  }
  R.rowok = (unsigned)__ballot(ok_l);  // lanes 0..NR-1 (NR <= 32)
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    It.ro[i] = rd_readlane64(ro_l, i);
    if (FLAT) {
      const gptr<const char> fb = (gptr<const char>)td.flat + rd_readlane64(fo_l, i) * 4;
#pragma unroll
      for (int k = 0; k < NC; ++k) R.F[i][k] = 1.0f / *(gptr<const float>)(fb + It.c[k] * 4u);
    }
  }
}

// The pixels of the item's rows in plane z.
template <int NR, int NC, int PIX>
__device__ __forceinline__ void rd_r2c_pix(RdR2cRaw<NR, NC>& R, const RdR2cItem<NR, NC>& It, const mm_tile_dev& td,
                                           int z) {
  const long long zo = (long long)z * td.img_sz;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    // uniform row pointers in the global address space: global loads with a
    // scalar base and a 32-bit lane offset
    constexpr int ES = PIX == MM_PIX_F32 ? 4 : PIX == MM_PIX_U16 ? 2 : 1;  // element bytes
    const gptr<const char> rb = (gptr<const char>)td.img + (zo + It.ro[i]) * ES;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      if (PIX == MM_PIX_F32) R.P[i][k] = *(gptr<const unsigned>)(rb + It.c[k] * 4u);
      else if (PIX == MM_PIX_U16) R.P[i][k] = (unsigned)*(gptr<const uint16_t>)(rb + It.c[k] * 2u);
      else {
        const unsigned u = *(gptr<const uint8_t>)(rb + It.c[k]);
        R.P[i][k] = u * u;
      }
    }
  }
}

template <int PR, int M, int NR, int NC, bool F32, bool FLAT>
__device__ __forceinline__ void rd_r2c_store(float2* __restrict__ bufA, const RdR2cRaw<NR, NC>& R,
                                             const RdCols<NC>& C) {
#pragma unroll
  for (int g = 0; g < NR / 2; ++g)
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      if (C.cs[k] < 0) continue;
      const unsigned ok = (R.colok >> k) & 1u;
      bufA[g * M * RdSpec<PR>::S + C.cs[k]] =
          make_float2(r2c_val_rcp<F32, FLAT>(R.P[2 * g][k], R.F[2 * g][k], ok & (R.rowok >> (2 * g))),
                      r2c_val_rcp<F32, FLAT>(R.P[2 * g + 1][k], R.F[2 * g + 1][k], ok & (R.rowok >> (2 * g + 1))));
    }
}

// Work items (tile, band of NR rows, chunk of zc planes); each block runs the
// planes of its item in turn, so the column and row maps and the flat values are
// read once per item (the flat is the same in every plane), and prefetches the
// next plane's pixels (the next item's first plane) during the current plane's
// stages.  zc = sz unless the batch has too few (tile, band) pairs to fill the GPU.
template <int PR, int M, int BS = 256>
__global__ __launch_bounds__(BS, 2) void k_rd_r2c(const mm_tile_dev* __restrict__ tiles, int ntiles, int T,
                                                       int nrows, float2* Sout, long long s_tstride, SpecLayout Ls,
                                                       RdTab tab, const float2* __restrict__ tw, int zc) {
  constexpr int G = rd_row_g<PR>(M, false), NC = rd_row_nc<PR>(M, BS), NR = 2 * G;
  static_assert(G * M * RdSpec<PR>::B <= BS, "in-place Rader stages: one item per thread");
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = PR * M;
  const RdLds L = rd_carve<PR>(smem, N, M, G, tw, tab);
  __syncthreads();  // slot_in
  const RdCols<NC> C = rd_cols<PR, M, NC, BS>(L.slot_in);
  const int sz = nrows / T, nband = (T + NR - 1) / NR, nch = (sz + zc - 1) / zc;
  const long long nitem = (long long)nband * ntiles * nch;
  RdR2cRaw<NR, NC> R;
  RdR2cItem<NR, NC> It;
  int tile = 0, x0 = 0, z0 = 0, z1 = 0, pix = MM_PIX_U16;
  bool flat = false;
  // the item's maps and flat values
  auto setup = [&](long long item) {
    const long long tb = item / nch;
    const int ch = (int)(item - tb * nch);
    tile = (int)(tb / nband);
    x0 = (int)(tb - (long long)tile * nband) * NR;
    z0 = ch * zc;
    z1 = min(sz, z0 + zc);
    const mm_tile_dev& td = tiles[tile];
    pix = td.img_type;
    flat = td.flat != nullptr;
    if (flat) rd_r2c_item<PR, M, NR, NC, true>(It, R, td, C, T, x0);
    else rd_r2c_item<PR, M, NR, NC, false>(It, R, td, C, T, x0);
  };
  auto fetch = [&](int z) {
    const mm_tile_dev& td = tiles[tile];
    if (pix == MM_PIX_F32) rd_r2c_pix<NR, NC, MM_PIX_F32>(R, It, td, z);
    else if (pix == MM_PIX_U8SQ) rd_r2c_pix<NR, NC, MM_PIX_U8SQ>(R, It, td, z);
    else rd_r2c_pix<NR, NC, MM_PIX_U16>(R, It, td, z);
  };
  auto put = [&]() {
    if (pix == MM_PIX_F32) {
      if (flat) rd_r2c_store<PR, M, NR, NC, true, true>(L.bufA, R, C);
      else rd_r2c_store<PR, M, NR, NC, true, false>(L.bufA, R, C);
    } else {
      if (flat) rd_r2c_store<PR, M, NR, NC, false, true>(L.bufA, R, C);
      else rd_r2c_store<PR, M, NR, NC, false, false>(L.bufA, R, C);
    }
  };
  long long item = xcd_remap(blockIdx.x, gridDim.x);
  if (item >= nitem) return;  // uniform over the block
  setup(item);
  fetch(z0);
  for (;;) {
    const int ctile = tile, cx0 = x0, cz1 = z1;
    for (int z = z0; z < cz1; ++z) {
      put();
      __syncthreads();
      if (z + 1 < cz1) {
        fetch(z + 1);
      } else if (item + gridDim.x < nitem) {
        setup(item + gridDim.x);
        fetch(z0);
      }
      rd_stages<PR, M, false, false>(L, G, RdMid{});
      float2* out = Sout + (long long)ctile * s_tstride + z * Ls.plane;
      auto rows = [&](int g, float2*& pa, float2*& pb) {
        const int xa = cx0 + 2 * g;
        pa = xa < T ? out + Ls.row(xa) : nullptr;
        pb = xa + 1 < T ? out + Ls.row(xa + 1) : nullptr;
      };
      if (Ls.blk == 6) {  // the M = 6 plan: constant block width, the division is a multiply-shift
        const int bstr = T * 6;
        pf_final_r2c<M>(PfShape(N, PR), (const float2*)L.bufB, L.tw, G, rows, [&](int n) {
          const int b = (int)(__umul24((unsigned)n, 43691u) >> 18);  // n / 6 for n < 2^15
          return __mul24(b, bstr) + (n - 6 * b);
        });
      } else if (Ls.blk) {
        const FastDiv fb(Ls.blk);
        const int bstr = T * Ls.blk;
        pf_final_r2c<M>(PfShape(N, PR), (const float2*)L.bufB, L.tw, G, rows, [&](int n) {
          int b, r;
          fb.divmod(n, b, r);
          return b * bstr + r;
        });
      } else {
        pf_final_r2c<M>(PfShape(N, PR), (const float2*)L.bufB, L.tw, G, rows);
      }
      __syncthreads();  // the final stage's reads before the next group's stores
    }
    item += gridDim.x;
    if (item >= nitem) break;
  }
}

// Where the C2R stores the owned part of each tile straight into the stitched
// volume (deconvolver.py:179-183, as k_stitch): rows and columns [ov, xe - ov) x
// [ov, ye - ov) of tile t go to out[z o_sz + (ox + x - ov) o_sx + oy + y - ov]
// inside (SX, SY).
struct C2rStitch {
  const mm_tile_dev* tiles;
  float* out;
  long long o_sz, o_sx;
  int SX, SY, ov;
};

// Rows (z, x) with x in [xlo, xlo + nx) only (nrows = sz * nx of them, taken
// in that order), output columns in [clo, chi) only.
// MARK: also the maxima threshold test (!(v <= thr), maxima.cu:19) of every
// stored output, as ww 64-bit words per row (bit c & 63 of word c >> 6 for
// column c; zero for the columns not stored) at marks[(tile * N + z) * N + x]
// (the row layout of the maxima workspace), for every computed row.
// STITCH: the stores go to the stitched volume (C2rStitch) instead of res.
template <int PR, int M, bool MARK = false, int BS = 256, bool STITCH = false>
__global__ __launch_bounds__(BS, BS == 128 && PR == 113 ? 4 : 2) void k_rd_c2r(const float2* Sin, long long s_tstride, SpecLayout Ls,
                                                       float* res, long long r_tstride, int ntiles, int nrows, int xlo,
                                                       int nx, int clo, int chi, RdTab tab,
                                                       const float2* __restrict__ tw,
                                                       unsigned long long* __restrict__ marks = nullptr,
                                                       float thr = 0.0f, int ww = 0, C2rStitch sti = C2rStitch{}) {
  constexpr int G = rd_row_g<PR>(M, true), NC = rd_row_nc<PR>(M, BS), NR = 2 * G;
  static_assert(G * M * RdSpec<PR>::B <= BS, "in-place Rader stages: one item per thread");
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = PR * M, H = N / 2 + 1;
  const RdLds L = rd_carve<PR>(smem, N, M, G, tw, tab);
  // MARK: the threshold words of the group's NR rows (after the Rader buffer)
  unsigned long long* wl = (unsigned long long*)L.extra;
  if (MARK)
    for (int i = threadIdx.x; i < NR * ww; i += blockDim.x) wl[i] = 0ull;
  __syncthreads();
  const RdCols<NC> C = rd_cols<PR, M, NC, BS>(L.slot_in);
  int kk[NC];  // in-row offset of the half-spectrum bin of column n (bin 0 for the absent columns n >= N)
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int b = C.n[k] >= N ? 0 : (C.n[k] >= H ? N - C.n[k] : C.n[k]);
    kk[k] = Ls.blk ? (b / Ls.blk) * N * Ls.blk + b % Ls.blk : b;
  }
  const int nrb = (nrows + NR - 1) / NR;
  const long long ngrp = (long long)nrb * ntiles;
  float2 RA[G][NC], RB[G][NC];  // bins kk of the two rows of each transform
  // row q of the window is (z, x) = (q / nx, xlo + q % nx); a group's rows cross
  // at most one plane boundary (NR < nx), so one division per group
  auto rowzx = [&](int q0, int z0, int i, int& z, int& x) {
    x = q0 - z0 * nx + i;
    z = z0;
    if (x >= nx) {
      x -= nx;
      ++z;
    }
    x += xlo;
  };
  auto fetch = [&](long long gi) {  // absent rows read the last row
    const int tile = (int)(gi / nrb);
    const int q0 = (int)(gi - (long long)tile * nrb) * NR;
    const int z0 = q0 / nx;
    const float2* in = Sin + (long long)tile * s_tstride;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int ia = min(2 * g, nrows - 1 - q0), ib = min(2 * g + 1, nrows - 1 - q0);
      int za, xa, zb, xb;
      rowzx(q0, z0, ia, za, xa);
      rowzx(q0, z0, ib, zb, xb);
      const float2* ra = in + za * Ls.plane + Ls.row(xa);
      const float2* rb = in + zb * Ls.plane + Ls.row(xb);
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        RA[g][k] = ra[kk[k]];
        RB[g][k] = rb[kk[k]];
      }
    }
  };
  long long grp = xcd_remap(blockIdx.x, gridDim.x);
  if (grp < ngrp) fetch(grp);
  for (; grp < ngrp; grp += gridDim.x) {
    const int tile = (int)(grp / nrb);
    const int q0 = (int)(grp - (long long)tile * nrb) * NR;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int q = q0 + 2 * g;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        if (C.cs[k] < 0) continue;
        const int n = C.n[k];
        float2 a = q < nrows ? RA[g][k] : make_float2(0.f, 0.f);
        float2 b = q + 1 < nrows ? RB[g][k] : make_float2(0.f, 0.f);
        if (n >= H) { a.y = -a.y; b.y = -b.y; }
        if (n == 0 || 2 * n == N) { a.y = 0.f; b.y = 0.f; }  // real part semantics of ifftn(...).real
        L.bufA[g * M * RdSpec<PR>::S + C.cs[k]] = make_float2(a.x - b.y, a.y + b.x);
      }
    }
    __syncthreads();
    if (grp + gridDim.x < ngrp) fetch(grp + gridDim.x);
    rd_stages<PR, M, false, false>(L, G, RdMid{});
    float* out = res + (long long)tile * r_tstride;
    const int z0 = q0 / nx;
    // STITCH: the owned rows and columns of this tile, and where they go
    int s_x1 = 0, s_y1 = 0, s_dx = 0, s_dy = 0;
    if (STITCH) {
      const mm_tile_dev& td = sti.tiles[tile];
      s_dx = td.ox - sti.ov;
      s_dy = td.oy - sti.ov;
      s_x1 = min(td.xe - sti.ov, sti.SX - s_dx);  // rows x in [ov, s_x1)
      s_y1 = min(td.ye - sti.ov, sti.SY - s_dy);  // columns c in [ov, s_y1)
    }
    pf_final<M, true>(PfShape(N, PR), (const float2*)L.bufB, L.tw, G, [&](int g, int j, const float2* v) {
      const int q = q0 + 2 * g;
      int za, xa, zb, xb;
      rowzx(q0, z0, 2 * g, za, xa);
      rowzx(q0, z0, 2 * g + 1, zb, xb);
      if (STITCH) {
        const bool oka = q < nrows && xa >= sti.ov && xa < s_x1, okb = q + 1 < nrows && xb >= sti.ov && xb < s_x1;
        float* sa = sti.out + za * sti.o_sz + (long long)(s_dx + xa) * sti.o_sx + s_dy + j;
        float* sb = sti.out + zb * sti.o_sz + (long long)(s_dx + xb) * sti.o_sx + s_dy + j;
#pragma unroll
        for (int r = 0; r < M; ++r) {
          const int c = j + r * PR;
          const bool cin = c >= sti.ov && c < s_y1;
          if (cin && oka) sa[r * PR] = v[r].x;
          if (cin && okb) sb[r * PR] = v[r].y;
        }
        return;
      }
      float* oa = out + ((long long)za * N + xa) * N + j;
      float* ob = out + ((long long)zb * N + xb) * N + j;
#pragma unroll
      for (int r = 0; r < M; ++r) {
        const int c = j + r * PR;
        const bool cin = c >= clo && c < chi;
        if (cin && q < nrows) oa[r * PR] = v[r].x;
        if (cin && q + 1 < nrows) ob[r * PR] = v[r].y;
        if (MARK && cin) {  // rare: LDS OR of the set bits
          const unsigned long long bit = 1ull << (c & 63);
          if (!(v[r].x <= thr)) atomicOr(wl + (2 * g) * ww + (c >> 6), bit);
          if (!(v[r].y <= thr)) atomicOr(wl + (2 * g + 1) * ww + (c >> 6), bit);
        }
      }
    });
    __syncthreads();  // the final stage's reads before the next group's stores
    if (MARK) {  // the group's words out; the thread that stores word i clears it for the next group
      for (int i = threadIdx.x; i < NR * ww; i += blockDim.x) {
        const int r = i / ww, q = q0 + r;
        if (q < nrows) {
          int z, x;
          rowzx(q0, z0, r, z, x);
          marks[(((long long)tile * (nrows / nx) + z) * N + x) * ww + (i - r * ww)] = wl[i];
        }
        wl[i] = 0ull;
      }
    }
  }
}

#define PF_SWITCH(M, ...)                     \
  switch (M) {                                \
    case 2: { constexpr int MM = 2; __VA_ARGS__; } break; \
    case 3: { constexpr int MM = 3; __VA_ARGS__; } break; \
    case 4: { constexpr int MM = 4; __VA_ARGS__; } break; \
    case 6: { constexpr int MM = 6; __VA_ARGS__; } break; \
    default: mm_set_error("prime-first FFT: radix %d", M); return MM_EINVAL; \
  }

static size_t pf_r2c_extra(int T, int G) { return 2 * (mm_align16((size_t)T * 4) + (size_t)4 * G * 8); }

// Grid of a persistent prime-first pass: as many blocks as fit on the chip at
// once (LDS bound), at most one per group.
static unsigned pf_grid(long long ngrp, size_t lds, int max_per_cu = 4) {
  long long per_cu = (long long)(MM_LDS_MAX / (lds ? lds : 1));
  if (per_cu < 1) per_cu = 1;
  if (per_cu > max_per_cu) per_cu = max_per_cu;
  const long long g = per_cu * 256;
  return (unsigned)(ngrp < g ? ngrp : g);
}

template <bool INV>
static int launch_pf_c2c(const mm_deconv* d, float2* data, long long estride, long long cstride, int ncols,
                         long long bstride, int nbatch, hipStream_t st) {
  const int G = d->pfG;
  const long long ngrp = (long long)((ncols + G - 1) / G) * nbatch;
  const size_t lds = pf_lds_bytes(d->pfS, d->pfM, G, 0);
  PF_SWITCH(d->pfM, hipLaunchKernelGGL((k_pf_c2c<INV, MM>), dim3(pf_grid(ngrp, lds)), dim3(256), lds, st, data,
                                       estride, cstride, ncols, bstride, nbatch, d->pfS, d->pT.tw, d->pT.coef, G));
  MM_LAUNCH_CHECK();
  return MM_OK;
}

// Rader plans: p = 113 with M in {2, 3, 4, 6}, p = 239 with M = 2 (PP, MM constant in CALL).
#define RD_SWITCH(P, M, ...)                                                         \
  if ((P) == 113) {                                                                  \
    constexpr int PP = 113;                                                          \
    PF_SWITCH(M, __VA_ARGS__);                                                       \
  } else if ((P) == 239 && (M) == 2) {                                               \
    constexpr int PP = 239, MM = 2;                                                  \
    __VA_ARGS__;                                                                     \
  } else {                                                                           \
    mm_set_error("Rader FFT: prime %d with radix %d", (int)(P), (int)(M));           \
    return MM_EINVAL;                                                                \
  }

template <bool INV>
// x transforms of every (tile, z) plane of the spectra in layout Ls (the
// groups of G columns are the layout's column blocks when it is blocked).
static int launch_rd_c2c(const mm_deconv* d, float2* data, const SpecLayout& Ls, int nbatch, int klo, int khi,
                         hipStream_t st) {
  const int G = d->pfG;
  const long long ngrp = (long long)((Ls.H + G - 1) / G) * nbatch;
  const long long es = Ls.blk ? Ls.blk : Ls.H;
  const long long gs = Ls.blk ? (long long)Ls.T * Ls.blk : G;
  RD_SWITCH(d->rdP, d->pfM, {
    const size_t lds = rd_lds_bytes<PP>(d->T, MM, G, 0);
    hipLaunchKernelGGL((k_rd_c2c<PP, INV, MM>), dim3(pf_grid(ngrp, lds)), dim3(256), lds, st, data, es, 1LL, Ls.H,
                       Ls.plane, nbatch, gs, INV ? d->rdI : d->rdF, d->pT.tw, G, klo, khi);
  });
  MM_LAUNCH_CHECK();
  return MM_OK;
}

template <int PR, int M>
static void launch_rd_r2c(const mm_tile_dev* td, int ntiles, int T, int nrows, float2* S, long long s_ts,
                          const SpecLayout& Ls, const RdTab& tab, const float2* tw, hipStream_t st) {
  constexpr int G = rd_row_g<PR>(M, false);
  const size_t lds = rd_lds_bytes<PR>(T, M, G, 0);
  // p = 113: 256 threads (4 waves per SIMD at 115 VGPRs); blocks of 128 (161 VGPRs, 3 waves per SIMD) measured
  // 1.45 -> 1.52 ms per 36-tile stack.  p = 239: the 4 transforms' in-place stages take 112 items, so blocks of
  // 128 leave fewer lanes idle (2.09 -> 1.92-1.98 ms per 100-tile T 478 stack)
  constexpr int BS = PR == 239 ? 128 : 256;
  // items of (tile, band) run all planes, unless they are too few to fill the GPU: then chunks of planes
  const int sz = nrows / T;
  const long long nb = (long long)((T + 2 * G - 1) / (2 * G)) * ntiles;
  const int want = (int)((2048 + nb - 1) / nb);
  const int nch = want < 1 ? 1 : (want > sz ? sz : want);
  const int zc = (sz + nch - 1) / nch;
  const long long nitem = nb * ((sz + zc - 1) / zc);
  hipLaunchKernelGGL((k_rd_r2c<PR, M, BS>), dim3(pf_grid(nitem, lds, 512 / BS * 2)), dim3(BS), lds, st, td, ntiles, T,
                     nrows, S, s_ts, Ls, tab, tw, zc);
}

template <int PR, int M>
static void launch_rd_c2r(const float2* S, long long s_ts, const SpecLayout& Ls, float* res, long long r_ts,
                          int ntiles, int sz, int lo, int hi, const RdTab& tab, const float2* tw, hipStream_t st,
                          unsigned long long* marks = nullptr, float thr = 0.0f, const C2rStitch* sti = nullptr) {
  constexpr int G = rd_row_g<PR>(M, true);
  // blocks of 128 threads when the group's in-place stages fit them (fewer idle lanes in stages 1 and 3)
  constexpr int BS = G * M * RdSpec<PR>::B <= 128 ? 128 : 256;
  const int ww = (PR * M + 63) / 64;
  const size_t lds = rd_lds_bytes<PR>(PR * M, M, G, marks ? (size_t)8 * 2 * G * ww : 0);
  const int nrows = sz * (hi - lo);
  const long long ngrp = (long long)((nrows + 2 * G - 1) / (2 * G)) * ntiles;
  const unsigned grid = pf_grid(ngrp, lds, 512 / BS * 2);
  if (sti)
    hipLaunchKernelGGL((k_rd_c2r<PR, M, false, BS, true>), dim3(grid), dim3(BS), lds, st, S, s_ts, Ls, res, r_ts,
                       ntiles, nrows, lo, hi - lo, lo, hi, tab, tw, nullptr, 0.0f, 0, *sti);
  else if (marks)
    hipLaunchKernelGGL((k_rd_c2r<PR, M, true, BS>), dim3(grid), dim3(BS), lds, st, S, s_ts, Ls, res, r_ts, ntiles,
                       nrows, lo, hi - lo, lo, hi, tab, tw, marks, thr, ww);
  else
    hipLaunchKernelGGL((k_rd_c2r<PR, M, false, BS>), dim3(grid), dim3(BS), lds, st, S, s_ts, Ls, res, r_ts, ntiles,
                       nrows, lo, hi - lo, lo, hi, tab, tw, nullptr, 0.0f, 0);
}

// ------------------------------------------------------------ Wiener spectra
// Centred PSF (sz, ts, ts) -> zero padded by (zpad, ov, ov) and rolled by
// (-shape)//2 (deconvolver.py:60-62): out[i] = padded[(i + ceil(n/2)) mod n].
__global__ void k_psf_place(const float* __restrict__ psf, int sz, int ts, int zpad, int ov, int Zp, int T,
                            float* __restrict__ out) {
  const long long total = (long long)Zp * T * T;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(i / ((long long)T * T));
    const int rem = (int)(i - (long long)p * T * T);
    const int x = rem / T, y = rem - (rem / T) * T;
    const int zz = (p + (Zp + 1) / 2) % Zp - zpad;
    const int xx = (x + (T + 1) / 2) % T - ov;
    const int yy = (y + (T + 1) / 2) % T - ov;
    float v = 0.0f;
    if (zz >= 0 && zz < sz && xx >= 0 && xx < ts && yy >= 0 && yy < ts)
      v = psf[((long long)zz * ts + xx) * ts + yy];
    out[i] = v;
  }
}

// W = conj(H) / (|H|^2 + beta |L|^2) * scale, |L|^2 in closed form for the
// Laplacian of deconvolver.py:24-36 (6 at the centre, -1 at z+-1, x+-1 and only
// y-1 on the last axis): |L|^2 = (6 - 2cos a - 2cos b - cos c)^2 + sin^2 c.
// Computed in float64, in place over the PSF spectrum.
//
// box_half = D > 0 also multiplies W by 1 - b(kx) b(ky), the transfer function
// of "subtract the separable box of width 2D + 1" (blur.py box_1d along x,
// then y, deconvolver.py:160-163) on the T-periodic tile:
//   b(k) = (1 / (2D + 1)) sum_{m=-D..D} cos(2 pi k m / T)
//        = sin(pi (2D + 1) k / T) / ((2D + 1) sin(pi k / T)),  b(0) = 1.
// The circular box equals the reference's clipped box at every voxel whose
// box stays inside the tile, so the inverse transform gives res - blur there.
__device__ __forceinline__ double box_tf(int k, int T, int D) {
  if (k == 0) return 1.0;
  const double n = 2.0 * D + 1.0;
  return sinpi(n * k / T) / (n * sinpi((double)k / T));
}

__global__ void k_wiener64(double2* __restrict__ Hs, int Zp, int T, int H, double beta, double scale, int box_half) {
  const long long total = (long long)Zp * T * H;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int kz = (int)(i / ((long long)T * H));
    const int rem = (int)(i - (long long)kz * T * H);
    const int kx = rem / H, ky = rem - (rem / H) * H;
    const double2 h = Hs[i];
    const double re = 6.0 - 2.0 * cospi(2.0 * kz / Zp) - 2.0 * cospi(2.0 * kx / T) - cospi(2.0 * ky / T);
    const double im = sinpi(2.0 * ky / T);
    const double den = h.x * h.x + h.y * h.y + beta * (re * re + im * im);
    const double sc = box_half > 0 ? scale * (1.0 - box_tf(kx, T, box_half) * box_tf(ky, T, box_half)) : scale;
    Hs[i] = make_double2(h.x / den * sc, -h.y / den * sc);
  }
}

__global__ void k_d2f(const double2* __restrict__ a, float2* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

// The same into the column-blocked layout Ls (planes of Ls.plane; the padding
// columns of the last block are zero).
__global__ void k_d2f_blocked(const double2* __restrict__ a, float2* __restrict__ b, int Zp, SpecLayout Ls) {
  const long long n = (long long)Zp * Ls.plane;
  const int bw = Ls.T * Ls.blk;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long p = i / Ls.plane;
    const int c = (int)(i - p * Ls.plane);
    const int kb = c / bw, x = (c - kb * bw) / Ls.blk;
    const int ky = kb * Ls.blk + (c - kb * bw - x * Ls.blk);
    float2 v = make_float2(0.f, 0.f);
    if (ky < Ls.H) {
      const double2 w = a[(p * Ls.T + x) * Ls.H + ky];
      v = make_float2((float)w.x, (float)w.y);
    }
    b[i] = v;
  }
}

// ------------------------------------------------------------ z-kernel pass
#define ZC_CH 12  // tiles per chunk: K is rebuilt once per chunk, the chunk's columns stream through LDS

// Padded planes that copy each data plane (4 slots, -1 empty), passed by value:
// uniform over the grid, read with scalar loads.
struct ZcSrc {
  int p[4 * 48];
};

// Thread (g, t) = (column, output plane).  Per chunk of CH tiles and per data
// plane s it forms K[t][s] = sum over the copies p of s of w[(zpad + t - p) mod Zp]
// from w in LDS (the same order of additions for every chunk, so K is
// identical), then adds K[t][s] x_j[s] to the CH accumulators.  Keeping the
// accumulators instead of a K row in registers leaves room for two blocks per CU.
template <int SZM, bool EXACT, int CH>
__global__ __launch_bounds__(640, 5) void k_tile_zconv(float2* S, long long s_tstride, int TH, int sz_rt, int zpad, int Zp,
                                                    ZcSrc src, const float2* __restrict__ Wz, long long w_stride,
                                                    const mm_tile_dev* __restrict__ tiles, int ntiles, int tpb,
                                                    int G_rt) {
  const int sz = EXACT ? SZM : sz_rt;
  const int G = EXACT ? 640 / SZM : G_rt;  // columns per block
  const int SP = EXACT ? SZM + 1 : sz_rt + 1;  // LDS stride of one column (odd: conflict-free across columns)
  extern __shared__ __align__(16) unsigned char smem[];
  float2* wl = (float2*)smem;            // [p][g]
  float2* xl = wl + (size_t)G * Zp;      // [j][g][s], column stride SP
  const int ncb = (TH + G - 1) / G;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = bid % ncb, tg = bid / ncb;
  const int c0 = cb * G;
  const int gmax = min(G, TH - c0);
  const int tid = threadIdx.x;
  const int g = tid % G, t = tid / G;
  const bool act = t < sz;
  const bool col = act && g < gmax;
  const int t0 = tg * tpb, t1 = min(ntiles, t0 + tpb);
  const float2* Wp = Wz + (long long)tiles[t0].psf * w_stride + c0;  // a group shares one PSF
  for (int i = tid; i < G * Zp; i += blockDim.x) {
    const int gg = i % G, p = i / G;
    wl[p * G + gg] = gg < gmax ? Wp[(long long)p * TH + gg] : make_float2(0.f, 0.f);
  }
  const long long cofs = (long long)t * TH + c0 + g;
  for (int cs = t0; cs < t1; cs += CH) {
    __syncthreads();  // wl is written / the previous chunk is no longer read from xl
    if (act) {  // the chunk's columns (the other resident block computes meanwhile)
      float2 v[CH];
#pragma unroll
      for (int j = 0; j < CH; ++j)
        v[j] = (col && cs + j < t1) ? S[(long long)(cs + j) * s_tstride + cofs] : make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 0; j < CH; ++j) xl[(j * G + g) * SP + t] = v[j];
    }
    __syncthreads();
    float2 y[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) y[j] = make_float2(0.f, 0.f);
    if (act) {
      // not fully unrolled: K[t][s] must stay per-plane work inside the chunk loop, not be
      // hoisted out of it as a whole register row
#pragma unroll(EXACT ? 2 : 1)
      for (int s = 0; s < sz; ++s) {
        {
          float2 k = make_float2(0.f, 0.f);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int p = src.p[s * 4 + q];
            if (p >= 0) {
              int idx = zpad + t - p;
              idx = idx < 0 ? idx + Zp : (idx >= Zp ? idx - Zp : idx);
              const float2 w = wl[idx * G + g];
              k.x += w.x;
              k.y += w.y;
            }
          }
#pragma unroll
          for (int j = 0; j < CH; ++j) {
            const float2 x = xl[(j * G + g) * SP + s];
            y[j].x = fmaf(k.x, x.x, y[j].x);
            y[j].x = fmaf(-k.y, x.y, y[j].x);
            y[j].y = fmaf(k.x, x.y, y[j].y);
            y[j].y = fmaf(k.y, x.x, y[j].y);
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < CH; ++j)
      if (col && cs + j < t1) S[(long long)(cs + j) * s_tstride + cofs] = y[j];
  }
}

// ------------------------------------------------------- z pass (FFT-80)
// Reflection padding with zpad = sz - 1 (the hyb setting), sz <= 40: the
// z-kernel matrix of a column is a Toeplitz part (the data copy) plus a Hankel
// part (the two reflections),
//   K[t][s] = w[t - s] + g[t + s] - [s = 0] w[t] - [s = sz - 1] w[t - sz + 1],
//   g[n] = w[n] + w[n - 2 (sz - 1)]   (indices of w mod Zp),
// and both are circular convolutions of length 80 without aliasing:
//   y = IDFT80(A X + G X~)[t] - c0[t] x[0] - c1[t] x[sz - 1],  t < sz,
// X = DFT80(x zero padded), X~[k] = X[-k], A = DFT80(a) / 80 with
// a[m mod 80] = w[m] (|m| < sz), G = DFT80(g) / 80, c0[t] = w[t],
// c1[t] = w[t - sz + 1].  A column's operator is ZF_OPS planes (A, G, c0, c1),
// built in float64 with W; a thread transforms one column of one tile in
// registers, reading the operator from LDS.
#define ZF_G 32      // columns per block
#define ZF_OPS 240   // operator planes: A[80], G[80], c0[40], c1[40]

template <int SZ>  // 0: depth sz at run time
__global__ __launch_bounds__(256, SZ ? 2 : 1) void k_tile_zfft(float2* S, long long s_tstride, int TH, int sz_rt,
                                                   const float2* __restrict__ Ops, long long ops_stride,
                                                   const mm_tile_dev* __restrict__ tiles, int ntiles, int tpb) {
  const int sz = SZ ? SZ : sz_rt;
  extern __shared__ __align__(16) float4 zsm[];
  float4* AG = zsm;              // [80][ZF_G]: (A[k], G[k])
  float4* CC = zsm + 80 * ZF_G;  // [40][ZF_G]: (c0[t], c1[t])
  const int ncb = (TH + ZF_G - 1) / ZF_G;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = bid % ncb, tg = bid / ncb;
  const int c0 = cb * ZF_G;
  const int gmax = min(ZF_G, TH - c0);
  const int t0 = tg * tpb, t1 = min(ntiles, t0 + tpb);
  const float2* Op = Ops + (long long)tiles[t0].psf * ops_stride + c0;  // a group shares one PSF
  for (int i = threadIdx.x; i < 120 * ZF_G; i += blockDim.x) {
    const int q = i / ZF_G, g = i - (i / ZF_G) * ZF_G;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (g < gmax) {
      const int pa = q < 80 ? q : 80 + q;        // A[k] | c0[t]
      const int pb = q < 80 ? 80 + q : 120 + q;  // G[k] | c1[t]
      const float2 a = Op[(long long)pa * TH + g], b = Op[(long long)pb * TH + g];
      v = make_float4(a.x, a.y, b.x, b.y);
    }
    zsm[i] = v;
  }
  __syncthreads();
  const int ntask = ZF_G * (t1 - t0);
  for (int task = threadIdx.x; task < ntask; task += blockDim.x) {
    const int g = task % ZF_G, j = t0 + task / ZF_G;
    if (g >= gmax) continue;
    float2* col = S + (long long)j * s_tstride + c0 + g;
    float2 x[40];
    {  // a running pointer: no per-plane offsets held in scalar registers
      const float2* p = col;
#pragma unroll
      for (int s = 0; s < 40; ++s) {
        x[s] = s < sz ? *p : make_float2(0.f, 0.f);
        p += TH;
        asm volatile("" : "+v"(p));
      }
    }
    const float2 xf = x[0];
    float2 xl = x[SZ ? SZ - 1 : 0];  // x[sz - 1]
    if (!SZ) {
#pragma unroll
      for (int s = 1; s < 40; ++s) xl = s == sz - 1 ? x[s] : xl;
    }
    float2 U[5][16];
    zf_fwd_stage1(x, U);
    auto AGk = [&](int k) { return AG[k * ZF_G + g]; };
    zf_middle<0>(U, AGk);
    zf_middle<1>(U, AGk);
    zf_middle<2>(U, AGk);
    float2 y[40];
    zf_inv_stage3(U, y);
    // the store addresses are formed here, not kept in registers from the loads
    asm volatile("" : "+v"(col));
#pragma unroll
    for (int t = 0; t < 40; ++t) {
      if (t < sz) {
        const float4 c = CC[t * ZF_G + g];
        *col = make_float2(y[t].x - (c.x * xf.x - c.y * xf.y) - (c.z * xl.x - c.w * xl.y),
                           y[t].y - (c.x * xf.y + c.y * xf.x) - (c.z * xl.y + c.w * xl.x));
      }
      col += TH;
      asm volatile("" : "+v"(col));
    }
  }
}

// The same z pass when every tile has its own PSF (tiles_per_block = 1: a grid
// of PSFs): nothing to share, so no LDS staging; a block takes 256 columns of
// one tile and every thread reads its column's operator straight from memory
// (coalesced across the block), so all 256 threads have work.  The operator
// reads bound this pass, so for the 40-plane tile the end-plane corrections are
// applied in the frequency domain (zf_middle_ends) and only the A and G planes,
// two thirds of the operator, are read.
template <int SZ>
__global__ __launch_bounds__(256, SZ ? 2 : 1) void k_tile_zfft_own(float2* S, long long s_tstride, int TH, int sz_rt,
                                                       const float2* __restrict__ Ops, long long ops_stride,
                                                       const mm_tile_dev* __restrict__ tiles, int ntiles) {
  const int sz = SZ ? SZ : sz_rt;
  const int ncb = (TH + 255) / 256;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = bid % ncb, j = bid / ncb;
  const int c = cb * 256 + threadIdx.x;
  if (j >= ntiles || c >= TH) return;
  const float2* Op = Ops + (long long)tiles[j].psf * ops_stride + c;
  float2* col = S + (long long)j * s_tstride + c;
  float2 x[40];
  {
    const float2* p = col;
#pragma unroll
    for (int s = 0; s < 40; ++s) {
      x[s] = s < sz ? *p : make_float2(0.f, 0.f);
      p += TH;
      asm volatile("" : "+v"(p));
    }
  }
  const float2 xf = x[0];
  float2 xl = x[SZ ? SZ - 1 : 0];  // x[sz - 1]
  if (!SZ) {
#pragma unroll
    for (int s = 1; s < 40; ++s) xl = s == sz - 1 ? x[s] : xl;
  }
  float2 U[5][16];
  zf_fwd_stage1(x, U);
  auto AGk = [&](int k) {
    const float2 a = Op[(long long)k * TH], b = Op[(long long)(80 + k) * TH];
    return make_float4(a.x, a.y, b.x, b.y);
  };
  if (SZ) {  // the end-plane corrections in the frequency domain: the operator's 160 (A, G) planes only
    zf_middle_ends<0, SZ - 1>(U, AGk, xf, xl);
    zf_middle_ends<1, SZ - 1>(U, AGk, xf, xl);
    zf_middle_ends<2, SZ - 1>(U, AGk, xf, xl);
  } else {
    zf_middle<0>(U, AGk);
    zf_middle<1>(U, AGk);
    zf_middle<2>(U, AGk);
  }
  float2 y[40];
  zf_inv_stage3(U, y);
  asm volatile("" : "+v"(col));
#pragma unroll
  for (int t = 0; t < 40; ++t) {
    if (t < sz) {
      if (SZ) {
        *col = y[t];
      } else {
        const float2 a = Op[(long long)(160 + t) * TH], b = Op[(long long)(200 + t) * TH];  // c0[t], c1[t]
        *col = make_float2(y[t].x - (a.x * xf.x - a.y * xf.y) - (b.x * xl.x - b.y * xl.y),
                           y[t].y - (a.x * xf.y + a.y * xf.x) - (b.x * xl.y + b.y * xl.x));
      }
    }
    col += TH;
    asm volatile("" : "+v"(col));
  }
}

static int launch_zfft(mm_deconv* d, float2* S, long long s_ts, const SpecLayout& Ls, const float2* Ops,
                       const mm_tile_dev* td, int ntiles, int tpb, hipStream_t st) {
  const int TH = (int)Ls.plane;
  if (tpb == 1) {  // one PSF per tile
    const long long nblk = (long long)((TH + 255) / 256) * ntiles;
    if (nblk > 0x7fffffffLL) { mm_set_error("z pass: too many blocks"); return MM_EINVAL; }
    if (d->sz == 40)
      hipLaunchKernelGGL(k_tile_zfft_own<40>, dim3((unsigned)nblk), dim3(256), 0, st, S, s_ts, TH, d->sz, Ops,
                         (long long)ZF_OPS * TH, td, ntiles);
    else
      hipLaunchKernelGGL(k_tile_zfft_own<0>, dim3((unsigned)nblk), dim3(256), 0, st, S, s_ts, TH, d->sz, Ops,
                         (long long)ZF_OPS * TH, td, ntiles);
    MM_LAUNCH_CHECK();
    return MM_OK;
  }
  const int ngroups = (ntiles + tpb - 1) / tpb;
  const long long nblk = (long long)((TH + ZF_G - 1) / ZF_G) * ngroups;
  if (nblk > 0x7fffffffLL) { mm_set_error("z pass: too many blocks"); return MM_EINVAL; }
  const size_t lds = sizeof(float4) * 120 * ZF_G;
  if (d->sz == 40)
    hipLaunchKernelGGL(k_tile_zfft<40>, dim3((unsigned)nblk), dim3(256), lds, st, S, s_ts, TH, d->sz, Ops,
                       (long long)ZF_OPS * TH, td, ntiles, tpb);
  else
    hipLaunchKernelGGL(k_tile_zfft<0>, dim3((unsigned)nblk), dim3(256), lds, st, S, s_ts, TH, d->sz, Ops,
                       (long long)ZF_OPS * TH, td, ntiles, tpb);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

// Operator planes of the FFT-80 z pass from the z-domain kernels w (Zp planes
// of T x H natural-layout columns, float64), written in layout Ls.  A block
// takes ZF_G natural columns, stages their w in LDS and forms all ZF_OPS
// planes of them.  The padding columns of a blocked layout are not written.
__global__ __launch_bounds__(256) void k_zfft_ops(const double2* __restrict__ w, int Zp, int sz, SpecLayout Ls,
                                                  float2* __restrict__ out) {
  extern __shared__ __align__(16) double2 zw[];  // [Zp][ZF_G]
  __shared__ double2 tw[80];                      // exp(-2 pi i m / 80)
  for (int m = threadIdx.x; m < 80; m += blockDim.x) {
    double sn, cs;
    sincospi(-2.0 * m / 80.0, &sn, &cs);
    tw[m] = make_double2(cs, sn);
  }
  const long long ncol = (long long)Ls.T * Ls.H;
  const long long c0 = (long long)blockIdx.x * ZF_G;
  const int gmax = (int)min((long long)ZF_G, ncol - c0);
  for (int i = threadIdx.x; i < Zp * ZF_G; i += blockDim.x) {
    const int p = i / ZF_G, g = i - (i / ZF_G) * ZF_G;
    zw[i] = g < gmax ? w[(long long)p * ncol + c0 + g] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ZF_OPS * ZF_G; i += blockDim.x) {
    const int q = i / ZF_G, g = i - (i / ZF_G) * ZF_G;
    if (g >= gmax) continue;
    auto W = [&](int m) { m %= Zp; if (m < 0) m += Zp; return zw[m * ZF_G + g]; };
    double re = 0, im = 0;
    if (q < 160) {  // A[k] or G[k]
      const int k = q % 80;
      const bool gpart = q >= 80;
      const int m0 = gpart ? 0 : -(sz - 1), m1 = gpart ? 2 * sz - 2 : sz - 1;
      for (int m = m0; m <= m1; ++m) {
        double2 v = W(m);
        if (gpart) {
          const double2 u = W(m - 2 * (sz - 1));
          v.x += u.x;
          v.y += u.y;
        }
        const double2 e = tw[(((m * k) % 80) + 80) % 80];
        re += v.x * e.x - v.y * e.y;
        im += v.x * e.y + v.y * e.x;
      }
      re /= 80.0;
      im /= 80.0;
    } else {  // c0[t] = w[t], c1[t] = w[t - sz + 1]
      const int t = q < 200 ? q - 160 : q - 200;
      const double2 v = t < sz ? W(q < 200 ? t : t - sz + 1) : make_double2(0.0, 0.0);
      re = v.x;
      im = v.y;
    }
    const long long c = c0 + g;
    const int x = (int)(c / Ls.H), ky = (int)(c - (long long)x * Ls.H);
    const long long o = Ls.blk ? (long long)(ky / Ls.blk) * Ls.T * Ls.blk + (long long)x * Ls.blk + ky % Ls.blk : c;
    out[(long long)q * Ls.plane + o] = make_float2((float)re, (float)im);
  }
}

// S and W in the same in-plane layout Ls (planes of TH columns).
static int launch_zconv(mm_deconv* d, float2* S, long long s_ts, const SpecLayout& Ls, const float2* Wz,
                        const mm_tile_dev* td, int ntiles, int tpb, hipStream_t st) {
  const int sz = d->sz, Zp = d->Zp, TH = (int)Ls.plane;
  const bool exact = sz == 40 || sz == 35 || sz == 20;
  const int G = exact ? 640 / sz : d->zG;
  const int ngroups = (ntiles + tpb - 1) / tpb;
  const long long nblk = (long long)((TH + G - 1) / G) * ngroups;
  const int threads = ((G * sz + 63) / 64) * 64;
  ZcSrc src;
  for (int i = 0; i < 4 * 48; ++i) src.p[i] = i < 4 * sz ? d->zsrc_h[i] : -1;
  const int ch = tpb >= ZC_CH ? ZC_CH : 1;  // one PSF per tile: no chunk of tiles shares K
  const size_t lds = sizeof(float2) * ((size_t)G * Zp + (size_t)ch * G * (sz + 1));
  if (lds > 163840) {
    mm_set_error("z pass: %zu bytes of LDS for %d planes, Zp %d", lds, sz, Zp);
    return MM_EINVAL;
  }
  const long long wstride = (long long)(d->zdual ? 2 : 1) * Zp * TH;  // planes of one PSF's operator
#define MM_ZCX(M, E)                                                                                            \
  do {                                                                                                          \
    if (ch == ZC_CH)                                                                                            \
      hipLaunchKernelGGL((k_tile_zconv<M, E, ZC_CH>), dim3((unsigned)nblk), dim3(threads), lds, st, S, s_ts, TH, \
                         sz, d->zpad, Zp, src, Wz, wstride, td, ntiles, tpb, G);                       \
    else                                                                                                        \
      hipLaunchKernelGGL((k_tile_zconv<M, E, 1>), dim3((unsigned)nblk), dim3(threads), lds, st, S, s_ts, TH, sz, \
                         d->zpad, Zp, src, Wz, wstride, td, ntiles, tpb, G);                           \
  } while (0)
#define MM_ZC(M) MM_ZCX(M, false)
  if (sz == 40) MM_ZCX(40, true);
  else if (sz == 35) MM_ZCX(35, true);
  else if (sz == 20) MM_ZCX(20, true);
  else if (sz <= 8) MM_ZC(8);
  else if (sz <= 16) MM_ZC(16);
  else if (sz <= 24) MM_ZC(24);
  else if (sz <= 32) MM_ZC(32);
  else if (sz <= 40) MM_ZC(40);
  else MM_ZC(48);
#undef MM_ZC
#undef MM_ZCX
  MM_LAUNCH_CHECK();
  return MM_OK;
}

// ------------------------------------------------------------------- host
static void host_zmap(int sz, int zpad, std::vector<int>& m) {
  const int Zp = sz + 2 * zpad;
  m.assign(Zp, -1);
  for (int i = 0; i < sz; ++i) m[zpad + i] = i;
  for (int j = 0; j < zpad; ++j) {  // 'in' about zpad while the upper pad is still zero
    const int s = 2 * zpad - j;
    if (s < Zp) m[j] = (s < zpad + sz) ? m[s] : -1;
  }
  const int top = zpad + sz - 1;  // 'out' about zpad+sz-1 reads the filled lower pad
  for (int j = top + 1; j < Zp; ++j) {
    const int s = 2 * top - j;
    if (s >= 0) m[j] = m[s];
  }
}

// Pass forms of the schedule (xy_form, z_form) -> the flags the launchers read.
// The Rader passes keep the spectra column-blocked by their x-pass group width,
// so that every x transform group is one contiguous run.
static void apply_schedule(mm_deconv* d) {
  d->rd = d->xy_form == MM_XY_RADER;
  d->pfM = d->xy_form == MM_XY_STOCKHAM ? 0 : d->pfM0;
  d->pfG = d->rd ? d->rdG : d->pfG0;
  d->zconv = d->z_form != MM_Z_FFT;
  d->zfft = d->z_form == MM_Z_DFT80;
  d->zdual = d->z_form == MM_Z_MATRIX;
  d->blk = d->rd ? d->pfG : 0;
  d->Hs = d->blk ? (d->H + d->blk - 1) / d->blk * d->blk : d->H;
}

extern "C" int mm_deconv_create(int sz, int tile_size, int overlap, int zpad, mm_deconv** out) {
  if (!out || sz < 1 || tile_size < 1 || overlap < 0 || zpad < 0) {
    mm_set_error("mm_deconv_create: need sz, tile_size >= 1 and overlap, zpad >= 0 (got %d %d %d %d)", sz,
                 tile_size, overlap, zpad);
    return MM_EINVAL;
  }
  mm_deconv* d = new mm_deconv();
  d->sz = sz; d->ts = tile_size; d->ov = overlap; d->zpad = zpad;
  d->T = tile_size + 2 * overlap;
  d->Zp = sz + 2 * zpad;
  d->H = d->T / 2 + 1;
  if (d->T < 3 || d->Zp < 3) {
    mm_set_error("mm_deconv_create: padded tile %dx%dx%d too small (the Laplacian needs 3 per axis)", d->Zp, d->T, d->T);
    delete d;
    return MM_EINVAL;
  }
  int rc;
  if ((rc = mm_get_plan_f32(d->T, &d->pT)) || (rc = mm_get_plan_f32(d->Zp, &d->pZ)) ||
      (rc = mm_get_plan_f64(d->T, &d->pT64)) || (rc = mm_get_plan_f64(d->Zp, &d->pZ64))) {
    delete d;
    return rc;
  }
  std::vector<int> zm;
  host_zmap(sz, zpad, zm);
  if (hipMalloc(&d->zmap_dev, sizeof(int) * d->Zp) != hipSuccess ||
      hipMemcpy(d->zmap_dev, zm.data(), sizeof(int) * d->Zp, hipMemcpyHostToDevice) != hipSuccess) {
    mm_set_error("mm_deconv_create: cannot upload the z map");
    delete d;
    return MM_ENOMEM;
  }
  // two blocks per CU (LDS <= 80 KB) for the T-length passes when the radix tables allow it
  const int budgetT = fft_lds_bytes(d->pT, 2, (size_t)sizeof(int) * d->T) <= 81920 ? 81920 : MM_LDS_MAX;
  d->gT_r2c = mm_fft_pick_g(d->pT, budgetT - (int)sizeof(int) * d->T, 0, 8);
  d->gT_x = mm_fft_pick_g(d->pT, budgetT, 0, 8);
  d->gZ = mm_fft_pick_g(d->pZ, 65536, 8 * d->Zp, 16);
  d->gT64_r2c = mm_fft_pick_g(d->pT64, 98304, 0, 8);
  d->gT64_x = mm_fft_pick_g(d->pT64, 98304, 0, 8);
  d->gZ64 = mm_fft_pick_g(d->pZ64, 98304, 0, 16);
  // prime-first passes when T = p * M with the matrix-product prime p
  d->pfM = 0;
  if (d->pT.coefR > 0 && d->T % d->pT.coefR == 0) {
    const int M = d->T / d->pT.coefR;
    if (M == 2 || M == 3 || M == 4 || M == 6) {
      const PfShape S(d->T, d->pT.coefR);
      int G = 256 / (M * S.nkt);
      while (G > 0 && (pf_lds_bytes(S, M, G, pf_r2c_extra(d->T, G)) > MM_LDS_MAX || G * M * S.h > PF_NI * 256)) --G;
      if (G > 0) { d->pfM = M; d->pfG = G; d->pfS = S; }
    }
  }
  d->pfM0 = d->pfM;
  d->pfG0 = d->pfG;
  // Rader form of the prime 113 (the full-FOV tile T = 678 = 113 * 6)
  d->rd = 0;
  d->rdP = 0;
  if (d->pfM && (d->pfS.p == 113 || (d->pfS.p == 239 && d->pfM == 2))) {
    const int M = d->pfM, p = d->pfS.p;
    const int B = p == 113 ? RdSpec<113>::B : RdSpec<239>::B;
    int G = 256 / (B * M);  // G M B <= 256: the in-place stages take one item per thread
    auto lds = [&](int g) { return p == 113 ? rd_lds_bytes<113>(d->T, M, g, 0) : rd_lds_bytes<239>(d->T, M, g, 0); };
    while (G > 0 && (lds(G) > 81920 || G * M * p > RD_NI * 256)) --G;
    if (G > 0 && (rc = mm_get_rader(p, &d->rdF, &d->rdI)) == MM_OK) {
      d->rd = 1;
      d->rdP = p;
      d->rdG = G;
    } else if (rc) {
      delete d;
      return rc;
    }
  }
  // z-kernel pass: the padded planes that copy each data plane
  d->zconv = sz <= 48;
  if (d->zconv) {
    std::vector<int> src(4 * sz, -1), cnt(sz, 0);
    for (int p = 0; p < d->Zp; ++p) {
      const int s = zm[p];
      if (s < 0) continue;
      if (cnt[s] >= 4) { d->zconv = 0; break; }
      src[4 * s + cnt[s]++] = p;
    }
    d->zG = 640 / sz;
    if (d->zG > 64) d->zG = 64;
    // LDS of a block: w of its columns and a chunk of ZC_CH tiles' columns
    if (sizeof(float2) * ((size_t)d->zG * d->Zp + (size_t)ZC_CH * d->zG * (sz + 1)) > 163840) d->zconv = 0;
    for (int i = 0; i < 4 * sz; ++i) d->zsrc_h[i] = src[i];
    if (d->zconv && (hipMalloc(&d->zsrc_dev, sizeof(int) * 4 * sz) != hipSuccess ||
                     hipMemcpy(d->zsrc_dev, src.data(), sizeof(int) * 4 * sz, hipMemcpyHostToDevice) != hipSuccess)) {
      mm_set_error("mm_deconv_create: cannot upload the z source table");
      delete d;
      return MM_ENOMEM;
    }
  }
  d->zfft = d->zconv && zpad == sz - 1 && sz <= 40;
  d->xy_ok = (1 << MM_XY_STOCKHAM) | (d->pfM0 ? 1 << MM_XY_PRIME_FIRST : 0) | (d->rd ? 1 << MM_XY_RADER : 0);
  d->z_ok = (1 << MM_Z_FFT) | (d->zconv ? 1 << MM_Z_MATRIX : 0) | (d->zfft ? 1 << MM_Z_DFT80 : 0);
  // default schedule: the fastest form of each pass the geometry admits (the z FFT form of the
  // DAPI tile, sz 40 / zpad 19, runs its length-78 transforms in registers and beats the matrix)
  d->xy_form = d->rd ? MM_XY_RADER : (d->pfM0 ? MM_XY_PRIME_FIRST : MM_XY_STOCKHAM);
  d->z_form = d->zfft ? MM_Z_DFT80 : (sz == 40 && zpad == 19) ? MM_Z_FFT : (d->zconv ? MM_Z_MATRIX : MM_Z_FFT);
  apply_schedule(d);
  *out = d;
  return MM_OK;
}

extern "C" int mm_deconv_schedule(const mm_deconv* d, int* xy, int* z, int* xy_mask, int* z_mask) {
  if (!d) { mm_set_error("mm_deconv_schedule: null plan"); return MM_EINVAL; }
  if (xy) *xy = d->xy_form;
  if (z) *z = d->z_form;
  if (xy_mask) *xy_mask = d->xy_ok;
  if (z_mask) *z_mask = d->z_ok;
  return MM_OK;
}

extern "C" int mm_deconv_set_schedule(mm_deconv* d, int xy, int z) {
  if (!d) { mm_set_error("mm_deconv_set_schedule: null plan"); return MM_EINVAL; }
  if (xy < 0 || xy > 2 || !((d->xy_ok >> xy) & 1) || z < 0 || z > 2 || !((d->z_ok >> z) & 1)) {
    mm_set_error("mm_deconv_set_schedule: forms (%d, %d) not admitted by the tile %dx%dx%d (masks %d, %d)", xy, z,
                 d->Zp, d->T, d->T, d->xy_ok, d->z_ok);
    return MM_EINVAL;
  }
  d->xy_form = xy;
  d->z_form = z;
  apply_schedule(d);
  return MM_OK;
}

extern "C" int mm_deconv_destroy(mm_deconv* d) {
  if (!d) return MM_OK;
  if (d->zmap_dev) (void)hipFree(d->zmap_dev);
  if (d->zsrc_dev) (void)hipFree(d->zsrc_dev);
  if (d->prof)
    for (int i = 0; i < 6; ++i) (void)hipEventDestroy(d->ev[i]);
  delete d;
  return MM_OK;
}

extern "C" int mm_deconv_dims(const mm_deconv* d, int* T, int* Zp, int* H) {
  if (!d) return MM_EINVAL;
  if (T) *T = d->T;
  if (Zp) *Zp = d->Zp;
  if (H) *H = d->H;
  return MM_OK;
}

extern "C" size_t mm_deconv_wiener_ws_bytes(const mm_deconv* d) {
  if (!d) return 0;
  const size_t vol = (size_t)d->Zp * d->T * d->T * sizeof(float);
  const size_t spec = (size_t)d->Zp * d->T * d->H * sizeof(double2);
  return ((vol + 255) / 256) * 256 + spec;
}

template <bool INV, typename C>
static int launch_c2c(C* data, long long estride, long long cstride, int ncols, long long bstride, int nbatch,
                      const FftPlanDev<C>& P, int G, double scale, hipStream_t st) {
  const int ncb = (ncols + G - 1) / G;
  const long long nblk = (long long)ncb * nbatch;
  if (nblk > 0x7fffffffLL) { mm_set_error("c2c grid too large"); return MM_EINVAL; }
  const size_t lds = fft_lds_bytes(P, G, 0);
  hipLaunchKernelGGL((k_fft_c2c<INV, C>), dim3((unsigned)nblk), dim3(256), lds, st, data, estride, cstride, ncols,
                     bstride, cstride == 1 ? 1 : 0, P, G, scale);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

static int build_wiener(mm_deconv* d, const float* psf_centered, double beta, int box_half, void* W_out, void* ws,
                        size_t ws_bytes, void* stream) {
  if (!d || !psf_centered || !W_out || !ws) { mm_set_error("mm_deconv_build_wiener: null argument"); return MM_EINVAL; }
  if (box_half < 0 || 2 * box_half + 1 > d->T) { mm_set_error("mm_deconv_build_wiener_box: box half-width %d", box_half); return MM_EINVAL; }
  if (ws_bytes < mm_deconv_wiener_ws_bytes(d)) { mm_set_error("mm_deconv_build_wiener: workspace too small"); return MM_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  const int T = d->T, Zp = d->Zp, H = d->H;
  const size_t vol = (size_t)Zp * T * T * sizeof(float);
  float* pv = (float*)ws;
  double2* Hs = (double2*)((char*)ws + ((vol + 255) / 256) * 256);
  hipLaunchKernelGGL(k_psf_place, dim3(2048), dim3(256), 0, st, psf_centered, d->sz, d->ts, d->zpad, d->ov, Zp, T, pv);
  MM_LAUNCH_CHECK();
  {
    const int G = d->gT64_r2c;
    const int nrows = Zp * T;
    const size_t lds = fft_lds_bytes(d->pT64, G, 0);
    hipLaunchKernelGGL((k_r2c_rows_plain<double2>), dim3((nrows + 2 * G - 1) / (2 * G)), dim3(256), lds, st, pv,
                       nrows, Hs, d->pT64, G);
    MM_LAUNCH_CHECK();
  }
  int rc;
  if ((rc = launch_c2c<false>(Hs, H, 1, H, (long long)T * H, Zp, d->pT64, d->gT64_x, 1.0, st))) return rc;
  if ((rc = launch_c2c<false>(Hs, (long long)T * H, 1, T * H, 0, 1, d->pZ64, d->gZ64, 1.0, st))) return rc;
  const double scale = 1.0 / ((double)Zp * T * T);
  hipLaunchKernelGGL(k_wiener64, dim3(4096), dim3(256), 0, st, Hs, Zp, T, H, beta, scale, box_half);
  MM_LAUNCH_CHECK();
  if (d->zdual) {  // W itself into planes [Zp, 2 Zp), for the z FFT form of one-PSF-per-tile batches
    const long long wplane = d->blk ? (long long)T * d->Hs : (long long)T * H;
    float2* Wf = (float2*)W_out + (long long)Zp * wplane;
    if (d->blk) {
      const SpecLayout Ls{T, H, d->blk, (long long)T * d->Hs};
      hipLaunchKernelGGL(k_d2f_blocked, dim3(4096), dim3(256), 0, st, Hs, Wf, Zp, Ls);
    } else {
      hipLaunchKernelGGL(k_d2f, dim3(4096), dim3(256), 0, st, Hs, Wf, (long long)Zp * T * H);
    }
    MM_LAUNCH_CHECK();
  }
  if (d->zconv)  // z-domain kernels w = unnormalised inverse z-DFT of W (the 1/Zp is already in W)
    if ((rc = launch_c2c<true>(Hs, (long long)T * H, 1, T * H, 0, 1, d->pZ64, d->gZ64, 1.0, st))) return rc;
  if (d->zfft) {
    const SpecLayout Ls{T, H, d->blk, (long long)T * d->Hs};
    if (d->blk) MM_CHECK(hipMemsetAsync(W_out, 0, sizeof(float2) * ZF_OPS * Ls.plane, st));  // padding columns
    const long long ncol = (long long)T * H;
    hipLaunchKernelGGL(k_zfft_ops, dim3((unsigned)((ncol + ZF_G - 1) / ZF_G)), dim3(256),
                       sizeof(double2) * Zp * ZF_G, st, Hs, Zp, d->sz, Ls, (float2*)W_out);
  } else if (d->blk) {
    const SpecLayout Ls{T, H, d->blk, (long long)T * d->Hs};
    hipLaunchKernelGGL(k_d2f_blocked, dim3(4096), dim3(256), 0, st, Hs, (float2*)W_out, Zp, Ls);
  } else {
    hipLaunchKernelGGL(k_d2f, dim3(4096), dim3(256), 0, st, Hs, (float2*)W_out, (long long)Zp * T * H);
  }
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_deconv_build_wiener(mm_deconv* d, const float* psf_centered, double beta, void* W_out, void* ws,
                                      size_t ws_bytes, void* stream) {
  return build_wiener(d, psf_centered, beta, 0, W_out, ws, ws_bytes, stream);
}

extern "C" int mm_deconv_build_wiener_box(mm_deconv* d, const float* psf_centered, double beta, int box_half,
                                          void* W_out, void* ws, size_t ws_bytes, void* stream) {
  return build_wiener(d, psf_centered, beta, box_half, W_out, ws, ws_bytes, stream);
}

extern "C" int mm_deconv_spec_layout(const mm_deconv* d, int* blk, int* Hs, int* wplanes) {
  if (!d) return MM_EINVAL;
  if (blk) *blk = d->blk;
  if (Hs) *Hs = d->Hs;
  if (wplanes) *wplanes = d->zfft ? ZF_OPS : (d->zdual ? 2 * d->Zp : d->Zp);
  return MM_OK;
}

extern "C" size_t mm_deconv_spec_bytes(const mm_deconv* d) {
  return d ? (size_t)d->sz * d->T * d->Hs * sizeof(float2) : 0;
}

// res rows and columns [lo, hi) of every plane (the Rader C2R computes only
// those; the other forms compute whole planes).
static int deconv_tiles(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W, int n_psf,
                        int tiles_per_block, void* spec_ws, float* res, int lo, int hi, void* stream,
                        unsigned long long* marks = nullptr, float thr = 0.0f, int* marked = nullptr,
                        const C2rStitch* sti = nullptr) {
  if (marked) *marked = 0;
  if (sti && !d->rd) { mm_set_error("mm_deconv_tiles_stitch: the plan's y passes are not the Rader form"); return MM_EINVAL; }
  if (!d || !tiles_dev || !W || !spec_ws || !res || ntiles < 0 || n_psf < 1 || lo < 0 || hi > d->T || lo >= hi) {
    mm_set_error("mm_deconv_tiles: bad argument");
    return MM_EINVAL;
  }
  // The C2R packs rows 2m and 2m + 1 (of z * T + x) into one complex transform,
  // and a row's rounding depends on its partner: the window keeps whole pairs
  // (T even: lo and hi even; T odd: pairs straddle planes, whole planes).
  // (the C2R also assumes a group of rows crosses at most one plane boundary:
  // windows narrower than 64 rows take whole planes)
  if ((d->T & 1) || hi - lo < 64) {
    lo = 0;
    hi = d->T;
  } else {
    lo &= ~1;
    hi = min(d->T, (hi + 1) & ~1);
  }
  if (ntiles == 0) return MM_OK;
  hipStream_t st = (hipStream_t)stream;
  const mm_tile_dev* td = (const mm_tile_dev*)tiles_dev;
  const int T = d->T, Zp = d->Zp, H = d->H, sz = d->sz;
  float2* S = (float2*)spec_ws;
#define MM_MARK(i) \
  if (d->prof) MM_CHECK(hipEventRecord(d->ev[i], st))
  MM_MARK(0);
  const SpecLayout Ls{T, H, d->blk, (long long)T * d->Hs};  // natural unless blk
  const long long s_ts = (long long)sz * Ls.plane;
  const int nrows = sz * T;
  if (d->rd) {  // 1. R2C along y of the padded rows (prime-first form, Rader prime stage)
    RD_SWITCH(d->rdP, d->pfM, (launch_rd_r2c<PP, MM>(td, ntiles, T, nrows, S, s_ts, Ls, d->rdF, d->pT.tw, st)));
    MM_LAUNCH_CHECK();
  } else if (d->pfM) {  // 1. R2C along y of the padded rows (prime-first form)
    const int G = d->pfG;
    const size_t lds = pf_lds_bytes(d->pfS, d->pfM, G, pf_r2c_extra(T, G));
    const long long ngrp = (long long)((nrows + 2 * G - 1) / (2 * G)) * ntiles;
    PF_SWITCH(d->pfM, hipLaunchKernelGGL(k_pf_r2c<MM>, dim3(pf_grid(ngrp, lds)), dim3(256), lds, st, td, ntiles, T,
                                         nrows, S, s_ts, d->pfS, d->pT.tw, d->pT.coef, G));
    MM_LAUNCH_CHECK();
  } else {  // 1. R2C along y of the padded rows
    const int G = d->gT_r2c;
    const size_t lds = fft_lds_bytes(d->pT, G, sizeof(int) * T);
    hipLaunchKernelGGL(k_tile_r2c, dim3((nrows + 2 * G - 1) / (2 * G), ntiles), dim3(256), lds, st, td, T, nrows, S,
                       s_ts, d->pT, G);
    MM_LAUNCH_CHECK();
  }
  MM_MARK(1);
  int rc;
  // 2. C2C along x, all tiles and planes at once
  if (d->rd) rc = launch_rd_c2c<false>(d, S, Ls, ntiles * sz, 0, T, st);
  else if (d->pfM) rc = launch_pf_c2c<false>(d, S, H, 1, H, (long long)T * H, ntiles * sz, st);
  else rc = launch_c2c<false>(S, H, 1, H, (long long)T * H, ntiles * sz, d->pT, d->gT_x, 1.0, st);
  if (rc) return rc;
  MM_MARK(2);
  const int tpb = tiles_per_block < 1 ? 1 : (tiles_per_block > ntiles ? ntiles : tiles_per_block);
  if (d->zfft) {  // 3. z pass by length-80 DFTs of each column
    if ((rc = launch_zfft(d, S, s_ts, Ls, (const float2*)W, td, ntiles, tpb, st))) return rc;
  } else if (d->zconv && !(d->zdual && tpb == 1)) {  // 3. z pass as the folded z-kernel matrix
    if ((rc = launch_zconv(d, S, s_ts, Ls, (const float2*)W, td, ntiles, tpb, st))) return rc;
  } else {  // 3. z pass with z FFTs and the Wiener multiply
    // (with one PSF per tile, no tile shares the z-kernel matrix the folded pass builds: the FFT form
    // reads W, Zp planes per tile, and costs a fraction of the matrix build)
    const int G = d->gZ;
    const int TH = (int)Ls.plane;
    const int ngroups = (ntiles + tpb - 1) / tpb;
    const long long nblk = (long long)((TH + G - 1) / G) * ngroups;
    const size_t lds = fft_lds_bytes(d->pZ, G, (size_t)8 * Zp * G + sizeof(int) * Zp);
    const float2* Wf = (const float2*)W + (d->zdual ? (long long)Zp * TH : 0);
    const long long wstride = (long long)(d->zdual ? 2 : 1) * Zp * TH;  // planes of one PSF's W
    if (sz == 40 && d->zpad == 19) {  // the DAPI tile: the length-78 transforms in registers
      const long long nb = (long long)((TH + 255) / 256) * ntiles;
      if (nb > 0x7fffffffLL) { mm_set_error("z pass: too many blocks"); return MM_EINVAL; }
      hipLaunchKernelGGL((k_tile_zfft78<40, 19>), dim3((unsigned)nb), dim3(256), 0, st, S, s_ts, TH, Wf, wstride, td,
                         ntiles);
    } else {
      // persistent blocks: a few rounds of the CUs' concurrent blocks
      const long long bpc = lds ? (long long)(MM_LDS_MAX / lds) : 8;
      const long long cap = 256 * (bpc < 1 ? 1 : (bpc > 8 ? 8 : bpc)) * 4;
      const long long nlaunch = nblk < cap ? nblk : cap;
      hipLaunchKernelGGL(k_tile_zwiener, dim3((unsigned)nlaunch), dim3(256), lds, st, S, s_ts, TH, sz, d->zpad,
                         d->zmap_dev, Wf, wstride, td, ntiles, tpb, d->pZ, G);
    }
    MM_LAUNCH_CHECK();
  }
  MM_MARK(3);
  // 4. inverse C2C along x (the Rader form stores only the rows [lo, hi) the C2R reads)
  if (d->rd) rc = launch_rd_c2c<true>(d, S, Ls, ntiles * sz, lo, hi, st);
  else if (d->pfM) rc = launch_pf_c2c<true>(d, S, H, 1, H, (long long)T * H, ntiles * sz, st);
  else rc = launch_c2c<true>(S, H, 1, H, (long long)T * H, ntiles * sz, d->pT, d->gT_x, 1.0, st);
  if (rc) return rc;
  MM_MARK(4);
  if (d->rd) {  // 5. C2R along y into res (prime-first form, Rader prime stage)
    RD_SWITCH(d->rdP, d->pfM,
              (launch_rd_c2r<PP, MM>(S, s_ts, Ls, res, (long long)sz * T * T, ntiles, sz, lo, hi, d->rdI, d->pT.tw, st,
                                     marks, thr, sti)));
    MM_LAUNCH_CHECK();
    if (marked && marks) *marked = 1;
  } else if (d->pfM) {  // 5. C2R along y into res (prime-first form)
    const int G = d->pfG;
    const size_t lds = pf_lds_bytes(d->pfS, d->pfM, G, 0);
    const long long ngrp = (long long)((nrows + 2 * G - 1) / (2 * G)) * ntiles;
    PF_SWITCH(d->pfM, hipLaunchKernelGGL(k_pf_c2r<MM>, dim3(pf_grid(ngrp, lds)), dim3(256), lds, st, S, s_ts, res,
                                         (long long)sz * T * T, ntiles, nrows, d->pfS, d->pT.tw, d->pT.coef, G));
    MM_LAUNCH_CHECK();
  } else {  // 5. C2R along y into res
    const int G = d->gT_r2c;
    const size_t lds = fft_lds_bytes(d->pT, G, 0);
    hipLaunchKernelGGL(k_tile_c2r, dim3((nrows + 2 * G - 1) / (2 * G), ntiles), dim3(256), lds, st, S, s_ts, res,
                       (long long)sz * T * T, nrows, d->pT, G);
    MM_LAUNCH_CHECK();
  }
  MM_MARK(5);
#undef MM_MARK
  return MM_OK;
}

extern "C" int mm_deconv_tiles(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W, int n_psf,
                               int tiles_per_block, void* spec_ws, float* res, void* stream) {
  return deconv_tiles(d, tiles_dev, ntiles, W, n_psf, tiles_per_block, spec_ws, res, 0, d ? d->T : 1, stream);
}

extern "C" int mm_deconv_tiles_window(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W,
                                      int n_psf, int tiles_per_block, void* spec_ws, float* res, int lo, int hi,
                                      void* stream) {
  return deconv_tiles(d, tiles_dev, ntiles, W, n_psf, tiles_per_block, spec_ws, res, lo, hi, stream);
}

int deconv_tiles_marked(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W, int n_psf,
                        int tiles_per_block, void* spec_ws, float* res, int lo, int hi, unsigned long long* marks,
                        float thr, int* marked, void* stream) {
  return deconv_tiles(d, tiles_dev, ntiles, W, n_psf, tiles_per_block, spec_ws, res, lo, hi, stream, marks, thr,
                      marked);
}

extern "C" int mm_deconv_set_profiling(mm_deconv* d, int on) {
  if (!d) return MM_EINVAL;
  if (on && !d->prof)
    for (int i = 0; i < 6; ++i) MM_CHECK(hipEventCreate(&d->ev[i]));
  if (!on && d->prof)
    for (int i = 0; i < 6; ++i) (void)hipEventDestroy(d->ev[i]);
  d->prof = on ? 1 : 0;
  return MM_OK;
}

// Durations (ms) of the five passes of the last profiled mm_deconv_tiles call
// (r2c, x forward, z Wiener, x inverse, c2r); waits for that call to finish.
extern "C" int mm_deconv_pass_times(mm_deconv* d, float* ms5) {
  if (!d || !ms5 || !d->prof) { mm_set_error("mm_deconv_pass_times: profiling is off"); return MM_EINVAL; }
  MM_CHECK(hipEventSynchronize(d->ev[5]));
  for (int i = 0; i < 5; ++i) MM_CHECK(hipEventElapsedTime(&ms5[i], d->ev[i], d->ev[i + 1]));
  return MM_OK;
}

// ----------------------------------------------------------- raw / stitch
__global__ void k_tile_raw(const mm_tile_dev* __restrict__ tiles, int T, int sz, float* __restrict__ raw) {
  const int tile = blockIdx.y;
  const mm_tile_dev td = tiles[tile];
  const long long n = (long long)sz * T * T;
  float* o = raw + (long long)tile * n;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(i / ((long long)T * T));
    const int rem = (int)(i - (long long)z * T * T);
    const int x = rem / T, y = rem - (rem / T) * T;
    o[i] = (x < td.xe && y < td.ye) ? tile_raw(td, T, z, x, y) : 0.0f;
  }
}

extern "C" int mm_deconv_tile_raw(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, float* raw, void* stream) {
  if (!d || !tiles_dev || !raw) { mm_set_error("mm_deconv_tile_raw: null argument"); return MM_EINVAL; }
  if (ntiles == 0) return MM_OK;
  hipLaunchKernelGGL(k_tile_raw, dim3(512, ntiles), dim3(256), 0, (hipStream_t)stream, (const mm_tile_dev*)tiles_dev,
                     d->T, d->sz, raw);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

// apply(): output[:, x:x+xe-2ov, y:y+ye-2ov] = tile[:, ov:xe-ov, ov:ye-ov] (deconvolver.py:179-183)
__global__ void k_stitch(const mm_tile_dev* __restrict__ tiles, int T, int sz, int ov, const float* __restrict__ res,
                         float* __restrict__ out, long long o_sz, long long o_sx, int SX, int SY) {
  const int tile = blockIdx.y;
  const mm_tile_dev td = tiles[tile];
  const int xw = td.xe - 2 * ov, yw = td.ye - 2 * ov;
  if (xw <= 0 || yw <= 0) return;
  const long long n = (long long)sz * xw * yw;
  const float* r = res + (long long)tile * sz * T * T;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int z = (int)(i / ((long long)xw * yw));
    const int rem = (int)(i - (long long)z * xw * yw);
    const int x = rem / yw, y = rem - (rem / yw) * yw;
    const int X = td.ox + x, Y = td.oy + y;
    if (X < SX && Y < SY) out[z * o_sz + (long long)X * o_sx + Y] = r[((long long)z * T + ov + x) * T + ov + y];
  }
}

extern "C" int mm_deconv_tiles_stitch(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W,
                                      int n_psf, int tiles_per_block, void* spec_ws, float* out, long long o_sz,
                                      long long o_sx, int SX, int SY, void* stream) {
  if (!d || !out) { mm_set_error("mm_deconv_tiles_stitch: null argument"); return MM_EINVAL; }
  const C2rStitch sti{(const mm_tile_dev*)tiles_dev, out, o_sz, o_sx, SX, SY, d->ov};
  // the stitch reads rows and columns [ov, T - ov) of every tile; res is not written
  return deconv_tiles(d, tiles_dev, ntiles, W, n_psf, tiles_per_block, spec_ws, out, d->ov, d->T - d->ov, stream,
                      nullptr, 0.0f, nullptr, &sti);
}

extern "C" int mm_deconv_stitch(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const float* res, float* out,
                                long long o_sz, long long o_sx, int SX, int SY, void* stream) {
  if (!d || !tiles_dev || !res || !out) { mm_set_error("mm_deconv_stitch: null argument"); return MM_EINVAL; }
  if (ntiles == 0) return MM_OK;
  hipLaunchKernelGGL(k_stitch, dim3(512, ntiles), dim3(256), 0, (hipStream_t)stream, (const mm_tile_dev*)tiles_dev,
                     d->T, d->sz, d->ov, res, out, o_sz, o_sx, SX, SY);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

// ------------------------------------------------- generic 3D FFT (C2C)
// In-place C2C over a contiguous (n0, n1, n2) complex64 volume along the axes
// selected by `axes_mask` (bit i = axis i); inverse transforms scale by 1/n.
// Forward transform of a grid whose values outside [0, m0) x [0, m1) (any index along axis 2) are zero: the
// axis-2 lines outside that box and the axis-1 lines of planes m0 .. n0 - 1 are zero before their pass and stay
// zero, so they are not transformed (a quarter and a half of the lines for a raster zero-padded to twice its size).
extern "C" int mm_fft3d_c2c_box(void* data, int n0, int n1, int n2, int m0, int m1, void* stream) {
  if (!data || n0 < 1 || n1 < 1 || n2 < 1 || m0 < 0 || m0 > n0 || m1 < 0 || m1 > n1) {
    mm_set_error("mm_fft3d_c2c_box: bad argument");
    return MM_EINVAL;
  }
  if (m0 == 0 || m1 == 0) return MM_OK;  // an all-zero grid transforms to zeros
  hipStream_t st = (hipStream_t)stream;
  float2* p = (float2*)data;
  const int n[3] = {n0, n1, n2};
  for (int ax = 2; ax >= 0; --ax) {
    FftPlanDev<float2> P;
    int rc = mm_get_plan_f32(n[ax], &P);
    if (rc) return rc;
    const int G = mm_fft_pick_g(P, 98304, 0, 16);
    if (ax == 2)
      rc = launch_c2c<false>(p, 1, n2, m1, (long long)n1 * n2, m0, P, G, 1.0, st);
    else if (ax == 1)
      rc = launch_c2c<false>(p, n2, 1, n2, (long long)n1 * n2, m0, P, G, 1.0, st);
    else
      rc = launch_c2c<false>(p, (long long)n1 * n2, 1, n1 * n2, 0, 1, P, G, 1.0, st);
    if (rc) return rc;
  }
  return MM_OK;
}

extern "C" int mm_fft3d_c2c(void* data, int n0, int n1, int n2, int axes_mask, int inverse, void* stream) {
  if (!data || n0 < 1 || n1 < 1 || n2 < 1) { mm_set_error("mm_fft3d_c2c: bad argument"); return MM_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  float2* p = (float2*)data;
  const int n[3] = {n0, n1, n2};
  for (int ax = 2; ax >= 0; --ax) {
    if (!(axes_mask & (1 << ax))) continue;
    FftPlanDev<float2> P;
    int rc = mm_get_plan_f32(n[ax], &P);
    if (rc) return rc;
    const double sc = inverse ? 1.0 / n[ax] : 1.0;
    if (ax == 2) {
      const int G = mm_fft_pick_g(P, 98304, 0, 16);
      rc = inverse ? launch_c2c<true>(p, 1, n2, n0 * n1, 0, 1, P, G, sc, st)
                   : launch_c2c<false>(p, 1, n2, n0 * n1, 0, 1, P, G, sc, st);
    } else if (ax == 1) {
      const int G = mm_fft_pick_g(P, 98304, 0, 16);
      rc = inverse ? launch_c2c<true>(p, n2, 1, n2, (long long)n1 * n2, n0, P, G, sc, st)
                   : launch_c2c<false>(p, n2, 1, n2, (long long)n1 * n2, n0, P, G, sc, st);
    } else {
      const int G = mm_fft_pick_g(P, 98304, 0, 16);
      rc = inverse ? launch_c2c<true>(p, (long long)n1 * n2, 1, n1 * n2, 0, 1, P, G, sc, st)
                   : launch_c2c<false>(p, (long long)n1 * n2, 1, n1 * n2, 0, 1, P, G, sc, st);
    }
    if (rc) return rc;
  }
  return MM_OK;
}
