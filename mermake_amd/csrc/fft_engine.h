// LDS-resident mixed-radix Stockham FFT engine for gfx950.
//
// A workgroup holds G independent transforms of length N in LDS (transform g at
// offset g*N) and runs every radix pass in LDS, ping-ponging between two
// buffers.  Radices 2, 3, 4, 5, 7, 8 are register butterflies; any other
// (odd prime) factor p runs as a symmetric direct DFT: the twiddled inputs are
// combined into sums/differences of mirrored pairs, so each output pair
// (k, p-k) costs (p-1)/2 real-by-complex multiply-adds per component instead of
// p complex ones.  Twiddles come from an accurate table exp(-2 pi i m / N)
// computed on the host in float64 (no device trigonometry).
//
// The element type is templated: float2 for the tile transforms, double2 for
// the one-off Wiener spectrum construction.
#pragma once
#include "common.h"

#define MM_FFT_MAXST 24

template <typename R>
struct cplx;
template <>
struct cplx<float> { typedef float2 T; };
template <>
struct cplx<double> { typedef double2 T; };

template <typename C>
struct FftPlanDev {
  int N;
  int nst;
  int radix[MM_FFT_MAXST];
  const C* tw;        // forward twiddles, N entries
  const float* coef;  // cos / sin matrices of the prime radix `coefR` (float32 plans), or nullptr
  int coefR;          // prime handled as a register-tiled matrix product (0: none)
  int coef_len;       // floats in `coef`: 2 * h * hp, h = (coefR-1)/2, hp = h rounded up to 8
};

template <typename C> __device__ __forceinline__ C mk(double x, double y);
template <> __device__ __forceinline__ float2 mk<float2>(double x, double y) { return make_float2((float)x, (float)y); }
template <> __device__ __forceinline__ double2 mk<double2>(double x, double y) { return make_double2(x, y); }

template <typename C> __device__ __forceinline__ C vadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename C> __device__ __forceinline__ C vsub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename C> __device__ __forceinline__ C vmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C, typename S> __device__ __forceinline__ C vscale(C a, S s) { C r; r.x = a.x * s; r.y = a.y * s; return r; }
template <typename C> __device__ __forceinline__ C vconj(C a) { C r; r.x = a.x; r.y = -a.y; return r; }
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename C> __device__ __forceinline__ C vmi(C a) {
  C r;
  if (INV) { r.x = -a.y; r.y = a.x; } else { r.x = a.y; r.y = -a.x; }
  return r;
}
template <bool INV, typename C> __device__ __forceinline__ C twl(const C* tw, int i) {
  C w = tw[i];
  if (INV) w.y = -w.y;
  return w;
}

// ---------------------------------------------------------------- index math
// Division by a block-uniform divisor d through a float reciprocal and one
// correction step: exact for 0 <= n < 2^22 (the estimate is off by at most 1).
struct FastDiv {
  int d;
  float inv;
  __device__ __forceinline__ explicit FastDiv(int dd) : d(dd), inv(1.0f / (float)dd) {}
  // q * d by the full-rate 24-bit multiply (both factors and the product are below 2^23)
  __device__ __forceinline__ int div(int n) const {
    int q = (int)((float)n * inv);
    const int r = n - __mul24(q, d);
    if (r >= d) ++q;
    else if (r < 0) --q;
    return q;
  }
  __device__ __forceinline__ void divmod(int n, int& q, int& r) const {
    q = (int)((float)n * inv);
    r = n - __mul24(q, d);
    if (r >= d) { ++q; r -= d; }
    else if (r < 0) { --q; r += d; }
  }
};

// ---------------------------------------------------------------- butterflies
template <bool INV, typename C>
__device__ __forceinline__ void bfly2(C* v) {
  C t = v[0];
  v[0] = vadd(t, v[1]);
  v[1] = vsub(t, v[1]);
}

template <bool INV, typename C>
__device__ __forceinline__ void bfly4(C* v) {
  C s02 = vadd(v[0], v[2]), d02 = vsub(v[0], v[2]);
  C s13 = vadd(v[1], v[3]), d13 = vmi<INV>(vsub(v[1], v[3]));
  v[0] = vadd(s02, s13);
  v[2] = vsub(s02, s13);
  v[1] = vadd(d02, d13);
  v[3] = vsub(d02, d13);
}

template <bool INV, typename C>
__device__ __forceinline__ void bfly8(C* v) {
  typedef decltype(v[0].x) S;
  C e[4] = {v[0], v[2], v[4], v[6]};
  C o[4] = {v[1], v[3], v[5], v[7]};
  bfly4<INV>(e);
  bfly4<INV>(o);
  const S h = (S)0.7071067811865476;
  // w8^1, w8^2, w8^3 (forward: exp(-2 pi i k / 8))
  C o1; o1.x = (o[1].x + (INV ? -o[1].y : o[1].y)) * h; o1.y = (o[1].y + (INV ? o[1].x : -o[1].x)) * h;
  C o2 = vmi<INV>(o[2]);
  C o3; o3.x = (-o[3].x + (INV ? -o[3].y : o[3].y)) * h; o3.y = (-o[3].y + (INV ? o[3].x : -o[3].x)) * h;
  v[0] = vadd(e[0], o[0]); v[4] = vsub(e[0], o[0]);
  v[1] = vadd(e[1], o1);   v[5] = vsub(e[1], o1);
  v[2] = vadd(e[2], o2);   v[6] = vsub(e[2], o2);
  v[3] = vadd(e[3], o3);   v[7] = vsub(e[3], o3);
}

template <bool INV, typename C>
__device__ __forceinline__ void bfly3(C* v) {
  typedef decltype(v[0].x) S;
  const S c = (S)0.8660254037844386;
  C s = vadd(v[1], v[2]);
  C d = vsub(v[1], v[2]);
  C x0 = vadd(v[0], s);
  C m; m.x = v[0].x - (S)0.5 * s.x; m.y = v[0].y - (S)0.5 * s.y;
  C id = vmi<INV>(vscale(d, c));
  v[0] = x0;
  v[1] = vadd(m, id);
  v[2] = vsub(m, id);
}

// symmetric odd DFT with literal cos/sin tables (R = 5, 7)
template <int R, bool INV, typename C>
__device__ __forceinline__ void bfly_odd(C* v, const double* cs, const double* sn) {
  typedef decltype(v[0].x) S;
  constexpr int H = (R - 1) / 2;
  C s[H], d[H];
#pragma unroll
  for (int r = 1; r <= H; ++r) { s[r - 1] = vadd(v[r], v[R - r]); d[r - 1] = vsub(v[r], v[R - r]); }
  C x0 = v[0];
#pragma unroll
  for (int r = 0; r < H; ++r) x0 = vadd(x0, s[r]);
  C out[R];
  out[0] = x0;
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    C A = v[0], B; B.x = 0; B.y = 0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int m = (r * k) % R;  // m in [1, R-1]; cos symmetric, sin antisymmetric
      const S c = (S)(m <= H ? cs[m - 1] : cs[R - m - 1]);
      const S q = (S)(m <= H ? sn[m - 1] : -sn[R - m - 1]);
      A.x += s[r - 1].x * c; A.y += s[r - 1].y * c;
      B.x += d[r - 1].x * q; B.y += d[r - 1].y * q;
    }
    C iB = vmi<INV>(B);
    out[k] = vadd(A, iB);
    out[R - k] = vsub(A, iB);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = out[r];
}

__device__ __constant__ static const double kCos5[2] = {0.30901699437494745, -0.8090169943749473};
__device__ __constant__ static const double kSin5[2] = {0.9510565162951535, 0.5877852522924732};
__device__ __constant__ static const double kCos7[3] = {0.6234898018587336, -0.22252093395631434, -0.900968867902419};
__device__ __constant__ static const double kSin7[3] = {0.7818314824680298, 0.9749279121818236, 0.43388373911755823};

template <int R, bool INV, typename C>
__device__ __forceinline__ void bfly(C* v) {
  if constexpr (R == 2) bfly2<INV>(v);
  else if constexpr (R == 3) bfly3<INV>(v);
  else if constexpr (R == 4) bfly4<INV>(v);
  else if constexpr (R == 5) bfly_odd<5, INV>(v, kCos5, kSin5);
  else if constexpr (R == 7) bfly_odd<7, INV>(v, kCos7, kSin7);
  else if constexpr (R == 8) bfly8<INV>(v);
}

// ---------------------------------------------------------------- passes
// Stockham step: butterfly j (of M = N/R) reads in[j + r*M], twiddles it by
// w_N^{(j mod Ns) * r * N/(Ns*R)}, and writes out[(j - j mod Ns)*R + j mod Ns + r*Ns].
template <int R, bool INV, typename C>
__device__ __forceinline__ void stage_small(const C* __restrict__ src, C* __restrict__ dst, const C* __restrict__ tw,
                                            int N, int M, int Ns, int tws, int G) {
  const int total = G * M;
  const FastDiv fM(M), fNs(Ns);
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    int g, j, q, k;
    fM.divmod(t, g, j);
    fNs.divmod(j, q, k);
    const C* in = src + g * N;
    C* out = dst + g * N;
    C v[R];
    v[0] = in[j];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = vmul(in[j + r * M], twl<INV>(tw, k * r * tws));
    bfly<R, INV>(v);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[base + r * Ns] = v[r];
  }
}

template <bool INV, typename C>
__device__ __forceinline__ void stage_prime(C* __restrict__ src, C* __restrict__ dst, const C* __restrict__ tw,
                                            int N, int R, int M, int Ns, int tws, int G) {
  typedef decltype(src[0].x) S;
  const int total = G * N;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {  // twiddle in place
    const int g = t / N;
    const int pos = t - g * N;
    const int r = pos / M;
    if (r > 0) {
      const int j = pos - r * M;
      src[t] = vmul(src[t], twl<INV>(tw, (j % Ns) * r * tws));
    }
  }
  __syncthreads();
  const int H = (R + 1) / 2;
  const int total2 = G * M * H;
  for (int t = threadIdx.x; t < total2; t += blockDim.x) {
    const int kk = t % H;
    const int rest = t / H;
    const int j = rest % M;
    const int g = rest / M;
    const C* in = src + g * N;
    C* out = dst + g * N;
    const int k = j % Ns;
    const int base = (j - k) * R + k;
    const C v0 = in[j];
    if (kk == 0) {
      C acc = v0;
      for (int r = 1; r < R; ++r) acc = vadd(acc, in[j + r * M]);
      out[base] = acc;
    } else {
      C A = v0, B;
      B.x = 0; B.y = 0;
      int m = 0;
      for (int r = 1; r < H; ++r) {
        m += kk;
        if (m >= R) m -= R;
        const C w = tw[m * M];  // exp(-2 pi i m / R) = (cos, -sin)
        const S c = w.x, q = -w.y;
        const C a1 = in[j + r * M], a2 = in[j + (R - r) * M];
        A.x += (a1.x + a2.x) * c;
        A.y += (a1.y + a2.y) * c;
        B.x += (a1.x - a2.x) * q;
        B.y += (a1.y - a2.y) * q;
      }
      const C iB = vmi<INV>(B);
      out[base + kk * Ns] = vadd(A, iB);
      out[base + (R - kk) * Ns] = vsub(A, iB);
    }
  }
}

// Odd prime radix R as a matrix product (float32).  Step 1 twiddles the inputs
// and replaces each mirrored pair (r, R-r) in place by its sum s_r and
// difference d_r.  Step 2: for a butterfly and a tile of 8 outputs k,
//   A_k = v0 + sum_r cos(2 pi r k / R) s_r,  B_k = sum_r sin(2 pi r k / R) d_r,
//   X_k = A_k -/+ i B_k,  X_{R-k} = A_k +/- i B_k,
// with the cos/sin matrices read from LDS (rows r, 8 consecutive k per row):
// 32 FMAs per 2 data loads and 4 vector coefficient loads.
template <bool INV>
__device__ __forceinline__ void stage_prime_mm(float2* __restrict__ src, float2* __restrict__ dst,
                                               const float2* __restrict__ tw, const float* __restrict__ coef,
                                               int N, int R, int M, int Ns, int tws, int G) {
  const int h = (R - 1) >> 1;
  const int hp = (h + 7) & ~7;
  const int total = G * M * h;
  const FastDiv fh(h), fM(M), fNs(Ns);
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    int rest, r, g, j, q, k;
    fh.divmod(t, rest, r);
    ++r;
    fM.divmod(rest, g, j);
    fNs.divmod(j, q, k);
    const int i1 = g * N + j + r * M, i2 = g * N + j + (R - r) * M;
    const float2 a = vmul(src[i1], twl<INV>(tw, k * r * tws));
    const float2 b = vmul(src[i2], twl<INV>(tw, k * (R - r) * tws));
    src[i1] = vadd(a, b);
    src[i2] = vsub(a, b);
  }
  __syncthreads();
  const int nkt = (h + 7) >> 3;
  const int total2 = G * M * (nkt + 1);
  const float* cm = coef;
  const float* sm = coef + h * hp;
  const FastDiv fkt(nkt + 1);
  for (int t = threadIdx.x; t < total2; t += blockDim.x) {
    int bfly, kt, g, j, q, kk;
    fkt.divmod(t, bfly, kt);
    fM.divmod(bfly, g, j);
    fNs.divmod(j, q, kk);
    const float2* in = src + g * N + j;
    float2* out = dst + g * N + (j - kk) * R + kk;
    const float2 v0 = in[0];
    if (kt == nkt) {
      float2 acc = v0;
      for (int r = 1; r <= h; ++r) acc = vadd(acc, in[r * M]);
      out[0] = acc;
      continue;
    }
    float2 A[8], B[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) { A[q] = v0; B[q] = make_float2(0.f, 0.f); }
    const int k0 = kt * 8;
    for (int r = 1; r <= h; ++r) {
      const float2 sv = in[r * M];
      const float2 dv = in[(R - r) * M];
      const float4 c0 = *(const float4*)(cm + (r - 1) * hp + k0);
      const float4 c1 = *(const float4*)(cm + (r - 1) * hp + k0 + 4);
      const float4 s0 = *(const float4*)(sm + (r - 1) * hp + k0);
      const float4 s1 = *(const float4*)(sm + (r - 1) * hp + k0 + 4);
      const float c[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
      const float q8[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        A[q].x += sv.x * c[q];
        A[q].y += sv.y * c[q];
        B[q].x += dv.x * q8[q];
        B[q].y += dv.y * q8[q];
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int kq = k0 + 1 + q;
      if (kq <= h) {
        const float2 iB = vmi<INV>(B[q]);
        out[kq * Ns] = vadd(A[q], iB);
        out[(R - kq) * Ns] = vsub(A[q], iB);
      }
    }
  }
}

// Runs all passes on the G transforms in `a` (scratch `b`); returns the buffer
// holding the result.  Must be called by every thread of the block; a
// __syncthreads() must separate the writes to `a` from this call.  `coef` is
// the LDS copy of P.coef (float32 plans with a matrix-product prime radix).
template <bool INV, typename C>
__device__ C* fft_lds(C* a, C* b, const C* __restrict__ tw, const FftPlanDev<C>& P, int G,
                      const float* coef = nullptr) {
  const int N = P.N;
  C* src = a;
  C* dst = b;
  int Ns = 1;
  for (int s = 0; s < P.nst; ++s) {
    const int R = P.radix[s];
    const int M = N / R;
    const int tws = N / (Ns * R);
    switch (R) {
      case 2: stage_small<2, INV>(src, dst, tw, N, M, Ns, tws, G); break;
      case 3: stage_small<3, INV>(src, dst, tw, N, M, Ns, tws, G); break;
      case 4: stage_small<4, INV>(src, dst, tw, N, M, Ns, tws, G); break;
      case 5: stage_small<5, INV>(src, dst, tw, N, M, Ns, tws, G); break;
      case 7: stage_small<7, INV>(src, dst, tw, N, M, Ns, tws, G); break;
      case 8: stage_small<8, INV>(src, dst, tw, N, M, Ns, tws, G); break;
      default:
        if constexpr (sizeof(C) == 8) {
          if (coef != nullptr && R == P.coefR) {
            stage_prime_mm<INV>(src, dst, tw, coef, N, R, M, Ns, tws, G);
            break;
          }
        }
        stage_prime<INV>(src, dst, tw, N, R, M, Ns, tws, G);
        break;
    }
    __syncthreads();
    C* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

template <typename C>
__device__ __forceinline__ void load_twiddles(C* lds_tw, const FftPlanDev<C>& P) {
  for (int i = threadIdx.x; i < P.N; i += blockDim.x) lds_tw[i] = P.tw[i];
}

// LDS layout of the FFT kernels: [twiddles N][coefficients][a: G*N][b: G*N][extra]
template <typename C>
struct FftLds {
  C* tw;
  float* coef;
  C* a;
  C* b;
  unsigned char* extra;
};

__host__ __device__ __forceinline__ size_t mm_align16(size_t x) { return (x + 15) & ~(size_t)15; }

template <typename C>
__host__ __device__ __forceinline__ int fft_coef_floats(const FftPlanDev<C>& P) {
  return (sizeof(C) == 8 && P.coef) ? ((P.coef_len + 3) & ~3) : 0;
}

// Dynamic LDS bytes of fft_lds_carve for G transforms plus `extra` bytes.
template <typename C>
__host__ __device__ __forceinline__ size_t fft_lds_bytes(const FftPlanDev<C>& P, int G, size_t extra) {
  return mm_align16((size_t)P.N * sizeof(C)) + mm_align16((size_t)fft_coef_floats(P) * 4) +
         mm_align16(2 * (size_t)G * P.N * sizeof(C)) + extra;
}

template <typename C>
__device__ __forceinline__ FftLds<C> fft_lds_carve(unsigned char* smem, const FftPlanDev<C>& P, int G) {
  FftLds<C> L;
  const int cl = fft_coef_floats(P);
  size_t off = 0;
  L.tw = (C*)(smem + off);
  off += mm_align16((size_t)P.N * sizeof(C));
  L.coef = (float*)(smem + off);
  off += mm_align16((size_t)cl * 4);
  L.a = (C*)(smem + off);
  L.b = L.a + (size_t)G * P.N;
  off += mm_align16(2 * (size_t)G * P.N * sizeof(C));
  L.extra = smem + off;
  for (int i = threadIdx.x; i < P.N; i += blockDim.x) L.tw[i] = P.tw[i];
  for (int i = threadIdx.x; i < cl; i += blockDim.x) L.coef[i] = i < P.coef_len ? P.coef[i] : 0.0f;
  if (cl == 0) L.coef = nullptr;
  return L;
}
