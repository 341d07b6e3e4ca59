// Length-80 DFTs in registers for the z pass with reflection padding
// zpad = sz - 1 (sz <= 40): one thread transforms one column of one tile.
//
// Forward, x zero from n = 40 on: with n = 16 n1 + n2 and k = k1 + 5 k2,
//   X[k] = sum_{n2} w80^{n2 k1} w16^{n2 k2} sum_{n1} x[n] w5^{n1 k1},
// i.e. 5-point DFTs over n1 (three inputs, two for n2 >= 8), twiddles, and
// 16-point DFTs over n2.  Inverse, outputs t < 40 only: with t = 16 t1 + t2,
// 16-point inverse DFTs over k2, twiddles w80^{-t2 k1}, and 5-point inverse
// DFTs over k1 keeping t1 < 3 (t1 < 2 for t2 >= 8).  Unnormalised both ways.
#pragma once
#include "fft_rader.h"

// w80^m = (cos, sin)(2 pi m / 80) for m < 64, rounded to float (forward twiddles
// are conj); constexpr, so that unrolled uses become literal operands.
static constexpr float kW80[64][2] = {
    {1.0f, 0.0f}, {0.9969173073768616f, 0.07845909893512726f},
    {0.9876883625984192f, 0.15643446147441864f}, {0.972369909286499f, 0.23344536125659943f},
    {0.9510565400123596f, 0.30901700258255005f}, {0.9238795042037964f, 0.3826834261417389f},
    {0.8910065293312073f, 0.45399048924446106f}, {0.8526401519775391f, 0.5224985480308533f},
    {0.80901700258255f, 0.5877852439880371f}, {0.760405957698822f, 0.649448037147522f},
    {0.7071067690849304f, 0.7071067690849304f}, {0.649448037147522f, 0.760405957698822f},
    {0.5877852439880371f, 0.80901700258255f}, {0.5224985480308533f, 0.8526401519775391f},
    {0.45399048924446106f, 0.8910065293312073f}, {0.3826834261417389f, 0.9238795042037964f},
    {0.30901700258255005f, 0.9510565400123596f}, {0.23344536125659943f, 0.972369909286499f},
    {0.15643446147441864f, 0.9876883625984192f}, {0.07845909893512726f, 0.9969173073768616f},
    {6.123234262925839e-17f, 1.0f}, {-0.07845909893512726f, 0.9969173073768616f},
    {-0.15643446147441864f, 0.9876883625984192f}, {-0.23344536125659943f, 0.972369909286499f},
    {-0.30901700258255005f, 0.9510565400123596f}, {-0.3826834261417389f, 0.9238795042037964f},
    {-0.45399048924446106f, 0.8910065293312073f}, {-0.5224985480308533f, 0.8526401519775391f},
    {-0.5877852439880371f, 0.80901700258255f}, {-0.649448037147522f, 0.760405957698822f},
    {-0.7071067690849304f, 0.7071067690849304f}, {-0.760405957698822f, 0.649448037147522f},
    {-0.80901700258255f, 0.5877852439880371f}, {-0.8526401519775391f, 0.5224985480308533f},
    {-0.8910065293312073f, 0.45399048924446106f}, {-0.9238795042037964f, 0.3826834261417389f},
    {-0.9510565400123596f, 0.30901700258255005f}, {-0.972369909286499f, 0.23344536125659943f},
    {-0.9876883625984192f, 0.15643446147441864f}, {-0.9969173073768616f, 0.07845909893512726f},
    {-1.0f, 1.2246468525851679e-16f}, {-0.9969173073768616f, -0.07845909893512726f},
    {-0.9876883625984192f, -0.15643446147441864f}, {-0.972369909286499f, -0.23344536125659943f},
    {-0.9510565400123596f, -0.30901700258255005f}, {-0.9238795042037964f, -0.3826834261417389f},
    {-0.8910065293312073f, -0.45399048924446106f}, {-0.8526401519775391f, -0.5224985480308533f},
    {-0.80901700258255f, -0.5877852439880371f}, {-0.760405957698822f, -0.649448037147522f},
    {-0.7071067690849304f, -0.7071067690849304f}, {-0.649448037147522f, -0.760405957698822f},
    {-0.5877852439880371f, -0.80901700258255f}, {-0.5224985480308533f, -0.8526401519775391f},
    {-0.45399048924446106f, -0.8910065293312073f}, {-0.3826834261417389f, -0.9238795042037964f},
    {-0.30901700258255005f, -0.9510565400123596f}, {-0.23344536125659943f, -0.972369909286499f},
    {-0.15643446147441864f, -0.9876883625984192f}, {-0.07845909893512726f, -0.9969173073768616f},
    {-1.8369701465288538e-16f, -1.0f}, {0.07845909893512726f, -0.9969173073768616f},
    {0.15643446147441864f, -0.9876883625984192f}, {0.23344536125659943f, -0.972369909286499f}};

// a * w80^{-m} (forward) or a * w80^{m} (inverse); m is a compile-time constant after unrolling
template <bool INV>
__device__ __forceinline__ float2 zf_tw(float2 a, int m) {
  if (m == 0) return a;
  const float c = kW80[m][0], s = INV ? kW80[m][1] : -kW80[m][1];
  return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// cos, sin of 2 pi / 5 and 4 pi / 5
#define ZF_C1 0.30901699437494745f
#define ZF_S1 0.9510565162951535f
#define ZF_C2 -0.8090169943749473f
#define ZF_S2 0.5877852522924732f

// 5-point DFT of (a0, a1, a2, 0, 0) (NIN = 3) or (a0, a1, 0, 0, 0) (NIN = 2).
template <bool INV, int NIN>
__device__ __forceinline__ void zf_dft5_in(float2 a0, float2 a1, float2 a2, float2* X) {
  // X_k = P_k -/+ i Q_k, X_{5-k} = P_k +/- i Q_k (forward: -i)
  float2 P1, Q1, P2, Q2;
  if (NIN == 3) {
    P1 = make_float2(a0.x + ZF_C1 * a1.x + ZF_C2 * a2.x, a0.y + ZF_C1 * a1.y + ZF_C2 * a2.y);
    Q1 = make_float2(ZF_S1 * a1.x + ZF_S2 * a2.x, ZF_S1 * a1.y + ZF_S2 * a2.y);
    P2 = make_float2(a0.x + ZF_C2 * a1.x + ZF_C1 * a2.x, a0.y + ZF_C2 * a1.y + ZF_C1 * a2.y);
    Q2 = make_float2(ZF_S2 * a1.x - ZF_S1 * a2.x, ZF_S2 * a1.y - ZF_S1 * a2.y);
    X[0] = make_float2(a0.x + a1.x + a2.x, a0.y + a1.y + a2.y);
  } else {
    P1 = make_float2(a0.x + ZF_C1 * a1.x, a0.y + ZF_C1 * a1.y);
    Q1 = make_float2(ZF_S1 * a1.x, ZF_S1 * a1.y);
    P2 = make_float2(a0.x + ZF_C2 * a1.x, a0.y + ZF_C2 * a1.y);
    Q2 = make_float2(ZF_S2 * a1.x, ZF_S2 * a1.y);
    X[0] = make_float2(a0.x + a1.x, a0.y + a1.y);
  }
  // -i Q = (Q.y, -Q.x); +i Q = (-Q.y, Q.x)
  const float sg = INV ? -1.0f : 1.0f;
  X[1] = make_float2(P1.x + sg * Q1.y, P1.y - sg * Q1.x);
  X[4] = make_float2(P1.x - sg * Q1.y, P1.y + sg * Q1.x);
  X[2] = make_float2(P2.x + sg * Q2.y, P2.y - sg * Q2.x);
  X[3] = make_float2(P2.x - sg * Q2.y, P2.y + sg * Q2.x);
}

// Outputs 0 .. NOUT - 1 (NOUT = 2, 3) of the 5-point DFT of v[0..4].
template <bool INV, int NOUT>
__device__ __forceinline__ void zf_dft5_out(const float2* v, float2* y) {
  const float2 s1 = vadd(v[1], v[4]), d1 = vsub(v[1], v[4]);
  const float2 s2 = vadd(v[2], v[3]), d2 = vsub(v[2], v[3]);
  const float sg = INV ? -1.0f : 1.0f;
  y[0] = make_float2(v[0].x + s1.x + s2.x, v[0].y + s1.y + s2.y);
  {  // k = 1: P = v0 + c1 s1 + c2 s2, Q = s1' d1 + s2' d2
    const float2 P = make_float2(v[0].x + ZF_C1 * s1.x + ZF_C2 * s2.x, v[0].y + ZF_C1 * s1.y + ZF_C2 * s2.y);
    const float2 Q = make_float2(ZF_S1 * d1.x + ZF_S2 * d2.x, ZF_S1 * d1.y + ZF_S2 * d2.y);
    y[1] = make_float2(P.x + sg * Q.y, P.y - sg * Q.x);
  }
  if (NOUT > 2) {  // k = 2: cos 8pi/5 = c1, sin 8pi/5 = -s1
    const float2 P = make_float2(v[0].x + ZF_C2 * s1.x + ZF_C1 * s2.x, v[0].y + ZF_C2 * s1.y + ZF_C1 * s2.y);
    const float2 Q = make_float2(ZF_S2 * d1.x - ZF_S1 * d2.x, ZF_S2 * d1.y - ZF_S1 * d2.y);
    y[2] = make_float2(P.x + sg * Q.y, P.y - sg * Q.x);
  }
}

// Forward stage 1 (5-point DFTs over n1 and the twiddles): U[k1][n2].
__device__ __forceinline__ void zf_fwd_stage1(const float2 (&x)[40], float2 (&U)[5][16]) {
#pragma unroll
  for (int n2 = 0; n2 < 16; ++n2) {
    float2 u[5];
    if (n2 < 8) zf_dft5_in<false, 3>(x[n2], x[16 + n2], x[32 + n2], u);
    else zf_dft5_in<false, 2>(x[n2], x[16 + n2], x[16 + n2], u);
#pragma unroll
    for (int k1 = 0; k1 < 5; ++k1) U[k1][n2] = zf_tw<false>(u[k1], n2 * k1);
  }
}

// Inverse stage 3 (twiddles and the kept outputs of the 5-point DFTs over k1):
// y[16 t1 + t2] from V[k1][t2].
__device__ __forceinline__ void zf_inv_stage3(const float2 (&V)[5][16], float2 (&y)[40]) {
#pragma unroll
  for (int t2 = 0; t2 < 16; ++t2) {
    float2 v[5], o[3];
#pragma unroll
    for (int k1 = 0; k1 < 5; ++k1) v[k1] = zf_tw<true>(V[k1][t2], t2 * k1);
    if (t2 < 8) {
      zf_dft5_out<true, 3>(v, o);
      y[t2] = o[0];
      y[16 + t2] = o[1];
      y[32 + t2] = o[2];
    } else {
      zf_dft5_out<true, 2>(v, o);
      y[t2] = o[0];
      y[16 + t2] = o[1];
    }
  }
}

// The middle of the z-pass operator for the groups k1 = a and b = (5 - a) % 5
// (X[k] and X[-k] lie in these two groups): forward 16-point DFTs of U[a],
// U[b] -> X, Y[k] = A[k] X[k] + G[k] X[-k] (AG(k) = (A, G)), inverse 16-point
// DFTs back into U[a], U[b] (then holding V).
template <int a, class AGf>
__device__ __forceinline__ void zf_middle(float2 (&U)[5][16], AGf AG) {
  constexpr int b = (5 - a) % 5;
  bfly16<false>(U[a]);
  if (b != a) bfly16<false>(U[b]);
  // k = a + 5 k2; -k mod 80 = b + 5 ((16 - k2 - (a ? 1 : 0)) mod 16)
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) {
    constexpr int adj = a ? 1 : 0;
    const int m2 = (32 - k2 - adj) % 16;  // X[-k] = U[b][m2]
    if (b == a && m2 < k2) continue;      // pair already done
    const int k = a + 5 * k2, kn = b + 5 * m2;
    const float2 Xk = U[a][k2], Xn = U[b][m2];
    const float4 p = AG(k);
    const float2 yk = make_float2(p.x * Xk.x - p.y * Xk.y + p.z * Xn.x - p.w * Xn.y,
                                  p.x * Xk.y + p.y * Xk.x + p.z * Xn.y + p.w * Xn.x);
    if (kn != k) {
      const float4 r = AG(kn);
      U[b][m2] = make_float2(r.x * Xn.x - r.y * Xn.y + r.z * Xk.x - r.w * Xk.y,
                             r.x * Xn.y + r.y * Xn.x + r.z * Xk.y + r.w * Xk.x);
    }
    U[a][k2] = yk;
  }
  bfly16<true>(U[a]);
  if (b != a) bfly16<true>(U[b]);
}

// a * w80^{-m} for any m < 80 (kW80 holds m < 64; w80^{-m} = -w80^{-(m - 40)})
__device__ __forceinline__ float2 zf_rot(float2 a, int m) {
  if (m < 40) return zf_tw<false>(a, m);
  const float2 r = zf_tw<false>(a, m - 40);
  return make_float2(-r.x, -r.y);
}

// zf_middle with the end-plane corrections folded into the frequency domain: the
// Toeplitz term acts on x without its planes 0 and SH = sz - 1,
//   Y[k] = A[k] (X[k] - x0 - xl w80^{-SH k}) + G[k] X[-k],
// so that IDFT80(Y)[t] is the whole operator and c0, c1 are not read (they are
// the taps A already holds: c0[t] = w[t], c1[t] = w[t - SH]).
template <int a, int SH, class AGf>
__device__ __forceinline__ void zf_middle_ends(float2 (&U)[5][16], AGf AG, float2 x0, float2 xl) {
  constexpr int b = (5 - a) % 5;
  bfly16<false>(U[a]);
  if (b != a) bfly16<false>(U[b]);
#pragma unroll
  for (int k2 = 0; k2 < 16; ++k2) {
    constexpr int adj = a ? 1 : 0;
    const int m2 = (32 - k2 - adj) % 16;  // X[-k] = U[b][m2]
    if (b == a && m2 < k2) continue;      // pair already done
    const int k = a + 5 * k2, kn = b + 5 * m2;
    const float2 Xk = U[a][k2], Xn = U[b][m2];
    const float2 ek = zf_rot(xl, (SH * k) % 80), en = zf_rot(xl, (SH * kn) % 80);
    const float2 Tk = make_float2(Xk.x - x0.x - ek.x, Xk.y - x0.y - ek.y);
    const float2 Tn = make_float2(Xn.x - x0.x - en.x, Xn.y - x0.y - en.y);
    const float4 p = AG(k);
    const float2 yk = make_float2(p.x * Tk.x - p.y * Tk.y + p.z * Xn.x - p.w * Xn.y,
                                  p.x * Tk.y + p.y * Tk.x + p.z * Xn.y + p.w * Xn.x);
    if (kn != k) {
      const float4 r = AG(kn);
      U[b][m2] = make_float2(r.x * Tn.x - r.y * Tn.y + r.z * Xk.x - r.w * Xk.y,
                             r.x * Tn.y + r.y * Tn.x + r.z * Xk.y + r.w * Xk.x);
    }
    U[a][k2] = yk;
  }
  bfly16<true>(U[a]);
  if (b != a) bfly16<true>(U[b]);
}

