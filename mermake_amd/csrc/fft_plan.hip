// FFT plan cache: radix schedule and float64-accurate twiddle tables.
#include "fft_plan.h"

#include <cmath>
#include <map>
#include <utility>
#include <mutex>
#include <vector>

namespace {

std::mutex g_plan_mu;
std::map<std::pair<int, int>, void*> g_tw;    // (device, N) -> twiddles (float2 or double2)
std::map<std::pair<int, int>, float*> g_coef;  // (device, R) -> cos/sin matrices of prime R

// Largest coefficient table kept in LDS (bytes): primes up to 241 (239 for T = 478).
const int kCoefMaxBytes = 122880;

// cos/sin(2 pi r k / R) for r, k in [1, h], rows r, row stride hp = h rounded up to 8.
float* coef_table(int dev, int R, int* len) {
  const int h = (R - 1) / 2, hp = (h + 7) & ~7;
  *len = 2 * h * hp;
  auto key = std::make_pair(dev, R);
  auto it = g_coef.find(key);
  if (it != g_coef.end()) return it->second;
  std::vector<float> c((size_t)*len, 0.0f);
  for (int r = 1; r <= h; ++r)
    for (int k = 1; k <= h; ++k) {
      const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)((r * k) % R) / R;
      c[(size_t)(r - 1) * hp + (k - 1)] = (float)cosl(a);
      c[(size_t)h * hp + (size_t)(r - 1) * hp + (k - 1)] = (float)sinl(a);
    }
  float* d = nullptr;
  if (hipMalloc(&d, sizeof(float) * c.size()) != hipSuccess) return nullptr;
  if (hipMemcpy(d, c.data(), sizeof(float) * c.size(), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  g_coef[key] = d;
  return d;
}

// Radix schedule: 8s first, then one 4 or 2 for the remaining power of two,
// then the odd primes in ascending order (3, 5, 7 are register butterflies,
// larger primes use the symmetric direct DFT pass).
bool factorize(int N, std::vector<int>& f) {
  f.clear();
  int n = N;
  while (n % 8 == 0) { f.push_back(8); n /= 8; }
  if (n % 4 == 0) { f.push_back(4); n /= 4; }
  if (n % 2 == 0) { f.push_back(2); n /= 2; }
  for (int p = 3; n > 1; p += 2) {
    while (n % p == 0) { f.push_back(p); n /= p; }
    if ((long long)p * p > n && n > 1) { f.push_back(n); n = 1; }
  }
  return (int)f.size() <= MM_FFT_MAXST;
}

template <typename C>
int get_plan(int N, FftPlanDev<C>* out, int tag) {
  if (N < 1 || N > 16384) {
    mm_set_error("FFT length %d outside [1, 16384]", N);
    return MM_EINVAL;
  }
  std::vector<int> f;
  if (!factorize(N, f)) {
    mm_set_error("FFT length %d has too many factors", N);
    return MM_EINVAL;
  }
  int dev = 0;
  MM_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto key = std::make_pair(dev * 2 + tag, N);
  auto it = g_tw.find(key);
  void* tw = nullptr;
  if (it == g_tw.end()) {
    std::vector<C> h(N);
    for (int k = 0; k < N; ++k) {
      // exp(-2 pi i k / N); reduce the angle to the first octant for accuracy
      long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)N;
      h[k].x = (decltype(h[k].x))cosl(a);
      h[k].y = (decltype(h[k].y))sinl(a);
    }
    MM_CHECK(hipMalloc(&tw, sizeof(C) * N));
    MM_CHECK(hipMemcpy(tw, h.data(), sizeof(C) * N, hipMemcpyHostToDevice));
    g_tw[key] = tw;
  } else {
    tw = it->second;
  }
  out->N = N;
  out->nst = (int)f.size();
  for (int i = 0; i < MM_FFT_MAXST; ++i) out->radix[i] = i < (int)f.size() ? f[i] : 0;
  out->tw = (const C*)tw;
  out->coef = nullptr;
  out->coefR = 0;
  out->coef_len = 0;
  if (sizeof(C) == 8) {  // float32 plans: the largest prime >= 11 whose matrices fit runs as a matrix product
    int best = 0;
    for (int r : f) {
      const int h = (r - 1) / 2, hp = (h + 7) & ~7;
      if (r >= 11 && r > best && 2 * h * hp * 4 <= kCoefMaxBytes) best = r;
    }
    if (best) {
      int len = 0;
      float* c = coef_table(dev, best, &len);
      if (!c) { mm_set_error("cannot upload the coefficient table of radix %d", best); return MM_ENOMEM; }
      out->coef = c;
      out->coefR = best;
      out->coef_len = len;
    }
  }
  return MM_OK;
}

std::map<std::pair<int, int>, void*> g_rader;  // (device, p) -> Rader tables

int inv_mod(int a, int m) {  // a^-1 mod m (gcd(a, m) = 1)
  a %= m;
  for (int x = 1; x < m; ++x)
    if (a * x % m == 1) return x;
  return 1;
}

}  // namespace

int mm_get_rader(int p, RdTab* fwd, RdTab* inv) {
  int A, B;
  if (p == 113) {
    A = RdSpec<113>::A;
    B = RdSpec<113>::B;
  } else if (p == 239) {
    A = RdSpec<239>::A;
    B = RdSpec<239>::B;
  } else {
    mm_set_error("Rader tables: no plan for the prime %d", p);
    return MM_EINVAL;
  }
  const int L = p - 1;
  int dev = 0;
  MM_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_plan_mu);
  void* buf = nullptr;
  auto it = g_rader.find({dev, p});
  if (it != g_rader.end()) {
    buf = it->second;
  } else {
    int gam = 0;  // smallest primitive root of p
    for (int c = 2; c < p && !gam; ++c) {
      int x = 1, ord = 0;
      do { x = x * c % p; ++ord; } while (x != 1);
      if (ord == p - 1) gam = c;
    }
    int ginv = 1;
    for (int e = 0; e < p - 2; ++e) ginv = ginv * gam % p;  // gam^(p-2) = gam^-1
    // Good-Thomas maps of L = A * B: n = B n1 + A n2, k = cA k1 + cB k2 (mod L)
    const int iBA = inv_mod(B, A), iAB = inv_mod(A, B);
    const int cA = B * iBA % L, cB = A * iAB % L;
    std::vector<int> pw(L), perm_out(L), slot_in(p, 0);
    int x = 1, y = 1;
    for (int a = 0; a < L; ++a) {
      pw[a] = x;                                              // gam^a
      slot_in[y] = (a * iAB) % B + B * ((a * iBA) % A);       // gam^-a -> slot n2 + B n1 of f_a
      x = x * gam % p;
      y = y * ginv % p;
    }
    slot_in[0] = L;  // x_0
    for (int m1 = 0; m1 < A; ++m1)
      for (int m2 = 0; m2 < B; ++m2) perm_out[m2 + B * m1] = pw[(B * m1 + A * m2) % L];
    const long double PI = 3.141592653589793238462643383279502884L;
    std::vector<float2> h(2 * L);
    for (int dir = 0; dir < 2; ++dir) {  // ghat = DFT_L(g) / L, g_n = exp(-/+ 2 pi i gam^n / p)
      const long double sg = dir ? 1.0L : -1.0L;
      for (int k1 = 0; k1 < A; ++k1)
        for (int k2 = 0; k2 < B; ++k2) {
          const int k = (cA * k1 + cB * k2) % L;
          long double re = 0, im = 0;
          for (int n = 0; n < L; ++n) {
            const long double ag = sg * 2.0L * PI * pw[n] / p;
            const long double aw = -2.0L * PI * (long double)((n * k) % L) / L;
            re += cosl(ag + aw);
            im += sinl(ag + aw);
          }
          h[dir * L + k1 + A * k2] = make_float2((float)(re / L), (float)(im / L));
        }
    }
    const size_t fbytes = sizeof(float2) * 2 * L, ibytes = sizeof(int) * (L + p);
    MM_CHECK(hipMalloc(&buf, fbytes + ibytes));
    std::vector<int> ints(perm_out);
    ints.insert(ints.end(), slot_in.begin(), slot_in.end());
    MM_CHECK(hipMemcpy(buf, h.data(), fbytes, hipMemcpyHostToDevice));
    MM_CHECK(hipMemcpy((char*)buf + fbytes, ints.data(), ibytes, hipMemcpyHostToDevice));
    g_rader[{dev, p}] = buf;
  }
  const float2* f = (const float2*)buf;
  const int* ip = (const int*)(f + 2 * L);
  *fwd = RdTab{f, ip, ip + L};
  *inv = RdTab{f + L, ip, ip + L};
  return MM_OK;
}

int mm_get_plan_f32(int N, FftPlanDev<float2>* out) { return get_plan<float2>(N, out, 0); }
int mm_get_plan_f64(int N, FftPlanDev<double2>* out) { return get_plan<double2>(N, out, 1); }

