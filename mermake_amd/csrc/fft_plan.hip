// FFT plan cache: radix schedule and float64-accurate twiddle tables.
#include "fft_plan.h"

#include <cmath>
#include <map>
#include <mutex>
#include <vector>

namespace {

std::mutex g_plan_mu;
std::map<std::pair<int, int>, void*> g_tw;  // (device, N) -> twiddles (float2 or double2)

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
  return MM_OK;
}

}  // namespace

int mm_get_plan_f32(int N, FftPlanDev<float2>* out) { return get_plan<float2>(N, out, 0); }
int mm_get_plan_f64(int N, FftPlanDev<double2>* out) { return get_plan<double2>(N, out, 1); }

int mm_fft_pick_g(int N, int esize, int budget, int extra_per_g, int gmax) {
  long long per = 2LL * N * esize + extra_per_g;
  long long g = ((long long)budget - (long long)N * esize) / per;
  if (g > gmax) g = gmax;
  if (g < 1) g = 1;
  return (int)g;
}
