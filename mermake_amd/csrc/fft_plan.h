// Host-side FFT plans: radix schedule + accurate twiddle table on the device.
#pragma once
#include "fft_rader.h"

// Returns a cached plan for length N (created on first use; the twiddle table
// stays on the current device until the process ends).  Thread safe.
int mm_get_plan_f32(int N, FftPlanDev<float2>* out);
int mm_get_plan_f64(int N, FftPlanDev<double2>* out);

// Rader tables of p = 113 on the current device (created on first use, kept
// until the process ends), for the forward and the inverse transform.
int mm_get_rader(int p, RdTab* fwd, RdTab* inv);  // p = 113 or 239

// Largest number of transforms G (at most gmax, at least 1) whose LDS
// (fft_lds_bytes with `extra_per_g` more bytes per transform) fits `budget`.
template <typename C>
int mm_fft_pick_g(const FftPlanDev<C>& P, int budget, int extra_per_g, int gmax) {
  int g = gmax;
  while (g > 1 && fft_lds_bytes(P, g, (size_t)extra_per_g * g) > (size_t)budget) --g;
  return g;
}

#define MM_LDS_MAX 163840
