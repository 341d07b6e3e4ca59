// Host-side FFT plans: radix schedule + accurate twiddle table on the device.
#pragma once
#include "fft_engine.h"

// Returns a cached plan for length N (created on first use; the twiddle table
// stays on the current device until the process ends).  Thread safe.
int mm_get_plan_f32(int N, FftPlanDev<float2>* out);
int mm_get_plan_f64(int N, FftPlanDev<double2>* out);

// Largest number of transforms G of length N (element size `esize`) that fit
// in `budget` bytes of LDS with two ping-pong buffers plus the twiddle table
// and `extra_per_g` additional bytes per transform; at least 1, at most gmax.
int mm_fft_pick_g(int N, int esize, int budget, int extra_per_g, int gmax);

// Dynamic LDS bytes for G transforms of length N.
static inline size_t mm_fft_lds_bytes(int N, int esize, int G, int extra_per_g) {
  return (size_t)N * esize + (size_t)G * (2 * (size_t)N * esize + extra_per_g);
}

#define MM_LDS_MAX 163840
