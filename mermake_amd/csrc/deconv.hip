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
//   3. k_tile_zwiener  per (kx,ky) column: gather the Zp padded planes through the
//                      z map, forward FFT, multiply by W, inverse FFT, keep the
//                      sz cropped planes (in place).  With one PSF, W is loaded
//                      once per column block and reused for every tile;
//   4. k_fft_c2c       inverse C2C along x;
//   5. k_tile_c2r      C2R along y (two rows per complex transform) -> res (sz,T,T).
// The 1/(Zp*T*T) normalisation of ifftn is folded into W.
#include "deconv.h"
#include "fft_plan.h"
#include "mermake_hip.h"

#include <vector>

static_assert(sizeof(mm_tile_dev) == sizeof(mm_tile_desc) && sizeof(mm_tile_desc) == 104,
              "tile descriptor layout");

// --------------------------------------------------------------- generic C2C
template <bool INV, typename C>
__global__ __launch_bounds__(256) void k_fft_c2c(C* data, long long estride, long long cstride, int ncols,
                                                 long long bstride, int gfast, FftPlanDev<C> P, int G,
                                                 double scale) {
  typedef decltype(data[0].x) S;
  extern __shared__ __align__(16) unsigned char smem[];
  C* tw = (C*)smem;
  C* a = tw + P.N;
  C* b = a + (size_t)G * P.N;
  const int N = P.N;
  const int ncb = (ncols + G - 1) / G;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = bid % ncb;
  const int batch = bid / ncb;
  load_twiddles(tw, P);
  C* base = data + (long long)batch * bstride + (long long)cb * G * cstride;
  const int gmax = min(G, ncols - cb * G);
  const int tot = G * N;
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int g, n;
    if (gfast) { g = t % G; n = t / G; } else { n = t % N; g = t / N; }
    C v;
    v.x = 0; v.y = 0;
    if (g < gmax) v = base[g * cstride + n * estride];
    a[g * N + n] = v;
  }
  __syncthreads();
  C* r = fft_lds<INV>(a, b, tw, P, G);
  const S sc = (S)scale;
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int g, n;
    if (gfast) { g = t % G; n = t / G; } else { n = t % N; g = t / N; }
    if (g < gmax) base[g * cstride + n * estride] = vscale(r[g * N + n], sc);
  }
}

// ---------------------------------------------------------------- R2C rows
// Row loaders return the real value at column y of row q (rows of one batch
// entry are numbered 0..nrows-1).
struct TileRowLoader {  // padded tile rows (z, x), z in [0,sz), x in [0,T)
  const mm_tile_dev* tiles;
  int T;
  __device__ __forceinline__ float operator()(const mm_tile_dev& td, int q, int y) const {
    const int z = q / T;
    const int p = q - z * T;
    const int r = pad_map(p, td.xs, td.xd, td.xfirst, T);
    const int c = pad_map(y, td.ys, td.yd, td.yfirst, T);
    if (r < 0 || c < 0) return 0.0f;
    return tile_src(td, z, r, c);
  }
};

template <typename C>
__global__ __launch_bounds__(256) void k_r2c_rows_plain(const float* vol, int nrows, C* out, FftPlanDev<C> P,
                                                        int G) {
  typedef decltype(out[0].x) S;
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = P.N, H = N / 2 + 1;
  C* tw = (C*)smem;
  C* a = tw + N;
  C* b = a + (size_t)G * N;
  const int q0 = blockIdx.x * 2 * G;
  load_twiddles(tw, P);
  S* af = (S*)a;
  for (int t = threadIdx.x; t < 2 * G * N; t += blockDim.x) {
    const int row = t / N, y = t - row * N, q = q0 + row;
    S v = 0;
    if (q < nrows) v = (S)vol[(long long)q * N + y];
    af[((row >> 1) * N + y) * 2 + (row & 1)] = v;
  }
  __syncthreads();
  C* r = fft_lds<false>(a, b, tw, P, G);
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
  float2* tw = (float2*)smem;
  float2* a = tw + N;
  float2* b = a + (size_t)G * N;
  const int tile = blockIdx.y;
  const int q0 = blockIdx.x * 2 * G;
  const mm_tile_dev td = tiles[tile];
  const TileRowLoader ld{tiles, T};
  load_twiddles(tw, P);
  float* af = (float*)a;
  for (int t = threadIdx.x; t < 2 * G * N; t += blockDim.x) {
    const int row = t / N, y = t - row * N, q = q0 + row;
    float v = 0.0f;
    if (q < nrows) v = ld(td, q, y);
    af[((row >> 1) * N + y) * 2 + (row & 1)] = v;
  }
  __syncthreads();
  float2* r = fft_lds<false>(a, b, tw, P, G);
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
// Columns c = kx*H + ky of the (sz, T*H) half spectrum.  Block = G adjacent
// columns x a group of `tpb` tiles.  W layout: (n_psf, Zp, T*H).
__global__ __launch_bounds__(256) void k_tile_zwiener(float2* S, long long s_tstride, int TH, int sz, int zpad,
                                                      const int* __restrict__ zmap,
                                                      const float2* __restrict__ W, long long w_stride,
                                                      const mm_tile_dev* __restrict__ tiles, int ntiles, int tpb,
                                                      FftPlanDev<float2> P, int G) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int Zp = P.N;
  float2* tw = (float2*)smem;
  float2* a = tw + Zp;
  float2* b = a + (size_t)G * Zp;
  float2* wl = b + (size_t)G * Zp;
  int* zm = (int*)(wl + (size_t)G * Zp);
  const int ncb = (TH + G - 1) / G;
  const unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
  const int cb = bid % ncb;
  const int tg = bid / ncb;
  const int c0 = cb * G;
  const int gmax = min(G, TH - c0);
  load_twiddles(tw, P);
  for (int i = threadIdx.x; i < Zp; i += blockDim.x) zm[i] = zmap[i];
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
    float2* r = fft_lds<false>(a, b, tw, P, G);
    for (int t = threadIdx.x; t < G * Zp; t += blockDim.x) r[t] = cmul(r[t], wl[t]);
    __syncthreads();
    float2* o = fft_lds<true>(r, r == a ? b : a, tw, P, G);
    for (int t = threadIdx.x; t < G * sz; t += blockDim.x) {
      const int g = t % G, p = t / G;
      if (g < gmax) St[(long long)p * TH + g] = o[g * Zp + p + zpad];
    }
  }
}

// ------------------------------------------------------------------ C2R rows
__global__ __launch_bounds__(256) void k_tile_c2r(const float2* S, long long s_tstride, float* res,
                                                  long long r_tstride, int nrows, FftPlanDev<float2> P, int G) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = P.N, H = N / 2 + 1;
  float2* tw = (float2*)smem;
  float2* a = tw + N;
  float2* b = a + (size_t)G * N;
  const int tile = blockIdx.y;
  const int q0 = blockIdx.x * 2 * G;
  load_twiddles(tw, P);
  const float2* in = S + (long long)tile * s_tstride;
  for (int t = threadIdx.x; t < G * N; t += blockDim.x) {
    const int g = t / N, k = t - g * N;
    const int q = q0 + 2 * g;
    const bool cj = k >= H;
    const int kk = cj ? N - k : k;
    float2 A = make_float2(0.f, 0.f), B = make_float2(0.f, 0.f);
    if (q < nrows) A = in[(long long)q * H + kk];
    if (q + 1 < nrows) B = in[(long long)(q + 1) * H + kk];
    if (cj) { A.y = -A.y; B.y = -B.y; }
    if (k == 0 || 2 * k == N) { A.y = 0.f; B.y = 0.f; }  // real part semantics of ifftn(...).real
    a[g * N + k] = make_float2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  float2* r = fft_lds<true>(a, b, tw, P, G);
  float* out = res + (long long)tile * r_tstride;
  for (int t = threadIdx.x; t < 2 * G * N; t += blockDim.x) {
    const int row = t / N, y = t - row * N, q = q0 + row;
    if (q < nrows) {
      const float2 v = r[(row >> 1) * N + y];
      out[(long long)q * N + y] = (row & 1) ? v.y : v.x;
    }
  }
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
__global__ void k_wiener(const double2* __restrict__ Hs, float2* __restrict__ W, int Zp, int T, int H, double beta,
                         double scale) {
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
    W[i] = make_float2((float)(h.x / den * scale), (float)(-h.y / den * scale));
  }
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
  d->gT_r2c = mm_fft_pick_g(d->T, 8, 65536, 0, 8);
  d->gT_x = mm_fft_pick_g(d->T, 8, 65536, 0, 8);
  d->gZ = mm_fft_pick_g(d->Zp, 8, 49152, 8 * d->Zp + 0, 16);
  d->gT64_r2c = mm_fft_pick_g(d->T, 16, 98304, 0, 8);
  d->gT64_x = mm_fft_pick_g(d->T, 16, 98304, 0, 8);
  d->gZ64 = mm_fft_pick_g(d->Zp, 16, 98304, 0, 16);
  *out = d;
  return MM_OK;
}

extern "C" int mm_deconv_destroy(mm_deconv* d) {
  if (!d) return MM_OK;
  if (d->zmap_dev) (void)hipFree(d->zmap_dev);
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
  const size_t lds = mm_fft_lds_bytes(P.N, sizeof(C), G, 0);
  hipLaunchKernelGGL((k_fft_c2c<INV, C>), dim3((unsigned)nblk), dim3(256), lds, st, data, estride, cstride, ncols,
                     bstride, cstride == 1 ? 1 : 0, P, G, scale);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_deconv_build_wiener(mm_deconv* d, const float* psf_centered, double beta, void* W_out, void* ws,
                                      size_t ws_bytes, void* stream) {
  if (!d || !psf_centered || !W_out || !ws) { mm_set_error("mm_deconv_build_wiener: null argument"); return MM_EINVAL; }
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
    const size_t lds = mm_fft_lds_bytes(T, 16, G, 0);
    hipLaunchKernelGGL((k_r2c_rows_plain<double2>), dim3((nrows + 2 * G - 1) / (2 * G)), dim3(256), lds, st, pv,
                       nrows, Hs, d->pT64, G);
    MM_LAUNCH_CHECK();
  }
  int rc;
  if ((rc = launch_c2c<false>(Hs, H, 1, H, (long long)T * H, Zp, d->pT64, d->gT64_x, 1.0, st))) return rc;
  if ((rc = launch_c2c<false>(Hs, (long long)T * H, 1, T * H, 0, 1, d->pZ64, d->gZ64, 1.0, st))) return rc;
  const double scale = 1.0 / ((double)Zp * T * T);
  hipLaunchKernelGGL(k_wiener, dim3(4096), dim3(256), 0, st, Hs, (float2*)W_out, Zp, T, H, beta, scale);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" size_t mm_deconv_spec_bytes(const mm_deconv* d) {
  return d ? (size_t)d->sz * d->T * d->H * sizeof(float2) : 0;
}

extern "C" int mm_deconv_tiles(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W, int n_psf,
                               int tiles_per_block, void* spec_ws, float* res, void* stream) {
  if (!d || !tiles_dev || !W || !spec_ws || !res || ntiles < 0 || n_psf < 1) {
    mm_set_error("mm_deconv_tiles: bad argument");
    return MM_EINVAL;
  }
  if (ntiles == 0) return MM_OK;
  hipStream_t st = (hipStream_t)stream;
  const mm_tile_dev* td = (const mm_tile_dev*)tiles_dev;
  const int T = d->T, Zp = d->Zp, H = d->H, sz = d->sz;
  float2* S = (float2*)spec_ws;
#define MM_MARK(i) \
  if (d->prof) MM_CHECK(hipEventRecord(d->ev[i], st))
  MM_MARK(0);
  const long long s_ts = (long long)sz * T * H;
  const int nrows = sz * T;
  {  // 1. R2C along y of the padded rows
    const int G = d->gT_r2c;
    const size_t lds = mm_fft_lds_bytes(T, 8, G, 0);
    hipLaunchKernelGGL(k_tile_r2c, dim3((nrows + 2 * G - 1) / (2 * G), ntiles), dim3(256), lds, st, td, T, nrows, S,
                       s_ts, d->pT, G);
    MM_LAUNCH_CHECK();
  }
  MM_MARK(1);
  int rc;
  // 2. C2C along x, all tiles and planes at once
  if ((rc = launch_c2c<false>(S, H, 1, H, (long long)T * H, ntiles * sz, d->pT, d->gT_x, 1.0, st))) return rc;
  MM_MARK(2);
  {  // 3. z pass with the Wiener multiply
    const int G = d->gZ;
    const int TH = T * H;
    const int tpb = tiles_per_block < 1 ? 1 : (tiles_per_block > ntiles ? ntiles : tiles_per_block);
    const int ngroups = (ntiles + tpb - 1) / tpb;
    const long long nblk = (long long)((TH + G - 1) / G) * ngroups;
    const size_t lds = mm_fft_lds_bytes(Zp, 8, G, 8 * Zp) + sizeof(int) * Zp;
    hipLaunchKernelGGL(k_tile_zwiener, dim3((unsigned)nblk), dim3(256), lds, st, S, s_ts, TH, sz, d->zpad,
                       d->zmap_dev, (const float2*)W, (long long)Zp * TH, td, ntiles, tpb, d->pZ, G);
    MM_LAUNCH_CHECK();
  }
  MM_MARK(3);
  // 4. inverse C2C along x
  if ((rc = launch_c2c<true>(S, H, 1, H, (long long)T * H, ntiles * sz, d->pT, d->gT_x, 1.0, st))) return rc;
  MM_MARK(4);
  {  // 5. C2R along y into res
    const int G = d->gT_r2c;
    const size_t lds = mm_fft_lds_bytes(T, 8, G, 0);
    hipLaunchKernelGGL(k_tile_c2r, dim3((nrows + 2 * G - 1) / (2 * G), ntiles), dim3(256), lds, st, S, s_ts, res,
                       (long long)sz * T * T, nrows, d->pT, G);
    MM_LAUNCH_CHECK();
  }
  MM_MARK(5);
#undef MM_MARK
  return MM_OK;
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
      const int G = mm_fft_pick_g(n2, 8, 65536, 0, 16);
      rc = inverse ? launch_c2c<true>(p, 1, n2, n0 * n1, 0, 1, P, G, sc, st)
                   : launch_c2c<false>(p, 1, n2, n0 * n1, 0, 1, P, G, sc, st);
    } else if (ax == 1) {
      const int G = mm_fft_pick_g(n1, 8, 65536, 0, 16);
      rc = inverse ? launch_c2c<true>(p, n2, 1, n2, (long long)n1 * n2, n0, P, G, sc, st)
                   : launch_c2c<false>(p, n2, 1, n2, (long long)n1 * n2, n0, P, G, sc, st);
    } else {
      const int G = mm_fft_pick_g(n0, 8, 65536, 0, 16);
      rc = inverse ? launch_c2c<true>(p, (long long)n1 * n2, 1, n1 * n2, 0, 1, P, G, sc, st)
                   : launch_c2c<false>(p, (long long)n1 * n2, 1, n1 * n2, 0, 1, P, G, sc, st);
    }
    if (rc) return rc;
  }
  return MM_OK;
}
