// 3D local maxima with deterministic raster-order compaction, delta_fit
// refinement, and the hyb keep-filter (gfx950).
//
// Reference: mermake/maxima.cu:1-145 (maxima test), :148-341 (fit),
// mermake/maxima.py:17-98 (driver), mermake/__main__.py:254-265 (keep-filter).
//
// Maxima are found in three passes that never use atomics, so the candidate
// list is in raster order and identical from run to run.  The unit is a row
// (z, x) of the volume, handled by one wavefront with the lanes along y:
//   k_max_mark     one bit per voxel (64-voxel ballot words per row) + row counts
//   k_max_scan     exclusive scan of the row counts of each volume
//   k_max_compact  (z, x, y) of each set bit at its row offset + rank, one thread per row
// The fit runs one wavefront per candidate: the lanes sample the delta_fit
// sphere, and the statistics are wave reductions (two-pass mean/variance).
#include "deconv.h"
#include "mermake_hip.h"

#include <mutex>
#include <vector>

struct VolDesc {
  const float* p;
  int D, H, W;
  long long sZ, sX;
  int w0, w1;  // candidates are searched at x and y in [w0, w1) only
};

struct OneVol {
  VolDesc d;
  __device__ __forceinline__ VolDesc operator()(int) const { return d; }
};

struct TileVols {  // view res[:, :xe, :ye] of every tile (deconvolver.py:170)
  const mm_tile_dev* tiles;
  const float* res;
  int sz, T;
  int w0 = 0, w1 = 0x7fffffff;  // candidate window of every tile
  __device__ __forceinline__ VolDesc operator()(int v) const {
    VolDesc d;
    d.p = res + (long long)v * sz * T * T;
    d.D = sz;
    d.H = tiles[v].xe;
    d.W = tiles[v].ye;
    d.sZ = (long long)T * T;
    d.sX = T;
    d.w0 = w0;
    d.w1 = w1;
    return d;
  }
};

#define MX_CHUNK 8  // row words whose centre loads are in flight together

__device__ __forceinline__ int refl1(int n, int N) {  // maxima.cu:44-52
  if (n < 0) return -n;
  if (n >= N) return 2 * N - n - 2;
  return n;
}

// Neighbourhood test of a voxel whose value c already passed the threshold.
__device__ __forceinline__ bool is_max_at(const VolDesc& V, int z, int x, int y, float c, float sign, int delta) {
  const int d2 = delta * delta;
  for (int dz = -delta; dz <= delta; ++dz) {
    const int nz = refl1(z + dz, V.D);
    const bool zok = nz >= 0 && nz < V.D;
    for (int dx = -delta; dx <= delta; ++dx) {
      const int nx = refl1(x + dx, V.H);
      const bool xok = nx >= 0 && nx < V.H;
      for (int dy = -delta; dy <= delta; ++dy) {
        if (dz == 0 && dx == 0 && dy == 0) continue;
        if (dz * dz + dx * dx + dy * dy > d2) continue;
        const int ny = refl1(y + dy, V.W);
        // neighbours still outside after the single reflection (axes shorter
        // than delta+1) are skipped instead of read out of bounds
        if (!zok || !xok || ny < 0 || ny >= V.W) continue;
        if (c < sign * V.p[nz * V.sZ + nx * V.sX + ny]) return false;
      }
    }
  }
  return true;
}

// Sphere offsets 1 <= dz^2 + dx^2 + dy^2 <= 25 ordered by distance, built at compile
// time: the sphere of delta <= 5 is the prefix of the kSphN[delta] nearest points.  A
// voxel is a maximum iff no neighbour of its sphere exceeds it, in whatever order they
// are tested; testing the nearest ones first rejects the non-maxima (whose larger
// neighbour is usually adjacent) after a few loads instead of scanning from the far
// side of the sphere.
#define MX_SPH_DMAX 5
#define MX_SPH_N 514
struct SphTab {
  signed char d[MX_SPH_N][4];  // dz, dx, dy, unused
  int n[MX_SPH_DMAX + 1];      // points with d^2 <= delta^2 (centre excluded)
};

constexpr SphTab make_sph_tab() {
  SphTab t{};
  int k = 0;
  for (int s = 1; s <= MX_SPH_DMAX * MX_SPH_DMAX; ++s) {
    for (int dz = -MX_SPH_DMAX; dz <= MX_SPH_DMAX; ++dz)
      for (int dx = -MX_SPH_DMAX; dx <= MX_SPH_DMAX; ++dx)
        for (int dy = -MX_SPH_DMAX; dy <= MX_SPH_DMAX; ++dy)
          if (dz * dz + dx * dx + dy * dy == s) {
            t.d[k][0] = (signed char)dz;
            t.d[k][1] = (signed char)dx;
            t.d[k][2] = (signed char)dy;
            ++k;
          }
    for (int r = 0; r <= MX_SPH_DMAX; ++r)
      if (s == r * r) t.n[r] = k;
  }
  return t;
}

__constant__ SphTab kSph = make_sph_tab();
static_assert(make_sph_tab().n[1] == 6 && make_sph_tab().n[3] == 122 && make_sph_tab().n[5] == MX_SPH_N,
              "sphere sizes of maxima.cu (6, 122, 514 neighbours at delta 1, 3, 5)");

// Interior voxel (the whole sphere inside the volume, no reflection): the sphere's
// element offsets are in LDS (soff, built per block), tested nearest first, four loads
// at a time.
__device__ __forceinline__ bool is_max_sph_in(const float* __restrict__ c0, const int* __restrict__ soff, int n,
                                              float c, float sign) {
  int i = 0;
  for (; i + 4 <= n; i += 4) {
    const float a0 = c0[soff[i]], a1 = c0[soff[i + 1]], a2 = c0[soff[i + 2]], a3 = c0[soff[i + 3]];
    if (c < sign * a0 || c < sign * a1 || c < sign * a2 || c < sign * a3) return false;
  }
  for (; i < n; ++i)
    if (c < sign * c0[soff[i]]) return false;
  return true;
}

// The rest of the sphere, neighbours [first, n), for the interior candidates of one wavefront that survived
// the first ones (bit l of pend: the voxel at row[l], centre value cs = sign * v in lane l): the whole wavefront
// tests one candidate at a time, each lane loading up to 8 of its neighbours at once, so a candidate costs one
// round of loads instead of (n - first) / 4 rounds of one lane's walk that the other lanes wait for.  Returns
// the mask of the candidates that are maxima.
#define MX_COOP_FIRST 8
static_assert(MX_SPH_N - MX_COOP_FIRST <= 8 * 64, "the rest of the sphere fits 8 loads per lane");
__device__ __forceinline__ unsigned long long is_max_sph_coop(const float* __restrict__ row, unsigned long long pend,
                                                              float cs, float sign, const int* __restrict__ soff,
                                                              int n) {
  const int lane = threadIdx.x & 63;
  const int nj = (n - MX_COOP_FIRST + 63) >> 6;  // uniform
  unsigned long long keep = 0;
  while (pend) {
    const int l = __ffsll((long long)pend) - 1;
    pend &= pend - 1;
    const float c = __shfl(cs, l);
    const float* c0 = row + l;
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int i = MX_COOP_FIRST + lane + 64 * j;
      if (j < nj) v[j] = c0[i < n ? soff[i] : 0];  // past the sphere: the centre, never larger than itself
    }
    bool big = false;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j < nj) big = big || c < sign * v[j];
    if (__ballot(big) == 0) keep |= 1ull << l;
  }
  return keep;
}

// The same for candidates near a border (bit l of pend: the voxel at (z, x, ybase + l)): each lane reflects its
// neighbours' coordinates as is_max_sph_border does and skips those still outside.
__device__ __forceinline__ unsigned long long is_max_sph_coop_border(const VolDesc& V, int z, int x, int ybase,
                                                                     unsigned long long pend, float cs, float sign,
                                                                     int n) {
  const int lane = threadIdx.x & 63;
  const int nj = (n - MX_COOP_FIRST + 63) >> 6;  // uniform
  unsigned long long keep = 0;
  while (pend) {
    const int l = __ffsll((long long)pend) - 1;
    pend &= pend - 1;
    const float c = __shfl(cs, l);
    const int y = ybase + l;
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int i = MX_COOP_FIRST + lane + 64 * j;
      if (j < nj) {
        long long o = -1;
        if (i < n) {
          const int nz = refl1(z + kSph.d[i][0], V.D), nx = refl1(x + kSph.d[i][1], V.H),
                    ny = refl1(y + kSph.d[i][2], V.W);
          if (nz >= 0 && nz < V.D && nx >= 0 && nx < V.H && ny >= 0 && ny < V.W)
            o = nz * V.sZ + nx * V.sX + ny;
        }
        v[j] = o >= 0 ? V.p[o] : -sign * INFINITY;  // outside: never larger than the centre
      }
    }
    bool big = false;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j < nj) big = big || c < sign * v[j];
    if (__ballot(big) == 0) keep |= 1ull << l;
  }
  return keep;
}

// Voxel near a border: the same sphere, nearest first, with the single reflection of
// maxima.cu:44-52 (neighbours still outside are skipped, as in is_max_at).
__device__ __forceinline__ bool is_max_sph_border(const VolDesc& V, int z, int x, int y, float c, float sign, int n) {
  for (int i = 0; i < n; ++i) {
    const int nz = refl1(z + kSph.d[i][0], V.D), nx = refl1(x + kSph.d[i][1], V.H), ny = refl1(y + kSph.d[i][2], V.W);
    if (nz < 0 || nz >= V.D || nx < 0 || nx >= V.H || ny < 0 || ny >= V.W) continue;
    if (c < sign * V.p[nz * V.sZ + nx * V.sX + ny]) return false;
  }
  return true;
}

// delta = 1 (the hyb setting): the 6 face neighbours, all loads issued at once.
// An axis too short for the single reflection leaves the neighbour out, as above.
__device__ __forceinline__ bool is_max_d1(const VolDesc& V, int z, int x, int y, float c, float sign) {
  const int nz[2] = {refl1(z - 1, V.D), refl1(z + 1, V.D)};
  const int nx[2] = {refl1(x - 1, V.H), refl1(x + 1, V.H)};
  const int ny[2] = {refl1(y - 1, V.W), refl1(y + 1, V.W)};
  const long long base = z * V.sZ + x * V.sX + y;
  long long o[6];
  bool ok[6];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    ok[s] = nz[s] >= 0 && nz[s] < V.D;
    o[s] = base + (nz[s] - z) * V.sZ;
    ok[2 + s] = nx[s] >= 0 && nx[s] < V.H;
    o[2 + s] = base + (nx[s] - x) * V.sX;
    ok[4 + s] = ny[s] >= 0 && ny[s] < V.W;
    o[4 + s] = base + (ny[s] - y);
  }
  float n[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) n[k] = V.p[ok[k] ? o[k] : base];  // the centre stands in for a missing neighbour
  bool m = true;
#pragma unroll
  for (int k = 0; k < 6; ++k) m = m && !(c < sign * n[k]);
  return m;
}

// Rows r = z hl + x of volume blockIdx.y (hl >= H: the workspace's rows per
// plane), four per block (one per wavefront); ww words of 64 voxels per row.
// Wavefront-only work: no block barrier.
template <class VS>
__global__ __launch_bounds__(256) void k_max_mark(VS vs, float sign, float thr, int delta,
                                                  unsigned long long* __restrict__ masks, int* __restrict__ rcnt,
                                                  int nrow_max, int hl, int ww) {
  const int v = blockIdx.y;
  const VolDesc V = vs(v);
  const int lane = threadIdx.x & 63;
  // element offsets of the sorted sphere (2 <= delta <= 5), for the interior voxels
  __shared__ int soff[MX_SPH_N];
  const bool sph = delta >= 2 && delta <= MX_SPH_DMAX;
  const int nsph = sph ? kSph.n[delta] : 0;
  for (int i = threadIdx.x; i < nsph; i += blockDim.x)
    soff[i] = (int)(kSph.d[i][0] * V.sZ + kSph.d[i][1] * V.sX + kSph.d[i][2]);
  if (sph) __syncthreads();
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= nrow_max) return;
  const long long ridx = (long long)v * nrow_max + row;
  int cnt = 0;
  const int z = row / hl, x = row - (row / hl) * hl;
  if (z < V.D && x < V.H && x >= V.w0 && x < V.w1) {  // rows outside the window keep a zero count
    const float* rp = V.p + z * V.sZ + x * V.sX;
    const bool zx_in = z >= delta && z < V.D - delta && x >= delta && x < V.H - delta;  // uniform
    unsigned long long* mw = masks + ridx * ww;
    const int y0 = max(V.w0, 0), y1 = min(V.W, V.w1);
    for (int k0 = 0; k0 < ww; k0 += MX_CHUNK) {
      float c[MX_CHUNK];
#pragma unroll
      for (int k = 0; k < MX_CHUNK; ++k) {  // the chunk's centre loads in flight before any test
        const int y = (k0 + k) * 64 + lane;
        c[k] = y >= y0 && y < y1 ? sign * rp[y] : 0.0f;
      }
#pragma unroll
      for (int k = 0; k < MX_CHUNK; ++k) {
        if (k0 + k >= ww) break;
        const int y = (k0 + k) * 64 + lane;
        // maxima.cu:19: v <= threshold is rejected (a NaN threshold rejects nothing)
        bool m = y >= y0 && y < y1 && !(c[k] <= thr);
        const bool interior = zx_in && y >= delta && y < V.W - delta;
        if (m) {
          if (delta == 1) {
            m = is_max_d1(V, z, x, y, c[k], sign);
          } else if (sph) {
            if (interior) m = is_max_sph_in(rp + y, soff, min(nsph, MX_COOP_FIRST), c[k], sign);
            else m = is_max_sph_border(V, z, x, y, c[k], sign, min(nsph, MX_COOP_FIRST));
          } else {
            m = is_max_at(V, z, x, y, c[k], sign, delta);
          }
        }
        if (sph && nsph > MX_COOP_FIRST) {  // survivors: the rest of the sphere, one at a time
          const unsigned long long pend = __ballot(m && interior), pendb = __ballot(m && !interior);
          unsigned long long ok = ~0ull;
          if (pend) ok &= is_max_sph_coop(rp + (k0 + k) * 64, pend, c[k], sign, soff, nsph) | ~pend;
          if (pendb) ok &= is_max_sph_coop_border(V, z, x, (k0 + k) * 64, pendb, c[k], sign, nsph) | ~pendb;
          m = m && ((ok >> lane) & 1ull);
        }
        const unsigned long long b = __ballot(m);
        if (lane == 0) mw[k0 + k] = b;
        cnt += __popcll(b);
      }
    }
  }
  if (lane == 0) rcnt[ridx] = cnt;
}

// Both signs of k_max_mark in one pass over the volume (the DAPI maxima of
// +buffer and -buffer, __main__.py:224-225): the centre values are read once,
// candidates of +v go to the workspace's volume 0 and those of -v to volume 1.
template <class VS>
__global__ __launch_bounds__(256) void k_max_mark_pm(VS vs, float thr, int delta,
                                                     unsigned long long* __restrict__ masks,
                                                     int* __restrict__ rcnt, int nrow_max, int hl, int ww) {
  const VolDesc V = vs(0);
  const int lane = threadIdx.x & 63;
  __shared__ int soff[MX_SPH_N];
  const bool sph = delta >= 2 && delta <= MX_SPH_DMAX;
  const int nsph = sph ? kSph.n[delta] : 0;
  for (int i = threadIdx.x; i < nsph; i += blockDim.x)
    soff[i] = (int)(kSph.d[i][0] * V.sZ + kSph.d[i][1] * V.sX + kSph.d[i][2]);
  if (sph) __syncthreads();
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= nrow_max) return;
  int cntp = 0, cntm = 0;
  const int z = row / hl, x = row - (row / hl) * hl;
  if (z < V.D && x < V.H) {
    const float* rp = V.p + z * V.sZ + x * V.sX;
    const bool zx_in = z >= delta && z < V.D - delta && x >= delta && x < V.H - delta;  // uniform
    unsigned long long* mwp = masks + (long long)row * ww;
    unsigned long long* mwm = masks + ((long long)nrow_max + row) * ww;
    for (int k0 = 0; k0 < ww; k0 += MX_CHUNK) {
      float c[MX_CHUNK];
#pragma unroll
      for (int k = 0; k < MX_CHUNK; ++k) {
        const int y = (k0 + k) * 64 + lane;
        c[k] = y < V.W ? rp[y] : 0.0f;
      }
#pragma unroll
      for (int k = 0; k < MX_CHUNK; ++k) {
        if (k0 + k >= ww) break;
        const int y = (k0 + k) * 64 + lane;
        const bool in = y < V.W;
        bool mp = in && !(c[k] <= thr);   // maxima.cu:19, for +v and for -v
        bool mm = in && !(-c[k] <= thr);
        const bool interior = zx_in && y >= delta && y < V.W - delta;
        const int n1 = min(nsph, MX_COOP_FIRST);
        if (mp) {
          if (delta == 1) mp = is_max_d1(V, z, x, y, c[k], 1.0f);
          else if (sph) mp = interior ? is_max_sph_in(rp + y, soff, n1, c[k], 1.0f)
                                      : is_max_sph_border(V, z, x, y, c[k], 1.0f, n1);
          else mp = is_max_at(V, z, x, y, c[k], 1.0f, delta);
        }
        if (mm) {
          if (delta == 1) mm = is_max_d1(V, z, x, y, -c[k], -1.0f);
          else if (sph) mm = interior ? is_max_sph_in(rp + y, soff, n1, -c[k], -1.0f)
                                      : is_max_sph_border(V, z, x, y, -c[k], -1.0f, n1);
          else mm = is_max_at(V, z, x, y, -c[k], -1.0f, delta);
        }
        if (sph && nsph > MX_COOP_FIRST) {  // survivors: the rest of the sphere, one at a time
          const float* wrow = rp + (k0 + k) * 64;
          const int yb = (k0 + k) * 64;
          const unsigned long long pp = __ballot(mp && interior), pm = __ballot(mm && interior);
          const unsigned long long ppb = __ballot(mp && !interior), pmb = __ballot(mm && !interior);
          unsigned long long okp = ~0ull, okm = ~0ull;
          if (pp) okp &= is_max_sph_coop(wrow, pp, c[k], 1.0f, soff, nsph) | ~pp;
          if (ppb) okp &= is_max_sph_coop_border(V, z, x, yb, ppb, c[k], 1.0f, nsph) | ~ppb;
          if (pm) okm &= is_max_sph_coop(wrow, pm, -c[k], -1.0f, soff, nsph) | ~pm;
          if (pmb) okm &= is_max_sph_coop_border(V, z, x, yb, pmb, -c[k], -1.0f, nsph) | ~pmb;
          mp = mp && ((okp >> lane) & 1ull);
          mm = mm && ((okm >> lane) & 1ull);
        }
        const unsigned long long bp = __ballot(mp), bm = __ballot(mm);
        if (lane == 0) {
          mwp[k0 + k] = bp;
          mwm[k0 + k] = bm;
        }
        cntp += __popcll(bp);
        cntm += __popcll(bm);
      }
    }
  }
  if (lane == 0) {
    rcnt[row] = cntp;
    rcnt[nrow_max + row] = cntm;
  }
}

// The same marks from threshold bits already in the masks (written by the blur
// that produced the volume: bit y of row r set iff !(sign v <= thr)): one thread
// per row turns every non-zero word into its maxima word in place, so only the
// candidates' neighbourhoods are read.  Zero words stay zero.
template <class VS>
__global__ __launch_bounds__(256) void k_max_mark_above(VS vs, float sign, int delta,
                                                        unsigned long long* __restrict__ masks,
                                                        int* __restrict__ rcnt, int nrow_max, int hl, int ww) {
  const int v = blockIdx.y;
  const VolDesc V = vs(v);
  __shared__ int soff[MX_SPH_N];
  const bool sph = delta >= 2 && delta <= MX_SPH_DMAX;
  const int nsph = sph ? kSph.n[delta] : 0;
  for (int i = threadIdx.x; i < nsph; i += blockDim.x)
    soff[i] = (int)(kSph.d[i][0] * V.sZ + kSph.d[i][1] * V.sX + kSph.d[i][2]);
  if (sph) __syncthreads();
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrow_max) return;
  const long long ridx = (long long)v * nrow_max + row;
  const int z = row / hl, x = row - (row / hl) * hl;
  int cnt = 0;
  if (z < V.D && x < V.H && x >= V.w0 && x < V.w1) {
    const float* rp = V.p + z * V.sZ + x * V.sX;
    const bool zx_in = z >= delta && z < V.D - delta && x >= delta && x < V.H - delta;
    unsigned long long* mw = masks + ridx * ww;
    const int y0 = max(V.w0, 0), y1 = min(V.W, V.w1);
    for (int k = 0; k < ww; ++k) {
      unsigned long long a = mw[k];
      if (!a) continue;
      unsigned long long m = 0;
      while (a) {
        const int b = __ffsll((long long)a) - 1;
        a &= a - 1;
        const int y = k * 64 + b;
        if (y < y0 || y >= y1) continue;
        const float c = sign * rp[y];
        bool mx;
        if (delta == 1) mx = is_max_d1(V, z, x, y, c, sign);
        else if (sph) mx = zx_in && y >= delta && y < V.W - delta ? is_max_sph_in(rp + y, soff, nsph, c, sign)
                                                                   : is_max_sph_border(V, z, x, y, c, sign, nsph);
        else mx = is_max_at(V, z, x, y, c, sign, delta);
        if (mx) m |= 1ull << b;
      }
      mw[k] = m;
      cnt += __popcll(m);
    }
  }
  rcnt[ridx] = cnt;
}

// exclusive scan of the block counts of volume blockIdx.x; total -> counts[v]
__global__ __launch_bounds__(1024) void k_max_scan(int* __restrict__ bcnt, int nblk, int* __restrict__ counts) {
  const int v = blockIdx.x;
  int* c = bcnt + (long long)v * nblk;
  __shared__ int wsum[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int carry = 0;
  for (int base = 0; base < nblk; base += blockDim.x) {
    const int i = base + threadIdx.x;
    const int x = i < nblk ? c[i] : 0;
    int s = x;
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(s, off, 64);
      if (lane >= off) s += y;
    }
    if (lane == 63) wsum[wave] = s;
    __syncthreads();
    int wpre = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
      if (w < wave) wpre += wsum[w];
      tot += wsum[w];
    }
    if (i < nblk) c[i] = carry + wpre + s - x;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[v] = carry;
}

// Row r of volume blockIdx.y writes its maxima at roff[r] + rank (raster order):
// one thread per row, since most rows hold none.
__global__ __launch_bounds__(256) void k_max_compact(const unsigned long long* __restrict__ masks,
                                                     const int* __restrict__ roff, const int* __restrict__ counts,
                                                     int nrow_max, int hl, int ww, int* __restrict__ coords,
                                                     int cap) {
  const int v = blockIdx.y;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrow_max) return;
  const long long ridx = (long long)v * nrow_max + row;
  int pos = roff[ridx];
  const int next = row + 1 < nrow_max ? roff[ridx + 1] : counts[v];
  if (next == pos || pos >= cap) return;  // no maxima in this row, or all beyond the cap
  const int z = row / hl, x = row - (row / hl) * hl;
  const unsigned long long* mw = masks + ridx * ww;
  int* o = coords + (long long)v * cap * 3;
  for (int k = 0; k < ww; ++k) {
    unsigned long long m = mw[k];
    while (m) {
      const int b = __ffsll((long long)m) - 1;
      m &= m - 1;
      o[pos * 3] = z;
      o[pos * 3 + 1] = x;
      o[pos * 3 + 2] = k * 64 + b;
      if (++pos >= cap) return;
    }
  }
}

// ------------------------------------------------------------------- fit
struct RawPlain {  // pixel type as mm_tile_desc.img_type: uint16, float32 or squared uint8
  const void* p;
  int type;
  long long sZ, sX;
  struct Ctx {};
  __device__ __forceinline__ Ctx ctx(int) const { return Ctx{}; }
  __device__ __forceinline__ float at(const Ctx&, int z, int x, int y) const {
    return pix_at(p, type, z * sZ + x * sX + y);
  }
};

struct RawTile {  // tile_pad[zpad:-zpad, :xend, :yend] after the flat re-multiply
  const mm_tile_dev* tiles;
  int T;
  typedef mm_tile_dev Ctx;
  __device__ __forceinline__ Ctx ctx(int v) const { return tiles[v]; }
  __device__ __forceinline__ float at(const Ctx& t, int z, int x, int y) const { return tile_raw(t, T, z, x, y); }
};

__device__ __forceinline__ float wsum(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wsumi(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wmin(float v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

#define MX_MAX_KERNEL_POINTS 600

// Sample k of the sphere for a candidate: reflected coordinates and validity
// (maxima.cu:190-206).  The 600-sample cap of maxima.cu:194 counts valid
// samples in loop order; `run` is the number of valid samples before this
// 64-sample chunk.
struct Sample {
  bool ok;
  int dz, dx, dy;
  float v, r, g;
};

template <class RS>
__device__ __forceinline__ Sample get_sample(const VolDesc& V, float sign, const RS& raw,
                                             const typename RS::Ctx& rc, const int* offs, int P, int k, int z0,
                                             int x0, int y0, float isz, float isxy, int& run) {
  Sample s;
  s.ok = false;
  s.dz = s.dx = s.dy = 0;
  s.v = s.r = s.g = 0.0f;
  int zz = 0, xx = 0, yy = 0;
  bool valid = false;
  if (k < P) {
    const int o = offs[k];
    s.dz = (o >> 16) - 128;
    s.dx = ((o >> 8) & 255) - 128;
    s.dy = (o & 255) - 128;
    zz = refl1(z0 + s.dz, V.D);
    xx = refl1(x0 + s.dx, V.H);
    yy = refl1(y0 + s.dy, V.W);
    valid = zz >= 0 && zz < V.D && xx >= 0 && xx < V.H && yy >= 0 && yy < V.W;
  }
  const unsigned long long b = __ballot(valid);
  const int lane = threadIdx.x & 63;
  const int before = run + __popcll(b & ((1ull << lane) - 1ull));
  run += __popcll(b);
  if (valid && before < MX_MAX_KERNEL_POINTS) {
    s.ok = true;
    s.v = sign * V.p[zz * V.sZ + xx * V.sX + yy];
    s.r = raw.at(rc, zz, xx, yy);
    s.g = expf(-((float)(s.dz * s.dz) * isz + (float)(s.dx * s.dx) * isxy + (float)(s.dy * s.dy) * isxy));
  }
  return s;
}

// NCH > 0: the candidate's samples (at most NCH chunks of 64) are gathered
// once and kept in registers for the four passes; NCH = 0 gathers them again
// in every pass (spheres of more than 9 chunks).  The same arithmetic in the
// same order either way.
template <class VS, class RS, int NCH>
__global__ __launch_bounds__(256) void k_fit(VS vs, float sign, RS raw, const int* __restrict__ coords,
                                             const int* __restrict__ counts, int cap, const int* __restrict__ offs,
                                             int P, float sigmaZ, float sigmaXY, float* __restrict__ out) {
  const int v = blockIdx.y;
  const VolDesc V = vs(v);
  const typename RS::Ctx rc = raw.ctx(v);
  const int n = min(counts[v], cap);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // reference expression dz*dz / (2*sigmaZ*sigmaZ) as a multiply by the reciprocal
  const float isz = 1.0f / (2.0f * sigmaZ * sigmaZ);
  const float isxy = 1.0f / (2.0f * sigmaXY * sigmaXY);
  const int nch = NCH ? NCH : (P + 63) / 64;
  for (int i = blockIdx.x * 4 + wave; i < n; i += gridDim.x * 4) {
    const int* c = coords + ((long long)v * cap + i) * 3;
    const int z0 = c[0], x0 = c[1], y0 = c[2];
    Sample S[NCH ? NCH : 1];
    int run = 0;
    auto sample = [&](int ch) -> Sample {
      if (NCH) return S[ch];
      return get_sample(V, sign, raw, rc, offs, P, ch * 64 + lane, z0, x0, y0, isz, isxy, run);
    };
    if (NCH) {
#pragma unroll
      for (int ch = 0; ch < (NCH ? NCH : 1); ++ch)
        S[ch] = get_sample(V, sign, raw, rc, offs, P, ch * 64 + lane, z0, x0, y0, isz, isxy, run);
    }
    // pass 1: count, sum G, min
    int cnt = 0;
    float sg = 0.f, mn = 3.0e38f;
    run = 0;
#pragma unroll
    for (int ch = 0; ch < nch; ++ch) {
      const Sample s = sample(ch);
      if (s.ok) { cnt += 1; sg += s.g; mn = fminf(mn, s.v); }
    }
    cnt = wsumi(cnt);
    float* o = out + ((long long)v * cap + i) * 8;
    if (cnt == 0) {  // maxima.cu:233 leaves the zero-initialised row
      if (lane < 8) o[lane] = 0.0f;
      continue;
    }
    sg = wsum(sg);
    const float bg = wmin(mn);
    const float fn = (float)cnt;
    const float mg = sg / fn;
    // pass 2: var G, sums of s and raw, centroid moments
    float vg = 0.f, ss = 0.f, sr = 0.f, mz = 0.f, mx = 0.f, my = 0.f;
    run = 0;
#pragma unroll
    for (int ch = 0; ch < nch; ++ch) {
      const Sample s = sample(ch);
      if (s.ok) {
        const float dg = s.g - mg;
        vg += dg * dg;
        const float w = s.v - bg;
        ss += w;
        sr += s.r;
        mz += (float)s.dz * w;
        mx += (float)s.dx * w;
        my += (float)s.dy * w;
      }
    }
    vg = wsum(vg); ss = wsum(ss); sr = wsum(sr);
    mz = wsum(mz); mx = wsum(mx); my = wsum(my);
    const float ms = ss / fn, mr = sr / fn;
    // pass 3: variances of s and raw
    float vs2 = 0.f, vr2 = 0.f;
    run = 0;
#pragma unroll
    for (int ch = 0; ch < nch; ++ch) {
      const Sample s = sample(ch);
      if (s.ok) {
        const float a = (s.v - bg) - ms, b = s.r - mr;
        vs2 += a * a;
        vr2 += b * b;
      }
    }
    vs2 = wsum(vs2); vr2 = wsum(vr2);
    const float stdg = sqrtf(vg / fn);  // no guard: 0/0 -> NaN as in maxima.cu:249-254
    const float sds = sqrtf(vs2 / fn), sdr = sqrtf(vr2 / fn);
    const float ds = sds > 0.f ? sds : 1.0f, dr = sdr > 0.f ? sdr : 1.0f;
    // pass 4: correlations
    float hn = 0.f, aa = 0.f;
    run = 0;
#pragma unroll
    for (int ch = 0; ch < nch; ++ch) {
      const Sample s = sample(ch);
      if (s.ok) {
        const float ng = (s.g - mg) / stdg;
        hn += ((s.v - bg) - ms) / ds * ng;
        aa += (s.r - mr) / dr * ng;
      }
    }
    hn = wsum(hn); aa = wsum(aa);
    if (lane == 0) {
      float zc = (float)z0, xc = (float)x0, yc = (float)y0;
      if (ss > 0.f) {  // centroid over the unreflected coordinates (maxima.cu:308-325)
        zc = (float)z0 + mz / ss;
        xc = (float)x0 + mx / ss;
        yc = (float)y0 + my / ss;
      }
      const long long ci = z0 * V.sZ + x0 * V.sX + y0;
      o[0] = zc;
      o[1] = xc;
      o[2] = yc;
      o[3] = bg;
      o[4] = aa / fn;
      o[5] = raw.at(rc, z0, x0, y0);
      o[6] = hn / fn;
      o[7] = sign * V.p[ci];
    }
  }
}

// k_fit with the samples of spheres of up to 2 (delta_fit <= 3) or 9 (delta_fit
// <= 5) chunks of 64 held in registers.
template <class VS, class RS>
static void launch_fit(const VS& vs, float sign, const RS& rs, const int* coords, const int* counts, int cap,
                       const int* offs, int P, float sigmaZ, float sigmaXY, float* out8, dim3 grid, hipStream_t st) {
  const int nch = (P + 63) / 64;
  if (nch <= 2)
    hipLaunchKernelGGL((k_fit<VS, RS, 2>), grid, dim3(256), 0, st, vs, sign, rs, coords, counts, cap, offs, P, sigmaZ,
                       sigmaXY, out8);
  else if (nch <= 9)
    hipLaunchKernelGGL((k_fit<VS, RS, 9>), grid, dim3(256), 0, st, vs, sign, rs, coords, counts, cap, offs, P, sigmaZ,
                       sigmaXY, out8);
  else
    hipLaunchKernelGGL((k_fit<VS, RS, 0>), grid, dim3(256), 0, st, vs, sign, rs, coords, counts, cap, offs, P, sigmaZ,
                       sigmaXY, out8);
}

// ------------------------------------------------------------- keep filter
// __main__.py:256-261: keep rows with ov <= fitted x,y < view extent - ov,
// then shift to stack coordinates (float32 adds), stacked in tile order.
__device__ __forceinline__ bool keep_row(const float* r, float lo, float hx, float hy) {
  return r[1] >= lo && r[2] >= lo && r[1] < hx && r[2] < hy;
}

__global__ __launch_bounds__(256) void k_keep_count(const float* __restrict__ fit, const int* __restrict__ counts,
                                                    int cap, const mm_tile_dev* __restrict__ tiles, int ov,
                                                    int* __restrict__ kept) {
  const int v = blockIdx.x;
  const int n = min(counts[v], cap);
  const float lo = (float)ov, hx = (float)(tiles[v].xe - ov), hy = (float)(tiles[v].ye - ov);
  int c = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) c += keep_row(fit + ((long long)v * cap + i) * 8, lo, hx, hy);
  __shared__ int red[256];
  red[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) kept[v] = red[0];
}

__global__ __launch_bounds__(256) void k_keep_write(const float* __restrict__ fit, const int* __restrict__ counts,
                                                    int cap, const mm_tile_dev* __restrict__ tiles, int ov,
                                                    const int* __restrict__ kept, int ntiles, float* __restrict__ Xh,
                                                    int out_cap, int* __restrict__ total,
                                                    int* __restrict__ max_count) {
  const int v = blockIdx.x;
  __shared__ int base_s;
  __shared__ int wc[4];
  if (threadIdx.x == 0) {
    int b = 0;
    for (int u = 0; u < v; ++u) b += kept[u];
    base_s = b;
    if (v == ntiles - 1) {
      *total = b + kept[v];
      if (max_count) {  // the largest candidate count: above cap, the batch is redone with a larger cap
        int m = 0;
        for (int u = 0; u < ntiles; ++u) m = max(m, counts[u]);
        *max_count = m;
      }
    }
  }
  __syncthreads();
  int base = base_s;
  const int n = min(counts[v], cap);
  const mm_tile_dev td = tiles[v];
  const float lo = (float)ov, hx = (float)(td.xe - ov), hy = (float)(td.ye - ov);
  const float offx = (float)(td.ox - ov), offy = (float)(td.oy - ov);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i0 = 0; i0 < n; i0 += blockDim.x) {
    const int i = i0 + threadIdx.x;
    const float* r = fit + ((long long)v * cap + i) * 8;
    const bool k = i < n && keep_row(r, lo, hx, hy);
    const unsigned long long b = __ballot(k);
    if (lane == 0) wc[wave] = __popcll(b);
    __syncthreads();
    int pre = 0, tot = 0;
    for (int w = 0; w < 4; ++w) { if (w < wave) pre += wc[w]; tot += wc[w]; }
    if (k) {
      const int pos = base + pre + __popcll(b & ((1ull << lane) - 1ull));
      if (pos < out_cap) {
        float* o = Xh + (long long)pos * 8;
        o[0] = r[0];
        o[1] = r[1] + offx;
        o[2] = r[2] + offy;
        for (int q = 3; q < 8; ++q) o[q] = r[q];
      }
    }
    base += tot;
    __syncthreads();
  }
}

// ------------------------------------------------------------------ host
extern "C" size_t mm_maxima_ws_bytes(int D, int H, int W, int nvol) {
  if (D <= 0 || H <= 0 || W <= 0 || nvol <= 0) return 256;
  const size_t ww = ((size_t)W + 63) / 64;
  return (size_t)nvol * D * H * (ww * 8 + 4) + 256;
}

// sphere offsets of delta_fit in dz -> dx -> dy order, packed (dz+128)<<16 | (dx+128)<<8 | (dy+128)
static int sphere_table(int dfit, int** dev_out, int* P_out) {
  static std::mutex mu;
  static int* cache[32] = {nullptr};
  static int cacheP[32] = {0};
  static int cache_dev[32] = {0};
  if (dfit < 0 || dfit > 31) { mm_set_error("delta_fit %d outside [0, 31]", dfit); return MM_EINVAL; }
  int dev = 0;
  MM_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if (!cache[dfit] || cache_dev[dfit] != dev) {
    std::vector<int> h;
    for (int dz = -dfit; dz <= dfit; ++dz)
      for (int dx = -dfit; dx <= dfit; ++dx)
        for (int dy = -dfit; dy <= dfit; ++dy)
          if (dz * dz + dx * dx + dy * dy <= dfit * dfit)
            h.push_back(((dz + 128) << 16) | ((dx + 128) << 8) | (dy + 128));
    int* d = nullptr;
    MM_CHECK(hipMalloc(&d, sizeof(int) * h.size()));
    MM_CHECK(hipMemcpy(d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    cache[dfit] = d;
    cacheP[dfit] = (int)h.size();
    cache_dev[dfit] = dev;
  }
  *dev_out = cache[dfit];
  *P_out = cacheP[dfit];
  return MM_OK;
}

// D, H, W: bounds of every volume's extents (the workspace layout).
// marked: the workspace's masks already hold the threshold bits of every row in
// the candidate window (k_blur_sub_xy with a mark output), rows z H + x.
template <class VS>
static int run_maxima(VS vs, int nvol, int D, int H, int W, float sign, float thr, int delta, int* coords, int cap,
                      int* counts, void* ws, size_t ws_bytes, hipStream_t st, bool marked = false) {
  if (delta < 0 || delta > 31) { mm_set_error("delta %d outside [0, 31]", delta); return MM_EINVAL; }
  if (nvol == 0) return MM_OK;
  if (ws_bytes < mm_maxima_ws_bytes(D, H, W, nvol)) { mm_set_error("maxima workspace too small"); return MM_EINVAL; }
  const int nrow = D * H;
  const int ww = W > 0 ? (W + 63) / 64 : 0;
  unsigned long long* masks = (unsigned long long*)ws;
  int* rcnt = (int*)(masks + (size_t)nvol * nrow * ww);
  const unsigned nbx = (unsigned)((nrow + 3) / 4);
  if (nbx == 0 || ww == 0) {  // empty volumes: zero counts without a memset node
    hipLaunchKernelGGL(k_max_scan, dim3(nvol), dim3(1024), 0, st, rcnt, 0, counts);
    MM_LAUNCH_CHECK();
    return MM_OK;
  }
  if (marked) {
    hipLaunchKernelGGL((k_max_mark_above<VS>), dim3((nrow + 255) / 256, nvol), dim3(256), 0, st, vs, sign, delta,
                       masks, rcnt, nrow, H, ww);
  } else {
    hipLaunchKernelGGL((k_max_mark<VS>), dim3(nbx, nvol), dim3(256), 0, st, vs, sign, thr, delta, masks, rcnt, nrow,
                       H, ww);
  }
  MM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_max_scan, dim3(nvol), dim3(1024), 0, st, rcnt, nrow, counts);
  MM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_max_compact, dim3((nrow + 255) / 256, nvol), dim3(256), 0, st, masks, rcnt, counts, nrow, H,
                     ww, coords, cap);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_local_maxima(const float* img, int D, int H, int W, long long sZ, long long sX, float sign,
                               float threshold, int delta, int* coords, int cap, int* count, void* ws, size_t ws_bytes,
                               void* stream) {
  if (!img || !coords || !count || D < 0 || H < 0 || W < 0) { mm_set_error("mm_local_maxima: bad argument"); return MM_EINVAL; }
  OneVol vs{VolDesc{img, D, H, W, sZ, sX, 0, 0x7fffffff}};
  return run_maxima(vs, 1, D, H, W, sign, threshold, delta, coords, cap, count, ws, ws_bytes,
                    (hipStream_t)stream);
}

extern "C" int mm_local_maxima_pm(const float* img, int D, int H, int W, long long sZ, long long sX,
                                  float threshold, int delta, int* coords, int cap, int* counts, void* ws,
                                  size_t ws_bytes, void* stream) {
  if (!img || !coords || !counts || !ws || D < 0 || H < 0 || W < 0 || cap < 0) {
    mm_set_error("mm_local_maxima_pm: bad argument");
    return MM_EINVAL;
  }
  if (delta < 0 || delta > 31) { mm_set_error("delta %d outside [0, 31]", delta); return MM_EINVAL; }
  if (ws_bytes < mm_maxima_ws_bytes(D, H, W, 2)) { mm_set_error("maxima workspace too small"); return MM_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  OneVol vs{VolDesc{img, D, H, W, sZ, sX, 0, 0x7fffffff}};
  const int nrow = D * H;
  const int ww = W > 0 ? (W + 63) / 64 : 0;
  unsigned long long* masks = (unsigned long long*)ws;
  int* rcnt = (int*)(masks + (size_t)2 * nrow * ww);
  if (nrow == 0 || ww == 0) {  // empty volume: zero counts without a memset node
    hipLaunchKernelGGL(k_max_scan, dim3(2), dim3(1024), 0, st, rcnt, 0, counts);
    MM_LAUNCH_CHECK();
    return MM_OK;
  }
  hipLaunchKernelGGL((k_max_mark_pm<OneVol>), dim3((unsigned)((nrow + 3) / 4)), dim3(256), 0, st, vs, threshold,
                     delta, masks, rcnt, nrow, H, ww);
  MM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_max_scan, dim3(2), dim3(1024), 0, st, rcnt, nrow, counts);
  MM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_max_compact, dim3((nrow + 255) / 256, 2), dim3(256), 0, st, masks, rcnt, counts, nrow, H, ww,
                     coords, cap);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_delta_fit(const float* img, int D, int H, int W, long long sZ, long long sX, float sign,
                            const void* raw, int raw_type, long long rZ, long long rX, const int* coords,
                            const int* count, int cap, int delta_fit, float sigmaZ, float sigmaXY, float* out8,
                            void* stream) {
  if (!img || !raw || !coords || !count || !out8) { mm_set_error("mm_delta_fit: null argument"); return MM_EINVAL; }
  if (raw_type != MM_PIX_U16 && raw_type != MM_PIX_F32 && raw_type != MM_PIX_U8SQ) {
    mm_set_error("mm_delta_fit: raw_type %d is not 0 (uint16), 1 (float32) or 2 (squared uint8)", raw_type);
    return MM_EINVAL;
  }
  int* offs = nullptr;
  int P = 0;
  int rc = sphere_table(delta_fit, &offs, &P);
  if (rc) return rc;
  OneVol vs{VolDesc{img, D, H, W, sZ, sX, 0, 0x7fffffff}};
  RawPlain rs{raw, raw_type, rZ, rX};
  int blocks = (cap + 3) / 4;
  if (blocks > 2048) blocks = 2048;
  if (blocks < 1) blocks = 1;
  launch_fit(vs, sign, rs, coords, count, cap, offs, P, sigmaZ, sigmaXY, out8, dim3(blocks, 1), (hipStream_t)stream);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_tiles_maxima(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const float* res,
                               float threshold, int delta, int* coords, int cap, int* counts, void* ws,
                               size_t ws_bytes, void* stream) {
  if (!d || !tiles_dev || !res) { mm_set_error("mm_tiles_maxima: bad argument"); return MM_EINVAL; }
  TileVols vs{(const mm_tile_dev*)tiles_dev, res, d->sz, d->T};
  return run_maxima(vs, ntiles, d->sz, d->T, d->T, 1.0f, threshold, delta, coords, cap, counts, ws, ws_bytes,
                    (hipStream_t)stream);
}

extern "C" int mm_tiles_maxima_window(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const float* res,
                                      float threshold, int delta, int lo, int hi, int* coords, int cap, int* counts,
                                      void* ws, size_t ws_bytes, void* stream) {
  if (!d || !tiles_dev || !res || lo < 0 || hi <= lo) {
    mm_set_error("mm_tiles_maxima_window: bad argument (lo %d hi %d)", lo, hi);
    return MM_EINVAL;
  }
  TileVols vs{(const mm_tile_dev*)tiles_dev, res, d->sz, d->T, lo, hi};
  return run_maxima(vs, ntiles, d->sz, d->T, d->T, 1.0f, threshold, delta, coords, cap, counts, ws, ws_bytes,
                    (hipStream_t)stream);
}

extern "C" int mm_tiles_deconv_maxima_window(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const void* W,
                                             int n_psf, int tiles_per_block, void* spec_ws, float* res, int dlo,
                                             int dhi, float threshold, int delta, int lo, int hi, int* coords, int cap,
                                             int* counts, void* ws, size_t ws_bytes, void* stream) {
  if (!d || !tiles_dev || !W || !spec_ws || !res || dlo < 0 || dhi > d->T || dlo >= dhi || lo < dlo || hi > dhi ||
      hi <= lo) {
    mm_set_error("mm_tiles_deconv_maxima_window: bad argument (deconvolution [%d, %d) maxima [%d, %d) T %d)", dlo, dhi,
                 lo, hi, d ? d->T : 0);
    return MM_EINVAL;
  }
  if (delta < 0 || delta > 31) { mm_set_error("delta %d outside [0, 31]", delta); return MM_EINVAL; }
  // the neighbour tests read out at [lo - delta, hi + delta): inside the computed window, or up to the
  // tile edge where the window reaches it (the tests reflect there)
  if ((dlo > 0 && lo - delta < dlo) || (dhi < d->T && hi + delta > dhi)) {
    mm_set_error("mm_tiles_deconv_maxima_window: maxima [%d, %d) with delta %d reads outside the window [%d, %d)",
                 lo, hi, delta, dlo, dhi);
    return MM_EINVAL;
  }
  if (ntiles == 0) return MM_OK;
  const int T = d->T, sz = d->sz;
  if (ws_bytes < mm_maxima_ws_bytes(sz, T, T, ntiles)) { mm_set_error("maxima workspace too small"); return MM_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  int marked = 0;
  int rc = deconv_tiles_marked(d, tiles_dev, ntiles, W, n_psf, tiles_per_block, spec_ws, res, dlo, dhi,
                               (unsigned long long*)ws, threshold, &marked, stream);
  if (rc) return rc;
  TileVols vs{(const mm_tile_dev*)tiles_dev, res, sz, T, lo, hi};
  return run_maxima(vs, ntiles, sz, T, T, 1.0f, threshold, delta, coords, cap, counts, ws, ws_bytes, st, marked != 0);
}

extern "C" int mm_tiles_blur_maxima_window(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles,
                                           const float* res, float* out, int blur_delta, int blo, int bhi,
                                           float threshold, int delta, int lo, int hi, int* coords, int cap,
                                           int* counts, void* ws, size_t ws_bytes, void* stream) {
  if (!d || !tiles_dev || !res || !out || blo < 0 || bhi > d->T || blo >= bhi || lo < blo || hi > bhi || hi <= lo) {
    mm_set_error("mm_tiles_blur_maxima_window: bad argument (blur [%d, %d) maxima [%d, %d) T %d)", blo, bhi, lo, hi,
                 d ? d->T : 0);
    return MM_EINVAL;
  }
  if (delta < 0 || delta > 31) { mm_set_error("delta %d outside [0, 31]", delta); return MM_EINVAL; }
  // the neighbour tests read out at [lo - delta, hi + delta): inside the blurred window, or up to
  // the tile edge where the window reaches it (the tests reflect there)
  if ((blo > 0 && lo - delta < blo) || (bhi < d->T && hi + delta > bhi)) {
    mm_set_error("mm_tiles_blur_maxima_window: maxima [%d, %d) with delta %d reads outside the blur window [%d, %d)",
                 lo, hi, delta, blo, bhi);
    return MM_EINVAL;
  }
  if (ntiles == 0) return MM_OK;
  const int T = d->T, sz = d->sz;
  if (ws_bytes < mm_maxima_ws_bytes(sz, T, T, ntiles)) { mm_set_error("maxima workspace too small"); return MM_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  // the masks of run_maxima's workspace, rows z T + x of each tile = rows of the blur's planes
  int rc = mm_blur_sub_mark(res, out, ntiles * sz, T, blur_delta, blo, bhi, 1.0f, threshold,
                            (unsigned long long*)ws, st);
  if (rc) return rc;
  TileVols vs{(const mm_tile_dev*)tiles_dev, out, sz, T, lo, hi};
  return run_maxima(vs, ntiles, sz, T, T, 1.0f, threshold, delta, coords, cap, counts, ws, ws_bytes, st, true);
}

extern "C" int mm_tiles_fit(mm_deconv* d, const mm_tile_desc* tiles_dev, int ntiles, const float* res,
                            const int* coords, const int* counts, int cap, int delta_fit, float sigmaZ, float sigmaXY,
                            float* out8, void* stream) {
  if (!d || !tiles_dev || !res) { mm_set_error("mm_tiles_fit: bad argument"); return MM_EINVAL; }
  if (ntiles == 0) return MM_OK;
  const int T = d->T;
  int* offs = nullptr;
  int P = 0;
  int rc = sphere_table(delta_fit, &offs, &P);
  if (rc) return rc;
  TileVols vs{(const mm_tile_dev*)tiles_dev, res, d->sz, T};
  RawTile rs{(const mm_tile_dev*)tiles_dev, T};
  int blocks = (cap + 3) / 4;
  if (blocks > 256) blocks = 256;
  launch_fit(vs, 1.0f, rs, coords, counts, cap, offs, P, sigmaZ, sigmaXY, out8, dim3(blocks, ntiles),
             (hipStream_t)stream);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

extern "C" int mm_tiles_keep(const mm_tile_desc* tiles_dev, int ntiles, int overlap, const float* fit,
                             const int* counts, int cap, int* kept_ws, float* Xh, int out_cap, int* total,
                             int* max_count, void* stream) {
  if (ntiles < 0 || (ntiles > 0 && (!tiles_dev || !fit || !counts || !kept_ws || !Xh || !total))) {
    mm_set_error("mm_tiles_keep: bad argument");
    return MM_EINVAL;
  }
  if (ntiles == 0) return MM_OK;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_keep_count, dim3(ntiles), dim3(256), 0, st, fit, counts, cap, (const mm_tile_dev*)tiles_dev,
                     overlap, kept_ws);
  MM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_keep_write, dim3(ntiles), dim3(256), 0, st, fit, counts, cap, (const mm_tile_dev*)tiles_dev,
                     overlap, kept_ws, ntiles, Xh, out_cap, total, max_count);
  MM_LAUNCH_CHECK();
  return MM_OK;
}

// ------------------------------------------------------- allocation-free maxima
// The workspace of mm_local_maxima_noalloc: a library-owned array that every GPU
// the library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol):
// it passes the device address hipGetSymbolAddress returns, kept per device.
static __device__ __align__(256) unsigned char scratch[MM_MAXIMA_SCRATCH_BYTES];

#define MM_SCRATCH_MAX_DEVICES 64

// Device address of `scratch` on the current device, looked up once per device.
static int scratch_address(void** out) {
  static std::mutex mu;
  static void* addr[MM_SCRATCH_MAX_DEVICES] = {nullptr};
  int dev = 0;
  MM_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MM_SCRATCH_MAX_DEVICES) {
    mm_set_error("mm_local_maxima_noalloc: device %d beyond the %d the library tracks", dev, MM_SCRATCH_MAX_DEVICES);
    return MM_EINVAL;
  }
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    MM_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      mm_set_error("mm_local_maxima_noalloc: no device address for the scratch array");
      return MM_EINVAL;
    }
    addr[dev] = p;
  }
  *out = addr[dev];
  return MM_OK;
}

extern "C" int mm_local_maxima_noalloc(const float* img, int D, int H, int W, long long sZ, long long sX, float sign,
                                       float threshold, int delta, int* coords, int cap, int* count, void* stream) {
  if (!img || !coords || !count || D < 0 || H < 0 || W < 0) {
    mm_set_error("mm_local_maxima_noalloc: bad argument");
    return MM_EINVAL;
  }
  const size_t need = mm_maxima_ws_bytes(D, H, W, 1);
  if (need > (size_t)MM_MAXIMA_SCRATCH_BYTES) {
    mm_set_error("mm_local_maxima_noalloc: a %d x %d x %d volume needs %zu workspace bytes, over the %zu of the "
                 "library's scratch array (use mm_local_maxima with a workspace)",
                 D, H, W, need, (size_t)MM_MAXIMA_SCRATCH_BYTES);
    return MM_ECAP;
  }
  void* ws = nullptr;
  int rc = scratch_address(&ws);
  if (rc) return rc;
  // every pass writes the workspace words and row counts it later reads: nothing has to start at zero
  OneVol vs{VolDesc{img, D, H, W, sZ, sX, 0, 0x7fffffff}};
  return run_maxima(vs, 1, D, H, W, sign, threshold, delta, coords, cap, count, ws, need, (hipStream_t)stream);
}
