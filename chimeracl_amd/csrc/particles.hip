// Particle kernels: coordinate push, cell keys + stable counting sort,
// longitudinal density profile.  gfx950, wave64.
//
// Sort design (replaces index_and_sum_in_cell + PyOpenCL cumsum + the
// atomic `sort` kernel, particles_generic.cl:88-126, 186-200):
//   1. zero the per-bin counters
//   2. key kernel: cell key per particle, run-aggregated atomic histogram
//      (one atomic per run of equal keys inside a wave), and the number of
//      out-of-domain particles per 256-particle block
//   3. one exclusive scan over [histogram | per-block out-of-domain counts]
//   4. scatter: valid keys take a slot from an atomic per-bin cursor
//      (run-aggregated, so each run lands contiguously in index order);
//      out-of-domain particles take a deterministic slot from the block scan
//   5. in-cell fix-up: a bin whose slots are not in ascending particle order
//      (runs from different waves raced) is sorted in place -- in registers
//      through global memory for <= 64 particles, in LDS by one workgroup
//      (bitonic) for up to 16384.
// The result is exactly numpy.argsort(key, kind='stable').
#include "chm_common.h"

#define SORT_BS 256
#define SCAN_ITEMS 8
#define SCAN_BS 256
#define SCAN_TILE (SCAN_BS * SCAN_ITEMS)
#define FIX_SMALL 64
#define FIX_BIG 16384

// ------------------------------------------------------------------ push
template <typename T>
__global__ void __launch_bounds__(256) k_push_x(T *__restrict__ x, T *__restrict__ y, T *__restrict__ z,
                                                const T *__restrict__ px, const T *__restrict__ py,
                                                const T *__restrict__ pz, const T *__restrict__ g_inv,
                                                T dt, int64_t Np) {
  int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= Np) return;
  // particles_generic.cl:143-151, same rounding order
  T dt_g = dt * g_inv[ip];
  T dx = px[ip] * dt_g;
  T dy = py[ip] * dt_g;
  T dz = pz[ip] * dt_g;
  x[ip] = x[ip] + dx;
  y[ip] = y[ip] + dy;
  z[ip] = z[ip] + dz;
}

template <typename T>
static int push_x(T *x, T *y, T *z, const T *px, const T *py, const T *pz, const T *g_inv, T dt,
                  int64_t Np, void *stream) {
  if (Np < 0) return CHM_EINVAL;
  if (Np == 0) return CHM_OK;
  k_push_x<T><<<chm_blocks(Np, 256), 256, 0, (hipStream_t)stream>>>(x, y, z, px, py, pz, g_inv, dt, Np);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_push_x_f64(double *x, double *y, double *z, const double *px, const double *py,
                              const double *pz, const double *g_inv, double dt, int64_t Np,
                              void *stream) {
  return push_x<double>(x, y, z, px, py, pz, g_inv, dt, Np, stream);
}
extern "C" int chm_push_x_f32(float *x, float *y, float *z, const float *px, const float *py,
                              const float *pz, const float *g_inv, float dt, int64_t Np,
                              void *stream) {
  return push_x<float>(x, y, z, px, py, pz, g_inv, dt, Np, stream);
}

// ------------------------------------------------------------------ keys
__device__ __forceinline__ double dev_sqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float dev_sqrt(float v) { return sqrtf(v); }
__device__ __forceinline__ double dev_floor(double v) { return floor(v); }
__device__ __forceinline__ float dev_floor(float v) { return floorf(v); }

// Cell key of index_and_sum_in_cell, particles_generic.cl:105-124.
template <typename T>
__device__ __forceinline__ uint32_t cell_key(T xp, T yp, T zp, int Nx_loc, int Nr_loc, T xmin,
                                             T dx_inv, T rmin, T dr_inv) {
  T r = dev_sqrt(yp * yp + zp * zp);
  int ix = (int)dev_floor((xp - xmin) * dx_inv);
  int ir = (int)dev_floor((r - rmin) * dr_inv);
  if (ix > 0 && ix < Nx_loc - 1 && ir < Nr_loc - 1 && ir >= 0) return (uint32_t)(ix + ir * Nx_loc);
  return (uint32_t)(Nr_loc * Nx_loc);
}

// Run structure of a wave: lane is a head if its key differs from the lane
// below.  Returns the head mask; `len` = run length for head lanes.
__device__ __forceinline__ unsigned long long wave_runs(uint32_t rkey, int lane, int *len,
                                                        int *head_lane) {
  uint32_t prev = (uint32_t)__shfl_up((int)rkey, 1);
  bool head = (lane == 0) || (prev != rkey);
  unsigned long long heads = __ballot(head);
  unsigned long long above = (lane == 63) ? 0ull : (heads & (~0ull << (lane + 1)));
  int next = above ? (__ffsll((long long)above) - 1) : 64;
  *len = next - lane;
  unsigned long long below = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
  *head_lane = 63 - __clzll((long long)below);
  return heads;
}

template <typename T>
__global__ void __launch_bounds__(SORT_BS) k_cell_key(const T *__restrict__ x, const T *__restrict__ y,
                                                      const T *__restrict__ z, int64_t Np, int Nx, int Nr,
                                                      T xmin, T dx_inv, T rmin, T dr_inv,
                                                      uint32_t *__restrict__ key_out,
                                                      uint32_t *__restrict__ hist,
                                                      uint32_t *__restrict__ ovf_cnt) {
  __shared__ uint32_t wcnt[SORT_BS / CHM_WAVE];
  const int Nx_loc = Nx - 1, Nr_loc = Nr - 1;
  const uint32_t ovf = (uint32_t)(Nr_loc * Nx_loc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t ip = (int64_t)blockIdx.x * SORT_BS + threadIdx.x;
  bool act = ip < Np;
  uint32_t key = ovf;
  if (act) {
    key = cell_key<T>(x[ip], y[ip], z[ip], Nx_loc, Nr_loc, xmin, dx_inv, rmin, dr_inv);
    key_out[ip] = key;
  }
  bool valid = act && key != ovf;
  uint32_t rkey = valid ? key : (0xFFFFFFFFu - (uint32_t)lane);
  int len, hl;
  wave_runs(rkey, lane, &len, &hl);
  if (valid && hl == lane) atomicAdd(&hist[key], (uint32_t)len);
  unsigned long long ob = __ballot(act && !valid);
  if (lane == 0) wcnt[wave] = (uint32_t)__popcll((long long)ob);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < SORT_BS / CHM_WAVE; w++) s += wcnt[w];
    ovf_cnt[blockIdx.x] = s;
  }
}

__global__ void k_zero_u32(uint32_t *__restrict__ a, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// ------------------------------------------------------------------ scan
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
  // 256 threads; returns the exclusive prefix of v over the block.
  __shared__ uint32_t wsum[SCAN_BS / CHM_WAVE];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = (uint32_t)__shfl_up((int)inc, d);
    if (lane >= d) inc += t;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint32_t wpre = 0, tot = 0;
  for (int w = 0; w < SCAN_BS / CHM_WAVE; w++) {
    uint32_t s = wsum[w];
    if (w < wave) wpre += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wpre + inc - v;
}

__global__ void __launch_bounds__(SCAN_BS) k_scan_reduce(const uint32_t *__restrict__ in, int64_t n,
                                                         uint32_t *__restrict__ bsum) {
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
  for (int i = 0; i < SCAN_ITEMS; i++) {
    int64_t k = base + (int64_t)i * SCAN_BS + threadIdx.x;
    if (k < n) s += in[k];
  }
  uint32_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_BS) k_scan_blocks(uint32_t *__restrict__ bsum, int64_t nb) {
  uint32_t carry = 0;
  for (int64_t base = 0; base < nb; base += SCAN_BS) {
    int64_t k = base + threadIdx.x;
    uint32_t v = k < nb ? bsum[k] : 0;
    uint32_t tot;
    uint32_t ex = block_excl_scan(v, &tot);
    if (k < nb) bsum[k] = carry + ex;
    carry += tot;
  }
}

// Exclusive scan of `in` (n entries) into out[0..n]; zeroes in[0..nzero).
__global__ void __launch_bounds__(SCAN_BS) k_scan_final(uint32_t *__restrict__ in, int64_t n,
                                                        int64_t nzero, const uint32_t *__restrict__ bsum,
                                                        uint32_t *__restrict__ out) {
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
  for (int i = 0; i < SCAN_ITEMS; i++) {
    int64_t k = base + i;
    v[i] = k < n ? in[k] : 0;
    s += v[i];
  }
  uint32_t tot;
  uint32_t pre = block_excl_scan(s, &tot) + bsum[blockIdx.x];
  for (int i = 0; i < SCAN_ITEMS; i++) {
    int64_t k = base + i;
    if (k < n) {
      out[k] = pre;
      if (k < nzero) in[k] = 0;
    }
    if (k == n - 1) out[n] = pre + v[i];
    pre += v[i];
  }
}

// ------------------------------------------------------------------ scatter
__global__ void __launch_bounds__(SORT_BS) k_scatter(const uint32_t *__restrict__ key, int64_t Np,
                                                     uint32_t nbins, uint32_t *__restrict__ cursor,
                                                     const uint32_t *__restrict__ pre,
                                                     uint32_t *__restrict__ sort_indx) {
  __shared__ uint32_t wcnt[SORT_BS / CHM_WAVE];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t ovf = nbins - 1;
  int64_t ip = (int64_t)blockIdx.x * SORT_BS + threadIdx.x;
  bool act = ip < Np;
  uint32_t k = act ? key[ip] : ovf;
  bool valid = act && k != ovf;
  uint32_t rkey = valid ? k : (0xFFFFFFFFu - (uint32_t)lane);
  int len, hl;
  wave_runs(rkey, lane, &len, &hl);
  uint32_t base = 0;
  if (valid && hl == lane) base = atomicAdd(&cursor[k], (uint32_t)len);
  base = (uint32_t)__shfl((int)base, hl);
  unsigned long long ob = __ballot(act && !valid);
  if (lane == 0) wcnt[wave] = (uint32_t)__popcll((long long)ob);
  __syncthreads();
  if (valid) {
    sort_indx[pre[k] + base + (uint32_t)(lane - hl)] = (uint32_t)ip;
  } else if (act) {
    uint32_t r = (uint32_t)__popcll((long long)(ob & ((1ull << lane) - 1)));
    for (int w = 0; w < wave; w++) r += wcnt[w];
    uint32_t slot = pre[ovf] + (pre[nbins + blockIdx.x] - pre[nbins]) + r;
    sort_indx[slot] = (uint32_t)ip;
  }
}

// ------------------------------------------------------------------ fix-up
__global__ void __launch_bounds__(256) k_fixup(const uint32_t *__restrict__ pre, uint32_t nbins,
                                               uint32_t Np, uint32_t *__restrict__ sort_indx,
                                               uint32_t *__restrict__ big_list,
                                               uint32_t *__restrict__ big_count,
                                               uint32_t *__restrict__ cell_offset,
                                               uint32_t *__restrict__ np_stay) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // the out-of-domain bin is counted per block, after the histogram, so the
  // scan's entry at nbins is Np_stay; the last offset is Np
  if (c < (int64_t)nbins) cell_offset[c] = pre[c];
  if (c == (int64_t)nbins) cell_offset[c] = Np;
  if (c == 0) *np_stay = pre[nbins - 1];
  if (c >= (int64_t)nbins - 1) return;
  uint32_t s = pre[c], e = pre[c + 1];
  if (e - s < 2) return;
  bool sorted = true;
  uint32_t prev = sort_indx[s];
  for (uint32_t i = s + 1; i < e; i++) {
    uint32_t v = sort_indx[i];
    if (v < prev) { sorted = false; break; }
    prev = v;
  }
  if (sorted) return;
  if (e - s <= FIX_SMALL) {
    for (uint32_t i = s + 1; i < e; i++) {
      uint32_t v = sort_indx[i];
      uint32_t j = i;
      while (j > s && sort_indx[j - 1] > v) {
        sort_indx[j] = sort_indx[j - 1];
        j--;
      }
      sort_indx[j] = v;
    }
  } else {
    uint32_t slot = atomicAdd(big_count, 1u);
    big_list[slot] = (uint32_t)c;
  }
}

__global__ void __launch_bounds__(1024) k_bigfix(const uint32_t *__restrict__ pre,
                                                 uint32_t *__restrict__ sort_indx,
                                                 const uint32_t *__restrict__ big_list,
                                                 const uint32_t *__restrict__ big_count,
                                                 uint32_t *__restrict__ flags) {
  __shared__ uint32_t buf[FIX_BIG];
  const uint32_t nbig = *big_count;
  for (uint32_t b = blockIdx.x; b < nbig; b += gridDim.x) {
    uint32_t c = big_list[b];
    uint32_t s = pre[c], n = pre[c + 1] - s;
    if (n > FIX_BIG) {
      if (threadIdx.x == 0) atomicOr(flags, 1u);
      continue;
    }
    uint32_t P = 1;
    while (P < n) P <<= 1;
    for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) buf[t] = t < n ? sort_indx[s + t] : 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t t = threadIdx.x; t < P; t += blockDim.x) {
          uint32_t q = t ^ j;
          if (q > t) {
            uint32_t a = buf[t], bq = buf[q];
            bool up = (t & k) == 0;
            if ((a > bq) == up) {
              buf[t] = bq;
              buf[q] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) sort_indx[s + t] = buf[t];
    __syncthreads();
  }
}

// ------------------------------------------------------------------ driver
struct SortWs {
  uint32_t *cnt, *pre, *bsum, *big_count, *flags, *big_list;
  int64_t nblk, ncnt, nscan;
};

static size_t sort_layout(int64_t Np, int64_t nbins, SortWs *w, void *base) {
  int64_t nblk = (Np + SORT_BS - 1) / SORT_BS;
  int64_t ncnt = nbins + nblk;
  int64_t nscan = (ncnt + SCAN_TILE - 1) / SCAN_TILE;
  size_t off = 0;
  char *b = (char *)base;
  size_t o_cnt = off; off += chm_align(ncnt * 4);
  size_t o_pre = off; off += chm_align((ncnt + 1) * 4);
  size_t o_bsum = off; off += chm_align(nscan * 4);
  size_t o_misc = off; off += 256;  // big_count, flags
  size_t o_big = off; off += chm_align(nbins * 4);
  if (w) {
    w->cnt = (uint32_t *)(b + o_cnt);
    w->pre = (uint32_t *)(b + o_pre);
    w->bsum = (uint32_t *)(b + o_bsum);
    w->big_count = (uint32_t *)(b + o_misc);
    w->flags = (uint32_t *)(b + o_misc + 4);
    w->big_list = (uint32_t *)(b + o_big);
    w->nblk = nblk;
    w->ncnt = ncnt;
    w->nscan = nscan;
  }
  return off;
}

extern "C" size_t chm_sort_workspace_bytes(int64_t Np, int64_t nbins) {
  return sort_layout(Np, nbins, nullptr, nullptr);
}

template <typename T>
static int index_sort(const T *x, const T *y, const T *z, int64_t Np, const chm_grid_t *g,
                      uint32_t *indx_in_cell, uint32_t *cell_offset, uint32_t *sort_indx,
                      uint32_t *np_stay, void *ws, size_t ws_bytes, void *stream) {
  if (!g || Np < 0 || Np > 0xFFFFFFF0ll || g->Nx < 4 || g->Nr < 3) return CHM_EINVAL;
  int64_t nbins = (int64_t)(g->Nx - 1) * (g->Nr - 1) + 1;
  SortWs w;
  if (sort_layout(Np, nbins, &w, ws) > ws_bytes || !ws) return CHM_EWORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  // zero the counters and the big-bin count (flags stay sticky)
  k_zero_u32<<<chm_blocks(nbins, 256), 256, 0, st>>>(w.cnt, nbins);
  k_zero_u32<<<1, 64, 0, st>>>(w.big_count, 1);
  if (Np > 0)
    k_cell_key<T><<<(unsigned)w.nblk, SORT_BS, 0, st>>>(x, y, z, Np, g->Nx, g->Nr, (T)g->xmin,
                                                        (T)g->dx_inv, (T)g->rmin, (T)g->dr_inv,
                                                        indx_in_cell, w.cnt, w.cnt + nbins);
  k_scan_reduce<<<(unsigned)w.nscan, SCAN_BS, 0, st>>>(w.cnt, w.ncnt, w.bsum);
  k_scan_blocks<<<1, SCAN_BS, 0, st>>>(w.bsum, w.nscan);
  k_scan_final<<<(unsigned)w.nscan, SCAN_BS, 0, st>>>(w.cnt, w.ncnt, nbins, w.bsum, w.pre);
  if (Np > 0)
    k_scatter<<<(unsigned)w.nblk, SORT_BS, 0, st>>>(indx_in_cell, Np, (uint32_t)nbins, w.cnt, w.pre,
                                                    sort_indx);
  k_fixup<<<chm_blocks(nbins + 1, 256), 256, 0, st>>>(w.pre, (uint32_t)nbins, (uint32_t)Np, sort_indx, w.big_list,
                                                      w.big_count, cell_offset, np_stay);
  k_bigfix<<<256, 1024, 0, st>>>(w.pre, sort_indx, w.big_list, w.big_count, w.flags);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_index_sort_f64(const double *x, const double *y, const double *z, int64_t Np,
                                  const chm_grid_t *g, uint32_t *indx_in_cell, uint32_t *cell_offset,
                                  uint32_t *sort_indx, uint32_t *np_stay, void *ws, size_t ws_bytes,
                                  void *stream) {
  return index_sort<double>(x, y, z, Np, g, indx_in_cell, cell_offset, sort_indx, np_stay, ws,
                            ws_bytes, stream);
}
extern "C" int chm_index_sort_f32(const float *x, const float *y, const float *z, int64_t Np,
                                  const chm_grid_t *g, uint32_t *indx_in_cell, uint32_t *cell_offset,
                                  uint32_t *sort_indx, uint32_t *np_stay, void *ws, size_t ws_bytes,
                                  void *stream) {
  return index_sort<float>(x, y, z, Np, g, indx_in_cell, cell_offset, sort_indx, np_stay, ws,
                           ws_bytes, stream);
}

extern "C" size_t chm_sort_flags_offset(int64_t Np, int64_t nbins) {
  char *base = (char *)(uintptr_t)4096;  // any aligned base; only the offset is used
  SortWs w;
  sort_layout(Np, nbins, &w, base);
  return (size_t)((char *)w.flags - base);
}

// ------------------------------------------------------------------ profile
// profile_by_interpolant, particles_generic.cl:6-30, constant extension
// outside the knots (where the reference reads past the knot array).
template <typename T>
__global__ void __launch_bounds__(256) k_dens_profile(const T *__restrict__ x, T *__restrict__ w,
                                                      int64_t Np, const double *__restrict__ xx,
                                                      const double *__restrict__ ff,
                                                      const double *__restrict__ dxm1, int n) {
  int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= Np) return;
  double xp = (double)x[ip];
  if (xp < xx[0]) xp = xx[0];
  if (xp > xx[n - 1]) xp = xx[n - 1];
  int ix = 0;
  if (xp > xx[0])
    for (ix = 0; ix < n - 2; ix++)
      if (xp > xx[ix] && xp <= xx[ix + 1]) break;
  double f_minus = ff[ix] * dxm1[ix];
  double f_plus = ff[ix + 1] * dxm1[ix];
  double f = f_minus * (xx[ix + 1] - xp) + f_plus * (xp - xx[ix]);
  w[ip] = (T)((double)w[ip] * f);
}

template <typename T>
static int dens_profile(const T *x, T *w, int64_t Np, const double *xx, const double *ff,
                        const double *dxm1, int n, void *stream) {
  if (Np < 0 || n < 2) return CHM_EINVAL;
  if (Np == 0) return CHM_OK;
  k_dens_profile<T><<<chm_blocks(Np, 256), 256, 0, (hipStream_t)stream>>>(x, w, Np, xx, ff, dxm1, n);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_dens_profile_f64(const double *x, double *w, int64_t Np, const double *xx,
                                    const double *ff, const double *dxm1, int32_t nknots,
                                    void *stream) {
  return dens_profile<double>(x, w, Np, xx, ff, dxm1, nknots, stream);
}
extern "C" int chm_dens_profile_f32(const float *x, float *w, int64_t Np, const double *xx,
                                    const double *ff, const double *dxm1, int32_t nknots,
                                    void *stream) {
  return dens_profile<float>(x, w, Np, xx, ff, dxm1, nknots, stream);
}
