// Particle kernels: coordinate push, cell keys + stable counting sort,
// longitudinal density profile.  gfx950, wave64.
//
// Sort design (replaces index_and_sum_in_cell + PyOpenCL cumsum + the
// atomic `sort` kernel, particles_generic.cl:88-126, 186-200):
//   1. key kernel: cell key per particle (optionally fused with the
//      coordinate push), no atomics
//   2. stable LSD radix sort of the keys with the particle index as value,
//      in one to three passes of 8..11-bit digits (two passes up to 22-bit
//      keys, i.e. 4M cells): per 4096-key tile an LDS digit histogram, one
//      scan over [digit][tile], then a scatter that ranks keys inside the
//      tile in index order (ballot-matched digits per wave, tagged per-wave
//      counts in LDS); per-tile counts are stored one row per tile and
//      scanned by chunks of tiles; the last pass also writes the sorted keys
//   3. cell_offset[c] = lower bound of c in the sorted keys (from the slots
//      where the sorted key changes), Np_stay
// The result is exactly numpy.argsort(key, kind='stable') and the
// reference's cell_offset; no global atomics, no order fix-up.
#include "chm_common.h"

#include <mutex>

#define SORT_BS 256
#define SCAN_ITEMS 8
#define SCAN_BS 256
#define SCAN_TILE (SCAN_BS * SCAN_ITEMS)
#define RS_BS 256
#define RS_ITEMS 16
#define RS_TILE (RS_BS * RS_ITEMS)
static_assert(RS_TILE <= 65535, "per-tile digit counts are 16-bit");

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

// ------------------------------------------------------------------ align
// data_align_dbl / data_align_int, particles_generic.cl:156-186:
// dst[i] = src[sort_indx[i]], i < n, for 8- or 4-byte elements.
template <typename E>
__global__ void __launch_bounds__(256) k_align(const uint32_t *__restrict__ idx, int64_t n,
                                               const E *__restrict__ src, E *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

extern "C" int chm_align_elems(const uint32_t *sort_indx, int64_t n, const void *src, void *dst,
                               int32_t elem_bytes, void *stream) {
  if (n < 0 || (elem_bytes != 4 && elem_bytes != 8) || (n > 0 && (!sort_indx || !src || !dst)) ||
      (src == dst && n > 0))
    return CHM_EINVAL;
  if (n == 0) return CHM_OK;
  hipStream_t st = (hipStream_t)stream;
  if (elem_bytes == 8)
    k_align<uint64_t><<<chm_blocks(n, 256), 256, 0, st>>>(sort_indx, n, (const uint64_t *)src,
                                                          (uint64_t *)dst);
  else
    k_align<uint32_t><<<chm_blocks(n, 256), 256, 0, st>>>(sort_indx, n, (const uint32_t *)src,
                                                          (uint32_t *)dst);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// Several arrays permuted by one launch: the slot's index is read once and
// the loads of all arrays are in flight together.
struct AlignSet {
  const void *src[CHM_ALIGN_MAXARR];
  void *dst[CHM_ALIGN_MAXARR];
  int eb[CHM_ALIGN_MAXARR];
  int narr;
};

__global__ void __launch_bounds__(256) k_align_multi(const uint32_t *__restrict__ idx, int64_t n,
                                                     AlignSet as) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = idx[i];
  uint64_t v[CHM_ALIGN_MAXARR];
#pragma unroll
  for (int a = 0; a < CHM_ALIGN_MAXARR; a++)
    if (a < as.narr)
      v[a] = as.eb[a] == 8 ? ((const uint64_t *)as.src[a])[p]
                           : (uint64_t)((const uint32_t *)as.src[a])[p];
#pragma unroll
  for (int a = 0; a < CHM_ALIGN_MAXARR; a++)
    if (a < as.narr) {
      if (as.eb[a] == 8) ((uint64_t *)as.dst[a])[i] = v[a];
      else ((uint32_t *)as.dst[a])[i] = (uint32_t)v[a];
    }
}

extern "C" int chm_align_multi(const uint32_t *sort_indx, int64_t n, int32_t narr,
                               const void *const *src, void *const *dst,
                               const int32_t *elem_bytes, void *stream) {
  if (n < 0 || narr < 1 || narr > CHM_ALIGN_MAXARR || !src || !dst || !elem_bytes)
    return CHM_EINVAL;
  AlignSet as = {};
  as.narr = narr;
  for (int a = 0; a < narr; a++) {
    if ((elem_bytes[a] != 4 && elem_bytes[a] != 8) || (n > 0 && (!src[a] || !dst[a])))
      return CHM_EINVAL;
    as.src[a] = src[a];
    as.dst[a] = dst[a];
    as.eb[a] = elem_bytes[a];
  }
  // no destination may be read by the same launch
  for (int a = 0; a < narr; a++)
    for (int b = 0; b < narr; b++)
      if (n > 0 && (dst[a] == src[b] || (a != b && dst[a] == dst[b]))) return CHM_EINVAL;
  if (n == 0) return CHM_OK;
  if (!sort_indx) return CHM_EINVAL;
  k_align_multi<<<chm_blocks(n, 256), 256, 0, (hipStream_t)stream>>>(sort_indx, n, as);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// ------------------------------------------------------------------ keys
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

// Wave-aggregated increment: lanes with act and equal consecutive values
// v form runs; each run's head adds its length to slot[v] with one atomic.
__device__ __forceinline__ void run_add(uint32_t *slot_base, bool act, uint32_t v, int lane) {
  const uint32_t rv = act ? v : (0xFFFFFFFFu - (uint32_t)lane);
  int len, hl;
  wave_runs(rv, lane, &len, &hl);
  if (act && hl == lane) atomicAdd(&slot_base[v], (uint32_t)len);
}

// With PUSH, the coordinate push of push_xyz (same rounding order) is applied
// first and the key is taken from the pushed position.  Keys only: the cell
// offsets come from the sorted keys, so this pass has no atomics and its cost
// does not depend on the particles' order.
template <typename T, bool PUSH>
__global__ void __launch_bounds__(SORT_BS) k_cell_key(T *__restrict__ x, T *__restrict__ y,
                                                      T *__restrict__ z, const T *__restrict__ px,
                                                      const T *__restrict__ py, const T *__restrict__ pz,
                                                      const T *__restrict__ g_inv, T dt, int64_t Np,
                                                      int Nx, int Nr, T xmin, T dx_inv, T rmin, T dr_inv,
                                                      uint32_t *__restrict__ key_out) {
  const int Nx_loc = Nx - 1, Nr_loc = Nr - 1;
  const int64_t ip = (int64_t)blockIdx.x * SORT_BS + threadIdx.x;
  if (ip >= Np) return;
  T xp = x[ip], yp = y[ip], zp = z[ip];
  if (PUSH) {
    T dt_g = dt * g_inv[ip];
    T dx = px[ip] * dt_g;
    T dy = py[ip] * dt_g;
    T dz = pz[ip] * dt_g;
    xp = xp + dx;
    yp = yp + dy;
    zp = zp + dz;
    x[ip] = xp;
    y[ip] = yp;
    z[ip] = zp;
  }
  key_out[ip] = cell_key<T>(xp, yp, zp, Nx_loc, Nr_loc, xmin, dx_inv, rmin, dr_inv);
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

// Exclusive scan of `in` (n entries) into out[0..n]; zeroes in[0..nzero).
// bsum holds the tile sums of k_scan_reduce; each block adds up those of
// the tiles before it.
__global__ void __launch_bounds__(SCAN_BS) k_scan_final(uint32_t *__restrict__ in, int64_t n,
                                                        int64_t nzero, const uint32_t *__restrict__ bsum,
                                                        uint32_t *__restrict__ out) {
  uint32_t bs = 0;
  for (int64_t i = threadIdx.x; i < (int64_t)blockIdx.x; i += SCAN_BS) bs += bsum[i];
  uint32_t bpre;
  block_excl_scan(bs, &bpre);
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
  for (int i = 0; i < SCAN_ITEMS; i++) {
    int64_t k = base + i;
    v[i] = k < n ? in[k] : 0;
    s += v[i];
  }
  uint32_t tot;
  uint32_t pre = block_excl_scan(s, &tot) + bpre;
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

// ------------------------------------------------------------------ offsets
// cell_offset[c] = first sorted slot whose key is >= c, c = 0..nbins, and
// Np_stay = cell_offset[nbins - 1], the start of the out-of-domain bin.
// Slot s starts the cells key[s-1] < c <= key[s] (key[-1] = -1,
// key[Np] = nbins); these intervals partition 0..nbins.  A slot writes the
// cells of its interval, at most OFF_RUN of them; the rest of a longer
// interval (a long stretch of empty cells, e.g. the invalid last row before
// the out-of-domain bin) goes to a short list that k_offsets_gaps fills.  If
// the list overflows, k_offsets_gaps recomputes every cell by bisection.
// gl[0] = entries, gl[1] = overflow; both cleared earlier in the same stream
// (k_rs_colsum of the first pass); entries (s, first cell, last cell) from
// gl[2].
#define OFF_RUN 32
#define OFF_LIST 1024

// OFF_PER consecutive slots per lane (two 16-byte key loads).
#define OFF_PER 8
__global__ void __launch_bounds__(256) k_offsets_bounds(const uint32_t *__restrict__ skey, int64_t Np,
                                                        int64_t nbins,
                                                        uint32_t *__restrict__ cell_offset,
                                                        uint32_t *__restrict__ np_stay,
                                                        uint32_t *__restrict__ gl) {
  const int64_t s0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * OFF_PER;
  if (s0 > Np) return;
  uint32_t kv[OFF_PER];
  if (s0 + OFF_PER <= Np) {
    const uint4 a = *(const uint4 *)(skey + s0), b = *(const uint4 *)(skey + s0 + 4);
    kv[0] = a.x; kv[1] = a.y; kv[2] = a.z; kv[3] = a.w;
    kv[4] = b.x; kv[5] = b.y; kv[6] = b.z; kv[7] = b.w;
  } else {
#pragma unroll
    for (int i = 0; i < OFF_PER; i++) kv[i] = s0 + i < Np ? skey[s0 + i] : 0u;
  }
  int64_t prev = s0 > 0 ? (int64_t)skey[s0 - 1] : -1;
#pragma unroll
  for (int i = 0; i < OFF_PER; i++) {
    const int64_t s = s0 + i;
    if (s > Np) break;
    const int64_t cur = s < Np ? (int64_t)kv[i] : nbins;
    if (cur > prev) {
      int64_t hi = cur;
      if (hi > prev + OFF_RUN) {
        hi = prev + OFF_RUN;
        const uint32_t e = atomicAdd(&gl[0], 1u);
        if (e < OFF_LIST) {
          gl[2 + 3 * e] = (uint32_t)s;
          gl[3 + 3 * e] = (uint32_t)(hi + 1);
          gl[4 + 3 * e] = (uint32_t)cur;
        } else {
          gl[1] = 1u;
        }
      }
      for (int64_t c = prev + 1; c <= hi; c++) {
        cell_offset[c] = (uint32_t)s;
        if (c == nbins - 1) *np_stay = (uint32_t)s;
      }
    }
    prev = cur;
  }
}

// Fixed grid; every lane strides over the listed cells (or, after an
// overflow of the list, over all cells with a bisection each).
__global__ void k_offsets_gaps(const uint32_t *__restrict__ skey, int64_t Np, int64_t nbins,
                               uint32_t *__restrict__ cell_offset, uint32_t *__restrict__ np_stay,
                               const uint32_t *__restrict__ gl) {
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nt = (int64_t)gridDim.x * blockDim.x;
  if (gl[1] != 0u) {
    for (int64_t c = t0; c <= nbins; c += nt) {
      int64_t lo = 0, hi = Np;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)skey[mid] < c) lo = mid + 1;
        else hi = mid;
      }
      cell_offset[c] = (uint32_t)lo;
      if (c == nbins - 1) *np_stay = (uint32_t)lo;
    }
    return;
  }
  const uint32_t ne = gl[0] < OFF_LIST ? gl[0] : OFF_LIST;
  for (uint32_t e = 0; e < ne; e++) {
    const uint32_t sv = gl[2 + 3 * e];
    const int64_t c0 = gl[3 + 3 * e], c1 = gl[4 + 3 * e];
    for (int64_t c = c0 + t0; c <= c1; c += nt) {
      cell_offset[c] = sv;
      if (c == nbins - 1) *np_stay = sv;
    }
  }
}

__global__ void k_list_clear(uint32_t *__restrict__ gl) { gl[threadIdx.x] = 0u; }

// ------------------------------------------------------------------ radix sort
// Digits of nbits <= 2^RB buckets; counts[tile * nd + d] (one contiguous
// row per tile), 16 bits: a tile holds RS_TILE = 4096 keys.
template <int RB>
__global__ void __launch_bounds__(RS_BS) k_rs_upsweep(const uint32_t *__restrict__ keys, int64_t n,
                                                      int shift, int nbits,
                                                      uint16_t *__restrict__ counts) {
  constexpr int NB = 1 << RB;
  __shared__ uint32_t h[NB];
  const int nd = 1 << nbits;
  const uint32_t mask = (uint32_t)nd - 1u;
  for (int d = threadIdx.x; d < nd; d += RS_BS) h[d] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  const int lane = threadIdx.x & 63;
  uint32_t kk[RS_ITEMS];
#pragma unroll
  for (int i = 0; i < RS_ITEMS; i++) {
    int64_t idx = base + (int64_t)i * RS_BS + threadIdx.x;
    kk[i] = idx < n ? keys[idx] : 0u;
  }
#pragma unroll
  for (int i = 0; i < RS_ITEMS; i++) {
    int64_t idx = base + (int64_t)i * RS_BS + threadIdx.x;
    const bool act = idx < n;
    run_add(h, act, (kk[i] >> shift) & mask, lane);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < nd; d += RS_BS) counts[(int64_t)blockIdx.x * nd + d] = (uint16_t)h[d];
}

// Lanes of the wave holding the same digit d (active lanes only).  Fast
// paths: one digit in the wave, or non-decreasing digits (equal digits are
// then contiguous); otherwise one ballot per digit bit.
__device__ __forceinline__ unsigned long long match_peers(uint32_t d, bool act, int lane,
                                                          int nbits) {
  const unsigned long long amask = __ballot(act);
  // active lanes are a prefix of the wave; lane 0 is active if any is
  const uint32_t d0 = (uint32_t)__shfl((int)d, 0);
  const uint32_t dprev = (uint32_t)__shfl_up((int)d, 1);
  if (__ballot(act && d == d0) == amask) return amask;
  if (__ballot(act && lane > 0 && dprev > d) == 0ull) {
    const unsigned long long heads = __ballot(act && (lane == 0 || dprev != d));
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int hl = 63 - __clzll((long long)(heads & upto));
    const unsigned long long above = heads & ~upto;
    const int nx = above ? (__ffsll((long long)above) - 1) : 64;
    const unsigned long long hi = (nx == 64) ? ~0ull : ((1ull << nx) - 1ull);
    return hi & ~((1ull << hl) - 1ull) & amask;
  }
  unsigned long long peers = amask;
  for (int b = 0; b < nbits; b++) {
    const bool bit = (d >> b) & 1u;
    const unsigned long long bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  return peers;
}

// Stable scatter of one tile.  Wave w owns the tile's items
// w*RS_ITEMS*64 .. (w+1)*RS_ITEMS*64 - 1, visited in index order (round r,
// lane l: item r*64 + l of the wave's range).  Pass 1: each wave counts its
// digits (one LDS add per run of equal digits).  The waves' counts then
// become their starting positions (tile offset + counts of the waves
// before).  Pass 2: each item goes to its wave's running position of its
// digit plus its rank among the equal digits of its round.  The tile's
// offsets (DPT digits per thread) are loaded with its keys, into registers.
// Three block barriers per tile; within a wave, LDS operations take effect
// in order.  Registers held to 4 waves per SIMD.  (Blocks that walk 2 or 4
// consecutive tiles, carrying their digit positions from tile to tile, read
// and write 2-4x fewer count and offset bytes but measured 6-11% slower on
// 66.8M keys: each block's tiles then run one after the other.)
template <int RB>
__global__ void __launch_bounds__(RS_BS) __attribute__((amdgpu_waves_per_eu(4))) k_rs_downsweep(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, int64_t n, int shift,
    int nbits, const uint32_t *__restrict__ offs, uint32_t *__restrict__ kout,
    uint32_t *__restrict__ vout) {
  constexpr int NB = 1 << RB;
  constexpr int NW = RS_BS / CHM_WAVE;
  constexpr int DPT = (NB + RS_BS - 1) / RS_BS;
  __shared__ uint32_t hw[NW][NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nd = 1 << nbits;
  const uint32_t mask = (uint32_t)nd - 1u;
  const unsigned long long lt = (1ull << lane) - 1ull;
  uint32_t pos[DPT];
#pragma unroll
  for (int i = 0; i < DPT; i++) {
    const int d = tid + i * RS_BS;
    pos[i] = d < nd ? offs[(int64_t)blockIdx.x * nd + d] : 0u;
    if (d < nd)
#pragma unroll
      for (int w = 0; w < NW; w++) hw[w][d] = 0;
  }
  const int64_t wbase = (int64_t)blockIdx.x * RS_TILE + (int64_t)wave * (RS_ITEMS * CHM_WAVE);
  uint32_t kk[RS_ITEMS], vv[RS_ITEMS], rc[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; r++) {
    const int64_t idx = wbase + r * CHM_WAVE + lane;
    const bool act = idx < n;
    kk[r] = act ? kin[idx] : 0u;
    vv[r] = act ? (vin ? vin[idx] : (uint32_t)idx) : 0u;
  }
  __syncthreads();
  uint32_t *h = hw[wave];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; r++) {
    const bool act = wbase + r * CHM_WAVE + lane < n;
    const uint32_t d = (kk[r] >> shift) & mask;
    const unsigned long long peers = match_peers(d, act, lane, nbits);
    const uint32_t rank = (uint32_t)__popcll((long long)(peers & lt));
    const uint32_t cnt = (uint32_t)__popcll((long long)peers);
    rc[r] = rank | (cnt << 8);
    if (act && rank == 0) atomicAdd(&h[d], cnt);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < DPT; i++) {
    const int d = tid + i * RS_BS;
    if (d < nd) {
      uint32_t run = pos[i];
#pragma unroll
      for (int w = 0; w < NW; w++) {
        const uint32_t c = hw[w][d];
        hw[w][d] = run;
        run += c;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; r++) {
    const bool act = wbase + r * CHM_WAVE + lane < n;
    const uint32_t d = (kk[r] >> shift) & mask;
    const uint32_t rank = rc[r] & 0xFFu, cnt = rc[r] >> 8;
    uint32_t dest = 0;
    if (act) dest = h[d] + rank;
    __builtin_amdgcn_wave_barrier();
    if (act && rank == 0) atomicAdd(&h[d], cnt);
    if (act) {
      if (kout) kout[dest] = kk[r];
      vout[dest] = vv[r];
    }
  }
}

// Exclusive offsets of the per-tile digit counts, digit-major:
// offs[t * nd + d] = sum of all counts of digits < d + counts of digit d in
// tiles < t.  Tiles are grouped in chunks of RS_TC: k_rs_colsum writes the
// chunk sums transposed (csumT[d * nch + ch]), a flat scan of csumT gives each
// chunk's starting offset, and k_rs_coloffs runs through the chunk's tiles.
#define RS_TC 16
__global__ void __launch_bounds__(256) k_rs_colsum(const uint16_t *__restrict__ counts, int64_t nblk,
                                                   int nbits, int64_t nch,
                                                   uint32_t *__restrict__ csumT,
                                                   uint32_t *__restrict__ clear) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;  // ch * nd + d
  if (clear && e < 2) clear[e] = 0u;
  const int nd = 1 << nbits;
  if (e >= nch * nd) return;
  const int64_t ch = e >> nbits;
  const int d = (int)(e & (nd - 1));
  const int64_t t0 = ch * RS_TC;
  uint32_t v[RS_TC];
#pragma unroll
  for (int i = 0; i < RS_TC; i++) v[i] = t0 + i < nblk ? counts[(t0 + i) * nd + d] : 0u;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < RS_TC; i++) s += v[i];
  csumT[(int64_t)d * nch + ch] = s;
}

__global__ void __launch_bounds__(256) k_rs_coloffs(const uint16_t *__restrict__ counts, int64_t nblk,
                                                    int nbits, int64_t nch,
                                                    const uint32_t *__restrict__ cpre,
                                                    uint32_t *__restrict__ offs) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int nd = 1 << nbits;
  if (e >= nch * nd) return;
  const int64_t ch = e >> nbits;
  const int d = (int)(e & (nd - 1));
  const int64_t t0 = ch * RS_TC;
  uint32_t v[RS_TC];
#pragma unroll
  for (int i = 0; i < RS_TC; i++) v[i] = t0 + i < nblk ? counts[(t0 + i) * nd + d] : 0u;
  uint32_t run = cpre[(int64_t)d * nch + ch];
#pragma unroll
  for (int i = 0; i < RS_TC; i++) {
    if (t0 + i < nblk) offs[(t0 + i) * nd + d] = run;
    run += v[i];
  }
}

// Pass plan for keys of `bits` bits: one or two 8-bit passes up to 16 bits,
// then two (up to 22 bits) or three passes of at most 11 bits.
struct RsPlan {
  int npass, rb, nb[3];
};
static RsPlan rs_plan(int bits) {
  RsPlan p = {};
  if (bits < 1) bits = 1;
  if (bits <= 16) {
    p.rb = 8;
    p.npass = (bits + 7) / 8;
  } else {
    p.rb = 11;
    p.npass = bits <= 22 ? 2 : 3;
  }
  int left = bits;
  for (int i = 0; i < p.npass; i++) {
    int w = (left + (p.npass - i) - 1) / (p.npass - i);
    p.nb[i] = w;
    left -= w;
  }
  return p;
}
static int key_bits(int64_t nbins) {
  int bits = 0;
  while (((int64_t)1 << bits) < nbins) bits++;  // keys are < nbins
  return bits < 1 ? 1 : bits;
}

// ------------------------------------------------------------------ driver
struct SortWs {
  uint16_t *counts;
  uint32_t *offs, *csum, *cpre, *bsum2, *kA, *kB, *vA, *vB, *trunc;
  int64_t nblk, nch;
};

static size_t sort_layout(int64_t Np, int64_t nbins, SortWs *w, void *base) {
  int64_t nblk = (Np + RS_TILE - 1) / RS_TILE;
  if (nblk < 1) nblk = 1;
  const int64_t nch = (nblk + RS_TC - 1) / RS_TC;
  const int64_t nd = (int64_t)1 << rs_plan(key_bits(nbins)).rb;
  const int64_t ncounts = nd * nblk, ncsum = nd * nch;
  const int64_t nscan2 = (ncsum + SCAN_TILE - 1) / SCAN_TILE;
  size_t off = 0;
  size_t o_counts = off; off += chm_align(ncounts * 2);
  size_t o_offs = off; off += chm_align(ncounts * 4);
  size_t o_csum = off; off += chm_align(ncsum * 4);
  size_t o_cpre = off; off += chm_align((ncsum + 1) * 4);
  size_t o_bsum2 = off; off += chm_align(nscan2 * 4);
  size_t o_k = off; off += 2 * chm_align(Np * 4);
  size_t o_v = off; off += 2 * chm_align(Np * 4);
  size_t o_t = off; off += chm_align((2 + 3 * OFF_LIST) * 4);
  if (w) {
    char *b = (char *)base;
    w->counts = (uint16_t *)(b + o_counts);
    w->offs = (uint32_t *)(b + o_offs);
    w->csum = (uint32_t *)(b + o_csum);
    w->cpre = (uint32_t *)(b + o_cpre);
    w->bsum2 = (uint32_t *)(b + o_bsum2);
    w->kA = (uint32_t *)(b + o_k);
    w->kB = (uint32_t *)(b + o_k + chm_align(Np * 4));
    w->vA = (uint32_t *)(b + o_v);
    w->vB = (uint32_t *)(b + o_v + chm_align(Np * 4));
    w->trunc = (uint32_t *)(b + o_t);
    w->nblk = nblk;
    w->nch = nch;
  }
  return off;
}

extern "C" size_t chm_sort_workspace_bytes(int64_t Np, int64_t nbins) {
  return sort_layout(Np, nbins, nullptr, nullptr);
}

// Stage 1: the cell keys (+ the push with PUSH).
template <typename T, bool PUSH>
static int sort_keys(T *x, T *y, T *z, const T *px, const T *py, const T *pz, const T *g_inv, T dt,
                     int64_t Np, const chm_grid_t *g, uint32_t *indx_in_cell, void *ws, size_t ws_bytes,
                     void *stream) {
  if (!g || Np < 0 || Np > 0x7FFFFFF0ll || g->Nx < 4 || g->Nr < 3) return CHM_EINVAL;
  int64_t nbins = (int64_t)(g->Nx - 1) * (g->Nr - 1) + 1;
  SortWs w;
  if (sort_layout(Np, nbins, &w, ws) > ws_bytes || !ws) return CHM_EWORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  if (Np > 0)
    k_cell_key<T, PUSH><<<chm_blocks(Np, SORT_BS), SORT_BS, 0, st>>>(
        x, y, z, px, py, pz, g_inv, dt, Np, g->Nx, g->Nr, (T)g->xmin, (T)g->dx_inv, (T)g->rmin,
        (T)g->dr_inv, indx_in_cell);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

static void scan_u32(const uint32_t *in, int64_t n, uint32_t *out, uint32_t *bsum, int64_t nscan,
                     hipStream_t st) {
  k_scan_reduce<<<(unsigned)nscan, SCAN_BS, 0, st>>>(in, n, bsum);
  k_scan_final<<<(unsigned)nscan, SCAN_BS, 0, st>>>((uint32_t *)in, n, 0, bsum, out);
}

// Stage 2: the stable radix sort of the keys of stage 1, then the offsets
// from the sorted keys.
static int sort_from_keys(int64_t Np, const chm_grid_t *g, const uint32_t *indx_in_cell,
                          uint32_t *cell_offset, uint32_t *sort_indx, uint32_t *np_stay, void *ws,
                          size_t ws_bytes, void *stream) {
  if (!g || Np < 0 || Np > 0x7FFFFFF0ll || g->Nx < 4 || g->Nr < 3) return CHM_EINVAL;
  int64_t nbins = (int64_t)(g->Nx - 1) * (g->Nr - 1) + 1;
  SortWs w;
  if (sort_layout(Np, nbins, &w, ws) > ws_bytes || !ws) return CHM_EWORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const uint32_t *skey = indx_in_cell;
  if (Np > 0) {
    const RsPlan plan = rs_plan(key_bits(nbins));
    const uint32_t *kin = indx_in_cell, *vin = nullptr;
    int shift = 0;
    for (int p = 0; p < plan.npass; p++) {
      const int nb = plan.nb[p];
      const bool last = p == plan.npass - 1;
      uint32_t *kout = (p & 1) ? w.kB : w.kA;
      uint32_t *vout = last ? sort_indx : ((p & 1) ? w.vB : w.vA);
      const int64_t ncs = ((int64_t)1 << nb) * w.nch;
      const int64_t nsc = (ncs + SCAN_TILE - 1) / SCAN_TILE;
      if (plan.rb == 8)
        k_rs_upsweep<8><<<(unsigned)w.nblk, RS_BS, 0, st>>>(kin, Np, shift, nb, w.counts);
      else
        k_rs_upsweep<11><<<(unsigned)w.nblk, RS_BS, 0, st>>>(kin, Np, shift, nb, w.counts);
      k_rs_colsum<<<chm_blocks(ncs, 256), 256, 0, st>>>(w.counts, w.nblk, nb, w.nch, w.csum,
                                                         p == 0 ? w.trunc : nullptr);
      scan_u32(w.csum, ncs, w.cpre, w.bsum2, nsc, st);
      k_rs_coloffs<<<chm_blocks(ncs, 256), 256, 0, st>>>(w.counts, w.nblk, nb, w.nch, w.cpre,
                                                          w.offs);
      if (plan.rb == 8)
        k_rs_downsweep<8><<<(unsigned)w.nblk, RS_BS, 0, st>>>(kin, vin, Np, shift, nb, w.offs,
                                                              kout, vout);
      else
        k_rs_downsweep<11><<<(unsigned)w.nblk, RS_BS, 0, st>>>(kin, vin, Np, shift, nb, w.offs,
                                                               kout, vout);
      kin = kout;
      vin = vout;
      shift += nb;
    }
    skey = kin;
  }
  if (Np == 0) k_list_clear<<<1, 2, 0, st>>>(w.trunc);  // no radix pass cleared it
  k_offsets_bounds<<<chm_blocks(Np / OFF_PER + 1, 256), 256, 0, st>>>(skey, Np, nbins, cell_offset,
                                                                       np_stay, w.trunc);
  k_offsets_gaps<<<256, 256, 0, st>>>(skey, Np, nbins, cell_offset, np_stay, w.trunc);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

template <typename T>
static int index_sort(const T *x, const T *y, const T *z, int64_t Np, const chm_grid_t *g,
                      uint32_t *indx_in_cell, uint32_t *cell_offset, uint32_t *sort_indx,
                      uint32_t *np_stay, void *ws, size_t ws_bytes, void *stream) {
  int rc = sort_keys<T, false>((T *)x, (T *)y, (T *)z, nullptr, nullptr, nullptr, nullptr, (T)0, Np, g,
                               indx_in_cell, ws, ws_bytes, stream);
  if (rc) return rc;
  return sort_from_keys(Np, g, indx_in_cell, cell_offset, sort_indx, np_stay, ws, ws_bytes, stream);
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

extern "C" int chm_push_x_keys_f64(double *x, double *y, double *z, const double *px, const double *py,
                                   const double *pz, const double *g_inv, double dt, int64_t Np,
                                   const chm_grid_t *g, uint32_t *indx_in_cell, void *ws,
                                   size_t ws_bytes, void *stream) {
  return sort_keys<double, true>(x, y, z, px, py, pz, g_inv, dt, Np, g, indx_in_cell, ws, ws_bytes,
                                 stream);
}
extern "C" int chm_push_x_keys_f32(float *x, float *y, float *z, const float *px, const float *py,
                                   const float *pz, const float *g_inv, float dt, int64_t Np,
                                   const chm_grid_t *g, uint32_t *indx_in_cell, void *ws,
                                   size_t ws_bytes, void *stream) {
  return sort_keys<float, true>(x, y, z, px, py, pz, g_inv, dt, Np, g, indx_in_cell, ws, ws_bytes,
                                stream);
}
// Where the last pass left the sorted keys in a sort workspace: the second
// key buffer after an even number of passes, the first after an odd number.
extern "C" const uint32_t *chm_sort_sorted_keys(const void *ws, int64_t Np, const chm_grid_t *g) {
  if (!ws || !g || Np <= 0 || Np > 0x7FFFFFF0ll || g->Nx < 4 || g->Nr < 3) return nullptr;
  const int64_t nbins = (int64_t)(g->Nx - 1) * (g->Nr - 1) + 1;
  SortWs w;
  sort_layout(Np, nbins, &w, (void *)ws);
  return (rs_plan(key_bits(nbins)).npass & 1) ? w.kA : w.kB;
}

extern "C" int chm_sort_from_keys(int64_t Np, const chm_grid_t *g, const uint32_t *indx_in_cell,
                                  uint32_t *cell_offset, uint32_t *sort_indx, uint32_t *np_stay,
                                  void *ws, size_t ws_bytes, void *stream) {
  return sort_from_keys(Np, g, indx_in_cell, cell_offset, sort_indx, np_stay, ws, ws_bytes, stream);
}

// ------------------------------------------------------------------ sort without allocation
// chm_index_sort_f64_noalloc takes its sort workspace from `scratch`, a fixed
// array in device memory (one copy per GPU, held until the process ends).
// In host code `scratch` only names the symbol: its device address on the
// current GPU comes from hipGetSymbolAddress, looked up once per device on
// the first call there, and only that address is handed to kernels.  The
// radix passes overwrite every word they read (tile counts, chunk sums, key
// and index buffers; the offsets' overflow list is cleared by the first pass
// or by k_list_clear), so nothing left by an earlier call is read.
static __device__ __align__(256) unsigned char scratch[CHM_SORT_SCRATCH_BYTES];

#define CHM_MAX_DEVICES 64
static std::mutex scratch_mu;
static void *scratch_dev[CHM_MAX_DEVICES];

static int scratch_ptr(void **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= CHM_MAX_DEVICES) return CHM_EUNSUPPORTED;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return CHM_EINVAL;
    scratch_dev[dev] = p;
  }
  *out = scratch_dev[dev];
  return CHM_OK;
}

extern "C" int chm_index_sort_f64_noalloc(const double *x, const double *y, const double *z,
                                          int64_t Np, const chm_grid_t *g, uint32_t *indx_in_cell,
                                          uint32_t *cell_offset, uint32_t *sort_indx,
                                          uint32_t *np_stay, void *stream) {
  if (!g || Np < 0 || Np > 0x7FFFFFF0ll || g->Nx < 4 || g->Nr < 3) return CHM_EINVAL;
  const int64_t nbins = (int64_t)(g->Nx - 1) * (g->Nr - 1) + 1;
  const size_t need = sort_layout(Np, nbins, nullptr, nullptr);
  if (need > (size_t)CHM_SORT_SCRATCH_BYTES) return CHM_ESCRATCH;
  void *ws = nullptr;
  int rc = scratch_ptr(&ws);
  if (rc) return rc;
  return index_sort<double>(x, y, z, Np, g, indx_in_cell, cell_offset, sort_indx, np_stay, ws, need,
                            stream);
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

// ------------------------------------------------------------------ domain fill
// Deterministic building blocks shared with the CPU oracle (which evaluates
// the same expressions in the same order, so the results are bit-identical):
//   rng_u01   splitmix64 (Steele, Lea, Flood 2014) of seed + (counter+1)*golden,
//             top 53 bits -> [0, 1)
//   det_sincos sin and cos from IEEE +, -, *, / only: Cody-Waite reduction by
//             pi/2 with the three-part constant of fdlibm (k*P1, k*P2 exact for
//             |k| < 2^20), Taylor polynomials to r^17 / r^18 on |r| <= pi/4.
__device__ __forceinline__ double rng_u01(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + (counter + 1ull) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ void det_sincos(double a, double *s_out, double *c_out) {
  const double P1 = 1.57079632673412561417e+00, P2 = 6.07710050630396597660e-11,
               P2T = 2.02226624879595063154e-21;
  const double k = floor(a * (2.0 / 3.14159265358979323846) + 0.5);
  const double r = ((a - k * P1) - k * P2) - k * P2T;
  const double r2 = r * r;
  double s = 1.0 / 355687428096000.0;                 // 1/17!
  s = (-1.0 / 1307674368000.0) + r2 * s;              // 1/15!
  s = (1.0 / 6227020800.0) + r2 * s;                  // 1/13!
  s = (-1.0 / 39916800.0) + r2 * s;                   // 1/11!
  s = (1.0 / 362880.0) + r2 * s;                      // 1/9!
  s = (-1.0 / 5040.0) + r2 * s;                       // 1/7!
  s = (1.0 / 120.0) + r2 * s;                         // 1/5!
  s = (-1.0 / 6.0) + r2 * s;                          // 1/3!
  s = r + r * (r2 * s);
  double c = -1.0 / 6402373705728000.0;               // -1/18!
  c = (1.0 / 20922789888000.0) + r2 * c;              // 1/16!
  c = (-1.0 / 87178291200.0) + r2 * c;                // -1/14!
  c = (1.0 / 479001600.0) + r2 * c;                   // 1/12!
  c = (-1.0 / 3628800.0) + r2 * c;                    // -1/10!
  c = (1.0 / 40320.0) + r2 * c;                       // 1/8!
  c = (-1.0 / 720.0) + r2 * c;                        // -1/6!
  c = (1.0 / 24.0) + r2 * c;                          // 1/4!
  c = (-0.5) + r2 * c;                                // -1/2!
  c = 1.0 + r2 * c;
  const int q = (int)((int64_t)k & 3);
  double sv, cv;
  if (q == 0) { sv = s; cv = c; }
  else if (q == 1) { sv = c; cv = -s; }
  else if (q == 2) { sv = -s; cv = -c; }
  else { sv = -c; cv = s; }
  *s_out = sv;
  *c_out = cv;
}

// log from IEEE operations and the exact frexp split, as the CPU oracle
// evaluates it: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(f),
// f = (m-1)/(m+1), series to f^19 (|f| <= 0.172); e ln2 with fdlibm's
// two-part ln2.
__device__ __forceinline__ double det_log(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  const double f = (m - 1.0) / (m + 1.0);
  const double f2 = f * f;
  double s = 1.0 / 19.0;
  s = (1.0 / 17.0) + f2 * s;
  s = (1.0 / 15.0) + f2 * s;
  s = (1.0 / 13.0) + f2 * s;
  s = (1.0 / 11.0) + f2 * s;
  s = (1.0 / 9.0) + f2 * s;
  s = (1.0 / 7.0) + f2 * s;
  s = (1.0 / 5.0) + f2 * s;
  s = (1.0 / 3.0) + f2 * s;
  const double lm = 2.0 * (f + f * (f2 * s));
  const double de = (double)e;
  return de * 6.93147180369123816490e-01 + (lm + de * 1.90821492927058770002e-10);
}

// Two independent standard normals from counters (c, c+1) (Box-Muller).
__device__ __forceinline__ void det_normal2(uint64_t seed, uint64_t c, double *z0, double *z1) {
  const double u1 = rng_u01(seed, c), u2 = rng_u01(seed, c + 1);
  const double rad = sqrt(-2.0 * det_log(1.0 - u1));
  double s, co;
  det_sincos((2.0 * 3.14159265358979323846) * u2, &s, &co);
  *z0 = rad * co;
  *z1 = rad * s;
}

// fill_grid, particles_generic.cl:33-84: per cell (ix, ir) of the local
// grids (xgrid [Nx_loc], rgrid), Nppc_th x Nppc_r x Nppc_x particles at the
// regular sub-positions, the theta set rotated by a per-cell uniform offset
// in [0, 2 pi) (theta_variator, particles_methods_cl.py:95-98, here from
// rng_u01(seed, counter0 + cell)); weights r * w0 (:103); momenta
// (:111-147): p_c, plus dp times a standard normal per particle and
// component when a spread dp is given (_fill_arr_randn(mu = p_c, sigma =
// dp)); particle ip draws its normals from counters counter_p + 4 ip .. + 2
// (two Box-Muller pairs, the fourth value unused); g_inv from the momenta.
template <typename T>
__global__ void __launch_bounds__(256)
    k_fill_domain(T *__restrict__ x, T *__restrict__ y, T *__restrict__ z, T *__restrict__ w,
                  T *__restrict__ px, T *__restrict__ py, T *__restrict__ pz,
                  T *__restrict__ g_inv, const double *__restrict__ xgrid,
                  const double *__restrict__ rgrid, int Nx, int64_t ncells, int npx, int npr,
                  int npt, double w0, double pxc, double pyc, double pzc, double dpx,
                  double dpy, double dpz, int moving, uint64_t seed, uint64_t counter0,
                  uint64_t counter_p) {
  const int64_t i_cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i_cell >= ncells) return;
  const int Nx_cell = Nx - 1;
  const int ir = (int)(i_cell / Nx_cell), ix = (int)(i_cell - (int64_t)ir * Nx_cell);
  int64_t ip = i_cell * (int64_t)(npx * npr * npt);
  const double xmin = xgrid[ix], rmin = rgrid[ir];
  const double thmin = (2.0 * 3.14159265358979323846) * rng_u01(seed, counter0 + (uint64_t)i_cell);
  const double Lx = xgrid[ix + 1] - xgrid[ix];
  const double Lr = rgrid[ir + 1] - rgrid[ir];
  const double dx = 1. / ((double)npx);
  const double dr = 1. / ((double)npr);
  const double dth = 2 * 3.14159265358979323846 / ((double)npt);
  const bool thermal = dpx != 0.0 || dpy != 0.0 || dpz != 0.0;
  double g = 1.0;
  if (pxc != 0.0 || pyc != 0.0 || pzc != 0.0) g = 1. / sqrt(1 + pxc * pxc + pyc * pyc + pzc * pzc);
  for (int it = 0; it < npt; it++) {
    const double th = thmin + it * dth;
    double sin_th, cos_th;
    det_sincos(th, &sin_th, &cos_th);
    for (int irr = 0; irr < npr; irr++) {
      const double rp = rmin + (0.5 + irr) * dr * Lr;
      const double rp_s = rp * sin_th;
      const double rp_c = rp * cos_th;
      for (int ixx = 0; ixx < npx; ixx++) {
        x[ip] = (T)(xmin + (0.5 + ixx) * dx * Lx);
        y[ip] = (T)rp_s;
        z[ip] = (T)rp_c;
        w[ip] = (T)(rp * w0);
        if (moving) {
          if (thermal) {
            double n0, n1, n2, n3;
            const uint64_t c = counter_p + 4ull * (uint64_t)ip;
            det_normal2(seed, c, &n0, &n1);
            det_normal2(seed, c + 2, &n2, &n3);
            const double ux = pxc + dpx * n0, uy = pyc + dpy * n1, uz = pzc + dpz * n2;
            px[ip] = (T)ux;
            py[ip] = (T)uy;
            pz[ip] = (T)uz;
            g_inv[ip] = (T)(1. / sqrt(1 + ux * ux + uy * uy + uz * uz));
          } else {
            px[ip] = (T)pxc;
            py[ip] = (T)pyc;
            pz[ip] = (T)pzc;
            g_inv[ip] = (T)g;
          }
        }
        ip += 1;
      }
    }
  }
}

template <typename T>
static int fill_domain(T *x, T *y, T *z, T *w, T *px, T *py, T *pz, T *g_inv, const double *xgrid,
                       const double *rgrid, int32_t Nx_loc, int64_t ncells, int32_t npx,
                       int32_t npr, int32_t npt, double w0, double pxc, double pyc, double pzc,
                       double dpx, double dpy, double dpz, uint64_t seed, uint64_t counter0,
                       void *stream) {
  if (!x || !y || !z || !w || !xgrid || !rgrid || Nx_loc < 2 || ncells < 0 || npx < 1 || npr < 1 ||
      npt < 1)
    return CHM_EINVAL;
  const int moving = px != nullptr;
  if (moving && (!py || !pz || !g_inv)) return CHM_EINVAL;
  if (ncells == 0) return CHM_OK;
  k_fill_domain<T><<<chm_blocks(ncells, 256), 256, 0, (hipStream_t)stream>>>(
      x, y, z, w, px, py, pz, g_inv, xgrid, rgrid, Nx_loc, ncells, npx, npr, npt, w0, pxc, pyc, pzc,
      dpx, dpy, dpz, moving, seed, counter0, counter0 + (uint64_t)ncells);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_fill_domain_f64(double *x, double *y, double *z, double *w, double *px,
                                   double *py, double *pz, double *g_inv, const double *xgrid,
                                   const double *rgrid, int32_t Nx_loc, int64_t ncells,
                                   int32_t npx, int32_t npr, int32_t npt, double w0, double pxc,
                                   double pyc, double pzc, double dpx, double dpy, double dpz,
                                   uint64_t seed, uint64_t counter0, void *stream) {
  return fill_domain<double>(x, y, z, w, px, py, pz, g_inv, xgrid, rgrid, Nx_loc, ncells, npx, npr,
                             npt, w0, pxc, pyc, pzc, dpx, dpy, dpz, seed, counter0, stream);
}
extern "C" int chm_fill_domain_f32(float *x, float *y, float *z, float *w, float *px, float *py,
                                   float *pz, float *g_inv, const double *xgrid,
                                   const double *rgrid, int32_t Nx_loc, int64_t ncells,
                                   int32_t npx, int32_t npr, int32_t npt, double w0, double pxc,
                                   double pyc, double pzc, double dpx, double dpy, double dpz,
                                   uint64_t seed, uint64_t counter0, void *stream) {
  return fill_domain<float>(x, y, z, w, px, py, pz, g_inv, xgrid, rgrid, Nx_loc, ncells, npx, npr,
                            npt, w0, pxc, pyc, pzc, dpx, dpy, dpz, seed, counter0, stream);
}

// ------------------------------------------------------------------ beam
// make_new_beam, particles_methods_cl.py:149-177: Gaussian positions and
// momenta, equal weights, g_inv from the momenta.  Particle p draws its six
// normals from counters counter0 + 6p .. + 5.
template <typename T>
__global__ void __launch_bounds__(256)
    k_fill_beam(T *__restrict__ x, T *__restrict__ y, T *__restrict__ z, T *__restrict__ px,
                T *__restrict__ py, T *__restrict__ pz, T *__restrict__ w, T *__restrict__ g_inv,
                int64_t Np, chm_beam_t b, uint64_t seed, uint64_t counter0) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  const uint64_t c = counter0 + 6ull * (uint64_t)p;
  double n[6];
  det_normal2(seed, c, &n[0], &n[1]);
  det_normal2(seed, c + 2, &n[2], &n[3]);
  det_normal2(seed, c + 4, &n[4], &n[5]);
  const double xv = b.c[0] + b.L[0] * n[0], yv = b.c[1] + b.L[1] * n[1],
               zv = b.c[2] + b.L[2] * n[2];
  const double ux = b.pc[0] + b.dp[0] * n[3], uy = b.pc[1] + b.dp[1] * n[4],
               uz = b.pc[2] + b.dp[2] * n[5];
  x[p] = (T)xv;
  y[p] = (T)yv;
  z[p] = (T)zv;
  px[p] = (T)ux;
  py[p] = (T)uy;
  pz[p] = (T)uz;
  w[p] = (T)b.w;
  g_inv[p] = (T)(1. / sqrt(1 + ux * ux + uy * uy + uz * uz));
}

template <typename T>
static int fill_beam(T *x, T *y, T *z, T *px, T *py, T *pz, T *w, T *g_inv, int64_t Np,
                     const chm_beam_t *b, uint64_t seed, uint64_t counter0, void *stream) {
  if (!b || Np < 0 || !x || !y || !z || !px || !py || !pz || !w || !g_inv) return CHM_EINVAL;
  if (Np == 0) return CHM_OK;
  k_fill_beam<T><<<chm_blocks(Np, 256), 256, 0, (hipStream_t)stream>>>(x, y, z, px, py, pz, w,
                                                                       g_inv, Np, *b, seed, counter0);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_fill_beam_f64(double *x, double *y, double *z, double *px, double *py,
                                 double *pz, double *w, double *g_inv, int64_t Np,
                                 const chm_beam_t *b, uint64_t seed, uint64_t counter0,
                                 void *stream) {
  return fill_beam<double>(x, y, z, px, py, pz, w, g_inv, Np, b, seed, counter0, stream);
}
extern "C" int chm_fill_beam_f32(float *x, float *y, float *z, float *px, float *py, float *pz,
                                 float *w, float *g_inv, int64_t Np, const chm_beam_t *b,
                                 uint64_t seed, uint64_t counter0, void *stream) {
  return fill_beam<float>(x, y, z, px, py, pz, w, g_inv, Np, b, seed, counter0, stream);
}
