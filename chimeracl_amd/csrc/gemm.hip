// Discrete Hankel transform products on the matrix cores (gfx950 MFMA).
//
//   C[M x N] = A[M x K] @ B[K x N]      (row-major, A = a DHT-type matrix)
//
// A is a per-mode real matrix (DHT, DHT_inv, dDHT+-) stored zero-padded to
// [round_up(M, 64)][lda], lda = round_up(K, 64), so A tiles load without
// bounds checks in 16-byte vectors.  B is either a real field [K][N] (mode 0),
// or a complex field [K][Nx] read as N = 2*Nx real columns, optionally formed
// on the fly as alpha*X + beta*Y with either operand read through the
// m = -1 map F_{-1}[k][x] = -conj(F_1[k][-x]) (get_m1,
// transformer_generic.cl:3-24).  This replaces Reikna's MatrixMul
// (transformer_methods_cl.py:458-472) and the axpy/get_m1 kernels that fed it.
//
// Three arithmetic forms, 256 threads, 64 x 64 output tile per block, wave w
// owning the 32 x 32 quadrant (w/2, w%2); the next K-step is loaded into
// registers while the current one is multiplied:
//   f64    : v_mfma_f64_16x16x4f64 (plain form)
//   f32    : v_mfma_f32_16x16x4f32 (mixed form, CHM_DHT_F32)
//   bf16x3 : v_mfma_f32_16x16x32_bf16 on split operands, A = Ah + Al,
//            B = Bh + Bl, C = Ah*Bh + (Ah*Bl + Al*Bh), the two small terms
//            accumulated apart and added at the end (mixed, CHM_DHT_BF16X3).
#include "gemm.h"

#include <mutex>
#include <type_traits>
#include <vector>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));
typedef unsigned short us4 __attribute__((ext_vector_type(4)));

#define BM 64
#define BN 64

// ---------------------------------------------------------------- B loader
// Elements B[k][n], B[k][n + 1] of an even column n (zero outside K x N):
// one 2-element load of the real field, or one complex element (re, im) of
// each operand.  bmode 0 takes the vector load when ldb is even (aligned).
// A thread's column is the same in every K-step: its complex-column indices
// in X and Y (through the m = -1 map where asked) come from bcols, once.
template <typename T> struct pair_t { T a, b; };
struct BColIdx { int x, y; };
__device__ __forceinline__ BColIdx bcols(const GemmBatch &gb, const GemmItem &it, int n) {
  const int Nx = gb.Nx, c = n >> 1;
  BColIdx r = {c, c};
  if (gb.bmode == 1 && Nx > 0) {
    if (it.remapX) r.x = (Nx - c) % Nx;
    if (it.remapY) r.y = (Nx - c) % Nx;
  }
  return r;
}
// The B mode is a template parameter of the product kernels.  BMODE 0 (a
// real field) loads each column pair straight into the values it stores.
// BMODE 1 (alpha X + beta Y of complex fields) keeps the loaded elements
// raw (load_b) and forms the B values only where it stores them to LDS
// (b_value): a block issues the next K-step's loads, multiplies the current
// one, and only then computes, so the loads stay in flight across the MFMAs
// (with the complex products formed right after the loads, each K-step
// waited for them: the mixed rot product took 597 us per step instead of
// 360 on the 4096 x 512 step).
template <typename T> struct braw_t {
  T x0, x1, y0, y1;
  bool ok;
};
template <typename T, int BMODE> struct BRaw { typedef pair_t<T> type; };
template <typename T> struct BRaw<T, 1> { typedef braw_t<T> type; };

template <typename T, int BMODE>
__device__ __forceinline__ typename BRaw<T, BMODE>::type load_b(const GemmBatch &gb,
                                                                const GemmItem &it, int k, int n,
                                                                BColIdx cx) {
  typedef T t2 __attribute__((ext_vector_type(2)));
  if constexpr (BMODE == 0) {
    pair_t<T> r = {(T)0, (T)0};
    if (k >= gb.K || n >= gb.N) return r;
    const T *p = (const T *)it.X + (int64_t)k * gb.ldb + n;
    if (n + 1 < gb.N && !(gb.ldb & 1)) {
      const t2 v = *(const t2 *)p;
      r.a = v.x;
      r.b = v.y;
    } else {
      r.a = p[0];
      if (n + 1 < gb.N) r.b = p[1];
    }
    return r;
  } else {
    braw_t<T> r = {(T)0, (T)0, (T)0, (T)0, false};
    if (k >= gb.K || n >= gb.N) return r;
    r.ok = true;
    const int Nx = gb.Nx;
    const t2 v = ((const t2 *)it.X)[(int64_t)k * Nx + cx.x];
    r.x0 = v.x;
    r.x1 = v.y;
    if (it.Y) {
      const t2 w = ((const t2 *)it.Y)[(int64_t)k * Nx + cx.y];
      r.y0 = w.x;
      r.y1 = w.y;
    }
    return r;
  }
}
template <typename T, int BMODE>
__device__ __forceinline__ pair_t<T> b_value(const GemmItem &it,
                                             const typename BRaw<T, BMODE>::type &v) {
  if constexpr (BMODE == 0) {
    return v;
  } else {
    pair_t<T> r = {(T)0, (T)0};
    if (!v.ok) return r;
    {
      const T a = it.remapX ? -v.x0 : v.x0, b = v.x1;  // -conj(F(-kx)): (-re, +im)
      r.a = (T)it.alpha.x * a - (T)it.alpha.y * b;
      r.b = (T)it.alpha.x * b + (T)it.alpha.y * a;
    }
    if (it.Y) {
      const T a = it.remapY ? -v.y0 : v.y0, b = v.y1;
      r.a += (T)it.beta.x * a - (T)it.beta.y * b;
      r.b += (T)it.beta.x * b + (T)it.beta.y * a;
    }
    return r;
  }
}

// ---------------------------------------------------------------- f64 / f32
// (registers held to 5 waves per SIMD: the accumulators stay in VGPRs)
template <typename T, int BMODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_gemm(GemmBatch gb) {
  constexpr int BK = 16;
  constexpr int AP = (sizeof(T) == 8) ? 8 : 16;  // row pad: stride = 128 B mod 256 B
  __shared__ T As[BK][BM + AP];
  __shared__ T Bs[BK][BN + AP];
  const GemmItem &it = gb.it[blockIdx.z];
  const int n0 = blockIdx.x * BN, m0 = blockIdx.y * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const T *A = (const T *)it.A;
  // A: thread loads row ar, k = ak..ak+3 (padded table, no checks)
  const int ar = tid >> 2, ak = (tid & 3) * 4;
  // B: thread owns columns 2*bp, 2*bp+1 and rows bk, bk+8
  const int bp = tid & 31, bk = tid >> 5;
  T ra[4];
  typename BRaw<T, BMODE>::type rb[2];
  const BColIdx bcx = bcols(gb, it, n0 + 2 * bp);
  auto gload = [&](int k0) {
    const T *pa = A + (int64_t)(m0 + ar) * gb.lda + k0 + ak;
#pragma unroll
    for (int q = 0; q < 4; q++) ra[q] = pa[q];
#pragma unroll
    for (int q = 0; q < 2; q++) rb[q] = load_b<T, BMODE>(gb, it, k0 + bk + 8 * q, n0 + 2 * bp, bcx);
  };
  typedef typename std::conditional<sizeof(T) == 8, d4, f4>::type acc_t;
  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (acc_t){0, 0, 0, 0};
  gload(0);
  for (int k0 = 0; k0 < gb.K; k0 += BK) {
#pragma unroll
    for (int q = 0; q < 4; q++) As[ak + q][ar] = ra[q];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const pair_t<T> b = b_value<T, BMODE>(it, rb[q]);
      Bs[bk + 8 * q][2 * bp] = b.a;
      Bs[bk + 8 * q][2 * bp + 1] = b.b;
    }
    __syncthreads();
    if (k0 + BK < gb.K) gload(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++) a[i] = As[kk + (lane >> 4)][32 * wr + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; j++) b[j] = Bs[kk + (lane >> 4)][32 * wc + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          if constexpr (sizeof(T) == 8)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  T *C = (T *)it.C;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        // f64 16x16x4: row = l/16 + 4r ; f32 16x16x4: row = 4*(l/16) + r
        int rloc = (sizeof(T) == 8) ? ((lane >> 4) + 4 * r) : (4 * (lane >> 4) + r);
        int row = m0 + 32 * wr + 16 * i + rloc;
        int col = n0 + 32 * wc + 16 * j + (lane & 15);
        if (row < gb.M && col < gb.N) C[(int64_t)row * gb.ldc + col] = acc[i][j][r];
      }
}

// ---------------------------------------------------------------- split-BF16
__device__ __forceinline__ void split_bf16(float v, unsigned short &hi, unsigned short &lo) {
  __bf16 h = (__bf16)v;
  __bf16 l = (__bf16)(v - (float)h);
  hi = __builtin_bit_cast(unsigned short, h);
  lo = __builtin_bit_cast(unsigned short, l);
}

template <int BK, int BMODE>
__global__ void __launch_bounds__(256) k_gemm_bf16x3(GemmBatch gb) {
  constexpr int KP = BK + 8;  // row stride 16-byte aligned, staggers banks
  constexpr int NA = BK / 32;  // 16-byte A vectors per thread (hi and lo)
  constexpr int NB = BK / 8;   // B column pairs per thread
  // operands stored k-contiguous so each lane reads 8 bf16 (16 B) at once
  __shared__ __attribute__((aligned(16))) unsigned short Ah[BM][KP], Al[BM][KP];
  __shared__ __attribute__((aligned(16))) unsigned short Bh[BN][KP], Bl[BN][KP];
  const GemmItem &it = gb.it[blockIdx.z];
  const int n0 = blockIdx.x * BN, m0 = blockIdx.y * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const unsigned short *Ahg = (const unsigned short *)it.A, *Alg = (const unsigned short *)it.A_lo;
  // A: thread loads row ar, k = ak .. ak + 8*NA - 1 (hi and lo)
  const int ar = tid >> 2, ak = (tid & 3) * 8 * NA;
  // B: thread owns columns 2*bp, 2*bp+1 and k = bk .. bk + NB - 1
  const int bp = tid & 31, bk = (tid >> 5) * NB;
  us8 rah[NA], ral[NA];
  typename BRaw<float, BMODE>::type rb[NB];
  const BColIdx bcx = bcols(gb, it, n0 + 2 * bp);
  auto gload = [&](int k0) {
    const int64_t o = (int64_t)(m0 + ar) * gb.lda + k0 + ak;
#pragma unroll
    for (int v = 0; v < NA; v++) {
      rah[v] = *(const us8 *)(Ahg + o + 8 * v);
      ral[v] = *(const us8 *)(Alg + o + 8 * v);
    }
#pragma unroll
    for (int q = 0; q < NB; q++) rb[q] = load_b<float, BMODE>(gb, it, k0 + bk + q, n0 + 2 * bp, bcx);
  };
  f4 acc[2][2], cor[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = cor[i][j] = (f4){0, 0, 0, 0};
  gload(0);
  for (int k0 = 0; k0 < gb.K; k0 += BK) {
#pragma unroll
    for (int v = 0; v < NA; v++) {
      *(us8 *)&Ah[ar][ak + 8 * v] = rah[v];
      *(us8 *)&Al[ar][ak + 8 * v] = ral[v];
    }
#pragma unroll
    for (int q0 = 0; q0 < NB; q0 += 4) {
      us4 bh0, bl0, bh1, bl1;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const pair_t<float> b = b_value<float, BMODE>(it, rb[q0 + q]);
        unsigned short h, l;
        split_bf16(b.a, h, l);
        bh0[q] = h;
        bl0[q] = l;
        split_bf16(b.b, h, l);
        bh1[q] = h;
        bl1[q] = l;
      }
      *(us4 *)&Bh[2 * bp][bk + q0] = bh0;
      *(us4 *)&Bl[2 * bp][bk + q0] = bl0;
      *(us4 *)&Bh[2 * bp + 1][bk + q0] = bh1;
      *(us4 *)&Bl[2 * bp + 1][bk + q0] = bl1;
    }
    __syncthreads();
    if (k0 + BK < gb.K) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      bf8 ah[2], al[2], bhv[2], blv[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
        const int row = 32 * wr + 16 * i + (lane & 15), k = ks + 8 * (lane >> 4);
        ah[i] = __builtin_bit_cast(bf8, *(const us8 *)&Ah[row][k]);
        al[i] = __builtin_bit_cast(bf8, *(const us8 *)&Al[row][k]);
      }
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const int col = 32 * wc + 16 * j + (lane & 15), k = ks + 8 * (lane >> 4);
        bhv[j] = __builtin_bit_cast(bf8, *(const us8 *)&Bh[col][k]);
        blv[j] = __builtin_bit_cast(bf8, *(const us8 *)&Bl[col][k]);
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bhv[j], acc[i][j], 0, 0, 0);
          cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], blv[j], cor[i][j], 0, 0, 0);
          cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bhv[j], cor[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  float *C = (float *)it.C;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = m0 + 32 * wr + 16 * i + 4 * (lane >> 4) + r;
        int col = n0 + 32 * wc + 16 * j + (lane & 15);
        if (row < gb.M && col < gb.N) C[(int64_t)row * gb.ldc + col] = acc[i][j][r] + cor[i][j][r];
      }
}

// ---------------------------------------------------------------- 128 x 128 tiles
// The split-BF16 product with a 128 x 128 output tile per block (each wave a
// 64 x 64 quadrant of 4 x 4 MFMA tiles), for launches with enough tiles to
// fill the GPU: four times the matrix-core work per LDS round trip and
// barrier pair, and half the operand loads per flop.
template <int BMODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_gemm128_bf16x3(GemmBatch gb) {
  constexpr int TM = 128, TN = 128, BK = 32, KP = BK + 8;
  __shared__ __attribute__((aligned(16))) unsigned short Ah[TM][KP], Al[TM][KP];
  __shared__ __attribute__((aligned(16))) unsigned short Bh[TN][KP], Bl[TN][KP];
  const GemmItem &it = gb.it[blockIdx.z];
  const int n0 = blockIdx.x * TN, m0 = blockIdx.y * TM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const unsigned short *Ahg = (const unsigned short *)it.A, *Alg = (const unsigned short *)it.A_lo;
  // A: thread loads row ar, k = ak .. ak + 15 (hi and lo), rows beyond the
  // 64-row padding of the table are not read
  const int ar = tid >> 1, ak = (tid & 1) * 16;
  const bool arow = m0 + ar < ((gb.M + 63) & ~63);
  // B: thread owns columns 2*bp, 2*bp+1 and k = bk .. bk + 7
  const int bp = tid & 63, bk = (tid >> 6) * 8;
  us8 rah[2], ral[2];
  typename BRaw<float, BMODE>::type rb[8];
  const BColIdx bcx = bcols(gb, it, n0 + 2 * bp);
  auto gload = [&](int k0) {
    const int64_t o = (int64_t)(m0 + ar) * gb.lda + k0 + ak;
#pragma unroll
    for (int v = 0; v < 2; v++) {
      rah[v] = arow ? *(const us8 *)(Ahg + o + 8 * v) : (us8){0, 0, 0, 0, 0, 0, 0, 0};
      ral[v] = arow ? *(const us8 *)(Alg + o + 8 * v) : (us8){0, 0, 0, 0, 0, 0, 0, 0};
    }
#pragma unroll
    for (int q = 0; q < 8; q++) rb[q] = load_b<float, BMODE>(gb, it, k0 + bk + q, n0 + 2 * bp, bcx);
  };
  f4 acc[4][4], cor[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = cor[i][j] = (f4){0, 0, 0, 0};
  gload(0);
  for (int k0 = 0; k0 < gb.K; k0 += BK) {
#pragma unroll
    for (int v = 0; v < 2; v++) {
      *(us8 *)&Ah[ar][ak + 8 * v] = rah[v];
      *(us8 *)&Al[ar][ak + 8 * v] = ral[v];
    }
    {
      us8 bh0, bl0, bh1, bl1;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const pair_t<float> b = b_value<float, BMODE>(it, rb[q]);
        unsigned short h, l;
        split_bf16(b.a, h, l);
        bh0[q] = h;
        bl0[q] = l;
        split_bf16(b.b, h, l);
        bh1[q] = h;
        bl1[q] = l;
      }
      *(us8 *)&Bh[2 * bp][bk] = bh0;
      *(us8 *)&Bl[2 * bp][bk] = bl0;
      *(us8 *)&Bh[2 * bp + 1][bk] = bh1;
      *(us8 *)&Bl[2 * bp + 1][bk] = bl1;
    }
    __syncthreads();
    if (k0 + BK < gb.K) gload(k0 + BK);
    const int kq = 8 * (lane >> 4);
    bf8 bhv[4], blv[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int col = 64 * wc + 16 * j + (lane & 15);
      bhv[j] = __builtin_bit_cast(bf8, *(const us8 *)&Bh[col][kq]);
      blv[j] = __builtin_bit_cast(bf8, *(const us8 *)&Bl[col][kq]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = 64 * wr + 16 * i + (lane & 15);
      const bf8 ah = __builtin_bit_cast(bf8, *(const us8 *)&Ah[row][kq]);
      const bf8 al = __builtin_bit_cast(bf8, *(const us8 *)&Al[row][kq]);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bhv[j], acc[i][j], 0, 0, 0);
        cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, blv[j], cor[i][j], 0, 0, 0);
        cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bhv[j], cor[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  float *C = (float *)it.C;
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = m0 + 64 * wr + 16 * i + 4 * (lane >> 4) + r;
        int col = n0 + 64 * wc + 16 * j + (lane & 15);
        if (row < gb.M && col < gb.N) C[(int64_t)row * gb.ldc + col] = acc[i][j][r] + cor[i][j][r];
      }
}

#ifndef CHM_BF16_BK
#define CHM_BF16_BK 32
#endif

// ---------------------------------------------------------------- launch timing
// Measurement aid: while enabled (chm_dht_timing(1)), every product launch
// on a stream that is not being captured is bracketed by two HIP events
// recorded on its own stream, kept with the launch's algorithmic flops
// (2 M K N per item: real FP32-equivalent work for the split-BF16 form).
struct TimedLaunch {
  hipEvent_t a, b;
  double flops;
  int kind;
};
static std::mutex t_mu;
static bool t_on = false;
static std::vector<TimedLaunch> t_launches;

static void t_clear() {
  for (auto &t : t_launches) {
    hipEventDestroy(t.a);
    hipEventDestroy(t.b);
  }
  t_launches.clear();
}

extern "C" int chm_dht_timing(int32_t enable) {
  std::lock_guard<std::mutex> lock(t_mu);
  t_clear();
  t_on = enable != 0;
  return CHM_OK;
}

extern "C" int64_t chm_dht_timing_collect(double *ms, double *flops, int32_t *kind, int64_t max) {
  std::lock_guard<std::mutex> lock(t_mu);
  int64_t n = 0;
  for (auto &t : t_launches) {
    if (n >= max) break;
    if (hipEventSynchronize(t.b) != hipSuccess) return -1;
    float v = 0.f;
    if (hipEventElapsedTime(&v, t.a, t.b) != hipSuccess) return -1;
    if (ms) ms[n] = v;
    if (flops) flops[n] = t.flops;
    if (kind) kind[n] = t.kind;
    n++;
  }
  t_clear();
  return n;
}

static bool t_begin(hipStream_t st, TimedLaunch *t) {
  std::lock_guard<std::mutex> lock(t_mu);
  if (!t_on) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
  if (hipEventCreate(&t->a) != hipSuccess) return false;
  if (hipEventCreate(&t->b) != hipSuccess) {
    hipEventDestroy(t->a);
    return false;
  }
  hipEventRecord(t->a, st);
  return true;
}

static void t_end(hipStream_t st, TimedLaunch *t) {
  hipEventRecord(t->b, st);
  std::lock_guard<std::mutex> lock(t_mu);
  t_launches.push_back(*t);
}

int chm_gemm_launch(const GemmBatch &gb_in, int nitems, int kind, hipStream_t st) {
  if (nitems <= 0) return CHM_OK;
  if (nitems > GEMM_MAXB || gb_in.M <= 0 || gb_in.K <= 0 || gb_in.N <= 0) return CHM_EINVAL;
  if (gb_in.lda < gemm_lda(gb_in.K)) return CHM_EINVAL;
  // A complex B operand taken as it is (alpha = 1, no second operand, no
  // m = -1 map) is a real [K][2 Nx] matrix: read it directly.
  const GemmBatch *pgb = &gb_in;
  GemmBatch plain;
  if (gb_in.bmode == 1 && gb_in.N == 2 * gb_in.Nx) {
    bool all = true;
    for (int i = 0; i < nitems && all; i++) {
      const GemmItem &it = gb_in.it[i];
      all = it.alpha.x == 1.0 && it.alpha.y == 0.0 && !it.Y && !it.remapX;
    }
    if (all) {
      plain = gb_in;
      plain.bmode = 0;
      plain.ldb = 2 * gb_in.Nx;
      pgb = &plain;
    }
  }
  const GemmBatch &gb = *pgb;
  // split-BF16: 128 x 128 tiles when they alone give every CU (256) at least
  // two blocks (measured at 4096 x 512: 1.41 -> 1.33 ms of products per
  // step).  The FP64 / FP32 products keep 64 x 64 tiles: the larger tile
  // measured slower there (3.11 -> 3.65 ms per step at the same size), its
  // 64 accumulators per lane allow only two waves per SIMD against seven.
  const int64_t big = (int64_t)((gb.N + 127) / 128) * ((gb.M + 127) / 128) * nitems;
  const bool use128 = big >= 512;
  dim3 grid((gb.N + BN - 1) / BN, (gb.M + BM - 1) / BM, nitems);
  dim3 grid128((gb.N + 127) / 128, (gb.M + 127) / 128, nitems);
  TimedLaunch tl;
  const bool timed = t_begin(st, &tl);
  const bool cplx = gb.bmode == 1;
  if (kind == GEMM_F64) {
    if (cplx) k_gemm<double, 1><<<grid, 256, 0, st>>>(gb);
    else k_gemm<double, 0><<<grid, 256, 0, st>>>(gb);
  } else if (kind == GEMM_F32) {
    if (cplx) k_gemm<float, 1><<<grid, 256, 0, st>>>(gb);
    else k_gemm<float, 0><<<grid, 256, 0, st>>>(gb);
  } else if (use128) {
    if (cplx) k_gemm128_bf16x3<1><<<grid128, 256, 0, st>>>(gb);
    else k_gemm128_bf16x3<0><<<grid128, 256, 0, st>>>(gb);
  } else {
    if (cplx) k_gemm_bf16x3<CHM_BF16_BK, 1><<<grid, 256, 0, st>>>(gb);
    else k_gemm_bf16x3<CHM_BF16_BK, 0><<<grid, 256, 0, st>>>(gb);
  }
  if (timed) {
    tl.flops = 2.0 * gb.M * (double)gb.K * gb.N * nitems;
    tl.kind = kind;
    t_end(st, &tl);
  }
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// Zero-padded copy of a row-major [M][K] double matrix into the layout the
// products read: [round_up(M, 64)][gemm_lda(K)] as double, float or split BF16.
__global__ void k_pad_table(const double *__restrict__ src, int M, int K, int Mp, int lda, double *d64,
                            float *f32, unsigned short *hi, unsigned short *lo) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)Mp * lda) return;
  int i = (int)(e / lda), k = (int)(e % lda);
  double v = (i < M && k < K) ? src[(int64_t)i * K + k] : 0.0;
  if (d64) d64[e] = v;
  if (f32) f32[e] = (float)v;
  if (hi) {
    __bf16 h = (__bf16)(float)v;
    __bf16 l = (__bf16)(float)(v - (double)(float)h);
    hi[e] = __builtin_bit_cast(unsigned short, h);
    lo[e] = __builtin_bit_cast(unsigned short, l);
  }
}

extern "C" size_t chm_dht_table_elems(int32_t M, int32_t K) {
  return (size_t)((M + BM - 1) / BM * BM) * gemm_lda(K);
}

extern "C" int chm_pad_table(const double *src, int32_t M, int32_t K, double *d64, float *f32,
                             uint16_t *bf16_hi, uint16_t *bf16_lo, void *stream) {
  if (!src || M <= 0 || K <= 0 || ((bf16_hi == nullptr) != (bf16_lo == nullptr))) return CHM_EINVAL;
  int Mp = (M + BM - 1) / BM * BM, lda = gemm_lda(K);
  int64_t n = (int64_t)Mp * lda;
  k_pad_table<<<chm_blocks(n, 256), 256, 0, (hipStream_t)stream>>>(src, M, K, Mp, lda, d64, f32,
                                                                   (unsigned short *)bf16_hi,
                                                                   (unsigned short *)bf16_lo);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// Plain real product for table construction and tests:
// C[M x N] = A[M x K] @ B[K x N]; A padded as chm_pad_table writes it
// ([round_up(M,64)][round_up(K,64)]), B and C row-major, contiguous.
extern "C" int chm_gemm_f64(const double *A, const double *B, double *C, int32_t M, int32_t K,
                            int32_t N, void *stream) {
  GemmBatch gb = {};
  gb.M = M;
  gb.K = K;
  gb.N = N;
  gb.lda = gemm_lda(K);
  gb.ldb = N;
  gb.ldc = N;
  gb.bmode = 0;
  gb.it[0].A = A;
  gb.it[0].X = B;
  gb.it[0].C = C;
  return chm_gemm_launch(gb, 1, GEMM_F64, (hipStream_t)stream);
}
