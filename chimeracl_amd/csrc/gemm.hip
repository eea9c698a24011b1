// Discrete Hankel transform products on the matrix cores (gfx950 MFMA).
//
//   C[M x N] = A[M x K] @ B[K x N]      (row-major, A = a DHT-type matrix)
//
// A is a per-mode real matrix (DHT, DHT_inv, dDHT+-) stored zero-padded to
// [round_up(M, 64)][lda], lda = round_up(K, 32), so A tiles load without
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
#include <type_traits>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

#define BM 64
#define BN 64

// ---------------------------------------------------------------- B loader
// Element B[k][n] (zero outside K x N).
template <typename T>
__device__ __forceinline__ T load_b(const GemmBatch &gb, const GemmItem &it, int k, int n) {
  if (k >= gb.K || n >= gb.N) return (T)0;
  if (gb.bmode == 0) return ((const T *)it.X)[(int64_t)k * gb.ldb + n];
  const int Nx = gb.Nx, x = n >> 1, h = n & 1;
  T re = 0, im = 0;
  {
    int xs = it.remapX ? ((Nx - x) % Nx) : x;
    const T *X = (const T *)it.X + 2 * ((int64_t)k * Nx + xs);
    T a = X[0], b = X[1];
    if (it.remapX) a = -a;  // -conj(F(-kx)): (-re, +im)
    re = (T)it.alpha.x * a - (T)it.alpha.y * b;
    im = (T)it.alpha.x * b + (T)it.alpha.y * a;
  }
  if (it.Y) {
    int xs = it.remapY ? ((Nx - x) % Nx) : x;
    const T *Y = (const T *)it.Y + 2 * ((int64_t)k * Nx + xs);
    T a = Y[0], b = Y[1];
    if (it.remapY) a = -a;
    re += (T)it.beta.x * a - (T)it.beta.y * b;
    im += (T)it.beta.x * b + (T)it.beta.y * a;
  }
  return h ? im : re;
}

// ---------------------------------------------------------------- f64 / f32
template <typename T>
__global__ void __launch_bounds__(256) k_gemm(GemmBatch gb) {
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
  // B: thread owns column bn and rows bk, bk+4, bk+8, bk+12
  const int bn = tid & 63, bk = tid >> 6;
  T ra[4], rb[4];
  auto gload = [&](int k0) {
    const T *pa = A + (int64_t)(m0 + ar) * gb.lda + k0 + ak;
#pragma unroll
    for (int q = 0; q < 4; q++) ra[q] = pa[q];
#pragma unroll
    for (int q = 0; q < 4; q++) rb[q] = load_b<T>(gb, it, k0 + bk + 4 * q, n0 + bn);
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
    for (int q = 0; q < 4; q++) Bs[bk + 4 * q][bn] = rb[q];
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

__global__ void __launch_bounds__(256) k_gemm_bf16x3(GemmBatch gb) {
  constexpr int BK = 32;
  constexpr int KP = BK + 8;  // row stride 80 B: 16-byte aligned, staggers banks
  // operands stored k-contiguous so each lane reads 8 bf16 (16 B) at once
  __shared__ __attribute__((aligned(16))) unsigned short Ah[BM][KP], Al[BM][KP];
  __shared__ __attribute__((aligned(16))) unsigned short Bh[BN][KP], Bl[BN][KP];
  const GemmItem &it = gb.it[blockIdx.z];
  const int n0 = blockIdx.x * BN, m0 = blockIdx.y * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const unsigned short *Ahg = (const unsigned short *)it.A, *Alg = (const unsigned short *)it.A_lo;
  // A: thread loads row ar, k = ak..ak+7 as one 16-byte vector (hi and lo)
  const int ar = tid >> 2, ak = (tid & 3) * 8;
  // B: thread owns column bn and k = bk..bk+7
  const int bn = tid & 63, bk = (tid >> 6) * 8;
  us8 rah, ral;
  float rb[8];
  auto gload = [&](int k0) {
    int64_t o = (int64_t)(m0 + ar) * gb.lda + k0 + ak;
    rah = *(const us8 *)(Ahg + o);
    ral = *(const us8 *)(Alg + o);
#pragma unroll
    for (int q = 0; q < 8; q++) rb[q] = load_b<float>(gb, it, k0 + bk + q, n0 + bn);
  };
  f4 acc[2][2], cor[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = cor[i][j] = (f4){0, 0, 0, 0};
  gload(0);
  for (int k0 = 0; k0 < gb.K; k0 += BK) {
    *(us8 *)&Ah[ar][ak] = rah;
    *(us8 *)&Al[ar][ak] = ral;
    us8 bh, bl;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      unsigned short h, l;
      split_bf16(rb[q], h, l);
      bh[q] = h;
      bl[q] = l;
    }
    *(us8 *)&Bh[bn][bk] = bh;
    *(us8 *)&Bl[bn][bk] = bl;
    __syncthreads();
    if (k0 + BK < gb.K) gload(k0 + BK);
    bf8 ah[2], al[2], bhv[2], blv[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      int row = 32 * wr + 16 * i + (lane & 15), k = 8 * (lane >> 4);
      ah[i] = __builtin_bit_cast(bf8, *(const us8 *)&Ah[row][k]);
      al[i] = __builtin_bit_cast(bf8, *(const us8 *)&Al[row][k]);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
      int col = 32 * wc + 16 * j + (lane & 15), k = 8 * (lane >> 4);
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
  dim3 grid((gb.N + BN - 1) / BN, (gb.M + BM - 1) / BM, nitems);
  if (kind == GEMM_F64) k_gemm<double><<<grid, 256, 0, st>>>(gb);
  else if (kind == GEMM_F32) k_gemm<float><<<grid, 256, 0, st>>>(gb);
  else k_gemm_bf16x3<<<grid, 256, 0, st>>>(gb);
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
// ([round_up(M,64)][round_up(K,32)]), B and C row-major, contiguous.
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
