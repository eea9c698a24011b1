// Discrete Hankel transform products on the matrix cores (gfx950 MFMA).
//
//   C[M x N] = A[M x K] @ B[K x N]      (row-major, A = a DHT-type matrix)
//
// A is a per-mode real matrix [Nk][Nk] (DHT, DHT_inv, dDHT+-).  B is either a
// real field [K][N] (mode 0), or a complex field [K][Nx] read as N = 2*Nx
// real columns, optionally formed on the fly as alpha*X + beta*Y with either
// operand read through the m = -1 map F_{-1}[k][x] = -conj(F_1[k][-x])
// (get_m1, transformer_generic.cl:3-24).  This replaces Reikna's MatrixMul
// (transformer_methods_cl.py:458-472) and the axpy/get_m1 kernels that fed it.
//
// Three arithmetic forms:
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

#define BM 64
#define BN 64
#define APAD 16

// ---------------------------------------------------------------- B loader
// Loads 4 consecutive real columns n..n+3 of row k of B (zero outside).
template <typename T>
__device__ __forceinline__ void load_b4(const GemmBatch &gb, const GemmItem &it, int k, int n, T v[4]) {
  if (k >= gb.K) {
    v[0] = v[1] = v[2] = v[3] = 0;
    return;
  }
  if (gb.bmode == 0) {
    const T *X = (const T *)it.X + (int64_t)k * gb.ldb;
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = (n + q < gb.N) ? X[n + q] : (T)0;
    return;
  }
  // complex: columns n, n+1 = (re, im) of x0 = n/2; n+2, n+3 of x0+1
  const int Nx = gb.Nx;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    int x = n / 2 + h;
    T re = 0, im = 0;
    if (x < Nx) {
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
    }
    v[2 * h] = re;
    v[2 * h + 1] = im;
  }
}

// ---------------------------------------------------------------- f64 / f32
// 256 threads, tile BM x BN, BK = 16; wave w owns rows 32*(w/2), cols 32*(w%2).
template <typename T>
__global__ void __launch_bounds__(256) k_gemm(GemmBatch gb) {
  constexpr int BK = 16;
  __shared__ T As[BK][BM + APAD];
  __shared__ T Bs[BK][BN + APAD];
  const GemmItem &it = gb.it[blockIdx.z];
  const int n0 = blockIdx.x * BN, m0 = blockIdx.y * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const T *A = (const T *)it.A;
  // global -> register staging
  const int ar = tid >> 2, ak = (tid & 3) * 4;  // A: row ar, k ak..ak+3
  const int bk = tid >> 4, bn = (tid & 15) * 4; // B: row bk, cols bn..bn+3
  T ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int row = m0 + ar, k = k0 + ak + q;
      ra[q] = (row < gb.M && k < gb.K) ? A[(int64_t)row * gb.lda + k] : (T)0;
    }
    load_b4<T>(gb, it, k0 + bk, n0 + bn, rb);
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
    for (int q = 0; q < 4; q++) Bs[bk][bn + q] = rb[q];
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
__device__ __forceinline__ void split_bf16(float v, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

__global__ void __launch_bounds__(256) k_gemm_bf16x3(GemmBatch gb) {
  constexpr int BK = 32;
  // operands stored k-contiguous so each lane reads 8 bf16 (16 B) at once
  __shared__ __attribute__((aligned(16))) __bf16 Ah[BM][BK + 8], Al[BM][BK + 8];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[BN][BK + 8], Bl[BN][BK + 8];
  const GemmItem &it = gb.it[blockIdx.z];
  const int n0 = blockIdx.x * BN, m0 = blockIdx.y * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const uint16_t *Ahg = (const uint16_t *)it.A, *Alg = (const uint16_t *)it.A_lo;
  f4 acc[2][2], cor[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = cor[i][j] = (f4){0, 0, 0, 0};
  const int ar = tid >> 2, ak = (tid & 3) * 8;  // A: row ar, k ak..ak+7
  for (int k0 = 0; k0 < gb.K; k0 += BK) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int row = m0 + ar, k = k0 + ak + q;
      bool ok = row < gb.M && k < gb.K;
      uint16_t h = ok ? Ahg[(int64_t)row * gb.lda + k] : 0, l = ok ? Alg[(int64_t)row * gb.lda + k] : 0;
      Ah[ar][ak + q] = __builtin_bit_cast(__bf16, h);
      Al[ar][ak + q] = __builtin_bit_cast(__bf16, l);
    }
#pragma unroll
    for (int s = 0; s < 2; s++) {
      int t = tid + 256 * s;
      int bk = t >> 4, bn = (t & 15) * 4;
      float v[4];
      load_b4<float>(gb, it, k0 + bk, n0 + bn, v);
#pragma unroll
      for (int q = 0; q < 4; q++) split_bf16(v[q], Bh[bn + q][bk], Bl[bn + q][bk]);
    }
    __syncthreads();
    bf8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      int row = 32 * wr + 16 * i + (lane & 15), k = 8 * (lane >> 4);
      ah[i] = *(const bf8 *)&Ah[row][k];
      al[i] = *(const bf8 *)&Al[row][k];
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
      int col = 32 * wc + 16 * j + (lane & 15), k = 8 * (lane >> 4);
      bh[j] = *(const bf8 *)&Bh[col][k];
      bl[j] = *(const bf8 *)&Bl[col][k];
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], cor[i][j], 0, 0, 0);
        cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], cor[i][j], 0, 0, 0);
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

int chm_gemm_launch(const GemmBatch &gb, int nitems, int kind, hipStream_t st) {
  if (nitems <= 0) return CHM_OK;
  if (nitems > GEMM_MAXB || gb.M <= 0 || gb.K <= 0 || gb.N <= 0) return CHM_EINVAL;
  dim3 grid((gb.N + BN - 1) / BN, (gb.M + BM - 1) / BM, nitems);
  if (kind == GEMM_F64) k_gemm<double><<<grid, 256, 0, st>>>(gb);
  else if (kind == GEMM_F32) k_gemm<float><<<grid, 256, 0, st>>>(gb);
  else k_gemm_bf16x3<<<grid, 256, 0, st>>>(gb);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// Plain real product for table construction and tests:
// C[M x N] = A[M x K] @ B[K x N], all double, row-major, contiguous.
extern "C" int chm_gemm_f64(const double *A, const double *B, double *C, int32_t M, int32_t K,
                            int32_t N, void *stream) {
  GemmBatch gb = {};
  gb.M = M;
  gb.K = K;
  gb.N = N;
  gb.lda = K;
  gb.ldb = N;
  gb.ldc = N;
  gb.bmode = 0;
  gb.it[0].A = A;
  gb.it[0].X = B;
  gb.it[0].C = C;
  return chm_gemm_launch(gb, 1, GEMM_F64, (hipStream_t)stream);
}
