// Batched DHT product interface shared by the transform and spectral files.
#pragma once
#include "chm_common.h"

#define GEMM_MAXB 24
#define GEMM_F64 0
#define GEMM_F32 1
#define GEMM_BF16X3 2

struct GemmItem {
  const void *A;    // [M][lda] (double / float / bf16 hi)
  const void *A_lo; // bf16 lo part (bf16x3 only)
  const void *X;    // B operand (bmode 0: real [K][ldb]; bmode 1: complex [K][Nx])
  const void *Y;    // optional second complex operand (bmode 1)
  void *C;          // [M][ldc]
  double2 alpha, beta;  // B = alpha*X + beta*Y (bmode 1)
  int remapX, remapY;   // read X / Y as F_{-1} = -conj(F_1(-kx))
};

struct GemmBatch {
  GemmItem it[GEMM_MAXB];
  int M, K, N;           // N = real columns of C (Nx for a real field, 2*Nx complex)
  int lda, ldb, ldc;     // leading dimensions in elements
  int bmode;             // 0 = real B, 1 = complex alpha*X + beta*Y
  int Nx;                // complex columns (bmode 1)
};

int chm_gemm_launch(const GemmBatch &gb, int nitems, int kind, hipStream_t st);

// Leading dimension of a padded DHT-type table (K rounded up to 64, a
// whole number of K-steps of every product).
static inline __host__ __device__ int gemm_lda(int K) { return (K + 63) / 64 * 64; }
