// Batched complex FFT of rows held in LDS: Stockham autosort, radix 8 with a
// final radix-4 or radix-2 pass, twiddles from a table exp(-2 pi i t / N).
// Used by the transforms (chm_transform_*) and by the fused damping kernel.
#pragma once
#include "chm_common.h"

template <typename T> struct C2 { typedef typename cplx<T>::type t; };

template <typename T2> __device__ __forceinline__ T2 c_add(T2 a, T2 b) { T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T2> __device__ __forceinline__ T2 c_sub(T2 a, T2 b) { T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename T2> __device__ __forceinline__ T2 c_mul(T2 a, T2 b) {
  T2 r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
// multiply by SIGN*i
template <int SIGN, typename T2> __device__ __forceinline__ T2 c_muli(T2 a) {
  T2 r;
  if (SIGN < 0) { r.x = a.y; r.y = -a.x; }
  else { r.x = -a.y; r.y = a.x; }
  return r;
}

template <int SIGN, typename T2> __device__ __forceinline__ void dft2(T2 *v) {
  T2 a = v[0], b = v[1];
  v[0] = c_add(a, b);
  v[1] = c_sub(a, b);
}

template <int SIGN, typename T2> __device__ __forceinline__ void dft4(T2 &v0, T2 &v1, T2 &v2, T2 &v3) {
  T2 t0 = c_add(v0, v2), t1 = c_sub(v0, v2), t2 = c_add(v1, v3), t3 = c_muli<SIGN>(c_sub(v1, v3));
  v0 = c_add(t0, t2);
  v2 = c_sub(t0, t2);
  v1 = c_add(t1, t3);
  v3 = c_sub(t1, t3);
}

template <int SIGN, typename T, typename T2> __device__ __forceinline__ void dft8(T2 *v) {
  const T h = (T)0.70710678118654752440084436210484903928483593768847;
  T2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  T2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<SIGN>(e0, e1, e2, e3);
  dft4<SIGN>(o0, o1, o2, o3);
  // o_k *= exp(SIGN i pi k / 4)
  T2 w1, w3;
  w1.x = h; w1.y = (T)SIGN * h;
  w3.x = -h; w3.y = (T)SIGN * h;
  o1 = c_mul(o1, w1);
  o2 = c_muli<SIGN>(o2);
  o3 = c_mul(o3, w3);
  v[0] = c_add(e0, o0); v[4] = c_sub(e0, o0);
  v[1] = c_add(e1, o1); v[5] = c_sub(e1, o1);
  v[2] = c_add(e2, o2); v[6] = c_sub(e2, o2);
  v[3] = c_add(e3, o3); v[7] = c_sub(e3, o3);
}

template <int R, int SIGN, typename T, typename T2> __device__ __forceinline__ void dftR(T2 *v) {
  if (R == 8) dft8<SIGN, T>(v);
  else if (R == 4) dft4<SIGN>(v[0], v[1], v[2], v[3]);
  else if (R == 2) dft2<SIGN>(v);
}

// One Stockham pass of radix R over RPB rows of length N in `buf`
// (row stride N), in place: every value is read before any is written.
template <typename T, int N, int R, int NS, int SIGN, int RPB, int NT>
__device__ __forceinline__ void stockham_pass(typename C2<T>::t *buf, const typename C2<T>::t *__restrict__ tw) {
  typedef typename C2<T>::t T2;
  constexpr int SLOTS = RPB * (N / R);
  constexpr int S = (SLOTS + NT - 1) / NT;
  T2 v[S][R];
#pragma unroll
  for (int s = 0; s < S; s++) {
    int slot = threadIdx.x + s * NT;
    if (slot < SLOTS) {
      int row = slot / (N / R), j = slot % (N / R);
      T2 *b = buf + row * N;
#pragma unroll
      for (int r = 0; r < R; r++) v[s][r] = b[j + r * (N / R)];
      if (NS > 1) {
        int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; r++) {
          int t = (r * k * (N / (NS * R))) % N;
          T2 w = tw[t];
          if (SIGN > 0) w.y = -w.y;
          v[s][r] = c_mul(v[s][r], w);
        }
      }
      dftR<R, SIGN, T>(v[s]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < S; s++) {
    int slot = threadIdx.x + s * NT;
    if (slot < SLOTS) {
      int row = slot / (N / R), j = slot % (N / R);
      int k = j % NS;
      int base = (j / NS) * NS * R + k;
      T2 *b = buf + row * N;
#pragma unroll
      for (int r = 0; r < R; r++) b[base + r * NS] = v[s][r];
    }
  }
  __syncthreads();
}

template <int N> struct Log2 { enum { v = 1 + Log2<N / 2>::v }; };
template <> struct Log2<1> { enum { v = 0 }; };

template <typename T, int N, int SIGN, int RPB, int NT, int NS>
__device__ __forceinline__ void fft_rest(typename C2<T>::t *buf, const typename C2<T>::t *__restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int REM = N / NS;
    constexpr int R = (REM % 8 == 0) ? 8 : ((REM % 4 == 0) ? 4 : 2);
    stockham_pass<T, N, R, NS, SIGN, RPB, NT>(buf, tw);
    fft_rest<T, N, SIGN, RPB, NT, NS * R>(buf, tw);
  }
}

// Unnormalised FFT (SIGN=-1) or inverse (SIGN=+1) of RPB rows in LDS.
// Caller has synchronised after filling buf; returns synchronised.
template <typename T, int N, int SIGN, int RPB, int NT>
__device__ __forceinline__ void fft_lds(typename C2<T>::t *buf, const typename C2<T>::t *__restrict__ tw) {
  fft_rest<T, N, SIGN, RPB, NT, 1>(buf, tw);
}

// ---------------------------------------------------------------- 4096 in registers
// 16-point DFT of v[0..15] in registers, natural order in and out (radix
// 4 x 4: n = a + 4b, k = c + 4d; DFT-4 over b, twiddle W16^(ac), DFT-4 over a).
template <int SIGN, typename T, typename T2> __device__ __forceinline__ void dft16(T2 *v) {
  const T c1 = (T)0.92387953251128675612818318939678828682241662586364;  // cos(pi/8)
  const T s1 = (T)0.38268343236508977172845998403039886676134456248563;  // sin(pi/8)
  const T h = (T)0.70710678118654752440084436210484903928483593768847;   // cos(pi/4)
  const T sg = (T)SIGN;
#pragma unroll
  for (int a = 0; a < 4; a++) dft4<SIGN>(v[a], v[a + 4], v[a + 8], v[a + 12]);
  // v[a + 4c] now holds t[a][c]; twiddles W16^(ac) = exp(SIGN i 2 pi ac / 16)
  T2 w;
  w.x = c1; w.y = sg * s1; v[1 + 4] = c_mul(v[1 + 4], w);       // a=1 c=1: e=1
  w.x = h;  w.y = sg * h;  v[1 + 8] = c_mul(v[1 + 8], w);       // a=1 c=2: e=2
                           v[2 + 4] = c_mul(v[2 + 4], w);       // a=2 c=1: e=2
  w.x = s1; w.y = sg * c1; v[1 + 12] = c_mul(v[1 + 12], w);     // a=1 c=3: e=3
                           v[3 + 4] = c_mul(v[3 + 4], w);       // a=3 c=1: e=3
  v[2 + 8] = c_muli<SIGN>(v[2 + 8]);                            // a=2 c=2: e=4
  w.x = -h; w.y = sg * h;  v[2 + 12] = c_mul(v[2 + 12], w);     // a=2 c=3: e=6
                           v[3 + 8] = c_mul(v[3 + 8], w);       // a=3 c=2: e=6
  w.x = -c1; w.y = -sg * s1; v[3 + 12] = c_mul(v[3 + 12], w);   // a=3 c=3: e=9
#pragma unroll
  for (int c = 0; c < 4; c++) dft4<SIGN>(v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]);
  // v[4c + d] holds V[c + 4d]: transpose to natural order
  T2 o[16];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int d = 0; d < 4; d++) o[c + 4 * d] = v[4 * c + d];
#pragma unroll
  for (int k = 0; k < 16; k++) v[k] = o[k];
}

// LDS elements (scalars) the 4096-point register FFT needs: 16 rows of 273
// (the row stride and the 17-element sub-stride put the 16 lanes of each
// quarter-wave on distinct banks); the real and imaginary parts of an
// exchange pass through the buffer one after the other, so it holds one
// scalar per element (35 KB in FP64, 17.5 KB in FP32).
#define FFT4096_S2 273
#define FFT4096_LDS (16 * FFT4096_S2)

// v[i] -> sbuf[wi(i)], then v[i] <- sbuf[ri(i)], real parts then imaginary
// parts, with the barriers that separate every write from the reads of
// other threads.
template <typename T, typename T2, typename WI, typename RI>
__device__ __forceinline__ void exchange16(T2 *v, T *sbuf, WI wi, RI ri) {
#pragma unroll
  for (int i = 0; i < 16; i++) sbuf[wi(i)] = v[i].x;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; i++) v[i].x = sbuf[ri(i)];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; i++) sbuf[wi(i)] = v[i].y;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 16; i++) v[i].y = sbuf[ri(i)];
  __syncthreads();
}

// Unnormalised FFT (SIGN = -1) or inverse (SIGN = +1) of one 4096-point row
// with 256 threads: thread t holds elements t + 256 j (j = 0..15) in v on
// entry and on exit.  Three 16-point stages in registers (4096 = 16^3,
// Cooley-Tukey n = n2 + 256 n1, k = k1 + 16 k2, then the 256-point part
// n2 = m2 + 16 m1, k2 = q1 + 16 q2), two exchanges through LDS.  tw[e] =
// exp(-2 pi i e / 4096).  Contains barriers: every thread of the block calls
// it; sbuf (FFT4096_LDS scalars) may be reused by the caller once it returns.
template <int SIGN, typename T>
__device__ __forceinline__ void fft4096_reg(typename C2<T>::t *v, T *sbuf,
                                            const typename C2<T>::t *__restrict__ tw) {
  typedef typename C2<T>::t T2;
  const int t = threadIdx.x;
  // Compiler-only barrier: the twiddles (and the caller's phase factors)
  // are loaded again in every call instead of being kept live across
  // calls, which spilled registers.
  __asm__ __volatile__("" ::: "memory");
  // stage A: DFT-16 over n1 = j, twiddle W4096^(n2 k1), n2 = t
  dft16<SIGN, T>(v);
#pragma unroll
  for (int k1 = 1; k1 < 16; k1++) {
    T2 w = tw[t * k1];
    if (SIGN > 0) w.y = -w.y;
    v[k1] = c_mul(v[k1], w);
  }
  // Y[n2][k1] -> thread (kb, m2) = (t >> 4, t & 15) gets Y[m2 + 16 m1][kb]
  const int kb = t >> 4, m2 = t & 15;
  exchange16<T>(v, sbuf, [&](int k1) { return k1 * FFT4096_S2 + t; },
                [&](int m1) { return kb * FFT4096_S2 + m2 + 16 * m1; });
  // stage B: DFT-16 over m1, twiddle W256^(m2 q1) = W4096^(16 m2 q1)
  dft16<SIGN, T>(v);
#pragma unroll
  for (int q1 = 1; q1 < 16; q1++) {
    T2 w = tw[16 * m2 * q1];
    if (SIGN > 0) w.y = -w.y;
    v[q1] = c_mul(v[q1], w);
  }
  // U[kb][m2][q1] -> thread (kc, qc) = (t & 15, t >> 4) gets U[kc][m][qc]
  const int kc = t & 15, qc = t >> 4;
  exchange16<T>(v, sbuf, [&](int q1) { return kb * FFT4096_S2 + q1 * 17 + m2; },
                [&](int m) { return kc * FFT4096_S2 + qc * 17 + m; });
  // stage C: DFT-16 over m2; output k = kc + 16 qc + 256 q2 = t + 256 q2
  dft16<SIGN, T>(v);
}
