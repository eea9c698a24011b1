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
