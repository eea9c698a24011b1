// Shared definitions of the chimeraCL MI355X kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/chimera_hip.h"

#define CHM_WAVE 64

// Error codes returned by every entry point (0 = success).  Positive values
// are hipError_t codes from a failed launch.
#define CHM_OK 0
#define CHM_EINVAL (-1)     // argument out of the supported range
#define CHM_EWORKSPACE (-2) // workspace too small
#define CHM_EUNSUPPORTED (-3)

#define CHM_CHECK_LAUNCH()                                   \
  do {                                                       \
    hipError_t e__ = hipGetLastError();                      \
    if (e__ != hipSuccess) return (int)e__;                  \
  } while (0)

template <typename T> struct cplx;
template <> struct cplx<double> { typedef double2 type; };
template <> struct cplx<float> { typedef float2 type; };

template <typename T>
__device__ __forceinline__ typename cplx<T>::type cmul(typename cplx<T>::type a,
                                                       typename cplx<T>::type b) {
  typename cplx<T>::type r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

static inline unsigned chm_blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static inline size_t chm_align(size_t n) { return (n + 255) & ~(size_t)255; }
