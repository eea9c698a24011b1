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

__device__ __forceinline__ double dev_sqrt(double v) { return sqrt(v); }
__device__ __forceinline__ float dev_sqrt(float v) { return sqrtf(v); }
__device__ __forceinline__ double dev_floor(double v) { return floor(v); }
__device__ __forceinline__ float dev_floor(float v) { return floorf(v); }

// Cell key of index_and_sum_in_cell, particles_generic.cl:105-124: the cell
// ix + ir*(Nx-1) of a valid cell, else the overflow bin (Nr-1)*(Nx-1).
template <typename T>
__device__ __forceinline__ uint32_t cell_key(T xp, T yp, T zp, int Nx_loc, int Nr_loc, T xmin,
                                             T dx_inv, T rmin, T dr_inv) {
  T r = dev_sqrt(yp * yp + zp * zp);
  int ix = (int)dev_floor((xp - xmin) * dx_inv);
  int ir = (int)dev_floor((r - rmin) * dr_inv);
  if (ix > 0 && ix < Nx_loc - 1 && ir < Nr_loc - 1 && ir >= 0) return (uint32_t)(ix + ir * Nx_loc);
  return (uint32_t)(Nr_loc * Nx_loc);
}

static inline unsigned chm_blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static inline size_t chm_align(size_t n) { return (n + 255) & ~(size_t)255; }
