// Particle <-> grid kernels: deterministic CIC deposition of rho and J into
// the azimuthal modes, axis post-processing, ghost-row warp, and the field
// gather + relativistic Boris push.  gfx950, wave64.
//
// Deposition (replaces the 4-colour depose_scalar / depose_vector launches,
// grid_deposit_m1.cl:20-327, grid_methods_cl.py:44-95) runs in two passes and
// uses no float atomics:
//   pass 1 (one lane per valid cell, particles staged through LDS): the
//          cell's particles, in sorted order, are summed into private
//          accumulators exactly as the reference kernel does; the 4 node
//          corners x components are written to a staging buffer
//          (component-major, coalesced).
//   pass 2 (one lane per node): each node adds the partial sums of its (up
//          to) four cells in the reference's colour order 0,1,2,3 -- the order
//          in which the reference's serialised colour launches add them.
// The result is bit-identical to the reference's deposit for a given
// in-cell order.
#include "chm_common.h"

template <typename T> __device__ __forceinline__ T dsqrt(T v);
template <> __device__ __forceinline__ double dsqrt<double>(double v) { return sqrt(v); }
template <> __device__ __forceinline__ float dsqrt<float>(float v) { return sqrtf(v); }
template <typename T> __device__ __forceinline__ T dfloor(T v);
template <> __device__ __forceinline__ double dfloor<double>(double v) { return floor(v); }
template <> __device__ __forceinline__ float dfloor<float>(float v) { return floorf(v); }

// a*b + c.  FP64: a product and a sum, each rounded (the reference's
// operations in its order, bit-exact against the oracle).  FP32 (the mixed
// form, compared with FP64 to a tolerance): one fused multiply-add.
template <typename T> __device__ __forceinline__ T madd(T a, T b, T c);
template <> __device__ __forceinline__ double madd<double>(double a, double b, double c) {
  return a * b + c;
}
template <> __device__ __forceinline__ float madd<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

struct PlanePtrs {
  void *m0[3];
  void *mc[3][CHM_MAXM];
};

// ------------------------------------------------------------------ pass 1
// One wave per group of CELLS consecutive cells.  Their particles occupy one
// contiguous range of sorted slots, which the wave stages into LDS in chunks
// of CAP slots with coalesced loads (sort_indx is close to the identity for a
// cell-major plasma, so the gathers are nearly coalesced too): all of a
// chunk's particle values are loaded first, then each lane computes its
// particles' CIC weights C_ij (weight included), mode factors
// ((y + i z)/r)^m and values exactly as the reference kernel does; then one
// lane per (cell, component, corner) sums its accumulators over the cell's
// particles sequentially, in sorted order.  Every accumulator sees the same
// operands in the same order as in the reference kernel, so the sums are
// bit-identical to it.  (Walking several groups per wave, with the next
// group's loads in flight while a group is summed, measured slower at C2:
// more groups per wave, slower.)
//
// KIND 0: rho (wp = w q); 1: J (wp = w g_inv q, values u); 2: field
// adjoints of the gather (wp = 1, values in ux/uy/uz, modes weighted 2 e_m).
// PUSH (KIND 1): the lane that stages a particle also applies the next
// coordinate push (x += u dt g_inv, push_x's operations) and writes the
// pushed position (to xo, which may be x) and its cell key (the keys of the
// following sort); the sums use the positions before the push.
template <typename T, int M, int KIND> struct DepCfg {
  static constexpr bool VEC = KIND != 0;
  static constexpr int NK = VEC ? 3 : 1;              // components
  static constexpr int TPC = 4 * NK;                   // lanes per cell
  static constexpr int CELLS = 64 / TPC;               // cells per wave
  // staged slots per chunk (measured at C2)
  static constexpr int CAP = sizeof(T) == 8 ? (VEC ? 192 : 128) : (VEC ? 192 : 256);
  // mode factors m >= 2 recomputed from the staged m = 1 pair by the summing
  // lanes, with the staging's operations (identical values), for less LDS
  // per wave (kind 2 stages doubled factors and keeps them all)
  static constexpr bool RECOMP = KIND != 2 && M >= 2;
  static constexpr int NF = RECOMP ? 1 : M;            // staged mode-factor pairs
  static constexpr int NV = 4 + 2 * NF + (VEC ? 3 : 0); // staged values per particle
};
// Lanes walk their own cell's slots, ~ppc apart: one pad element per 32
// spreads them over the LDS banks.
#define DEP_SW(t) ((t) + ((t) >> 5))

template <typename T, int M, int KIND, bool PUSH>
__global__ void __launch_bounds__(64)
    k_cell_sums(const uint32_t *__restrict__ sort_indx, const uint32_t *__restrict__ cell_offset,
                const T *x, const T *y, const T *z,
                const T *__restrict__ ux, const T *__restrict__ uy, const T *__restrict__ uz,
                const T *__restrict__ g_inv, const T *__restrict__ w, T charge, int Nx, int Nr,
                T xmin, T dx_inv, T rmin, T dr_inv, T *__restrict__ part, T dt,
                uint32_t *__restrict__ key_out, T *xo, T *yo, T *zo) {
  typedef DepCfg<T, M, KIND> Cfg;
  constexpr bool VEC = Cfg::VEC;
  constexpr int CAP = Cfg::CAP, NV = Cfg::NV, TPC = Cfg::TPC, CELLS = Cfg::CELLS;
  constexpr int MM = M > 0 ? M : 1;
  constexpr int VE = 4, VU = 4 + 2 * Cfg::NF;  // first mode-factor / value slot
  constexpr int R = (CAP + 63) / 64;
  constexpr int NQ = 1 + 2 * M;
  __shared__ T st[NV][CAP + CAP / 32];
  const int Nx_cell = Nx - 1, Nr_cell = Nr - 1;
  const int64_t ncell = (int64_t)Nx_cell * Nr_cell;
  const int lane = threadIdx.x;
  const int task = lane % TPC, k = task >> 2, corner = task & 3;
  // this wave's cells [c0, c1)
  const int64_t c0 = (int64_t)blockIdx.x * CELLS;
  const int64_t c1 = (c0 + CELLS < ncell) ? c0 + CELLS : ncell;

  uint32_t pix[R];  // the chunk's sorted-slot indices
  auto load_idx = [&](uint32_t (&dst)[R], uint32_t s, uint32_t e) {
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
      const uint32_t t = lane + 64 * rr;
      dst[rr] = (t < CAP && s + t < e) ? sort_indx[s + t] : 0u;
    }
  };
  // particle values of the chunk whose indices are in pix
  T lx[R], ly[R], lz[R], lw[R], lg[R], lu[R][3];
  auto load_data = [&](uint32_t n) {
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
      const uint32_t t = lane + 64 * rr;
      lx[rr] = ly[rr] = lz[rr] = lw[rr] = lg[rr] = 0;
      lu[rr][0] = lu[rr][1] = lu[rr][2] = 0;
      if (t < n) {
        const uint32_t p = pix[rr];
        lx[rr] = x[p];
        ly[rr] = y[p];
        lz[rr] = z[p];
        if (KIND != 2) lw[rr] = w[p];
        if (KIND == 1) lg[rr] = g_inv[p];
        if (VEC) {
          lu[rr][0] = ux[p];
          lu[rr][1] = uy[p];
          lu[rr][2] = uz[p];
        }
      }
    }
  };
  // the chunk's weights, mode factors and values into LDS (and the push)
  auto stage = [&](uint32_t n) {
#pragma unroll
    for (int rr = 0; rr < R; rr++) {
      const uint32_t t = lane + 64 * rr;
      if (t >= n) continue;
      const uint32_t q = DEP_SW(t);
      const T xp = lx[rr], yp = ly[rr], zp = lz[rr];
      T wp;
      if (KIND == 1) {
        const T gp = lg[rr];
        wp = lw[rr] * gp * charge;
        if (PUSH) {
          const uint32_t p = pix[rr];
          const T dt_g = dt * gp;
          const T xn = xp + lu[rr][0] * dt_g, yn = yp + lu[rr][1] * dt_g,
                  zn = zp + lu[rr][2] * dt_g;
          xo[p] = xn;
          yo[p] = yn;
          zo[p] = zn;
          key_out[p] = cell_key<T>(xn, yn, zn, Nx - 1, Nr - 1, xmin, dx_inv, rmin, dr_inv);
        }
      } else if (KIND == 0) {
        wp = lw[rr] * charge;
      } else {
        wp = (T)1;
      }
      const T rp = dsqrt<T>(yp * yp + zp * zp);
      T rp_inv;
      if (VEC) {
        rp_inv = 0;
        if (rp > 0) rp_inv = (T)1. / rp;
      } else {
        rp_inv = (T)1. / rp;
      }
      // ((y + i z)/r)^m, m = 1..M
      T er[MM + 1], ei[MM + 1];
      er[1] = yp * rp_inv;
      ei[1] = zp * rp_inv;
#pragma unroll
      for (int m = 2; m <= M; m++) {
        er[m] = er[m - 1] * er[1] - ei[m - 1] * ei[1];
        ei[m] = er[m - 1] * ei[1] + ei[m - 1] * er[1];
      }
      if (KIND == 2) {
#pragma unroll
        for (int m = 1; m <= M; m++) {
          er[m] = (T)2 * er[m];
          ei[m] = (T)2 * ei[m];
        }
      }
      // the particle's cell is the one its key names
      const T fx = (xp - xmin) * dx_inv, fr = (rp - rmin) * dr_inv;
      const T sX1 = fx - dfloor<T>(fx), sR1 = fr - dfloor<T>(fr);
      T sX0 = (T)1.0 - sX1;
      const T sR0 = (T)1.0 - sR1;
      T sXw1 = sX1;
      sX0 *= wp;
      sXw1 *= wp;
      st[0][q] = sR0 * sX0;
      st[1][q] = sR0 * sXw1;
      st[2][q] = sR1 * sX0;
      st[3][q] = sR1 * sXw1;
#pragma unroll
      for (int m = 1; m <= Cfg::NF; m++) {
        st[VE + 2 * (m - 1)][q] = er[m];
        st[VE + 2 * (m - 1) + 1][q] = ei[m];
      }
      if (VEC) {
#pragma unroll
        for (int kk = 0; kk < 3; kk++) st[VU + kk][q] = lu[rr][kk];
      }
    }
  };

  const uint32_t S0 = cell_offset[c0], S1 = cell_offset[c1];
  const int64_t c = c0 + lane / TPC;
  bool act = lane < CELLS * TPC && c < c1;
  uint32_t ip0 = 0, ip1 = 0;
  if (act) {
    const int ir = (int)(c / Nx_cell), ix = (int)(c - (int64_t)ir * Nx_cell);
    act = ix > 0 && ix < Nx_cell - 1 && ir < Nr_cell - 1;
    ip0 = cell_offset[c];
    ip1 = cell_offset[c + 1];
  }
  T a0 = 0, ar[MM], ai[MM];
#pragma unroll
  for (int m = 0; m < MM; m++) ar[m] = ai[m] = 0;
  for (uint32_t cs = S0; cs < S1; cs += CAP) {
    const uint32_t n = (S1 - cs < CAP) ? S1 - cs : CAP;
    load_idx(pix, cs, S1);
    load_data(n);
    stage(n);
    __syncthreads();
    if (act) {
      const uint32_t lo = ip0 > cs ? ip0 : cs, hi = ip1 < cs + n ? ip1 : cs + n;
#pragma unroll 4
      for (uint32_t ip = lo; ip < hi; ip++) {
        const uint32_t t = DEP_SW(ip - cs);
        const T C = st[corner][t];
        // the particle's mode factors ((y + i z)/r)^m
        T fr[MM + 1], fi[MM + 1];
#pragma unroll
        for (int m = 1; m <= Cfg::NF; m++) {
          fr[m] = st[VE + 2 * (m - 1)][t];
          fi[m] = st[VE + 2 * (m - 1) + 1][t];
        }
#pragma unroll
        for (int m = Cfg::NF + 1; m <= M; m++) {
          fr[m] = fr[m - 1] * fr[1] - fi[m - 1] * fi[1];
          fi[m] = fr[m - 1] * fi[1] + fi[m - 1] * fr[1];
        }
        if (VEC) {
          const T u = st[VU + k][t];
          const T jproj = C * u;
          a0 = madd(C, u, a0);
#pragma unroll
          for (int m = 1; m <= M; m++) {
            ar[m - 1] = madd(jproj, fr[m], ar[m - 1]);
            ai[m - 1] = madd(jproj, fi[m], ai[m - 1]);
          }
        } else {
          a0 += C;
#pragma unroll
          for (int m = 1; m <= M; m++) {
            ar[m - 1] = madd(C, fr[m], ar[m - 1]);
            ai[m - 1] = madd(C, fi[m], ai[m - 1]);
          }
        }
      }
    }
    __syncthreads();
  }
  if (!act || ip0 == ip1) return;
  // staging layout: part[((k*(1+2M) + q)*4 + corner) * ncell + c]
  part[((int64_t)(k * NQ) * 4 + corner) * ncell + c] = a0;
#pragma unroll
  for (int m = 1; m <= M; m++) {
    part[((int64_t)(k * NQ + 2 * m - 1) * 4 + corner) * ncell + c] = ar[m - 1];
    part[((int64_t)(k * NQ + 2 * m) * 4 + corner) * ncell + c] = ai[m - 1];
  }
}

// The coordinate push + cell key of the particles a PUSH deposit (or a
// pre-pushing gather) does not visit: sorted slots from Np_stay (=
// cell_offset[ncell], the start of the out-of-domain bin) to Np; reads x,
// writes xo (may be x).  Grid-stride over a fixed grid, as Np_stay is known on
// the device only.
template <typename T>
__global__ void __launch_bounds__(256)
    k_push_keys_tail(const uint32_t *__restrict__ sort_indx, const uint32_t *__restrict__ cell_offset,
                     int64_t ncell, int64_t Np, const T *x, const T *y, const T *z,
                     const T *__restrict__ ux, const T *__restrict__ uy,
                     const T *__restrict__ uz, const T *__restrict__ g_inv, T dt, int Nx, int Nr,
                     T xmin, T dx_inv, T rmin, T dr_inv, uint32_t *__restrict__ key_out, T *xo,
                     T *yo, T *zo) {
  const int64_t np_stay = cell_offset[ncell];
  for (int64_t slot = np_stay + (int64_t)blockIdx.x * 256 + threadIdx.x; slot < Np;
       slot += (int64_t)gridDim.x * 256) {
    const uint32_t p = sort_indx[slot];
    const T dt_g = dt * g_inv[p];
    const T xn = x[p] + ux[p] * dt_g, yn = y[p] + uy[p] * dt_g, zn = z[p] + uz[p] * dt_g;
    xo[p] = xn;
    yo[p] = yn;
    zo[p] = zn;
    key_out[p] = cell_key<T>(xn, yn, zn, Nx - 1, Nr - 1, xmin, dx_inv, rmin, dr_inv);
  }
}

// ------------------------------------------------------------------ pass 2
// The node's (up to) four cells are found first (their offsets loaded
// together); then, component by component, the partial sums of all four are
// loaded together before they are added in colour order: the two cells of a
// column pair are read from the same lines of a corner plane (even cells by
// one, odd by the other), so the lines are still in the L2 for the second
// read instead of being fetched from HBM twice.
template <typename T, int M, int NK>
__global__ void __launch_bounds__(256)
    k_node_add(const uint32_t *__restrict__ cell_offset, const T *__restrict__ part, int Nx, int Nr,
               PlanePtrs planes) {
  const int Nx_cell = Nx - 1, Nr_cell = Nr - 1;
  const int64_t ncell = (int64_t)Nx_cell * Nr_cell;
  int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= (int64_t)(Nr - 1) * Nx) return;
  int ir = (int)(node / Nx), ix = (int)(node - (int64_t)ir * Nx);
  constexpr int NQ = 1 + 2 * M;
  int64_t cc[4];
  int corner[4];
  bool use[4];
#pragma unroll
  for (int col = 0; col < 4; col++) {
    const int px = col & 1, pr = col >> 1;
    const int cx = ((ix & 1) == px) ? ix : ix - 1;
    const int cr = ((ir & 1) == pr) ? ir : ir - 1;
    use[col] = cx > 0 && cx < Nx_cell - 1 && cr >= 0 && cr < Nr_cell - 1;
    cc[col] = use[col] ? cx + (int64_t)cr * Nx_cell : 0;
    corner[col] = 2 * (ir - cr) + (ix - cx);
  }
  uint32_t lo[4], hi[4];
#pragma unroll
  for (int col = 0; col < 4; col++) {
    lo[col] = hi[col] = 0u;
    if (use[col]) {
      lo[col] = cell_offset[cc[col]];
      hi[col] = cell_offset[cc[col] + 1];
    }
  }
  bool any = false;
#pragma unroll
  for (int col = 0; col < 4; col++) {
    use[col] = use[col] && lo[col] != hi[col];
    any = any || use[col];
  }
  if (!any) return;
  T acc[NK][NQ];
#pragma unroll
  for (int k = 0; k < NK; k++) {
    acc[k][0] = ((const T *)planes.m0[k])[node];
#pragma unroll
    for (int m = 1; m <= M; m++) {
      const T *pc = (const T *)planes.mc[k][m - 1];
      acc[k][2 * m - 1] = pc[2 * node];
      acc[k][2 * m] = pc[2 * node + 1];
    }
  }
#pragma unroll
  for (int k = 0; k < NK; k++) {
    T v[4][NQ];
#pragma unroll
    for (int col = 0; col < 4; col++)
#pragma unroll
      for (int q = 0; q < NQ; q++)
        v[col][q] = use[col] ? part[((int64_t)(k * NQ + q) * 4 + corner[col]) * ncell + cc[col]] : (T)0;
#pragma unroll
    for (int col = 0; col < 4; col++)
      if (use[col])
#pragma unroll
        for (int q = 0; q < NQ; q++) acc[k][q] = acc[k][q] + v[col][q];
  }
#pragma unroll
  for (int k = 0; k < NK; k++) {
    ((T *)planes.m0[k])[node] = acc[k][0];
#pragma unroll
    for (int m = 1; m <= M; m++) {
      T *pc = (T *)planes.mc[k][m - 1];
      pc[2 * node] = acc[k][2 * m - 1];
      pc[2 * node + 1] = acc[k][2 * m];
    }
  }
}

extern "C" size_t chm_deposit_workspace_bytes(int32_t Nx, int32_t Nr, int32_t M, int32_t vector) {
  size_t ncell = (size_t)(Nx - 1) * (Nr - 1);
  size_t nq = (size_t)(vector ? 3 : 1) * (1 + 2 * M) * 4;
  return chm_align(nq * ncell * sizeof(double));
}

template <typename T>
static void launch_push_tail(const uint32_t *sort_indx, const uint32_t *cell_offset, int64_t ncell,
                             int64_t Np, const T *x, const T *y, const T *z, const T *ux,
                             const T *uy, const T *uz, const T *g_inv, T dt, const chm_grid_t *g,
                             uint32_t *keys, T *xo, T *yo, T *zo, hipStream_t st) {
  const int64_t nt = chm_blocks(Np, 256);
  k_push_keys_tail<T><<<(unsigned)(nt < 1024 ? (nt < 1 ? 1 : nt) : 1024), 256, 0, st>>>(
      sort_indx, cell_offset, ncell, Np, x, y, z, ux, uy, uz, g_inv, dt, g->Nx, g->Nr, (T)g->xmin,
      (T)g->dx_inv, (T)g->rmin, (T)g->dr_inv, keys, xo, yo, zo);
}

// Push of a PUSH deposit: time step and key output.
template <typename T> struct DepPush {
  T dt;
  int64_t Np;
  uint32_t *keys;  // nullptr: no push
  T *xo, *yo, *zo; // pushed positions (may be the input positions)
};

template <typename T, int M, int KIND>
static void launch_deposit(const uint32_t *sort_indx, const uint32_t *cell_offset, const T *x,
                           const T *y, const T *z, const T *ux, const T *uy, const T *uz,
                           const T *g_inv, const T *w, T charge, const chm_grid_t *g,
                           const PlanePtrs &pp, T *part, bool pass1, hipStream_t st,
                           const DepPush<T> &push) {
  int64_t ncell = (int64_t)(g->Nx - 1) * (g->Nr - 1);
  const unsigned nb = chm_blocks(ncell, DepCfg<T, M, KIND>::CELLS);
  if (pass1 && KIND == 1 && push.keys) {
    k_cell_sums<T, M, KIND, true><<<nb, 64, 0, st>>>(
        sort_indx, cell_offset, x, y, z, ux, uy, uz, g_inv, w, charge, g->Nx, g->Nr, (T)g->xmin,
        (T)g->dx_inv, (T)g->rmin, (T)g->dr_inv, part, push.dt, push.keys, push.xo, push.yo,
        push.zo);
    launch_push_tail<T>(sort_indx, cell_offset, ncell, push.Np, x, y, z, ux, uy, uz, g_inv,
                        push.dt, g, push.keys, push.xo, push.yo, push.zo, st);
  } else if (pass1) {
    k_cell_sums<T, M, KIND, false><<<nb, 64, 0, st>>>(
        sort_indx, cell_offset, x, y, z, ux, uy, uz, g_inv, w, charge, g->Nx, g->Nr, (T)g->xmin,
        (T)g->dx_inv, (T)g->rmin, (T)g->dr_inv, part, (T)0, nullptr, nullptr, nullptr, nullptr);
  }
  int64_t nnode = (int64_t)(g->Nr - 1) * g->Nx;
  k_node_add<T, M, KIND ? 3 : 1><<<chm_blocks(nnode, 256), 256, 0, st>>>(cell_offset, part, g->Nx,
                                                                        g->Nr, pp);
}

template <typename T, int KIND>
static int deposit(const uint32_t *sort_indx, const uint32_t *cell_offset, const T *x, const T *y,
                   const T *z, const T *ux, const T *uy, const T *uz, const T *g_inv, const T *w,
                   int32_t charge, const chm_grid_t *g, const chm_planes_t *flds, void *ws,
                   size_t ws_bytes, void *stream, bool pass1 = true,
                   DepPush<T> push = DepPush<T>{(T)0, 0, nullptr, nullptr, nullptr, nullptr}) {
  if (!g || !flds || g->M < 0 || g->M > CHM_MAXM || g->Nx < 4 || g->Nr < 3) return CHM_EINVAL;
  if (chm_deposit_workspace_bytes(g->Nx, g->Nr, g->M, KIND != 0) > ws_bytes || !ws) return CHM_EWORKSPACE;
  PlanePtrs pp = {};
  for (int k = 0; k < (KIND ? 3 : 1); k++) {
    pp.m0[k] = flds[k].m0;
    for (int m = 0; m < g->M; m++) pp.mc[k][m] = flds[k].mc[m];
  }
  // int8 charge, as the reference passes it (grid_methods_cl.py:64, 89)
  T q = (T)(int8_t)charge;
  hipStream_t st = (hipStream_t)stream;
  T *part = (T *)ws;
  switch (g->M) {
#define CASE(MM)                                                                                   \
  case MM:                                                                                         \
    launch_deposit<T, MM, KIND>(sort_indx, cell_offset, x, y, z, ux, uy, uz, g_inv, w, q, g, pp,   \
                               part, pass1, st, push);                                             \
    break;
    CASE(0) CASE(1) CASE(2) CASE(3) CASE(4)
#undef CASE
  }
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_deposit_scalar_f64(const uint32_t *sort_indx, const uint32_t *cell_offset,
                                      const double *x, const double *y, const double *z,
                                      const double *w, int32_t charge, const chm_grid_t *g,
                                      const chm_planes_t *rho, void *ws, size_t ws_bytes,
                                      void *stream) {
  return deposit<double, 0>(sort_indx, cell_offset, x, y, z, nullptr, nullptr, nullptr, nullptr,
                                w, charge, g, rho, ws, ws_bytes, stream);
}
extern "C" int chm_deposit_scalar_f32(const uint32_t *sort_indx, const uint32_t *cell_offset,
                                      const float *x, const float *y, const float *z,
                                      const float *w, int32_t charge, const chm_grid_t *g,
                                      const chm_planes_t *rho, void *ws, size_t ws_bytes,
                                      void *stream) {
  return deposit<float, 0>(sort_indx, cell_offset, x, y, z, nullptr, nullptr, nullptr, nullptr,
                               w, charge, g, rho, ws, ws_bytes, stream);
}
extern "C" int chm_deposit_vector_f64(const uint32_t *sort_indx, const uint32_t *cell_offset,
                                      const double *x, const double *y, const double *z,
                                      const double *ux, const double *uy, const double *uz,
                                      const double *g_inv, const double *w, int32_t charge,
                                      const chm_grid_t *g, const chm_planes_t *J, void *ws,
                                      size_t ws_bytes, void *stream) {
  return deposit<double, 1>(sort_indx, cell_offset, x, y, z, ux, uy, uz, g_inv, w, charge, g, J,
                               ws, ws_bytes, stream);
}
extern "C" int chm_deposit_vector_f32(const uint32_t *sort_indx, const uint32_t *cell_offset,
                                      const float *x, const float *y, const float *z,
                                      const float *ux, const float *uy, const float *uz,
                                      const float *g_inv, const float *w, int32_t charge,
                                      const chm_grid_t *g, const chm_planes_t *J, void *ws,
                                      size_t ws_bytes, void *stream) {
  return deposit<float, 1>(sort_indx, cell_offset, x, y, z, ux, uy, uz, g_inv, w, charge, g, J,
                              ws, ws_bytes, stream);
}

template <typename T>
static int deposit_vector_push(const uint32_t *sort_indx, const uint32_t *cell_offset,
                               const T *xin, const T *yin, const T *zin, T *x, T *y, T *z,
                               const T *ux, const T *uy, const T *uz, const T *g_inv, const T *w,
                               int32_t charge, T dt, int64_t Np, uint32_t *indx_in_cell,
                               const chm_grid_t *g, const chm_planes_t *J, void *ws,
                               size_t ws_bytes, void *stream) {
  if (!indx_in_cell || !xin || !yin || !zin || !x || !y || !z || Np < 0 || Np > 0x7FFFFFF0ll)
    return CHM_EINVAL;
  if (Np == 0) return CHM_OK;
  return deposit<T, 1>(sort_indx, cell_offset, xin, yin, zin, ux, uy, uz, g_inv, w, charge, g, J,
                       ws, ws_bytes, stream, true, DepPush<T>{dt, Np, indx_in_cell, x, y, z});
}
extern "C" int chm_deposit_vector_push_f64(const uint32_t *sort_indx, const uint32_t *cell_offset,
                                           const double *xin, const double *yin,
                                           const double *zin, double *x, double *y, double *z,
                                           const double *ux, const double *uy, const double *uz,
                                           const double *g_inv, const double *w, int32_t charge,
                                           double dt, int64_t Np, uint32_t *indx_in_cell,
                                           const chm_grid_t *g, const chm_planes_t *J, void *ws,
                                           size_t ws_bytes, void *stream) {
  return deposit_vector_push<double>(sort_indx, cell_offset, xin, yin, zin, x, y, z, ux, uy, uz,
                                     g_inv, w, charge, dt, Np, indx_in_cell, g, J, ws, ws_bytes,
                                     stream);
}
extern "C" int chm_deposit_vector_push_f32(const uint32_t *sort_indx, const uint32_t *cell_offset,
                                           const float *xin, const float *yin, const float *zin,
                                           float *x, float *y, float *z, const float *ux,
                                           const float *uy, const float *uz, const float *g_inv,
                                           const float *w, int32_t charge, float dt, int64_t Np,
                                           uint32_t *indx_in_cell, const chm_grid_t *g,
                                           const chm_planes_t *J, void *ws, size_t ws_bytes,
                                           void *stream) {
  return deposit_vector_push<float>(sort_indx, cell_offset, xin, yin, zin, x, y, z, ux, uy, uz,
                                    g_inv, w, charge, dt, Np, indx_in_cell, g, J, ws, ws_bytes,
                                    stream);
}

// Scatter of the gather's per-particle field adjoints (three components,
// e.g. those of E_p) onto the adjoint planes: planes += sum_p C_ij v_p
// (mode 0) and C_ij v_p 2 e_m (modes m >= 1), the transpose of the gather's
// interpolation.  Same two passes and workspace as the vector deposit.
extern "C" int chm_gather_adj_scatter_f64(const uint32_t *sort_indx, const uint32_t *cell_offset,
                                          const double *x, const double *y, const double *z,
                                          const double *v0, const double *v1, const double *v2,
                                          const chm_grid_t *g, const chm_planes_t *out, void *ws,
                                          size_t ws_bytes, void *stream) {
  return deposit<double, 2>(sort_indx, cell_offset, x, y, z, v0, v1, v2, nullptr, nullptr, 1, g,
                            out, ws, ws_bytes, stream);
}
extern "C" int chm_gather_adj_scatter_f32(const uint32_t *sort_indx, const uint32_t *cell_offset,
                                          const float *x, const float *y, const float *z,
                                          const float *v0, const float *v1, const float *v2,
                                          const chm_grid_t *g, const chm_planes_t *out, void *ws,
                                          size_t ws_bytes, void *stream) {
  return deposit<float, 2>(sort_indx, cell_offset, x, y, z, v0, v1, v2, nullptr, nullptr, 1, g,
                           out, ws, ws_bytes, stream);
}

// Second pass alone: adds the per-cell sums a previous deposit call left in
// ws (for a species whose particles have not changed since).
extern "C" int chm_deposit_add_staged_f64(const uint32_t *cell_offset, const chm_grid_t *g,
                                          const chm_planes_t *flds, int32_t vector, void *ws,
                                          size_t ws_bytes, void *stream) {
  if (vector)
    return deposit<double, 1>(nullptr, cell_offset, nullptr, nullptr, nullptr, nullptr, nullptr,
                                 nullptr, nullptr, nullptr, 0, g, flds, ws, ws_bytes, stream, false);
  return deposit<double, 0>(nullptr, cell_offset, nullptr, nullptr, nullptr, nullptr, nullptr,
                                nullptr, nullptr, nullptr, 0, g, flds, ws, ws_bytes, stream, false);
}
extern "C" int chm_deposit_add_staged_f32(const uint32_t *cell_offset, const chm_grid_t *g,
                                          const chm_planes_t *flds, int32_t vector, void *ws,
                                          size_t ws_bytes, void *stream) {
  if (vector)
    return deposit<float, 1>(nullptr, cell_offset, nullptr, nullptr, nullptr, nullptr, nullptr,
                                nullptr, nullptr, nullptr, 0, g, flds, ws, ws_bytes, stream, false);
  return deposit<float, 0>(nullptr, cell_offset, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nullptr, 0, g, flds, ws, ws_bytes, stream, false);
}

// ------------------------------------------------------------------ post-processing
#define MAX_PLANES 64
struct PlaneList {
  void *p[MAX_PLANES];
  int cplx[MAX_PLANES];
};

// treat_axis (row1 -= row0) then *= dV_inv[ir] (grid_generic.cl:4-60); the
// lane of row 1 also rescales row 0 so that row 0 is read before it changes.
template <typename T>
__global__ void k_postproc(PlaneList pl, const double *__restrict__ dV_inv, int Nx, int Nr) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int pi = blockIdx.y;
  int64_t n = (int64_t)(Nr - 1) * Nx;
  if (e >= n) return;
  int ir = 1 + (int)(e / Nx), ix = (int)(e % Nx);
  T *a = (T *)pl.p[pi];
  int nc = pl.cplx[pi] ? 2 : 1;
  T dvi = (T)dV_inv[ir];
  for (int h = 0; h < nc; h++) {
    int64_t i1 = ((int64_t)ir * Nx + ix) * nc + h;
    if (ir == 1) {
      int64_t i0 = (int64_t)ix * nc + h;
      T a0 = a[i0];
      T v = a[i1] - a0;
      a[i1] = v * dvi;
      a[i0] = a0 * (T)dV_inv[0];
    } else {
      a[i1] = a[i1] * dvi;
    }
  }
}

// warp_axis_m0_d / warp_axis_m1plus_c, grid_generic.cl:63-86
template <typename T> __global__ void k_warp(PlaneList pl, int Nx) {
  int ix = blockIdx.x * blockDim.x + threadIdx.x;
  int pi = blockIdx.y;
  if (ix >= Nx) return;
  T *a = (T *)pl.p[pi];
  if (pl.cplx[pi]) {
    a[2 * ix] = -a[2 * (Nx + ix)];
    a[2 * ix + 1] = -a[2 * (Nx + ix) + 1];
  } else {
    a[ix] = a[Nx + ix];
  }
}

static int fill_plane_list(PlaneList *pl, const chm_planes_t *flds, int nflds, int M) {
  int n = 0;
  for (int f = 0; f < nflds; f++) {
    if (n >= MAX_PLANES) return -1;
    pl->p[n] = flds[f].m0;
    pl->cplx[n++] = 0;
    for (int m = 0; m < M; m++) {
      if (n >= MAX_PLANES) return -1;
      pl->p[n] = flds[f].mc[m];
      pl->cplx[n++] = 1;
    }
  }
  return n;
}

template <typename T>
static int postproc(const chm_planes_t *flds, int32_t nflds, const double *dV_inv, int32_t Nx,
                    int32_t Nr, int32_t M, void *stream) {
  if (!flds || nflds <= 0 || M < 0 || M > CHM_MAXM || Nr < 2) return CHM_EINVAL;
  PlaneList pl;
  int np = fill_plane_list(&pl, flds, nflds, M);
  if (np < 0) return CHM_EINVAL;
  int64_t n = (int64_t)(Nr - 1) * Nx;
  dim3 grid(chm_blocks(n, 256), np);
  k_postproc<T><<<grid, 256, 0, (hipStream_t)stream>>>(pl, dV_inv, Nx, Nr);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

template <typename T>
static int warp(const chm_planes_t *flds, int32_t nflds, int32_t Nx, int32_t Nr, int32_t M,
                void *stream) {
  if (!flds || nflds <= 0 || M < 0 || M > CHM_MAXM || Nr < 2) return CHM_EINVAL;
  PlaneList pl;
  int np = fill_plane_list(&pl, flds, nflds, M);
  if (np < 0) return CHM_EINVAL;
  dim3 grid(chm_blocks(Nx, 256), np);
  k_warp<T><<<grid, 256, 0, (hipStream_t)stream>>>(pl, Nx);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// All planes of nflds fields set to zero in one launch (the set_to(.., 0)
// calls before a deposit, grid.py:23-54).
template <typename T> __global__ void k_zero_planes(PlaneList pl, int64_t n_real) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int pi = blockIdx.y;
  const int64_t n = pl.cplx[pi] ? 2 * n_real : n_real;
  if (e < n) ((T *)pl.p[pi])[e] = (T)0;
}

template <typename T>
static int zero_planes(const chm_planes_t *flds, int32_t nflds, int32_t Nx, int32_t Nr, int32_t M,
                       void *stream) {
  if (!flds || nflds <= 0 || M < 0 || M > CHM_MAXM || Nr < 2 || Nx < 1) return CHM_EINVAL;
  PlaneList pl;
  const int np = fill_plane_list(&pl, flds, nflds, M);
  if (np < 0) return CHM_EINVAL;
  const int64_t n = (int64_t)Nr * Nx;
  dim3 grid(chm_blocks(2 * n, 256), np);
  k_zero_planes<T><<<grid, 256, 0, (hipStream_t)stream>>>(pl, n);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_zero_planes_f64(const chm_planes_t *flds, int32_t nflds, int32_t Nx, int32_t Nr,
                                   int32_t M, void *stream) {
  return zero_planes<double>(flds, nflds, Nx, Nr, M, stream);
}
extern "C" int chm_zero_planes_f32(const chm_planes_t *flds, int32_t nflds, int32_t Nx, int32_t Nr,
                                   int32_t M, void *stream) {
  return zero_planes<float>(flds, nflds, Nx, Nr, M, stream);
}

extern "C" int chm_postproc_deposit_f64(const chm_planes_t *flds, int32_t nflds,
                                        const double *dV_inv, int32_t Nx, int32_t Nr, int32_t M,
                                        void *stream) {
  return postproc<double>(flds, nflds, dV_inv, Nx, Nr, M, stream);
}
extern "C" int chm_postproc_deposit_f32(const chm_planes_t *flds, int32_t nflds,
                                        const double *dV_inv, int32_t Nx, int32_t Nr, int32_t M,
                                        void *stream) {
  return postproc<float>(flds, nflds, dV_inv, Nx, Nr, M, stream);
}
extern "C" int chm_warp_axis_f64(const chm_planes_t *flds, int32_t nflds, int32_t Nx, int32_t Nr,
                                 int32_t M, void *stream) {
  return warp<double>(flds, nflds, Nx, Nr, M, stream);
}
extern "C" int chm_warp_axis_f32(const chm_planes_t *flds, int32_t nflds, int32_t Nx, int32_t Nr,
                                 int32_t M, void *stream) {
  return warp<float>(flds, nflds, Nx, Nr, M, stream);
}

// ------------------------------------------------------------------ gather + push
struct FieldPtrs {
  const void *e0[3], *b0[3];
  const void *ec[3][CHM_MAXM], *bc[3][CHM_MAXM];
};

// Relativistic Boris rotation, grid_deposit_m1.cl:472-507, in the
// reference's operation order (madd keeps it exact in FP64).  Updates u,
// returns the new 1/gamma.
template <typename T>
__device__ __forceinline__ T boris(T (&u)[3], const T (&e_p)[3], const T (&b_p)[3], T dt_2) {
  T um[3], up[3], u0[3], t[3], s[3];
#pragma unroll
  for (int k = 0; k < 3; k++) um[k] = madd(dt_2, e_p[k], u[k]);
  T g = (T)1. / dsqrt<T>(madd(um[2], um[2], madd(um[1], um[1], madd(um[0], um[0], (T)1.))));
#pragma unroll
  for (int k = 0; k < 3; k++) t[k] = dt_2 * b_p[k] * g;
  const T t2 = (T)2. / madd(t[2], t[2], madd(t[1], t[1], madd(t[0], t[0], (T)1.)));
#pragma unroll
  for (int k = 0; k < 3; k++) s[k] = t[k] * t2;
  u0[0] = madd(-um[2], t[1], madd(um[1], t[2], um[0]));
  u0[1] = madd(um[2], t[0], madd(-um[0], t[2], um[1]));
  u0[2] = madd(-um[1], t[0], madd(um[0], t[1], um[2]));
  up[0] = madd(-u0[2], s[1], madd(u0[1], s[2], um[0]));
  up[1] = madd(u0[2], s[0], madd(-u0[0], s[2], um[1]));
  up[2] = madd(-u0[1], s[0], madd(u0[0], s[1], um[2]));
#pragma unroll
  for (int k = 0; k < 3; k++) u[k] = madd(dt_2, e_p[k], up[k]);
  return (T)1. / dsqrt<T>(madd(u[2], u[2], madd(u[1], u[1], madd(u[0], u[0], (T)1.))));
}

// gather_and_push, grid_deposit_m1.cl:330-511 (m0: grid_deposit_m0.cl:280-427).
// One lane per particle index p; the reference's per-slot launch visits the
// same set {p < Np_stay, p in a valid cell} and each particle's update only
// reads its own data, so the results are identical.
template <typename T, int M>
__global__ void __launch_bounds__(256)
    k_gather_push(const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
                  T *__restrict__ px, T *__restrict__ py, T *__restrict__ pz,
                  T *__restrict__ g_inv, int64_t Np, const uint32_t *__restrict__ np_stay,
                  const uint32_t *__restrict__ pid, T factor_push, int Nx, int Nr, T xmin,
                  T dx_inv, T rmin, T dr_inv, FieldPtrs fp) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  if (!((pid ? (int64_t)pid[p] : p) < (int64_t)(*np_stay))) return;
  const int Nx_cell = Nx - 1, Nr_cell = Nr - 1;
  T xp = x[p], yp = y[p], zp = z[p];
  T rp = dsqrt<T>(yp * yp + zp * zp);
  int ix = (int)dfloor<T>((xp - xmin) * dx_inv);
  int ir = (int)dfloor<T>((rp - rmin) * dr_inv);
  if (!(ix > 0 && ix < Nx_cell - 1 && ir >= 0 && ir < Nr_cell - 1)) return;
  int64_t igrid = ix + (int64_t)ir * Nx;
  T u[3] = {px[p], py[p], pz[p]};
  T dt_2 = (T)0.5 * factor_push;
  T rp_inv = (T)1. / rp;
  T sX1 = (xp - xmin) * dx_inv - (T)ix;
  T sX0 = (T)1.0 - sX1;
  T sR1 = (rp - rmin) * dr_inv - (T)ir;
  T sR0 = (T)1.0 - sR1;
  // conj(((y + i z)/r)^m) = ((y - i z)/r)^m, m = 1..M
  T gr[M > 0 ? M + 1 : 2], gi[M > 0 ? M + 1 : 2];
  gr[1] = yp * rp_inv;
  gi[1] = -zp * rp_inv;
#pragma unroll
  for (int m = 2; m <= M; m++) {
    gr[m] = gr[m - 1] * gr[1] - gi[m - 1] * gi[1];
    gi[m] = gr[m - 1] * gi[1] + gi[m - 1] * gr[1];
  }
  T C[2][2];
  C[0][0] = sR0 * sX0;
  C[0][1] = sR0 * sX1;
  C[1][0] = sR1 * sX0;
  C[1][1] = sR1 * sX1;
  T e_p[3] = {0, 0, 0}, b_p[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) {
        int64_t id = igrid + j + (int64_t)Nx * i;
        e_p[k] = madd(C[i][j], ((const T *)fp.e0[k])[id], e_p[k]);
        b_p[k] = madd(C[i][j], ((const T *)fp.b0[k])[id], b_p[k]);
#pragma unroll
        for (int m = 1; m <= M; m++) {
          const T *e = (const T *)fp.ec[k][m - 1];
          const T *b = (const T *)fp.bc[k][m - 1];
          T er = 2 * e[2 * id], ei = 2 * e[2 * id + 1];
          T br = 2 * b[2 * id], bi = 2 * b[2 * id + 1];
          e_p[k] = madd(C[i][j] * er, gr[m], e_p[k]);
          e_p[k] = madd(-(C[i][j] * ei), gi[m], e_p[k]);
          b_p[k] = madd(C[i][j] * br, gr[m], b_p[k]);
          b_p[k] = madd(-(C[i][j] * bi), gi[m], b_p[k]);
        }
      }
  const T g_p_inv = boris<T>(u, e_p, b_p, dt_2);
  px[p] = u[0];
  py[p] = u[1];
  pz[p] = u[2];
  g_inv[p] = g_p_inv;
}

// The same gather + push visiting sorted slots ip < Np_stay with
// p = sort_indx[ip] (the reference kernel's own iteration, grid_deposit_m1.cl:
// 330-350, skipping p >= Np_stay).  A block of GS_NT threads takes
// GS_NT * IT consecutive sorted slots, a short run of consecutive cells; when
// those cells lie in one grid row and span at most
// MAXN - 1 cells, the E and B values of the two node rows under them (NV
// values per node, modes pre-doubled) are staged into LDS once, one node per
// lane with all of its loads in flight together, and every lane reads its 72
// (M = 1) field values from there instead of from 18 separate planes.
// Identical operands and operation order: the results equal k_gather_push
// bit for bit.
#define GS_NT 256
template <typename T, int M> struct GatherCfg {
  // two sorted slots per thread (their loads in flight together); registers
  // held to a number of waves per SIMD measured on the 2048 x 256 step (M = 1:
  // 5 FP64, 4 at M = 0, 6 FP32, no spills) and the 4096 x 512 step (M = 2:
  // 4 FP64, a few spilled words, 4% faster than the unbounded 3 waves; 6 FP32
  // for the pre-pushing gather, 13% faster than 4-5); no bound above M = 2
  static constexpr int IT = 2;
  static constexpr int wpe(bool pre) {
    return M <= 1 ? (sizeof(T) == 8 ? (M == 1 ? 5 : 4) : 6)
                  : (M == 2 ? (sizeof(T) == 8 ? 4 : (pre ? 6 : 1)) : 1);
  }
  static constexpr int NQ = 1 + 2 * M;
  static constexpr int NV = 6 * NQ;  // E and B, 3 components, all modes
  static constexpr int MN = 20480 / (2 * NV * (int)sizeof(T));
  static constexpr int MAXN = MN > GS_NT / 2 ? GS_NT / 2 : (MN < 4 ? 4 : MN);
};

template <typename T, int M>
__device__ __forceinline__ T field_val(const FieldPtrs &fp, int v, int64_t node) {
  typedef GatherCfg<T, M> Cfg;
  const int f = v / (3 * Cfg::NQ), r = v - f * 3 * Cfg::NQ;
  const int k = r / Cfg::NQ, q = r - k * Cfg::NQ;
  if (q == 0) return ((const T *)(f ? fp.b0[k] : fp.e0[k]))[node];
  const int m = (q + 1) >> 1;
  return ((const T *)(f ? fp.bc[k][m - 1] : fp.ec[k][m - 1]))[2 * node + ((q - 1) & 1)];
}

// E_p, B_p of one particle from its 4 nodes (value v of node (i, j) from
// ld(v, i, j), values ordered as in GatherCfg), in k_gather_push's order.
// DBL: the loader already returns the mode values doubled (exact).
template <typename T, int M, bool DBL, typename LD>
__device__ __forceinline__ void gather_sum(LD ld, const T (&C)[2][2], const T *gr, const T *gi,
                                           T *e_p, T *b_p) {
  const T two = DBL ? (T)1 : (T)2;
  constexpr int NQ = 1 + 2 * M;
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) {
        e_p[k] = madd(C[i][j], ld(k * NQ, i, j), e_p[k]);
        b_p[k] = madd(C[i][j], ld(3 * NQ + k * NQ, i, j), b_p[k]);
#pragma unroll
        for (int m = 1; m <= M; m++) {
          T er = ld(k * NQ + 2 * m - 1, i, j), ei = ld(k * NQ + 2 * m, i, j);
          T br = ld(3 * NQ + k * NQ + 2 * m - 1, i, j), bi = ld(3 * NQ + k * NQ + 2 * m, i, j);
          if (!DBL) {
            er = two * er;
            ei = two * ei;
            br = two * br;
            bi = two * bi;
          }
          e_p[k] = madd(C[i][j] * er, gr[m], e_p[k]);
          e_p[k] = madd(-(C[i][j] * ei), gi[m], e_p[k]);
          b_p[k] = madd(C[i][j] * br, gr[m], b_p[k]);
          b_p[k] = madd(-(C[i][j] * bi), gi[m], b_p[k]);
        }
      }
}

// All NV values of one node into dst (GatherCfg order, modes doubled): the
// loads are issued together, then written to LDS.
template <typename T, int M>
__device__ __forceinline__ void stage_node(const FieldPtrs &fp, int64_t node, T *dst) {
  constexpr int NQ = GatherCfg<T, M>::NQ, NV = GatherCfg<T, M>::NV;
  T val[NV];
#pragma unroll
  for (int f = 0; f < 2; f++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int vb = (3 * f + k) * NQ;
      val[vb] = ((const T *)(f ? fp.b0[k] : fp.e0[k]))[node];
#pragma unroll
      for (int m = 1; m <= M; m++) {
        const T *pc = (const T *)(f ? fp.bc[k][m - 1] : fp.ec[k][m - 1]);
        val[vb + 2 * m - 1] = pc[2 * node];
        val[vb + 2 * m] = pc[2 * node + 1];
      }
    }
#pragma unroll
  for (int v = 0; v < NV; v++) dst[v] = (v % NQ) ? (T)2 * val[v] : val[v];
}

// The next step's first coordinate push, applied by a PRE gather to each
// particle it visits (push_x's operations, from the momenta and 1/gamma the
// particle leaves the step with), into separate position arrays, with the
// cell key of the pushed position.
template <typename T> struct PrePush {
  T dt;
  T *x, *y, *z;
  uint32_t *keys;
};

template <typename T>
__device__ __forceinline__ void pre_push(const PrePush<T> &pp, uint32_t p, T xp, T yp, T zp,
                                         const T (&u)[3], T g, int Nx, int Nr, T xmin, T dx_inv,
                                         T rmin, T dr_inv) {
  const T dt_g = pp.dt * g;
  const T xn = xp + u[0] * dt_g, yn = yp + u[1] * dt_g, zn = zp + u[2] * dt_g;
  pp.x[p] = xn;
  pp.y[p] = yn;
  pp.z[p] = zn;
  pp.keys[p] = cell_key<T>(xn, yn, zn, Nx - 1, Nr - 1, xmin, dx_inv, rmin, dr_inv);
}

template <typename T, int M, bool PRE, bool SK>
__global__ void __launch_bounds__(GS_NT) __attribute__((amdgpu_waves_per_eu(GatherCfg<T, M>::wpe(PRE))))
    k_gather_sorted(const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
                    T *__restrict__ px, T *__restrict__ py, T *__restrict__ pz,
                    T *__restrict__ g_inv, const uint32_t *__restrict__ sort_indx,
                    const uint32_t *__restrict__ np_stay_p, const uint32_t *__restrict__ pid,
                    T factor_push, int Nx, int Nr, T xmin, T dx_inv, T rmin, T dr_inv,
                    FieldPtrs fp, PrePush<T> pp, const uint32_t *__restrict__ skey) {
  typedef GatherCfg<T, M> Cfg;
  constexpr int NV = Cfg::NV, MAXN = Cfg::MAXN, IT = Cfg::IT;
  __shared__ T fs[2 * MAXN * NV];
  __shared__ int cab[4];
  const int64_t np_stay = (int64_t)(*np_stay_p);
  const int64_t s0 = (int64_t)blockIdx.x * (GS_NT * IT);
  if (s0 >= np_stay) return;  // uniform over the block
  const int nlive = (int)((np_stay - s0 < GS_NT * IT) ? np_stay - s0 : GS_NT * IT);
  // IT slots per thread (slot s0 + q*GS_NT + tid): all their loads in flight
  // together
  uint32_t p[IT];
  T xp[IT], yp[IT], zp[IT], rp[IT], u[IT][3];
  int ix[IT], ir[IT];
#pragma unroll
  for (int q = 0; q < IT; q++) {
    const int64_t slot = s0 + q * GS_NT + threadIdx.x;
    p[q] = slot < np_stay ? sort_indx[slot] : 0u;
  }
  // The block's cells: with the sorted keys (SK), from the keys of its
  // first and last slot (key = ix + ir (Nx - 1) for the valid cells below
  // Np_stay), so that the field staging loads go out together with the
  // particle loads; otherwise from the first and last particle's position.
  // The two node rows under the cells are staged when the cells lie in one
  // row and span at most MAXN - 1 of them.
  int ixA = 0, irA = 0, ncol = 0;
  bool staged = false;
  if constexpr (SK) {
    if (threadIdx.x == 0) {
      const uint32_t c0 = skey[s0], c1 = skey[s0 + nlive - 1];
      const uint32_t nxl = (uint32_t)(Nx - 1);
      cab[0] = (int)(c0 % nxl);
      cab[1] = (int)(c0 / nxl);
      cab[2] = (int)(c1 % nxl);
      cab[3] = (int)(c1 / nxl);
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < IT; q++) {
    xp[q] = yp[q] = zp[q] = rp[q] = 0;
    u[q][0] = u[q][1] = u[q][2] = 0;
    ix[q] = ir[q] = 0;
    if (s0 + q * GS_NT + threadIdx.x < np_stay) {
      xp[q] = x[p[q]];
      yp[q] = y[p[q]];
      zp[q] = z[p[q]];
      u[q][0] = px[p[q]];
      u[q][1] = py[p[q]];
      u[q][2] = pz[p[q]];
    }
  }
  if constexpr (SK) {
    // (the range is also checked against the grid: keys that no longer
    // match the positions cannot make the staging read outside it)
    ixA = cab[0];
    irA = cab[1];
    ncol = cab[2] - ixA + 2;
    staged = cab[3] == irA && ncol >= 2 && ncol <= MAXN && irA >= 0 && irA + 1 < Nr &&
             ixA >= 0 && ixA + ncol <= Nx;
    if (staged && (int)threadIdx.x < 2 * ncol) {
      const int i = (int)threadIdx.x >= ncol ? 1 : 0;
      const int64_t node = (int64_t)(ixA + (int)threadIdx.x - i * ncol) + (int64_t)(irA + i) * Nx;
      stage_node<T, M>(fp, node, fs + threadIdx.x * NV);
    }
  }
#pragma unroll
  for (int q = 0; q < IT; q++) {
    if (s0 + q * GS_NT + threadIdx.x < np_stay) {
      rp[q] = dsqrt<T>(yp[q] * yp[q] + zp[q] * zp[q]);
      ix[q] = (int)dfloor<T>((xp[q] - xmin) * dx_inv);
      ir[q] = (int)dfloor<T>((rp[q] - rmin) * dr_inv);
    }
  }
  if constexpr (!SK) {
    // slots below Np_stay hold particles of valid cells, in cell order: the
    // first and the last live slot bound the block's cells
    if (threadIdx.x == 0) {
      cab[0] = ix[0];
      cab[1] = ir[0];
    }
#pragma unroll
    for (int q = 0; q < IT; q++)
      if (q * GS_NT + (int)threadIdx.x == nlive - 1) {
        cab[2] = ix[q];
        cab[3] = ir[q];
      }
    __syncthreads();
    ixA = cab[0];
    irA = cab[1];
    ncol = cab[2] - ixA + 2;  // node columns under the block's cells
    staged = cab[3] == irA && ncol <= MAXN;
    if (staged && (int)threadIdx.x < 2 * ncol) {
      const int i = (int)threadIdx.x >= ncol ? 1 : 0;
      const int64_t node = (int64_t)(ixA + (int)threadIdx.x - i * ncol) + (int64_t)(irA + i) * Nx;
      stage_node<T, M>(fp, node, fs + threadIdx.x * NV);
    }
  }
  __syncthreads();
  const T dt_2 = (T)0.5 * factor_push;
#pragma unroll
  for (int q = 0; q < IT; q++) {
    if (!(s0 + q * GS_NT + threadIdx.x < np_stay)) continue;
    T(&uq)[3] = u[q];
    if (!((int64_t)(pid ? pid[p[q]] : p[q]) < np_stay)) {
      // skipped by the reference's Np_stay test: momenta unchanged
      if (PRE)
        pre_push<T>(pp, p[q], xp[q], yp[q], zp[q], uq, g_inv[p[q]], Nx, Nr, xmin, dx_inv, rmin,
                    dr_inv);
      continue;
    }
    const int64_t igrid = ix[q] + (int64_t)ir[q] * Nx;
    const int lc = ix[q] - ixA;
    T rp_inv = (T)1. / rp[q];
    T sX1 = (xp[q] - xmin) * dx_inv - (T)ix[q];
    T sX0 = (T)1.0 - sX1;
    T sR1 = (rp[q] - rmin) * dr_inv - (T)ir[q];
    T sR0 = (T)1.0 - sR1;
    T gr[M > 0 ? M + 1 : 2], gi[M > 0 ? M + 1 : 2];
    gr[1] = yp[q] * rp_inv;
    gi[1] = -zp[q] * rp_inv;
#pragma unroll
    for (int m = 2; m <= M; m++) {
      gr[m] = gr[m - 1] * gr[1] - gi[m - 1] * gi[1];
      gi[m] = gr[m - 1] * gi[1] + gi[m - 1] * gr[1];
    }
    T C[2][2];
    C[0][0] = sR0 * sX0;
    C[0][1] = sR0 * sX1;
    C[1][0] = sR1 * sX0;
    C[1][1] = sR1 * sX1;
    T e_p[3] = {0, 0, 0}, b_p[3] = {0, 0, 0};
    // (with the keys, a particle outside the staged cells, only possible if
    // the positions moved after the sort, reads the planes directly)
    bool from_lds = staged;
    if constexpr (SK) from_lds = staged && ir[q] == irA && lc >= 0 && lc <= ncol - 2;
    if (from_lds) {
      const T *f0 = fs + lc * NV;
      gather_sum<T, M, true>([&](int v, int i, int j) { return f0[(i * ncol + j) * NV + v]; }, C,
                             gr, gi, e_p, b_p);
    } else {
      gather_sum<T, M, false>([&](int v, int i, int j) {
        return field_val<T, M>(fp, v, igrid + j + (int64_t)Nx * i);
      }, C, gr, gi, e_p, b_p);
    }
    const T g_p_inv = boris<T>(uq, e_p, b_p, dt_2);
    px[p[q]] = uq[0];
    py[p[q]] = uq[1];
    pz[p[q]] = uq[2];
    g_inv[p[q]] = g_p_inv;
    if (PRE)
      pre_push<T>(pp, p[q], xp[q], yp[q], zp[q], uq, g_p_inv, Nx, Nr, xmin, dx_inv, rmin, dr_inv);
  }
}

template <typename T>
static int gather_push(const T *x, const T *y, const T *z, T *px, T *py, T *pz, T *g_inv,
                       int64_t Np, const uint32_t *sort_indx, const uint32_t *np_stay,
                       const uint32_t *pid, T factor_push, const chm_grid_t *g,
                       const chm_planes_t *E, const chm_planes_t *B, void *stream,
                       const uint32_t *skey, const uint32_t *cell_offset = nullptr,
                       PrePush<T> pp = PrePush<T>{(T)0, nullptr, nullptr, nullptr, nullptr}) {
  const bool pre = pp.keys != nullptr;
  if (pre && (!sort_indx || !cell_offset || !pp.x || !pp.y || !pp.z)) return CHM_EINVAL;
  if (!g || !E || !B || g->M < 0 || g->M > CHM_MAXM || Np < 0) return CHM_EINVAL;
  if (Np == 0) return CHM_OK;
  FieldPtrs fp = {};
  for (int k = 0; k < 3; k++) {
    fp.e0[k] = E[k].m0;
    fp.b0[k] = B[k].m0;
    for (int m = 0; m < g->M; m++) {
      fp.ec[k][m] = E[k].mc[m];
      fp.bc[k][m] = B[k].mc[m];
    }
  }
  hipStream_t st = (hipStream_t)stream;
  unsigned nb = chm_blocks(Np, 256);
  // The sorted keys are used in FP64 only: the FP32 kernel with the key path
  // needs more registers than its bound allows (spills) and measured slower
  // at 4096 x 512, M = 2 (1.30 against 1.13 ms per call; FP64 2.08 against
  // 2.17)
  const bool sk = skey != nullptr && sizeof(T) == 8;
#define GATHER_SORTED(MM, PREV, SKV)                                                               \
  k_gather_sorted<T, MM, PREV, SKV><<<chm_blocks(Np, GS_NT * GatherCfg<T, MM>::IT), GS_NT, 0, st>>>( \
      x, y, z, px, py, pz, g_inv, sort_indx, np_stay, pid, factor_push, g->Nx, g->Nr, (T)g->xmin,  \
      (T)g->dx_inv, (T)g->rmin, (T)g->dr_inv, fp, pp, skey)
  switch (g->M) {
#define CASE(MM)                                                                                   \
  case MM:                                                                                         \
    if (pre && sk)                                                                                 \
      GATHER_SORTED(MM, true, true);                                                               \
    else if (pre)                                                                                  \
      GATHER_SORTED(MM, true, false);                                                              \
    else if (sort_indx && sk)                                                                      \
      GATHER_SORTED(MM, false, true);                                                              \
    else if (sort_indx)                                                                            \
      GATHER_SORTED(MM, false, false);                                                             \
    else                                                                                           \
      k_gather_push<T, MM><<<nb, 256, 0, st>>>(x, y, z, px, py, pz, g_inv, Np, np_stay, pid,       \
                                               factor_push, g->Nx, g->Nr, (T)g->xmin,              \
                                               (T)g->dx_inv, (T)g->rmin, (T)g->dr_inv, fp);        \
    break;
    CASE(0) CASE(1) CASE(2) CASE(3) CASE(4)
#undef CASE
#undef GATHER_SORTED
  }
  if (pre) {
    // the particles of the out-of-domain bin, which the gather does not visit
    const int64_t ncell = (int64_t)(g->Nx - 1) * (g->Nr - 1);
    launch_push_tail<T>(sort_indx, cell_offset, ncell, Np, x, y, z, px, py, pz, g_inv, pp.dt, g,
                        pp.keys, pp.x, pp.y, pp.z, st);
  }
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_gather_push_f64(const double *x, const double *y, const double *z, double *px,
                                   double *py, double *pz, double *g_inv, int64_t Np,
                                   const uint32_t *sort_indx, const uint32_t *np_stay,
                                   const uint32_t *sorted_keys, double factor_push,
                                   const chm_grid_t *g, const chm_planes_t *E,
                                   const chm_planes_t *B, void *stream) {
  return gather_push<double>(x, y, z, px, py, pz, g_inv, Np, sort_indx, np_stay, nullptr,
                             factor_push, g, E, B, stream, sorted_keys);
}
extern "C" int chm_gather_push_f32(const float *x, const float *y, const float *z, float *px,
                                   float *py, float *pz, float *g_inv, int64_t Np,
                                   const uint32_t *sort_indx, const uint32_t *np_stay,
                                   const uint32_t *sorted_keys, float factor_push,
                                   const chm_grid_t *g, const chm_planes_t *E,
                                   const chm_planes_t *B, void *stream) {
  return gather_push<float>(x, y, z, px, py, pz, g_inv, Np, sort_indx, np_stay, nullptr,
                            factor_push, g, E, B, stream, sorted_keys);
}
extern "C" int chm_gather_push_ids_f64(const double *x, const double *y, const double *z,
                                       double *px, double *py, double *pz, double *g_inv,
                                       int64_t Np, const uint32_t *sort_indx,
                                       const uint32_t *np_stay, const uint32_t *pid,
                                       const uint32_t *sorted_keys, double factor_push,
                                       const chm_grid_t *g, const chm_planes_t *E,
                                       const chm_planes_t *B, void *stream) {
  return gather_push<double>(x, y, z, px, py, pz, g_inv, Np, sort_indx, np_stay, pid, factor_push,
                             g, E, B, stream, sorted_keys);
}
extern "C" int chm_gather_push_ids_f32(const float *x, const float *y, const float *z, float *px,
                                       float *py, float *pz, float *g_inv, int64_t Np,
                                       const uint32_t *sort_indx, const uint32_t *np_stay,
                                       const uint32_t *pid, const uint32_t *sorted_keys,
                                       float factor_push, const chm_grid_t *g,
                                       const chm_planes_t *E, const chm_planes_t *B,
                                       void *stream) {
  return gather_push<float>(x, y, z, px, py, pz, g_inv, Np, sort_indx, np_stay, pid, factor_push,
                            g, E, B, stream, sorted_keys);
}
extern "C" int chm_gather_push_pre_f64(const double *x, const double *y, const double *z,
                                       double *px, double *py, double *pz, double *g_inv,
                                       int64_t Np, const uint32_t *sort_indx,
                                       const uint32_t *cell_offset, const uint32_t *np_stay,
                                       const uint32_t *pid, const uint32_t *sorted_keys,
                                       double factor_push, double dt, double *xh, double *yh,
                                       double *zh, uint32_t *indx_in_cell, const chm_grid_t *g,
                                       const chm_planes_t *E, const chm_planes_t *B,
                                       void *stream) {
  if (!indx_in_cell) return CHM_EINVAL;
  return gather_push<double>(x, y, z, px, py, pz, g_inv, Np, sort_indx, np_stay, pid, factor_push,
                             g, E, B, stream, sorted_keys, cell_offset,
                             PrePush<double>{dt, xh, yh, zh, indx_in_cell});
}
extern "C" int chm_gather_push_pre_f32(const float *x, const float *y, const float *z, float *px,
                                       float *py, float *pz, float *g_inv, int64_t Np,
                                       const uint32_t *sort_indx, const uint32_t *cell_offset,
                                       const uint32_t *np_stay, const uint32_t *pid,
                                       const uint32_t *sorted_keys, float factor_push, float dt,
                                       float *xh, float *yh, float *zh, uint32_t *indx_in_cell,
                                       const chm_grid_t *g, const chm_planes_t *E,
                                       const chm_planes_t *B, void *stream) {
  if (!indx_in_cell) return CHM_EINVAL;
  return gather_push<float>(x, y, z, px, py, pz, g_inv, Np, sort_indx, np_stay, pid, factor_push,
                            g, E, B, stream, sorted_keys, cell_offset,
                            PrePush<float>{dt, xh, yh, zh, indx_in_cell});
}
