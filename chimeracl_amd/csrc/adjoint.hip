// Reverse-mode (adjoint) kernels of the particle side of the PIC step, for
// the differentiable form: gradients of a scalar objective through the
// half-step pushes, the J and rho deposits, the field gather + Boris push,
// the deposit post-processing, the ghost-row warp and the density profile.
//
// Forward maps differentiated (per particle p, CIC weights in the cell of p):
//   push_x        x' = x + u * g_inv * dt                   (particles_generic.cl:129-153)
//   deposit rho   rho_m[node] += w q C_ij e_m               (grid_deposit_m1.cl:20-152)
//   deposit J     J_k,m[node] += w g_inv q u_k C_ij e_m     (grid_deposit_m1.cl:171-327)
//   gather        F_p = sum C_ij (F_0 + 2 Re(F_m conj(e_m)))
//   Boris         u' = B(u, E_p, B_p)                       (grid_deposit_m1.cl:472-507)
// with C_ij = sR_i sX_j and e_m = ((y + i z)/r)^m, e_0 = 1.
//
// Adjoint conventions: a real variable v has adjoint dL/dv; a complex
// variable z = a + i b has adjoint dL/da + i dL/db.  Adjoints of particle
// quantities are SoA arrays indexed like the particles; every kernel adds
// into them (one lane per particle, no atomics), except where a forward
// value is overwritten (the gather's momenta), whose adjoint is replaced.
// The field adjoints of the gather are scattered back to the grid by the
// deposit machinery (chm_gather_adj_scatter_*, deposit.hip), so no float
// atomics are used anywhere.
#include "chm_common.h"

template <typename T> __device__ __forceinline__ T a_sqrt(T v);
template <> __device__ __forceinline__ double a_sqrt<double>(double v) { return sqrt(v); }
template <> __device__ __forceinline__ float a_sqrt<float>(float v) { return sqrtf(v); }
template <typename T> __device__ __forceinline__ T a_floor(T v);
template <> __device__ __forceinline__ double a_floor<double>(double v) { return floor(v); }
template <> __device__ __forceinline__ float a_floor<float>(float v) { return floorf(v); }

struct VecPlanes {
  const void *m0[3];
  const void *mc[3][CHM_MAXM];
};

// ---------------------------------------------------------------- helpers
// CIC interpolation of the node values f_ij = P0 + fac * sum_m Re(P_m conj(e_m))
// around a particle, with its x and r derivatives and the weighted complex
// node sums S_m = sum_ij C_ij P_m (the derivative with respect to e_m).
template <typename T, int M>
__device__ __forceinline__ void interp(const T *p0, const void *const *pc, int64_t igrid, int Nx,
                                       T sX0, T sX1, T sR0, T sR1, const T *er, const T *ei,
                                       T fac, T dx_inv, T dr_inv, T &V, T &dVx, T &dVr, T *Sr,
                                       T *Si) {
  T f[2][2];
  const T sR[2] = {sR0, sR1}, sX[2] = {sX0, sX1};
#pragma unroll
  for (int m = 0; m < (M > 0 ? M : 1); m++) Sr[m] = Si[m] = 0;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int64_t id = igrid + j + (int64_t)Nx * i;
      T v = p0[id];
      const T C = sR[i] * sX[j];
#pragma unroll
      for (int m = 1; m <= M; m++) {
        const T *pm = (const T *)pc[m - 1];
        T br = pm[2 * id], bi = pm[2 * id + 1];
        v += fac * (br * er[m] + bi * ei[m]);
        Sr[m - 1] += C * br;
        Si[m - 1] += C * bi;
      }
      f[i][j] = v;
    }
  V = sR0 * (sX0 * f[0][0] + sX1 * f[0][1]) + sR1 * (sX0 * f[1][0] + sX1 * f[1][1]);
  dVx = dx_inv * (sR0 * (f[0][1] - f[0][0]) + sR1 * (f[1][1] - f[1][0]));
  dVr = dr_inv * (sX0 * (f[1][0] - f[0][0]) + sX1 * (f[1][1] - f[0][1]));
}

// e_m = ((y + i z)/r)^m, m = 1..M
template <typename T, int M>
__device__ __forceinline__ void mode_powers(T yp, T zp, T rp_inv, T *er, T *ei) {
  er[0] = 1;
  ei[0] = 0;
  er[1] = yp * rp_inv;
  ei[1] = zp * rp_inv;
#pragma unroll
  for (int m = 2; m <= M; m++) {
    er[m] = er[m - 1] * er[1] - ei[m - 1] * ei[1];
    ei[m] = er[m - 1] * ei[1] + ei[m - 1] * er[1];
  }
}

// Adds the y, z adjoints of r = |(y, z)| (adjoint rb) and of e_m (complex
// adjoints ebr + i ebi, m = 1..M) for a particle at r > 0.
template <typename T, int M>
__device__ __forceinline__ void yz_adjoint(T yp, T zp, T rp_inv, const T *er, const T *ei,
                                           const T *ebr, const T *ebi, T rb, T &yb, T &zb) {
  // e_m = e_1^m: adjoint of e_1 is sum_m m conj(e_{m-1}) ebar_m
  T a = 0, b = 0;
#pragma unroll
  for (int m = 1; m <= M; m++) {
    a += (T)m * (er[m - 1] * ebr[m - 1] + ei[m - 1] * ebi[m - 1]);
    b += (T)m * (er[m - 1] * ebi[m - 1] - ei[m - 1] * ebr[m - 1]);
  }
  // e_1 = (y/r, z/r): d(y/r)/dy = z^2/r^3, d(y/r)/dz = d(z/r)/dy = -yz/r^3,
  // d(z/r)/dz = y^2/r^3
  const T ri3 = rp_inv * rp_inv * rp_inv;
  const T q = (a * zp - b * yp) * ri3;
  yb += q * zp + rb * yp * rp_inv;
  zb += -q * yp + rb * zp * rp_inv;
}

// ---------------------------------------------------------------- push_x
template <typename T>
__global__ void __launch_bounds__(256)
    k_push_x_adj(const T *__restrict__ xb, const T *__restrict__ yb, const T *__restrict__ zb,
                 const T *__restrict__ ux, const T *__restrict__ uy, const T *__restrict__ uz,
                 const T *__restrict__ g_inv, T *__restrict__ uxb, T *__restrict__ uyb,
                 T *__restrict__ uzb, T *__restrict__ gb, T dt, int64_t Np) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  const T a = xb[p], b = yb[p], c = zb[p], g = g_inv[p];
  const T dtg = dt * g;
  uxb[p] += a * dtg;
  uyb[p] += b * dtg;
  uzb[p] += c * dtg;
  gb[p] += dt * (a * ux[p] + b * uy[p] + c * uz[p]);
}

template <typename T>
static int push_x_adj(const T *xb, const T *yb, const T *zb, const T *ux, const T *uy, const T *uz,
                      const T *g_inv, T *uxb, T *uyb, T *uzb, T *gb, T dt, int64_t Np,
                      void *stream) {
  if (Np < 0) return CHM_EINVAL;
  if (Np == 0) return CHM_OK;
  k_push_x_adj<T><<<chm_blocks(Np, 256), 256, 0, (hipStream_t)stream>>>(
      xb, yb, zb, ux, uy, uz, g_inv, uxb, uyb, uzb, gb, dt, Np);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_push_x_adj_f64(const double *xb, const double *yb, const double *zb,
                                  const double *ux, const double *uy, const double *uz,
                                  const double *g_inv, double *uxb, double *uyb, double *uzb,
                                  double *gb, double dt, int64_t Np, void *stream) {
  return push_x_adj<double>(xb, yb, zb, ux, uy, uz, g_inv, uxb, uyb, uzb, gb, dt, Np, stream);
}
extern "C" int chm_push_x_adj_f32(const float *xb, const float *yb, const float *zb,
                                  const float *ux, const float *uy, const float *uz,
                                  const float *g_inv, float *uxb, float *uyb, float *uzb,
                                  float *gb, float dt, int64_t Np, void *stream) {
  return push_x_adj<float>(xb, yb, zb, ux, uy, uz, g_inv, uxb, uyb, uzb, gb, dt, Np, stream);
}

// ---------------------------------------------------------------- deposits
// Adjoint of depose_scalar (VEC = false: rho += w q C e_m) and
// depose_vector (VEC = true: J_k += w g_inv q u_k C e_m) for one particle,
// given the adjoint planes Pb of the deposited field (after the adjoint of
// the axis post-processing).  Every particle in a valid cell was deposited.
struct DepAdjOut {
  void *xb, *yb, *zb, *wb;
  void *ub[3], *gb;
};

template <typename T, int M, bool VEC>
__global__ void __launch_bounds__(256)
    k_deposit_adj(const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
                  const T *__restrict__ ux, const T *__restrict__ uy, const T *__restrict__ uz,
                  const T *__restrict__ g_inv, const T *__restrict__ w, T charge, int64_t Np,
                  int Nx, int Nr, T xmin, T dx_inv, T rmin, T dr_inv, VecPlanes Pb, DepAdjOut o) {
  constexpr int NK = VEC ? 3 : 1;
  constexpr int MM = M > 0 ? M : 1;
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  const int Nx_cell = Nx - 1, Nr_cell = Nr - 1;
  const T xp = x[p], yp = y[p], zp = z[p];
  const T rp = a_sqrt<T>(yp * yp + zp * zp);
  const int ix = (int)a_floor<T>((xp - xmin) * dx_inv);
  const int ir = (int)a_floor<T>((rp - rmin) * dr_inv);
  if (!(ix > 0 && ix < Nx_cell - 1 && ir >= 0 && ir < Nr_cell - 1)) return;
  const int64_t igrid = ix + (int64_t)ir * Nx;
  T rp_inv;
  if (VEC) rp_inv = rp > 0 ? (T)1 / rp : (T)0;
  else rp_inv = (T)1 / rp;
  T er[MM + 1], ei[MM + 1];
  mode_powers<T, M>(yp, zp, rp_inv, er, ei);
  const T sX1 = (xp - xmin) * dx_inv - (T)ix, sX0 = (T)1 - sX1;
  const T sR1 = (rp - rmin) * dr_inv - (T)ir, sR0 = (T)1 - sR1;
  const T wv = w[p];
  T u[3] = {1, 0, 0}, g = 1;
  if (VEC) {
    u[0] = ux[p];
    u[1] = uy[p];
    u[2] = uz[p];
    g = g_inv[p];
  }
  const T wp = wv * g * charge;
  T Q = 0, QX = 0, QR = 0, ebr[MM], ebi[MM];
#pragma unroll
  for (int m = 0; m < MM; m++) ebr[m] = ebi[m] = 0;
#pragma unroll
  for (int k = 0; k < NK; k++) {
    T V, dVx, dVr, Sr[MM], Si[MM];
    interp<T, M>((const T *)Pb.m0[k], Pb.mc[k], igrid, Nx, sX0, sX1, sR0, sR1, er, ei, (T)1,
                 dx_inv, dr_inv, V, dVx, dVr, Sr, Si);
    if (VEC) ((T *)o.ub[k])[p] += wp * V;
    Q += u[k] * V;
    QX += u[k] * dVx;
    QR += u[k] * dVr;
#pragma unroll
    for (int m = 0; m < M; m++) {
      ebr[m] += u[k] * Sr[m];
      ebi[m] += u[k] * Si[m];
    }
  }
  ((T *)o.wb)[p] += g * charge * Q;
  if (VEC) ((T *)o.gb)[p] += wv * charge * Q;
  ((T *)o.xb)[p] += wp * QX;
  if (rp > 0) {
#pragma unroll
    for (int m = 0; m < M; m++) {
      ebr[m] *= wp;
      ebi[m] *= wp;
    }
    T yb = 0, zb = 0;
    yz_adjoint<T, M>(yp, zp, rp_inv, er, ei, ebr, ebi, wp * QR, yb, zb);
    ((T *)o.yb)[p] += yb;
    ((T *)o.zb)[p] += zb;
  }
}

template <typename T, bool VEC>
static int deposit_adj(const T *x, const T *y, const T *z, const T *ux, const T *uy, const T *uz,
                       const T *g_inv, const T *w, int32_t charge, int64_t Np, const chm_grid_t *g,
                       const chm_planes_t *Pb, T *xb, T *yb, T *zb, T *uxb, T *uyb, T *uzb,
                       T *gb, T *wb, void *stream) {
  if (!g || !Pb || g->M < 0 || g->M > CHM_MAXM || Np < 0 || g->Nx < 4 || g->Nr < 3) return CHM_EINVAL;
  if (Np == 0) return CHM_OK;
  VecPlanes vp = {};
  for (int k = 0; k < (VEC ? 3 : 1); k++) {
    vp.m0[k] = Pb[k].m0;
    for (int m = 0; m < g->M; m++) vp.mc[k][m] = Pb[k].mc[m];
  }
  DepAdjOut o = {};
  o.xb = xb;
  o.yb = yb;
  o.zb = zb;
  o.wb = wb;
  o.ub[0] = uxb;
  o.ub[1] = uyb;
  o.ub[2] = uzb;
  o.gb = gb;
  T q = (T)(int8_t)charge;
  hipStream_t st = (hipStream_t)stream;
  unsigned nb = chm_blocks(Np, 256);
  switch (g->M) {
#define CASE(MM)                                                                                   \
  case MM:                                                                                         \
    k_deposit_adj<T, MM, VEC><<<nb, 256, 0, st>>>(x, y, z, ux, uy, uz, g_inv, w, q, Np, g->Nx,     \
                                                  g->Nr, (T)g->xmin, (T)g->dx_inv, (T)g->rmin,     \
                                                  (T)g->dr_inv, vp, o);                            \
    break;
    CASE(0) CASE(1) CASE(2) CASE(3) CASE(4)
#undef CASE
  }
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_deposit_scalar_adj_f64(const double *x, const double *y, const double *z,
                                          const double *w, int32_t charge, int64_t Np,
                                          const chm_grid_t *g, const chm_planes_t *rho_b,
                                          double *xb, double *yb, double *zb, double *wb,
                                          void *stream) {
  return deposit_adj<double, false>(x, y, z, nullptr, nullptr, nullptr, nullptr, w, charge, Np, g,
                                    rho_b, xb, yb, zb, nullptr, nullptr, nullptr, nullptr, wb,
                                    stream);
}
extern "C" int chm_deposit_scalar_adj_f32(const float *x, const float *y, const float *z,
                                          const float *w, int32_t charge, int64_t Np,
                                          const chm_grid_t *g, const chm_planes_t *rho_b,
                                          float *xb, float *yb, float *zb, float *wb,
                                          void *stream) {
  return deposit_adj<float, false>(x, y, z, nullptr, nullptr, nullptr, nullptr, w, charge, Np, g,
                                   rho_b, xb, yb, zb, nullptr, nullptr, nullptr, nullptr, wb,
                                   stream);
}
extern "C" int chm_deposit_vector_adj_f64(const double *x, const double *y, const double *z,
                                          const double *ux, const double *uy, const double *uz,
                                          const double *g_inv, const double *w, int32_t charge,
                                          int64_t Np, const chm_grid_t *g, const chm_planes_t *J_b,
                                          double *xb, double *yb, double *zb, double *uxb,
                                          double *uyb, double *uzb, double *gb, double *wb,
                                          void *stream) {
  return deposit_adj<double, true>(x, y, z, ux, uy, uz, g_inv, w, charge, Np, g, J_b, xb, yb, zb,
                                   uxb, uyb, uzb, gb, wb, stream);
}
extern "C" int chm_deposit_vector_adj_f32(const float *x, const float *y, const float *z,
                                          const float *ux, const float *uy, const float *uz,
                                          const float *g_inv, const float *w, int32_t charge,
                                          int64_t Np, const chm_grid_t *g, const chm_planes_t *J_b,
                                          float *xb, float *yb, float *zb, float *uxb, float *uyb,
                                          float *uzb, float *gb, float *wb, void *stream) {
  return deposit_adj<float, true>(x, y, z, ux, uy, uz, g_inv, w, charge, Np, g, J_b, xb, yb, zb,
                                  uxb, uyb, uzb, gb, wb, stream);
}

// ---------------------------------------------------------------- gather + push
struct GatherAdjOut {
  void *ub[3], *gb, *xb, *yb, *zb;
  void *fb[6];  // per-particle adjoints of E_p (0..2) and B_p (3..5)
};

// Adjoint of the gather + Boris push of one particle.  In: the adjoints of
// the pushed momenta and g_inv; out: the adjoint of the momenta before the
// push (g_inv before the push is overwritten, so its adjoint becomes 0), the
// position adjoints (added) and the per-particle field adjoints (written,
// zero for particles the push skips).
template <typename T, int M>
__global__ void __launch_bounds__(256)
    k_gather_push_adj(const T *__restrict__ x, const T *__restrict__ y, const T *__restrict__ z,
                      const T *__restrict__ ux, const T *__restrict__ uy, const T *__restrict__ uz,
                      int64_t Np, const uint32_t *__restrict__ np_stay, T factor_push, int Nx,
                      int Nr, T xmin, T dx_inv, T rmin, T dr_inv, VecPlanes E, VecPlanes B,
                      GatherAdjOut o) {
  constexpr int MM = M > 0 ? M : 1;
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Np) return;
  const int Nx_cell = Nx - 1, Nr_cell = Nr - 1;
  const T xp = x[p], yp = y[p], zp = z[p];
  const T rp = a_sqrt<T>(yp * yp + zp * zp);
  const int ix = (int)a_floor<T>((xp - xmin) * dx_inv);
  const int ir = (int)a_floor<T>((rp - rmin) * dr_inv);
  const bool act = p < (int64_t)(*np_stay) && ix > 0 && ix < Nx_cell - 1 && ir >= 0 &&
                   ir < Nr_cell - 1;
  if (!act) {
#pragma unroll
    for (int k = 0; k < 6; k++) ((T *)o.fb[k])[p] = 0;
    return;
  }
  const int64_t igrid = ix + (int64_t)ir * Nx;
  const T rp_inv = (T)1 / rp;
  T er[MM + 1], ei[MM + 1];
  mode_powers<T, M>(yp, zp, rp_inv, er, ei);
  const T sX1 = (xp - xmin) * dx_inv - (T)ix, sX0 = (T)1 - sX1;
  const T sR1 = (rp - rmin) * dr_inv - (T)ir, sR0 = (T)1 - sR1;
  T Fv[6], Fx[6], Fr[6], FSr[6][MM], FSi[6][MM];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const VecPlanes &P = k < 3 ? E : B;
    const int c = k < 3 ? k : k - 3;
    interp<T, M>((const T *)P.m0[c], P.mc[c], igrid, Nx, sX0, sX1, sR0, sR1, er, ei, (T)2, dx_inv,
                 dr_inv, Fv[k], Fx[k], Fr[k], FSr[k], FSi[k]);
  }
  // forward Boris (grid_deposit_m1.cl:472-507), intermediates kept
  const T h = (T)0.5 * factor_push;
  const T u[3] = {ux[p], uy[p], uz[p]};
  T um[3], t[3], s[3], u0[3], up[3], uf[3];
#pragma unroll
  for (int k = 0; k < 3; k++) um[k] = u[k] + h * Fv[k];
  const T gi = (T)1 / a_sqrt<T>((T)1 + um[0] * um[0] + um[1] * um[1] + um[2] * um[2]);
#pragma unroll
  for (int k = 0; k < 3; k++) t[k] = h * Fv[3 + k] * gi;
  const T t2 = (T)2 / ((T)1 + t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
#pragma unroll
  for (int k = 0; k < 3; k++) s[k] = t[k] * t2;
  u0[0] = um[0] + um[1] * t[2] - um[2] * t[1];
  u0[1] = um[1] - um[0] * t[2] + um[2] * t[0];
  u0[2] = um[2] + um[0] * t[1] - um[1] * t[0];
  up[0] = um[0] + u0[1] * s[2] - u0[2] * s[1];
  up[1] = um[1] - u0[0] * s[2] + u0[2] * s[0];
  up[2] = um[2] + u0[0] * s[1] - u0[1] * s[0];
#pragma unroll
  for (int k = 0; k < 3; k++) uf[k] = up[k] + h * Fv[k];
  const T gf = (T)1 / a_sqrt<T>((T)1 + uf[0] * uf[0] + uf[1] * uf[1] + uf[2] * uf[2]);
  // reverse
  T ufb[3] = {((T *)o.ub[0])[p], ((T *)o.ub[1])[p], ((T *)o.ub[2])[p]};
  const T gfb = ((T *)o.gb)[p];
  const T cg = -gfb * gf * gf * gf;  // d g/d u = -g^3 u
#pragma unroll
  for (int k = 0; k < 3; k++) ufb[k] += cg * uf[k];
  // up = um + u0 x s   (c = a x b:  abar = b x cbar, bbar = cbar x a)
  T umb[3], u0b[3], sb[3], tb[3];
  u0b[0] = s[1] * ufb[2] - s[2] * ufb[1];
  u0b[1] = s[2] * ufb[0] - s[0] * ufb[2];
  u0b[2] = s[0] * ufb[1] - s[1] * ufb[0];
  sb[0] = ufb[1] * u0[2] - ufb[2] * u0[1];
  sb[1] = ufb[2] * u0[0] - ufb[0] * u0[2];
  sb[2] = ufb[0] * u0[1] - ufb[1] * u0[0];
  // u0 = um + um x t
  umb[0] = ufb[0] + u0b[0] + (t[1] * u0b[2] - t[2] * u0b[1]);
  umb[1] = ufb[1] + u0b[1] + (t[2] * u0b[0] - t[0] * u0b[2]);
  umb[2] = ufb[2] + u0b[2] + (t[0] * u0b[1] - t[1] * u0b[0]);
  tb[0] = u0b[1] * um[2] - u0b[2] * um[1];
  tb[1] = u0b[2] * um[0] - u0b[0] * um[2];
  tb[2] = u0b[0] * um[1] - u0b[1] * um[0];
  // s = t * t2, t2 = 2/(1+|t|^2):  dt2/dt = -t2^2 t
  const T st = sb[0] * t[0] + sb[1] * t[1] + sb[2] * t[2];
#pragma unroll
  for (int k = 0; k < 3; k++) tb[k] += t2 * sb[k] - t2 * t2 * st * t[k];
  // t = h B gi
  T Fb[6];
  T gib = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    Fb[3 + k] = h * gi * tb[k];
    gib += h * Fv[3 + k] * tb[k];
  }
  // gi = (1 + |um|^2)^(-1/2)
  const T cm = -gib * gi * gi * gi;
#pragma unroll
  for (int k = 0; k < 3; k++) umb[k] += cm * um[k];
  // um = u + h E ; uf = up + h E
#pragma unroll
  for (int k = 0; k < 3; k++) Fb[k] = h * (ufb[k] + umb[k]);
#pragma unroll
  for (int k = 0; k < 3; k++) ((T *)o.ub[k])[p] = umb[k];
  ((T *)o.gb)[p] = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) ((T *)o.fb[k])[p] = Fb[k];
  // position adjoints through the interpolation
  T QX = 0, QR = 0, ebr[MM], ebi[MM];
#pragma unroll
  for (int m = 0; m < MM; m++) ebr[m] = ebi[m] = 0;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    QX += Fb[k] * Fx[k];
    QR += Fb[k] * Fr[k];
#pragma unroll
    for (int m = 0; m < M; m++) {
      ebr[m] += (T)2 * Fb[k] * FSr[k][m];
      ebi[m] += (T)2 * Fb[k] * FSi[k][m];
    }
  }
  ((T *)o.xb)[p] += QX;
  T yb = 0, zb = 0;
  yz_adjoint<T, M>(yp, zp, rp_inv, er, ei, ebr, ebi, QR, yb, zb);
  ((T *)o.yb)[p] += yb;
  ((T *)o.zb)[p] += zb;
}

template <typename T>
static int gather_push_adj(const T *x, const T *y, const T *z, const T *ux, const T *uy,
                           const T *uz, int64_t Np, const uint32_t *np_stay, T factor_push,
                           const chm_grid_t *g, const chm_planes_t *E, const chm_planes_t *B,
                           T *uxb, T *uyb, T *uzb, T *gb, T *xb, T *yb, T *zb, T *const *fb,
                           void *stream) {
  if (!g || !E || !B || !fb || g->M < 0 || g->M > CHM_MAXM || Np < 0) return CHM_EINVAL;
  if (Np == 0) return CHM_OK;
  VecPlanes ve = {}, vb = {};
  for (int k = 0; k < 3; k++) {
    ve.m0[k] = E[k].m0;
    vb.m0[k] = B[k].m0;
    for (int m = 0; m < g->M; m++) {
      ve.mc[k][m] = E[k].mc[m];
      vb.mc[k][m] = B[k].mc[m];
    }
  }
  GatherAdjOut o = {};
  o.ub[0] = uxb;
  o.ub[1] = uyb;
  o.ub[2] = uzb;
  o.gb = gb;
  o.xb = xb;
  o.yb = yb;
  o.zb = zb;
  for (int k = 0; k < 6; k++) {
    if (!fb[k]) return CHM_EINVAL;
    o.fb[k] = fb[k];
  }
  hipStream_t st = (hipStream_t)stream;
  unsigned nb = chm_blocks(Np, 256);
  switch (g->M) {
#define CASE(MM)                                                                                    \
  case MM:                                                                                          \
    k_gather_push_adj<T, MM><<<nb, 256, 0, st>>>(x, y, z, ux, uy, uz, Np, np_stay, factor_push,     \
                                                 g->Nx, g->Nr, (T)g->xmin, (T)g->dx_inv,            \
                                                 (T)g->rmin, (T)g->dr_inv, ve, vb, o);              \
    break;
    CASE(0) CASE(1) CASE(2) CASE(3) CASE(4)
#undef CASE
  }
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_gather_push_adj_f64(const double *x, const double *y, const double *z,
                                       const double *ux, const double *uy, const double *uz,
                                       int64_t Np, const uint32_t *np_stay, double factor_push,
                                       const chm_grid_t *g, const chm_planes_t *E,
                                       const chm_planes_t *B, double *uxb, double *uyb,
                                       double *uzb, double *gb, double *xb, double *yb, double *zb,
                                       double *const *fb, void *stream) {
  return gather_push_adj<double>(x, y, z, ux, uy, uz, Np, np_stay, factor_push, g, E, B, uxb, uyb,
                                 uzb, gb, xb, yb, zb, fb, stream);
}
extern "C" int chm_gather_push_adj_f32(const float *x, const float *y, const float *z,
                                       const float *ux, const float *uy, const float *uz,
                                       int64_t Np, const uint32_t *np_stay, float factor_push,
                                       const chm_grid_t *g, const chm_planes_t *E,
                                       const chm_planes_t *B, float *uxb, float *uyb, float *uzb,
                                       float *gb, float *xb, float *yb, float *zb,
                                       float *const *fb, void *stream) {
  return gather_push_adj<float>(x, y, z, ux, uy, uz, Np, np_stay, factor_push, g, E, B, uxb, uyb,
                                uzb, gb, xb, yb, zb, fb, stream);
}

// ---------------------------------------------------------------- plane adjoints
#define ADJ_MAX_PLANES 64
struct AdjPlaneList {
  void *p[ADJ_MAX_PLANES];
  int cplx[ADJ_MAX_PLANES];
};

static int adj_plane_list(AdjPlaneList *pl, const chm_planes_t *flds, int nflds, int M) {
  int n = 0;
  for (int f = 0; f < nflds; f++) {
    if (n + 1 + M > ADJ_MAX_PLANES) return -1;
    pl->p[n] = flds[f].m0;
    pl->cplx[n++] = 0;
    for (int m = 0; m < M; m++) {
      pl->p[n] = flds[f].mc[m];
      pl->cplx[n++] = 1;
    }
  }
  return n;
}

// Adjoint of the axis post-processing (row1 = (row1 - row0) * dV_inv[1],
// row0 = row0 * dV_inv[0], row_i = row_i * dV_inv[i]):
//   row0 <- dV_inv[0] row0 - dV_inv[1] row1,  row_i <- dV_inv[i] row_i.
template <typename T>
__global__ void k_postproc_adj(AdjPlaneList pl, const double *__restrict__ dV_inv, int Nx, int Nr) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int pi = blockIdx.y;
  if (e >= (int64_t)(Nr - 1) * Nx) return;
  int ir = 1 + (int)(e / Nx), ix = (int)(e % Nx);
  T *a = (T *)pl.p[pi];
  int nc = pl.cplx[pi] ? 2 : 1;
  T dvi = (T)dV_inv[ir];
  for (int h = 0; h < nc; h++) {
    int64_t i1 = ((int64_t)ir * Nx + ix) * nc + h;
    T v = a[i1];
    a[i1] = v * dvi;
    if (ir == 1) {
      int64_t i0 = (int64_t)ix * nc + h;
      a[i0] = a[i0] * (T)dV_inv[0] - v * dvi;
    }
  }
}

// Adjoint of the ghost-row warp (row0 = row1 for m = 0, -row1 for m >= 1):
//   row1 += +-row0, row0 <- 0.
template <typename T> __global__ void k_warp_adj(AdjPlaneList pl, int Nx) {
  int ix = blockIdx.x * blockDim.x + threadIdx.x;
  int pi = blockIdx.y;
  if (ix >= Nx) return;
  T *a = (T *)pl.p[pi];
  if (pl.cplx[pi]) {
    a[2 * (Nx + ix)] -= a[2 * ix];
    a[2 * (Nx + ix) + 1] -= a[2 * ix + 1];
    a[2 * ix] = 0;
    a[2 * ix + 1] = 0;
  } else {
    a[Nx + ix] += a[ix];
    a[ix] = 0;
  }
}

template <typename T>
static int postproc_adj(const chm_planes_t *flds, int32_t nflds, const double *dV_inv, int32_t Nx,
                        int32_t Nr, int32_t M, void *stream) {
  if (!flds || nflds <= 0 || M < 0 || M > CHM_MAXM || Nr < 2 || !dV_inv) return CHM_EINVAL;
  AdjPlaneList pl;
  int np = adj_plane_list(&pl, flds, nflds, M);
  if (np < 0) return CHM_EINVAL;
  dim3 grid(chm_blocks((int64_t)(Nr - 1) * Nx, 256), np);
  k_postproc_adj<T><<<grid, 256, 0, (hipStream_t)stream>>>(pl, dV_inv, Nx, Nr);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

template <typename T>
static int warp_adj(const chm_planes_t *flds, int32_t nflds, int32_t Nx, int32_t Nr, int32_t M,
                    void *stream) {
  if (!flds || nflds <= 0 || M < 0 || M > CHM_MAXM || Nr < 2) return CHM_EINVAL;
  AdjPlaneList pl;
  int np = adj_plane_list(&pl, flds, nflds, M);
  if (np < 0) return CHM_EINVAL;
  dim3 grid(chm_blocks(Nx, 256), np);
  k_warp_adj<T><<<grid, 256, 0, (hipStream_t)stream>>>(pl, Nx);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_postproc_deposit_adj_f64(const chm_planes_t *flds, int32_t nflds,
                                            const double *dV_inv, int32_t Nx, int32_t Nr,
                                            int32_t M, void *stream) {
  return postproc_adj<double>(flds, nflds, dV_inv, Nx, Nr, M, stream);
}
extern "C" int chm_postproc_deposit_adj_f32(const chm_planes_t *flds, int32_t nflds,
                                            const double *dV_inv, int32_t Nx, int32_t Nr,
                                            int32_t M, void *stream) {
  return postproc_adj<float>(flds, nflds, dV_inv, Nx, Nr, M, stream);
}
extern "C" int chm_warp_axis_adj_f64(const chm_planes_t *flds, int32_t nflds, int32_t Nx,
                                     int32_t Nr, int32_t M, void *stream) {
  return warp_adj<double>(flds, nflds, Nx, Nr, M, stream);
}
extern "C" int chm_warp_axis_adj_f32(const chm_planes_t *flds, int32_t nflds, int32_t Nx,
                                     int32_t Nr, int32_t M, void *stream) {
  return warp_adj<float>(flds, nflds, Nx, Nr, M, stream);
}

// ---------------------------------------------------------------- density profile
// Adjoint of w = w_base * f(x) with f piecewise linear through the knots
// (profile_by_interpolant, particles_generic.cl:6-30, constant extension):
//   knot_grad[j] += sum_p wbar_p * w_base_p * df(x_p)/dff_j.
// Pass 1: each block sums its particles into NKN per-lane accumulators,
// reduces them over the block in a fixed order and writes one row of
// partial sums; pass 2 adds the rows in block order.  Deterministic.
#define PROF_BLOCKS 512
#define PROF_NT 256

template <typename T, int NKN>
__global__ void __launch_bounds__(PROF_NT)
    k_profile_adj(const T *__restrict__ x, const T *__restrict__ w_base,
                  const T *__restrict__ w_bar, int64_t Np, const double *__restrict__ xx,
                  const double *__restrict__ dxm1, int n, double *__restrict__ partial) {
  double acc[NKN];
#pragma unroll
  for (int k = 0; k < NKN; k++) acc[k] = 0;
  for (int64_t p = (int64_t)blockIdx.x * PROF_NT + threadIdx.x; p < Np;
       p += (int64_t)gridDim.x * PROF_NT) {
    double xp = (double)x[p];
    if (xp < xx[0]) xp = xx[0];
    if (xp > xx[n - 1]) xp = xx[n - 1];
    int ix = 0;
    if (xp > xx[0])
      for (ix = 0; ix < n - 2; ix++)
        if (xp > xx[ix] && xp <= xx[ix + 1]) break;
    const double c = (double)w_bar[p] * (double)w_base[p] * dxm1[ix];
    const double c0 = c * (xx[ix + 1] - xp), c1 = c * (xp - xx[ix]);
#pragma unroll
    for (int k = 0; k < NKN; k++) {
      if (k == ix) acc[k] += c0;
      if (k == ix + 1) acc[k] += c1;
    }
  }
  __shared__ double red[PROF_NT / 64][NKN];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NKN; k++) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NKN; k += PROF_NT) {
    double v = 0;
    for (int wv = 0; wv < PROF_NT / 64; wv++) v += red[wv][k];
    partial[(int64_t)blockIdx.x * NKN + k] = v;
  }
}

template <int NKN>
__global__ void k_profile_sum(const double *__restrict__ partial, int nblk, int n,
                              double *__restrict__ grad) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v = 0;
  for (int b = 0; b < nblk; b++) v += partial[(int64_t)b * NKN + k];
  grad[k] += v;
}

extern "C" size_t chm_profile_adj_workspace_bytes(int32_t nknots) {
  return chm_align((size_t)PROF_BLOCKS * 128 * sizeof(double));
}

template <typename T>
static int profile_adj(const T *x, const T *w_base, const T *w_bar, int64_t Np, const double *xx,
                       const double *dxm1, int32_t n, double *grad, void *ws, size_t ws_bytes,
                       void *stream) {
  if (Np < 0 || n < 2 || !grad) return CHM_EINVAL;
  if (n > 128) return CHM_EUNSUPPORTED;
  if (!ws || ws_bytes < chm_profile_adj_workspace_bytes(n)) return CHM_EWORKSPACE;
  if (Np == 0) return CHM_OK;
  hipStream_t st = (hipStream_t)stream;
  double *part = (double *)ws;
  int nblk = (int)((Np + PROF_NT - 1) / PROF_NT);
  if (nblk > PROF_BLOCKS) nblk = PROF_BLOCKS;
#define RUN(NK)                                                                               \
  k_profile_adj<T, NK><<<nblk, PROF_NT, 0, st>>>(x, w_base, w_bar, Np, xx, dxm1, n, part);    \
  k_profile_sum<NK><<<1, 128, 0, st>>>(part, nblk, n, grad);
  if (n <= 16) { RUN(16) }
  else if (n <= 32) { RUN(32) }
  else if (n <= 64) { RUN(64) }
  else { RUN(128) }
#undef RUN
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_dens_profile_adj_f64(const double *x, const double *w_base,
                                        const double *w_bar, int64_t Np, const double *xx,
                                        const double *dxm1, int32_t nknots, double *knot_grad,
                                        void *ws, size_t ws_bytes, void *stream) {
  return profile_adj<double>(x, w_base, w_bar, Np, xx, dxm1, nknots, knot_grad, ws, ws_bytes,
                             stream);
}
extern "C" int chm_dens_profile_adj_f32(const float *x, const float *w_base, const float *w_bar,
                                        int64_t Np, const double *xx, const double *dxm1,
                                        int32_t nknots, double *knot_grad, void *ws,
                                        size_t ws_bytes, void *stream) {
  return profile_adj<float>(x, w_base, w_bar, Np, xx, dxm1, nknots, knot_grad, ws, ws_bytes,
                            stream);
}
