/*
 * CPU FP64 oracle for the particle side of the chimeraCL PIC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithm in
 * plain C so that the HIP kernels can be checked against it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product path never calls it.
 *
 * Built with -O2 -ffp-contract=off and no fast-math, so every multiply and add
 * is rounded on its own exactly as the IEEE-754 expression is written.  The
 * operation order of each expression follows the reference kernels line by
 * line (cited per function) so that the HIP kernels, built the same way, are
 * bit-exact against it.
 *
 * Layouts: particle arrays are SoA [Np] f64.  A real-space plane is [Nr][Nx]
 * row-major; mode 0 is f64, modes m=1..M are complex f64 stored interleaved
 * (re, im) in one array [M][Nr][Nx][2].  Row 0 is the ghost row r = -dr/2.
 *
 * Modes m >= 2 have no reference kernel (the reference ships kernels for M=0
 * and M=1 only: chimeraCL/methods/grid_methods_cl.py:21-23).  They use the
 * deposit weight ((y + i z)/r)^m and the gather weight conj(.) of the same,
 * which for m=1 is exactly grid_deposit_m1.cl:116-117 and :407-408.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* push_xyz, particles_generic.cl:129-153 */
void orc_push_xyz(double *x, double *y, double *z, const double *px,
                  const double *py, const double *pz, const double *g_inv,
                  double dt, int64_t Np) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < Np; ip++) {
    double dt_g = dt * g_inv[ip];
    double dx = px[ip] * dt_g;
    double dy = py[ip] * dt_g;
    double dz = pz[ip] * dt_g;
    x[ip] = x[ip] + dx;
    y[ip] = y[ip] + dy;
    z[ip] = z[ip] + dz;
  }
}

/* index_and_sum_in_cell, particles_generic.cl:88-126.
 * key = ix + ir*(Nx-1) for a valid cell, else (Nr-1)*(Nx-1). */
void orc_cell_keys(const double *x, const double *y, const double *z,
                   int64_t Np, int32_t Nx, double xmin, double dx_inv,
                   int32_t Nr, double rmin, double dr_inv, uint32_t *key) {
  const int Nx_loc = Nx - 1, Nr_loc = Nr - 1;
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < Np; ip++) {
    double r = sqrt(y[ip] * y[ip] + z[ip] * z[ip]);
    int ix = (int)floor((x[ip] - xmin) * dx_inv);
    int ir = (int)floor((r - rmin) * dr_inv);
    if (ix > 0 && ix < Nx_loc - 1 && ir < Nr_loc - 1 && ir >= 0)
      key[ip] = (uint32_t)(ix + ir * Nx_loc);
    else
      key[ip] = (uint32_t)(Nr_loc * Nx_loc);
  }
}

/* Histogram + cumsum + scatter of index_sort, particles_methods_cl.py:225-261,
 * with the canonical stable in-cell order (ascending particle index), i.e.
 * numpy.argsort(key, kind='stable').  cell_offset has nbins+1 entries:
 * cell_offset[0] = 0, cell_offset[c+1] = number of particles with key <= c.
 * (The reference's array is one longer; its [-2] entry, Np_stay, is
 * cell_offset[nbins-1] here, the count of particles in valid cells.) */
void orc_stable_sort(const uint32_t *key, int64_t Np, int64_t nbins,
                     uint32_t *cell_offset, uint32_t *sort_indx) {
  uint32_t *cur = (uint32_t *)calloc((size_t)nbins, sizeof(uint32_t));
  memset(cell_offset, 0, (size_t)(nbins + 1) * sizeof(uint32_t));
  for (int64_t ip = 0; ip < Np; ip++) cell_offset[key[ip] + 1]++;
  for (int64_t c = 0; c < nbins; c++) cell_offset[c + 1] += cell_offset[c];
  for (int64_t ip = 0; ip < Np; ip++) {
    uint32_t c = key[ip];
    sort_indx[cell_offset[c] + cur[c]] = (uint32_t)ip;
    cur[c]++;
  }
  free(cur);
}

/* One cell's CIC weights, grid_deposit_m1.cl:113-130 (deposit form). */
static inline void cic_weights(double xp, double rp, double wp, double xmin,
                               double dx_inv, double rmin, double dr_inv,
                               int ix, int ir, double C[2][2]) {
  double sX1 = (xp - xmin) * dx_inv - ix;
  double sX0 = 1.0 - sX1;
  double sR1 = (rp - rmin) * dr_inv - ir;
  double sR0 = 1.0 - sR1;
  sX0 *= wp;
  sX1 *= wp;
  C[0][0] = sR0 * sX0;
  C[0][1] = sR0 * sX1;
  C[1][0] = sR1 * sX0;
  C[1][1] = sR1 * sX1;
}

/* e_m = e_{m-1} * e_1 (complex), e_1 given. */
static inline void mode_powers(double e1r, double e1i, int M, double *er,
                               double *ei) {
  er[1] = e1r;
  ei[1] = e1i;
  for (int m = 2; m <= M; m++) {
    er[m] = er[m - 1] * e1r - ei[m - 1] * e1i;
    ei[m] = er[m - 1] * e1i + ei[m - 1] * e1r;
  }
}

#define MAXM 8

/*
 * depose_scalar, grid_deposit_m1.cl:20-152 (m0: grid_deposit_m0.cl:20-129)
 * launched as in grid_methods_cl.py:44-69: four colour passes over the
 * (Nr//2)*(Nx//2) 4-cell groups; within a pass each valid non-empty cell sums
 * its particles in sort order into private accumulators and then adds them to
 * its four nodes.  rho_m0 [Nr][Nx]; rho_mc [M][Nr][Nx][2].
 */
void orc_deposit_scalar(const uint32_t *sort_indx, const uint32_t *cell_offset,
                        const double *x, const double *y, const double *z,
                        const double *w, int32_t charge, int32_t Nx,
                        double xmin, double dx_inv, int32_t Nr, double rmin,
                        double dr_inv, int32_t M, double *rho_m0,
                        double *rho_mc) {
  const int Nx_cell = Nx - 1, Nr_cell = Nr - 1, Nx_2 = Nx / 2;
  const int64_t ngroups = (int64_t)(Nr / 2) * (Nx / 2);
  const int64_t plane = (int64_t)Nr * Nx;
  const double q = (double)(int8_t)charge; /* int8 charge, grid_methods_cl.py:64 */
  for (int off = 0; off < 4; off++) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t g = 0; g < ngroups; g++) {
      int ir = (int)(g / Nx_2), ix = (int)(g - (int64_t)ir * Nx_2);
      ix *= 2;
      ir *= 2;
      if (off == 1) ix += 1;
      else if (off == 2) ir += 1;
      else if (off == 3) { ix += 1; ir += 1; }
      if (!(ix > 0 && ix < Nx_cell - 1 && ir < Nr_cell - 1)) continue;
      int64_t icell = ix + (int64_t)ir * Nx_cell;
      int64_t igrid = ix + (int64_t)ir * Nx;
      uint32_t ip0 = cell_offset[icell], ip1 = cell_offset[icell + 1];
      if (ip0 == ip1) continue;
      double a0[2][2] = {{0, 0}, {0, 0}};
      double ar[MAXM + 1][2][2], ai[MAXM + 1][2][2];
      memset(ar, 0, sizeof(ar));
      memset(ai, 0, sizeof(ai));
      for (uint32_t ip = ip0; ip < ip1; ip++) {
        uint32_t p = sort_indx[ip];
        double xp = x[p], yp = y[p], zp = z[p];
        double wp = w[p] * q;
        double rp = sqrt(yp * yp + zp * zp);
        double rp_inv = 1. / rp;
        double er[MAXM + 1], ei[MAXM + 1];
        mode_powers(yp * rp_inv, zp * rp_inv, M, er, ei);
        double C[2][2];
        cic_weights(xp, rp, wp, xmin, dx_inv, rmin, dr_inv, ix, ir, C);
        for (int i = 0; i < 2; i++)
          for (int j = 0; j < 2; j++) {
            a0[i][j] += C[i][j];
            for (int m = 1; m <= M; m++) {
              ar[m][i][j] += C[i][j] * er[m];
              ai[m][i][j] += C[i][j] * ei[m];
            }
          }
      }
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          int64_t id = igrid + j + (int64_t)Nx * i;
          rho_m0[id] = rho_m0[id] + a0[i][j];
          for (int m = 1; m <= M; m++) {
            double *c = rho_mc + 2 * ((m - 1) * plane + id);
            c[0] = c[0] + ar[m][i][j];
            c[1] = c[1] + ai[m][i][j];
          }
        }
    }
  }
}

/*
 * depose_vector, grid_deposit_m1.cl:171-327 (m0: grid_deposit_m0.cl:148-277),
 * launched as in grid_methods_cl.py:71-95.  J_m0 [3][Nr][Nx];
 * J_mc [3][M][Nr][Nx][2] (component-major).  wp = w*g_inv*charge;
 * rp_inv = 0 at r = 0 (grid_deposit_m1.cl:280-281).
 */
void orc_deposit_vector(const uint32_t *sort_indx, const uint32_t *cell_offset,
                        const double *x, const double *y, const double *z,
                        const double *ux, const double *uy, const double *uz,
                        const double *g_inv, const double *w, int32_t charge,
                        int32_t Nx, double xmin, double dx_inv, int32_t Nr,
                        double rmin, double dr_inv, int32_t M, double *J_m0,
                        double *J_mc) {
  const int Nx_cell = Nx - 1, Nr_cell = Nr - 1, Nx_2 = Nx / 2;
  const int64_t ngroups = (int64_t)(Nr / 2) * (Nx / 2);
  const int64_t plane = (int64_t)Nr * Nx;
  const double q = (double)(int8_t)charge;
  for (int off = 0; off < 4; off++) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t g = 0; g < ngroups; g++) {
      int ir = (int)(g / Nx_2), ix = (int)(g - (int64_t)ir * Nx_2);
      ix *= 2;
      ir *= 2;
      if (off == 1) ix += 1;
      else if (off == 2) ir += 1;
      else if (off == 3) { ix += 1; ir += 1; }
      if (!(ix > 0 && ix < Nx_cell - 1 && ir < Nr_cell - 1)) continue;
      int64_t icell = ix + (int64_t)ir * Nx_cell;
      int64_t igrid = ix + (int64_t)ir * Nx;
      uint32_t ip0 = cell_offset[icell], ip1 = cell_offset[icell + 1];
      if (ip0 == ip1) continue;
      double a0[3][2][2];
      double ar[3][MAXM + 1][2][2], ai[3][MAXM + 1][2][2];
      memset(a0, 0, sizeof(a0));
      memset(ar, 0, sizeof(ar));
      memset(ai, 0, sizeof(ai));
      for (uint32_t ip = ip0; ip < ip1; ip++) {
        uint32_t p = sort_indx[ip];
        double xp = x[p], yp = y[p], zp = z[p];
        double jp[3] = {ux[p], uy[p], uz[p]};
        double wp = w[p] * g_inv[p] * q;
        double rp = sqrt(yp * yp + zp * zp);
        double rp_inv = 0;
        if (rp > 0) rp_inv = 1. / rp;
        double er[MAXM + 1], ei[MAXM + 1];
        mode_powers(yp * rp_inv, zp * rp_inv, M, er, ei);
        double C[2][2];
        cic_weights(xp, rp, wp, xmin, dx_inv, rmin, dr_inv, ix, ir, C);
        for (int k = 0; k < 3; k++)
          for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) {
              double jproj = C[i][j] * jp[k];
              a0[k][i][j] += C[i][j] * jp[k];
              for (int m = 1; m <= M; m++) {
                ar[k][m][i][j] += jproj * er[m];
                ai[k][m][i][j] += jproj * ei[m];
              }
            }
      }
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          int64_t id = igrid + j + (int64_t)Nx * i;
          for (int k = 0; k < 3; k++) {
            double *r0 = J_m0 + k * plane;
            r0[id] = r0[id] + a0[k][i][j];
          }
          for (int k = 0; k < 3; k++)
            for (int m = 1; m <= M; m++) {
              double *c = J_mc + 2 * (((int64_t)k * M + (m - 1)) * plane + id);
              c[0] = c[0] + ar[k][m][i][j];
              c[1] = c[1] + ai[k][m][i][j];
            }
        }
    }
  }
}

/*
 * gather_and_push, grid_deposit_m1.cl:330-511 (m0: grid_deposit_m0.cl:280-427),
 * launched per sorted slot as in grid_methods_cl.py:168-192.  The result of
 * each particle does not depend on the slot order, so this loops over the
 * slots exactly as the reference does.  pid (may be NULL) holds each
 * particle's index in the unreordered numbering.  factor_push is Particles
 * Args['FactorPush'] = 2*pi*dt*q/m (particles.py:90); the kernel halves it.
 * E_m0/B_m0 [3][Nr][Nx]; E_mc/B_mc [3][M][Nr][Nx][2].
 */
void orc_gather_push(const double *x, const double *y, const double *z,
                     double *px, double *py, double *pz, double *g_inv,
                     const uint32_t *sort_indx, const uint32_t *pid, int64_t Np,
                     int64_t Np_stay,
                     double factor_push, int32_t Nx, double xmin,
                     double dx_inv, int32_t Nr, double rmin, double dr_inv,
                     int32_t M, const double *E_m0, const double *B_m0,
                     const double *E_mc, const double *B_mc) {
  const int Nx_cell = Nx - 1, Nr_cell = Nr - 1;
  const int64_t plane = (int64_t)Nr * Nx;
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < Np; ip++) {
    uint32_t p = sort_indx[ip];
    /* K5: the skip compares the particle's index without reordering
     * (pid, when the storage has been reordered) with Np_stay */
    if (!((int64_t)(pid ? pid[p] : p) < Np_stay)) continue;
    double xp = x[p], yp = y[p], zp = z[p];
    double rp = sqrt(yp * yp + zp * zp);
    int ix = (int)floor((xp - xmin) * dx_inv);
    int ir = (int)floor((rp - rmin) * dr_inv);
    if (!(ix > 0 && ix < Nx_cell - 1 && ir < Nr_cell - 1)) continue;
    int64_t igrid = ix + (int64_t)ir * Nx;
    double u[3] = {px[p], py[p], pz[p]};
    double dt_2 = 0.5 * factor_push;
    double rp_inv = 1. / rp;
    double sX1 = (xp - xmin) * dx_inv - ix;
    double sX0 = 1.0 - sX1;
    double sR1 = (rp - rmin) * dr_inv - ir;
    double sR0 = 1.0 - sR1;
    double gr[MAXM + 1], gi[MAXM + 1];
    mode_powers(yp * rp_inv, -zp * rp_inv, M, gr, gi);
    double C[2][2];
    C[0][0] = sR0 * sX0;
    C[0][1] = sR0 * sX1;
    C[1][0] = sR1 * sX0;
    C[1][1] = sR1 * sX1;
    double e_p[3] = {0, 0, 0}, b_p[3] = {0, 0, 0};
    for (int k = 0; k < 3; k++)
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          int64_t id = igrid + j + (int64_t)Nx * i;
          e_p[k] += C[i][j] * E_m0[k * plane + id];
          b_p[k] += C[i][j] * B_m0[k * plane + id];
          for (int m = 1; m <= M; m++) {
            const double *e = E_mc + 2 * (((int64_t)k * M + (m - 1)) * plane + id);
            const double *b = B_mc + 2 * (((int64_t)k * M + (m - 1)) * plane + id);
            double er = 2 * e[0], ei = 2 * e[1];
            double br = 2 * b[0], bi = 2 * b[1];
            e_p[k] += C[i][j] * er * gr[m];
            e_p[k] -= C[i][j] * ei * gi[m];
            b_p[k] += C[i][j] * br * gr[m];
            b_p[k] -= C[i][j] * bi * gi[m];
          }
        }
    /* Boris rotation, grid_deposit_m1.cl:472-507 */
    double um[3], up[3], u0[3], t[3], s[3];
    for (int k = 0; k < 3; k++) um[k] = u[k] + dt_2 * e_p[k];
    double g_p_inv = 1. / sqrt(1. + um[0] * um[0] + um[1] * um[1] + um[2] * um[2]);
    for (int k = 0; k < 3; k++) t[k] = dt_2 * b_p[k] * g_p_inv;
    double t2 = 2. / (1. + t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    for (int k = 0; k < 3; k++) s[k] = t[k] * t2;
    u0[0] = um[0] + um[1] * t[2] - um[2] * t[1];
    u0[1] = um[1] - um[0] * t[2] + um[2] * t[0];
    u0[2] = um[2] + um[0] * t[1] - um[1] * t[0];
    up[0] = um[0] + u0[1] * s[2] - u0[2] * s[1];
    up[1] = um[1] - u0[0] * s[2] + u0[2] * s[0];
    up[2] = um[2] + u0[0] * s[1] - u0[1] * s[0];
    for (int k = 0; k < 3; k++) u[k] = up[k] + dt_2 * e_p[k];
    g_p_inv = 1. / sqrt(1. + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    px[p] = u[0];
    py[p] = u[1];
    pz[p] = u[2];
    g_inv[p] = g_p_inv;
  }
}

/* profile_by_interpolant, particles_generic.cl:6-30.  Inside [xx[0], xx[n-1]]
 * the factor is the reference's expression.  Outside, where the reference
 * search falls through and reads ff[n] out of bounds, x is clamped to the end
 * knots (constant extension). */
void orc_profile_by_interpolant(const double *x, double *w, int64_t Np,
                                const double *xx, const double *ff,
                                const double *dxm1, int32_t nknots) {
#pragma omp parallel for schedule(static)
  for (int64_t ip = 0; ip < Np; ip++) {
    double xp = x[ip];
    if (xp < xx[0]) xp = xx[0];
    if (xp > xx[nknots - 1]) xp = xx[nknots - 1];
    int ix = 0;
    if (xp > xx[0])
      for (ix = 0; ix < nknots - 2; ix++)
        if (xp > xx[ix] && xp <= xx[ix + 1]) break;
    double f_minus = ff[ix] * dxm1[ix];
    double f_plus = ff[ix + 1] * dxm1[ix];
    w[ip] *= f_minus * (xx[ix + 1] - xp) + f_plus * (xp - xx[ix]);
  }
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
