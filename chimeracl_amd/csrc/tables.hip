// FP64 construction of every spectral table on the GPU.
//
//   kx      = 2 pi fftfreq(Nx, dx)                     (transformer.py:38)
//   kr_m    = j_{m,k} / R_period, k = 1..Nr-1, m = 0..M+1 (transformer.py:40-45)
//             Bessel roots: McMahon's asymptotic start, then Newton on J_m.
//   DHT_inv = J_m(r_i kr_{m,k}), r_i = Rgrid[1:]      (transformer.py:77-82)
//   DHT     = DHT_inv^{-1} (Gauss-Jordan, partial pivoting; the reference's
//             pinv equals the inverse for these full-rank matrices)
//   dDHT+-  = DHT @ [kr_{m+-1}/2 J_m(r_i kr_{m+-1,k})] (transformer.py:89-92)
//   w, Poisson 1/w^2, smoothing filter                 (transformer.py:47-57)
//   cos(w dt), w sin(w dt), 1/w^2                      (solver.py:41-52)
//   phase exp(-+ i kx Xmin)                            (transformer_generic.cl:28-55)
//   damping profile                                    (solver_methods_cl.py:32-41)
//   FFT twiddles exp(-2 pi i t / Nx)
#include "gemm.h"

#define PI_D 3.14159265358979323846

// J_n(x), n >= 0, x >= 0: power series below x = 2, forward recurrence from
// J_0, J_1 above (stable there for the small orders used here).
__device__ double bessel_jn(int n, double x) {
  if (x < 2.0) {
    double h = 0.5 * x, h2 = h * h;
    double t = 1.0;
    for (int k = 1; k <= n; k++) t *= h / (double)k;
    double s = t;
    for (int k = 1; k < 40; k++) {
      t *= -h2 / ((double)k * (double)(k + n));
      s += t;
      if (fabs(t) < 1e-18 * fabs(s)) break;
    }
    return s;
  }
  double j0v = j0(x);
  if (n == 0) return j0v;
  double j1v = j1(x);
  for (int i = 1; i < n; i++) {
    double j2 = (2.0 * i) / x * j1v - j0v;
    j0v = j1v;
    j1v = j2;
  }
  return j1v;
}

// k-th positive root (k >= 1) of J_m.
__device__ double bessel_root(int m, int k) {
  double mu = 4.0 * m * m;
  double b = (k + 0.5 * m - 0.25) * PI_D;
  double e = 8.0 * b;
  double x = b - (mu - 1) / e - 4 * (mu - 1) * (7 * mu - 31) / (3 * e * e * e) -
             32 * (mu - 1) * (83 * mu * mu - 982 * mu + 3779) / (15 * e * e * e * e * e);
  for (int it = 0; it < 50; it++) {
    double f = bessel_jn(m, x);
    double fp = (m == 0) ? -bessel_jn(1, x) : (bessel_jn(m - 1, x) - (double)m / x * f);
    double dx = f / fp;
    x -= dx;
    if (fabs(dx) < 1e-16 * x) break;
  }
  return x;
}

__global__ void k_kx(double *kx, int Nx, double dx) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Nx) return;
  // numpy.fft.fftfreq: integers [0..(n-1)/2, -(n/2)..-1] times 1/(n d)
  int half = (Nx - 1) / 2 + 1;
  double fi = i < half ? (double)i : (double)(i - Nx);
  double val = 1.0 / ((double)Nx * dx);
  kx[i] = (2.0 * PI_D) * (fi * val);
}

__global__ void k_roots(double *kr, int nm, int Nk, double R_period) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nm * Nk) return;
  int m = e / Nk, k = e % Nk;
  kr[e] = bessel_root(m, k + 1) / R_period;
}

// out[i][k] = scale(k) * J_m(r_i kr[k]); scale = 1 or kr[k]/2
__global__ void k_jmat(double *out, const double *kr, int m, int Nk, double rmin, double dr, int half_kr) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)Nk * Nk) return;
  int i = (int)(e / Nk), k = (int)(e % Nk);
  double r = rmin + dr * (double)(i + 1);
  double v = bessel_jn(m, r * kr[k]);
  out[e] = half_kr ? (0.5 * kr[k]) * v : v;
}

// Gauss-Jordan inverse of the n x n matrix a (row-major) into inv, one
// workgroup, partial pivoting; aug is [n][2n] scratch.
__global__ void __launch_bounds__(1024) k_gauss_jordan(const double *a, double *inv, double *aug, int n) {
  __shared__ double fac[1024];
  __shared__ double rv[1024];
  __shared__ int ri[1024];
  const int t = threadIdx.x, nt = blockDim.x, w2 = 2 * n;
  for (int64_t e = t; e < (int64_t)n * w2; e += nt) {
    int i = (int)(e / w2), j = (int)(e % w2);
    aug[e] = j < n ? a[(int64_t)i * n + j] : (j - n == i ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int p = 0; p < n; p++) {
    // pivot search over rows p..n-1 (ties -> lowest row)
    double best = -1.0;
    int bi = p;
    for (int i = p + t; i < n; i += nt) {
      double v = fabs(aug[(int64_t)i * w2 + p]);
      if (v > best) { best = v; bi = i; }
    }
    rv[t] = best;
    ri[t] = bi;
    __syncthreads();
    for (int s = nt / 2; s > 0; s >>= 1) {
      if (t < s) {
        double o = rv[t + s];
        int oi = ri[t + s];
        if (o > rv[t] || (o == rv[t] && oi < ri[t])) { rv[t] = o; ri[t] = oi; }
      }
      __syncthreads();
    }
    int piv = ri[0];
    __syncthreads();
    if (piv != p)
      for (int j = t; j < w2; j += nt) {
        double u = aug[(int64_t)p * w2 + j];
        aug[(int64_t)p * w2 + j] = aug[(int64_t)piv * w2 + j];
        aug[(int64_t)piv * w2 + j] = u;
      }
    __syncthreads();
    double d = aug[(int64_t)p * w2 + p];
    __syncthreads();
    for (int j = t; j < w2; j += nt) aug[(int64_t)p * w2 + j] /= d;
    __syncthreads();
    for (int i = t; i < n; i += nt) fac[i] = aug[(int64_t)i * w2 + p];
    __syncthreads();
    for (int64_t e = t; e < (int64_t)n * w2; e += nt) {
      int i = (int)(e / w2), j = (int)(e % w2);
      if (i != p) aug[e] -= fac[i] * aug[(int64_t)p * w2 + j];
    }
    __syncthreads();
  }
  for (int64_t e = t; e < (int64_t)n * n; e += nt) {
    int i = (int)(e / n), j = (int)(e % n);
    inv[e] = aug[(int64_t)i * w2 + n + j];
  }
}

__global__ void k_wtables(const double *kx, const double *kr, int Nx, int Nk, double dt, double *w,
                          double *c1, double *c2, double *c3, double *smooth) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)Nk * Nx) return;
  int i = (int)(e / Nx), j = (int)(e % Nx);
  double kxv = kx[j], krv = kr[i];
  double ww = sqrt(kxv * kxv + krv * krv);
  w[e] = ww;
  c1[e] = cos(ww * dt);
  c2[e] = sin(ww * dt) * ww;
  c3[e] = 1. / (ww * ww);
  // kx.max() of fftfreq order is kx[(Nx-1)/2]; kr.max() is the last root
  double kxmax = kx[(Nx - 1) / 2], krmax = kr[Nk - 1];
  double sx = sin(0.5 * PI_D * kxv / kxmax), sr = sin(0.5 * PI_D * krv / krmax);
  smooth[e] = (1 - sx * sx) * (1 - sr * sr);
}

__global__ void k_phase_tw(const double *kx, int Nx, double xmin, double2 *pm, double2 *pp, double2 *tw) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Nx) return;
  double a = xmin * kx[i];
  double c = cos(a), s = sin(a);
  pm[i] = make_double2(c, -s);
  pp[i] = make_double2(c, s);
  double sn, cs;
  sincospi(2.0 * (double)i / (double)Nx, &sn, &cs);
  tw[i] = make_double2(cs, -sn);
}

__global__ void k_damp_profile(double *f, int N) {
  int z = blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= 2 * N) return;
  double zs = 3. * (z - N + 1) / (N + 1);
  double s = sin(0.5 * PI_D * zs);
  double v = ((z < 4. * N / 3) && (z >= N)) ? s * s : 0.0;
  if (z >= 4. * N / 3) v += 1.0;
  f[z] = v;
}

extern "C" size_t chm_tables_workspace_bytes(int32_t Nx, int32_t Nr, int32_t M) {
  size_t Nk = (size_t)(Nr - 1);
  return chm_align(Nk * 2 * Nk * 8) + 2 * chm_align(Nk * Nk * 8) +
         chm_align(chm_dht_table_elems((int32_t)Nk, (int32_t)Nk) * 8);
}

extern "C" int chm_build_tables_f64(int32_t Nx, int32_t Nr, int32_t M, double dx, double dr, double dt,
                                    double xmin, int32_t damp_cells, const chm_tables_out_t *out, void *ws,
                                    size_t ws_bytes, void *stream) {
  if (!out || Nx < 2 || Nr < 3 || M < 0 || M > CHM_MAXM) return CHM_EINVAL;
  if (chm_tables_workspace_bytes(Nx, Nr, M) > ws_bytes || !ws) return CHM_EWORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int Nk = Nr - 1;
  const double rmin = -0.5 * dr;
  const double R_period = (rmin + dr * (double)(Nr - 1)) + 0.5 * dr;
  char *b = (char *)ws;
  double *aug = (double *)b;
  double *Jp = (double *)(b + chm_align((size_t)Nk * 2 * Nk * 8));
  double *Jm = (double *)((char *)Jp + chm_align((size_t)Nk * Nk * 8));
  double *Dpad = (double *)((char *)Jm + chm_align((size_t)Nk * Nk * 8));
  k_kx<<<chm_blocks(Nx, 256), 256, 0, st>>>(out->kx, Nx, dx);
  k_roots<<<chm_blocks((M + 2) * Nk, 64), 64, 0, st>>>(out->kr, M + 2, Nk, R_period);
  int64_t nn = (int64_t)Nk * Nk;
  for (int m = 0; m <= M; m++) {
    const double *kr0 = out->kr + (size_t)m * Nk;
    const double *krp = out->kr + (size_t)(m + 1) * Nk;
    const double *krm = out->kr + (size_t)(m > 0 ? m - 1 : 1) * Nk;
    k_jmat<<<chm_blocks(nn, 256), 256, 0, st>>>(out->DHT_inv[m], kr0, m, Nk, rmin, dr, 0);
    k_gauss_jordan<<<1, 1024, 0, st>>>(out->DHT_inv[m], out->DHT[m], aug, Nk);
    k_jmat<<<chm_blocks(nn, 256), 256, 0, st>>>(Jp, krp, m, Nk, rmin, dr, 1);
    k_jmat<<<chm_blocks(nn, 256), 256, 0, st>>>(Jm, krm, m, Nk, rmin, dr, 1);
    int rc = chm_pad_table(out->DHT[m], Nk, Nk, Dpad, nullptr, nullptr, nullptr, stream);
    if (rc) return rc;
    rc = chm_gemm_f64(Dpad, Jp, out->dDp[m], Nk, Nk, Nk, stream);
    if (rc) return rc;
    rc = chm_gemm_f64(Dpad, Jm, out->dDm[m], Nk, Nk, Nk, stream);
    if (rc) return rc;
    k_wtables<<<chm_blocks((int64_t)Nk * Nx, 256), 256, 0, st>>>(out->kx, kr0, Nx, Nk, dt, out->w[m], out->c1[m],
                                                                 out->c2[m], out->c3[m], out->smooth[m]);
  }
  k_phase_tw<<<chm_blocks(Nx, 256), 256, 0, st>>>(out->kx, Nx, xmin, (double2 *)out->phase_minus,
                                                  (double2 *)out->phase_plus, (double2 *)out->twiddle);
  if (damp_cells > 0 && out->damp)
    k_damp_profile<<<chm_blocks(2 * damp_cells, 256), 256, 0, st>>>(out->damp, damp_cells);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// ---------------------------------------------------------------- conversions
__global__ void k_convert(const double *src, int64_t n, float *f32, uint16_t *hi, uint16_t *lo) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = src[i];
  if (f32) f32[i] = (float)v;
  if (hi) {
    __bf16 h = (__bf16)(float)v;
    __bf16 l = (__bf16)(float)(v - (double)(float)h);
    hi[i] = __builtin_bit_cast(uint16_t, h);
    lo[i] = __builtin_bit_cast(uint16_t, l);
  }
}

extern "C" int chm_convert_table(const double *src, int64_t n, float *f32, uint16_t *bf16_hi,
                                 uint16_t *bf16_lo, void *stream) {
  if (!src || n < 0 || ((bf16_hi == nullptr) != (bf16_lo == nullptr))) return CHM_EINVAL;
  if (n == 0) return CHM_OK;
  k_convert<<<chm_blocks(n, 256), 256, 0, (hipStream_t)stream>>>(src, n, f32, bf16_hi, bf16_lo);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// ---------------------------------------------------------------- frame
// get_phase_minus / get_phase_plus (transformer_generic.cl:28-55) for a new
// window origin xmin (moving window, frame.py:22-30).
__global__ void k_phases(const double *kx, int Nx, double xmin, double2 *pm, double2 *pp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Nx) return;
  double a = xmin * kx[i];
  double c = cos(a), s = sin(a);
  pm[i] = make_double2(c, -s);
  pp[i] = make_double2(c, s);
}

extern "C" int chm_phase_tables_f64(const double *kx, int32_t Nx, double xmin, void *phase_minus,
                                    void *phase_plus, void *stream) {
  if (!kx || !phase_minus || !phase_plus || Nx < 1) return CHM_EINVAL;
  k_phases<<<chm_blocks(Nx, 256), 256, 0, (hipStream_t)stream>>>(kx, Nx, xmin, (double2 *)phase_minus,
                                                                 (double2 *)phase_plus);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}
