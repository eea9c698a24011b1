// Spectral side of the step: Fourier-Bessel transforms (DHT on MFMA + batched
// row FFT), spectral differential operators, PSATD advance and edge damping.
//
// Transforms (transform_field, transformer_methods_cl.py:38-64, 290-455):
//   forward  F_fb[m] = phase_minus * FFT_x( DHT_m @ F_m[1:] )   (* smooth)
//   backward F_m[1:] = DHT_inv_m @ Re?( IFFT_x( phase_plus * F_fb[m] ) )
//   ('Re' for m = 0 only; 'half' mode drops the DHT product).
// The reference's copy/cast/phase launches are folded into the FFT kernel's
// load and store; the DHT products for all fields of one mode run as one
// batched MFMA launch.
#include "fft.h"
#include "gemm.h"

#define FFT_NT 256
#define MAXP 32

template <typename T> struct FftPlane {
  const T *in;
  T *out;
  const T *table; // optional real [nrows][N] post-multiplier
  int in_cplx, out_cplx;
};
template <typename T> struct FftBatch {
  FftPlane<T> p[MAXP];
  int nrows;
  const T *phase_in;  // optional complex [N] pre-multiplier
  const T *phase_out; // optional complex [N] post-multiplier
  const T *tw;        // complex [N]: exp(-2 pi i t / N)
  T scale;
};

template <int N> struct FftCfg {
  static constexpr int RPB = (2048 / N) > 0 ? (2048 / N) : 1;
};

template <typename T, int N, int SIGN>
__global__ void __launch_bounds__(FFT_NT) k_fft_rows(FftBatch<T> fb) {
  typedef typename C2<T>::t T2;
  constexpr int RPB = FftCfg<N>::RPB;
  __shared__ T2 buf[RPB * N];
  const FftPlane<T> &pl = fb.p[blockIdx.y];
  const int row0 = blockIdx.x * RPB;
  const T2 *phin = (const T2 *)fb.phase_in, *phout = (const T2 *)fb.phase_out;
  for (int e = threadIdx.x; e < RPB * N; e += FFT_NT) {
    int row = row0 + e / N, x = e % N;
    T2 v;
    v.x = 0;
    v.y = 0;
    if (row < fb.nrows) {
      if (pl.in_cplx) v = ((const T2 *)pl.in)[(int64_t)row * N + x];
      else v.x = pl.in[(int64_t)row * N + x];
      if (phin) v = c_mul(v, phin[x]);
    }
    buf[e] = v;
  }
  __syncthreads();
  fft_lds<T, N, SIGN, RPB, FFT_NT>(buf, (const T2 *)fb.tw);
  for (int e = threadIdx.x; e < RPB * N; e += FFT_NT) {
    int row = row0 + e / N, x = e % N;
    if (row >= fb.nrows) continue;
    T2 v = buf[e];
    v.x *= fb.scale;
    v.y *= fb.scale;
    if (phout) v = c_mul(v, phout[x]);
    if (pl.table) {
      T t = pl.table[(int64_t)row * N + x];
      v.x = t * v.x;
      v.y = t * v.y;
    }
    if (pl.out_cplx) ((T2 *)pl.out)[(int64_t)row * N + x] = v;
    else pl.out[(int64_t)row * N + x] = v.x;
  }
}

// Rows of 4096: the register-resident FFT (fft4096_reg), one row per block
// of 256 threads, thread t handling elements t + 256 j; same loads, phases,
// scale, table and stores as k_fft_rows.  VAR >= 0 fixes the plane's form at
// compile time (bits: 1 complex input, 2 complex output, 4 phase on load,
// 8 phase on store, 16 table), so that only that form's registers are
// allocated: in FP64 the form-generic kernel (VAR = -1) holds 200 VGPRs (2
// waves per SIMD), the backward transform's forms 128 and the forward ones
// 92 (4 waves, the LDS limit).
template <typename T, int SIGN, int VAR>
__global__ void __launch_bounds__(256) k_fft_rows4096(FftBatch<T> fb) {
  typedef typename C2<T>::t T2;
  constexpr int N = 4096;
  constexpr bool GEN = VAR < 0;
  __shared__ T sbuf[FFT4096_LDS];
  const FftPlane<T> &pl = fb.p[blockIdx.y];
  const int row = blockIdx.x;
  const int64_t off = (int64_t)row * N;
  const T2 *phin = (const T2 *)fb.phase_in, *phout = (const T2 *)fb.phase_out;
  const bool in_c = GEN ? pl.in_cplx != 0 : (VAR & 1) != 0;
  const bool out_c = GEN ? pl.out_cplx != 0 : (VAR & 2) != 0;
  const bool use_pin = GEN ? phin != nullptr : (VAR & 4) != 0;
  const bool use_pout = GEN ? phout != nullptr : (VAR & 8) != 0;
  const bool use_tab = GEN ? pl.table != nullptr : (VAR & 16) != 0;
  T2 v[16];
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int x = threadIdx.x + 256 * j;
    T2 a;
    a.x = 0;
    a.y = 0;
    if (in_c) a = ((const T2 *)pl.in)[off + x];
    else a.x = pl.in[off + x];
    if (use_pin) a = c_mul(a, phin[x]);
    v[j] = a;
  }
  fft4096_reg<SIGN, T>(v, sbuf, (const T2 *)fb.tw);
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int x = threadIdx.x + 256 * j;
    T2 a = v[j];
    a.x *= fb.scale;
    a.y *= fb.scale;
    if (use_pout) a = c_mul(a, phout[x]);
    if (use_tab) {
      T tb = pl.table[off + x];
      a.x = tb * a.x;
      a.y = tb * a.y;
    }
    if (out_c) ((T2 *)pl.out)[off + x] = a;
    else pl.out[off + x] = a.x;
  }
}

// One launch of the 4096-point kernel for planes of one form (VAR bits as
// above); the forms the transforms use have their own kernels, others take
// the form-generic one.
template <typename T, int SIGN>
static void launch_fft4096(const FftBatch<T> &fb, int nplanes, int var, hipStream_t st) {
  const dim3 g(fb.nrows, nplanes);
  if constexpr (SIGN < 0) {
    switch (var) {
    case 2 | 8: k_fft_rows4096<T, SIGN, 2 | 8><<<g, 256, 0, st>>>(fb); return;
    case 1 | 2 | 8: k_fft_rows4096<T, SIGN, 1 | 2 | 8><<<g, 256, 0, st>>>(fb); return;
    case 2 | 8 | 16: k_fft_rows4096<T, SIGN, 2 | 8 | 16><<<g, 256, 0, st>>>(fb); return;
    case 1 | 2 | 8 | 16: k_fft_rows4096<T, SIGN, 1 | 2 | 8 | 16><<<g, 256, 0, st>>>(fb); return;
    default: break;
    }
  } else {
    switch (var) {
    case 1 | 4: k_fft_rows4096<T, SIGN, 1 | 4><<<g, 256, 0, st>>>(fb); return;
    case 1 | 2 | 4: k_fft_rows4096<T, SIGN, 1 | 2 | 4><<<g, 256, 0, st>>>(fb); return;
    default: break;
    }
  }
  k_fft_rows4096<T, SIGN, -1><<<g, 256, 0, st>>>(fb);
}

template <typename T, int SIGN>
static int launch_fft(const FftBatch<T> &fb, int nplanes, int N, hipStream_t st) {
  if (nplanes <= 0) return CHM_OK;
  switch (N) {
#define CASE(NN)                                                                              \
  case NN: {                                                                                  \
    dim3 g((fb.nrows + FftCfg<NN>::RPB - 1) / FftCfg<NN>::RPB, nplanes);                      \
    k_fft_rows<T, NN, SIGN><<<g, FFT_NT, 0, st>>>(fb);                                        \
    break;                                                                                    \
  }
    CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048)
#undef CASE
  case 4096: {
    // planes grouped by form, one launch per form
    const int pv = (fb.phase_in ? 4 : 0) | (fb.phase_out ? 8 : 0);
    auto form = [&](int i) {
      return (fb.p[i].in_cplx ? 1 : 0) | (fb.p[i].out_cplx ? 2 : 0) | pv | (fb.p[i].table ? 16 : 0);
    };
    bool done[MAXP] = {};
    for (int i = 0; i < nplanes; i++) {
      if (done[i]) continue;
      const int var = form(i);
      FftBatch<T> sub = fb;
      int n = 0;
      for (int j = i; j < nplanes; j++)
        if (!done[j] && form(j) == var) {
          sub.p[n++] = fb.p[j];
          done[j] = true;
        }
      launch_fft4096<T, SIGN>(sub, n, var, st);
    }
    break;
  }
  default:
    return CHM_EUNSUPPORTED;
  }
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

static bool fft_size_ok(int N) { return N >= 16 && N <= 4096 && (N & (N - 1)) == 0; }

// ---------------------------------------------------------------- helpers
static int gemm_kind(const chm_spectral_t *sp, int elem) {
  if (elem == 8) return GEMM_F64;
  return sp->dht_kind == CHM_DHT_BF16X3 ? GEMM_BF16X3 : GEMM_F32;
}

template <typename T> static int elem_of() { return (int)sizeof(T); }

// ---------------------------------------------------------------- transforms
extern "C" size_t chm_transform_workspace_bytes(int32_t Nx, int32_t Nr, int32_t M, int32_t nflds,
                                                int32_t elem_bytes) {
  return chm_align((size_t)nflds * (M + 1) * (Nr - 1) * Nx * 2 * elem_bytes);
}

template <typename T>
static int transform(const chm_spectral_t *sp, const chm_planes_t *real, const chm_fb_t *fb,
                     int32_t nflds, int32_t dir, int32_t mode, int32_t post_smooth, void *ws,
                     size_t ws_bytes, void *stream) {
  if (!sp || !real || !fb || nflds <= 0 || sp->M < 0 || sp->M > CHM_MAXM) return CHM_EINVAL;
  const int Nx = sp->Nx, Nk = sp->Nr - 1, M = sp->M;
  if (!fft_size_ok(Nx) || Nk < 1) return CHM_EUNSUPPORTED;
  if (nflds * (M + 1) > MAXP) return CHM_EINVAL;
  const bool full = mode == 0;
  size_t plane_elems = (size_t)Nk * Nx * 2;
  if (full && (chm_transform_workspace_bytes(Nx, sp->Nr, M, nflds, sizeof(T)) > ws_bytes || !ws))
    return CHM_EWORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  T *scratch = (T *)ws;
  auto scr = [&](int f, int m) { return scratch + ((size_t)f * (M + 1) + m) * plane_elems; };
  auto rows1 = [&](int f, int m) -> T * {  // row 1 of the real-space plane
    if (m == 0) return (T *)real[f].m0 + Nx;
    return (T *)real[f].mc[m - 1] + 2 * Nx;
  };
  const int kind = gemm_kind(sp, sizeof(T));
  int rc;
  FftBatch<T> b = {};
  b.nrows = Nk;
  b.tw = (const T *)sp->twiddle;
  if (dir == 0) {
    if (full) {
      for (int m = 0; m <= M; m++) {
        GemmBatch gb = {};
        gb.M = Nk;
        gb.K = Nk;
        gb.N = m == 0 ? Nx : 2 * Nx;
        gb.lda = gemm_lda(Nk);
        gb.ldb = gb.N;
        gb.ldc = gb.N;
        gb.bmode = 0;
        for (int f = 0; f < nflds; f++) {
          gb.it[f].A = sp->DHT[m];
          gb.it[f].A_lo = sp->DHT_lo[m];
          gb.it[f].X = rows1(f, m);
          gb.it[f].C = scr(f, m);
        }
        if ((rc = chm_gemm_launch(gb, nflds, kind, st))) return rc;
      }
    }
    int np = 0;
    for (int f = 0; f < nflds; f++)
      for (int m = 0; m <= M; m++) {
        FftPlane<T> &p = b.p[np++];
        p.in = full ? scr(f, m) : rows1(f, m);
        p.in_cplx = m > 0;
        p.out = (T *)fb[f].m[m];
        p.out_cplx = 1;
        p.table = post_smooth ? (const T *)sp->smooth[m] : nullptr;
      }
    b.phase_out = (const T *)sp->phase_minus;
    b.scale = (T)1;
    return launch_fft<T, -1>(b, np, Nx, st);
  }
  // dir 1
  int np = 0;
  for (int f = 0; f < nflds; f++)
    for (int m = 0; m <= M; m++) {
      FftPlane<T> &p = b.p[np++];
      p.in = (const T *)fb[f].m[m];
      p.in_cplx = 1;
      p.out = full ? scr(f, m) : rows1(f, m);
      p.out_cplx = m > 0;
      p.table = nullptr;
    }
  b.phase_in = (const T *)sp->phase_plus;
  b.scale = (T)1 / (T)Nx;
  if ((rc = launch_fft<T, 1>(b, np, Nx, st))) return rc;
  if (full) {
    for (int m = 0; m <= M; m++) {
      GemmBatch gb = {};
      gb.M = Nk;
      gb.K = Nk;
      gb.N = m == 0 ? Nx : 2 * Nx;
      gb.lda = gemm_lda(Nk);
      gb.ldb = gb.N;
      gb.ldc = gb.N;
      gb.bmode = 0;
      for (int f = 0; f < nflds; f++) {
        gb.it[f].A = sp->DHT_inv[m];
        gb.it[f].A_lo = sp->DHT_inv_lo[m];
        gb.it[f].X = scr(f, m);
        gb.it[f].C = rows1(f, m);
      }
      if ((rc = chm_gemm_launch(gb, nflds, kind, st))) return rc;
    }
  }
  return CHM_OK;
}

extern "C" int chm_transform_f64(const chm_spectral_t *sp, const chm_planes_t *real,
                                 const chm_fb_t *fb, int32_t nflds, int32_t dir, int32_t mode,
                                 int32_t post_smooth, void *ws, size_t ws_bytes, void *stream) {
  return transform<double>(sp, real, fb, nflds, dir, mode, post_smooth, ws, ws_bytes, stream);
}
extern "C" int chm_transform_f32(const chm_spectral_t *sp, const chm_planes_t *real,
                                 const chm_fb_t *fb, int32_t nflds, int32_t dir, int32_t mode,
                                 int32_t post_smooth, void *ws, size_t ws_bytes, void *stream) {
  return transform<float>(sp, real, fb, nflds, dir, mode, post_smooth, ws, ws_bytes, stream);
}

// ---------------------------------------------------------------- table multiply
struct TabPlanes {
  void *p[MAXP];
  const void *t[MAXP];
};

template <typename T> __global__ void k_mult_table(TabPlanes tp, int64_t n) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  T *z = (T *)tp.p[blockIdx.y];
  T x = ((const T *)tp.t[blockIdx.y])[e];
  // mult_elementwise_d2c, generic.cl:47-58
  z[2 * e] = x * z[2 * e];
  z[2 * e + 1] = x * z[2 * e + 1];
}

template <typename T>
static int mult_table(const chm_spectral_t *sp, const chm_fb_t *flds, int32_t nflds, int32_t which,
                      void *stream) {
  if (!sp || !flds || nflds <= 0 || nflds * (sp->M + 1) > MAXP) return CHM_EINVAL;
  TabPlanes tp = {};
  int np = 0;
  for (int m = 0; m <= sp->M; m++)
    for (int f = 0; f < nflds; f++) {
      tp.p[np] = flds[f].m[m];
      tp.t[np++] = which == 0 ? sp->smooth[m] : sp->c3[m];
    }
  int64_t n = (int64_t)(sp->Nr - 1) * sp->Nx;
  dim3 g(chm_blocks(n, 256), np);
  k_mult_table<T><<<g, 256, 0, (hipStream_t)stream>>>(tp, n);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_mult_table_f64(const chm_spectral_t *sp, const chm_fb_t *flds, int32_t nflds,
                                  int32_t which, void *stream) {
  return mult_table<double>(sp, flds, nflds, which, stream);
}
extern "C" int chm_mult_table_f32(const chm_spectral_t *sp, const chm_fb_t *flds, int32_t nflds,
                                  int32_t which, void *stream) {
  return mult_table<float>(sp, flds, nflds, which, stream);
}

// ---------------------------------------------------------------- grad / div / rot
// Scratch: up to 4 complex [Nk][Nx] planes per mode.
extern "C" size_t chm_spectral_workspace_bytes(int32_t Nx, int32_t Nr, int32_t M, int32_t elem_bytes) {
  return chm_align((size_t)4 * (M + 1) * (Nr - 1) * Nx * 2 * elem_bytes);
}

struct OpPlanes {
  // per mode m: inputs / outputs / products
  const void *in[3][CHM_MAXM + 1];   // op inputs at mode m (comp 0..2)
  void *out[3][CHM_MAXM + 1];        // outputs
  void *prev[3][CHM_MAXM + 1];       // grad: old output copy target (nullable)
  const void *P[4][CHM_MAXM + 1];    // products (nullable)
  const void *tab[CHM_MAXM + 1];     // rot: Poisson table (nullable)
};

// op 0 grad: in[0] = scl; out x,y,z; P0 = D- S_{m-1}, P1 = D+ S_{m+1}
// op 1 div:  in = V x,y,z; out[0]; P0 = D-(-iVz - Vy)_{m-1}, P1 = D+(-iVz + Vy)_{m+1}
// op 2 rot:  in = G x,y,z; out x,y,z; P0 = D-(-Gz + iGy), P1 = D- Gx, P2 = D+(Gz + iGy), P3 = D+ Gx
template <typename T, int OP>
__global__ void k_assemble(OpPlanes op, const T *__restrict__ kx, int Nx, int64_t n) {
  typedef typename C2<T>::t T2;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (e >= n) return;
  int ix = (int)(e % Nx);
  T k = kx[ix];
  auto ld = [&](const void *p) -> T2 {
    T2 z;
    z.x = 0;
    z.y = 0;
    if (p) z = ((const T2 *)p)[e];
    return z;
  };
  auto st = [&](void *p, T2 v) { ((T2 *)p)[e] = v; };
  // ab_dot_x(a, kx, x): z = kx * (a*x) (generic.cl:82-99)
  auto kdot = [&](T ar, T ai, T2 x) {
    T2 z;
    z.x = k * (ar * x.x - ai * x.y);
    z.y = k * (ar * x.y + ai * x.x);
    return z;
  };
  if (OP == 0) {
    if (op.prev[0][m])
      for (int c = 0; c < 3; c++) st(op.prev[c][m], ld(op.out[c][m]));
    T2 s = ld(op.in[0][m]);
    T2 P = ld(op.P[0][m]), Q = ld(op.P[1][m]);
    T2 x = kdot(0, 1, s);
    T2 y, z;
    y.x = -P.x + Q.x;
    y.y = -P.y + Q.y;
    // -i*P - i*Q
    z.x = P.y + Q.y;
    z.y = -P.x - Q.x;
    st(op.out[0][m], x);
    st(op.out[1][m], y);
    st(op.out[2][m], z);
  } else if (OP == 1) {
    T2 vx = ld(op.in[0][m]);
    T2 P = ld(op.P[0][m]), Q = ld(op.P[1][m]);
    T2 o = kdot(0, 1, vx);
    o.x = o.x + P.x + Q.x;
    o.y = o.y + P.y + Q.y;
    st(op.out[0][m], o);
  } else {
    T2 gy = ld(op.in[1][m]), gz = ld(op.in[2][m]);
    T2 P1 = ld(op.P[0][m]), P2 = ld(op.P[1][m]), Q1 = ld(op.P[2][m]), Q2 = ld(op.P[3][m]);
    T2 bx, by, bz;
    bx.x = P1.x + Q1.x;
    bx.y = P1.y + Q1.y;
    by = kdot(0, -1, gz);
    // -i*(P2 + Q2)
    by.x += P2.y + Q2.y;
    by.y += -P2.x - Q2.x;
    bz = kdot(0, 1, gy);
    bz.x += P2.x - Q2.x;
    bz.y += P2.y - Q2.y;
    if (op.tab[m]) {
      T t = ((const T *)op.tab[m])[e];
      bx.x *= t; bx.y *= t;
      by.x *= t; by.y *= t;
      bz.x *= t; bz.y *= t;
    }
    st(op.out[0][m], bx);
    st(op.out[1][m], by);
    st(op.out[2][m], bz);
  }
}

template <typename T>
static int spectral_op(int OP, const chm_spectral_t *sp, const chm_fb_t *in, const chm_fb_t *out,
                       const chm_fb_t *prev, int32_t poisson, void *ws, size_t ws_bytes,
                       void *stream) {
  if (!sp || !in || !out || sp->M < 0 || sp->M > CHM_MAXM) return CHM_EINVAL;
  const int Nx = sp->Nx, Nk = sp->Nr - 1, M = sp->M;
  if (chm_spectral_workspace_bytes(Nx, sp->Nr, M, sizeof(T)) > ws_bytes || !ws) return CHM_EWORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const size_t pe = (size_t)Nk * Nx * 2;
  T *scr = (T *)ws;
  OpPlanes op = {};
  GemmBatch gb = {};
  gb.M = Nk;
  gb.K = Nk;
  gb.N = 2 * Nx;
  gb.Nx = Nx;
  gb.lda = gemm_lda(Nk);
  gb.ldc = 2 * Nx;
  gb.bmode = 1;
  int ni = 0, ns = 0;
  const int kind = gemm_kind(sp, sizeof(T));
  int rc;
  auto item = [&](const void *A, const void *Alo, const void *X, double2 a, int rx, const void *Y,
                  double2 b, int ry) -> const void * {
    if (ni == GEMM_MAXB) {
      if ((rc = chm_gemm_launch(gb, ni, kind, st))) return nullptr;
      ni = 0;
    }
    GemmItem &it = gb.it[ni++];
    it.A = A;
    it.A_lo = Alo;
    it.X = X;
    it.Y = Y;
    it.alpha = a;
    it.beta = b;
    it.remapX = rx;
    it.remapY = ry;
    it.C = scr + (size_t)(ns++) * pe;
    return it.C;
  };
  const double2 one = {1, 0}, mone = {-1, 0}, I = {0, 1}, mI = {0, -1}, zero = {0, 0};
  const int ncomp_in = (OP == 0) ? 1 : 3;
  for (int m = 0; m <= M; m++)
    for (int c = 0; c < ncomp_in; c++) op.in[c][m] = in[c].m[m];
  for (int m = 0; m <= M; m++)
    for (int c = 0; c < (OP == 1 ? 1 : 3); c++) {
      op.out[c][m] = out[c].m[m];
      if (prev) op.prev[c][m] = prev[c].m[m];
    }
  rc = CHM_OK;
  for (int m = 0; m <= M; m++) {
    const bool lower = m > 0 || M > 0;  // an m-1 neighbour exists (m = 0 uses F_{-1})
    const int mm = m > 0 ? m - 1 : 1;   // source mode of the m-1 term
    const int rm = m > 0 ? 0 : 1;       // read it through the F_{-1} map
    if (OP == 0) {
      if (lower) op.P[0][m] = item(sp->dDm[m], sp->dDm_lo[m], in[0].m[mm], one, rm, nullptr, zero, 0);
      if (m < M) op.P[1][m] = item(sp->dDp[m], sp->dDp_lo[m], in[0].m[m + 1], one, 0, nullptr, zero, 0);
    } else if (OP == 1) {
      if (lower) op.P[0][m] = item(sp->dDm[m], sp->dDm_lo[m], in[2].m[mm], mI, rm, in[1].m[mm], mone, rm);
      if (m < M) op.P[1][m] = item(sp->dDp[m], sp->dDp_lo[m], in[2].m[m + 1], mI, 0, in[1].m[m + 1], one, 0);
    } else {
      if (lower) {
        op.P[0][m] = item(sp->dDm[m], sp->dDm_lo[m], in[2].m[mm], mone, rm, in[1].m[mm], I, rm);
        op.P[1][m] = item(sp->dDm[m], sp->dDm_lo[m], in[0].m[mm], one, rm, nullptr, zero, 0);
      }
      if (m < M) {
        op.P[2][m] = item(sp->dDp[m], sp->dDp_lo[m], in[2].m[m + 1], one, 0, in[1].m[m + 1], I, 0);
        op.P[3][m] = item(sp->dDp[m], sp->dDp_lo[m], in[0].m[m + 1], one, 0, nullptr, zero, 0);
      }
      if (poisson) op.tab[m] = sp->c3[m];
    }
    if (rc) return rc;
  }
  if ((rc = chm_gemm_launch(gb, ni, kind, st))) return rc;
  int64_t n = (int64_t)Nk * Nx;
  dim3 g(chm_blocks(n, 256), M + 1);
  if (OP == 0) k_assemble<T, 0><<<g, 256, 0, st>>>(op, (const T *)sp->kx, Nx, n);
  else if (OP == 1) k_assemble<T, 1><<<g, 256, 0, st>>>(op, (const T *)sp->kx, Nx, n);
  else k_assemble<T, 2><<<g, 256, 0, st>>>(op, (const T *)sp->kx, Nx, n);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_field_grad_f64(const chm_spectral_t *sp, const chm_fb_t *scl, const chm_fb_t *out,
                                  const chm_fb_t *prev, void *ws, size_t ws_bytes, void *stream) {
  return spectral_op<double>(0, sp, scl, out, prev, 0, ws, ws_bytes, stream);
}
extern "C" int chm_field_grad_f32(const chm_spectral_t *sp, const chm_fb_t *scl, const chm_fb_t *out,
                                  const chm_fb_t *prev, void *ws, size_t ws_bytes, void *stream) {
  return spectral_op<float>(0, sp, scl, out, prev, 0, ws, ws_bytes, stream);
}
extern "C" int chm_field_div_f64(const chm_spectral_t *sp, const chm_fb_t *vec, const chm_fb_t *out,
                                 void *ws, size_t ws_bytes, void *stream) {
  return spectral_op<double>(1, sp, vec, out, nullptr, 0, ws, ws_bytes, stream);
}
extern "C" int chm_field_div_f32(const chm_spectral_t *sp, const chm_fb_t *vec, const chm_fb_t *out,
                                 void *ws, size_t ws_bytes, void *stream) {
  return spectral_op<float>(1, sp, vec, out, nullptr, 0, ws, ws_bytes, stream);
}
extern "C" int chm_field_rot_f64(const chm_spectral_t *sp, const chm_fb_t *in, const chm_fb_t *out,
                                 int32_t poisson, void *ws, size_t ws_bytes, void *stream) {
  return spectral_op<double>(2, sp, in, out, nullptr, poisson, ws, ws_bytes, stream);
}
extern "C" int chm_field_rot_f32(const chm_spectral_t *sp, const chm_fb_t *in, const chm_fb_t *out,
                                 int32_t poisson, void *ws, size_t ws_bytes, void *stream) {
  return spectral_op<float>(2, sp, in, out, nullptr, poisson, ws, ws_bytes, stream);
}

// ---------------------------------------------------------------- PSATD
struct SolvPlanes {
  void *e[3][CHM_MAXM + 1], *g[3][CHM_MAXM + 1];
  const void *j[3][CHM_MAXM + 1], *n0[3][CHM_MAXM + 1], *n1[3][CHM_MAXM + 1];
  const void *c1[CHM_MAXM + 1], *c2[CHM_MAXM + 1], *c3[CHM_MAXM + 1];
  int c64;  // the c tables are double (always for T = double; optional for float)
};

// advance_e_g_m, solver_ms_pic.cl:187-213 (same expression order)
template <typename T>
__device__ __forceinline__ void psatd(T c1, T c2, T c3, T dt_inv, T e0, T g0, T j0, T n0, T n1,
                                      T &e1, T &g1) {
  const T pi2 = (T)(2 * 3.14159265358979323846);
  j0 *= pi2;
  n0 *= pi2;
  n1 *= pi2;
  e1 = c1 * e0 + c2 * c3 * (g0 - j0) + c3 * (c1 * n0 - n1 - (n0 - n1) * dt_inv * c2 * c3);
  g1 = -c2 * e0 + c1 * (g0 - j0) + j0 + c3 * (dt_inv * ((T)1. - c1) * (n0 - n1) - c2 * n0);
}

// PSATD of one complex element: with double coefficient tables (FP64, or
// the FP32 form given FP64 tables) the update runs in double and the result
// is rounded to T once.
template <typename T>
__device__ __forceinline__ void psatd_elem(const SolvPlanes &sv, int m, int c, int64_t id, T dt_inv,
                                           typename C2<T>::t e0, typename C2<T>::t g0,
                                           typename C2<T>::t &e1, typename C2<T>::t &g1) {
  typedef typename C2<T>::t T2;
  T2 j0 = ((const T2 *)sv.j[c][m])[id], n0 = ((const T2 *)sv.n0[c][m])[id], n1 = ((const T2 *)sv.n1[c][m])[id];
  if (sizeof(T) == 4 && sv.c64) {
    const double c1 = ((const double *)sv.c1[m])[id], c2 = ((const double *)sv.c2[m])[id],
                 c3 = ((const double *)sv.c3[m])[id];
    double ex, gx, ey, gy;
    psatd<double>(c1, c2, c3, (double)dt_inv, e0.x, g0.x, j0.x, n0.x, n1.x, ex, gx);
    psatd<double>(c1, c2, c3, (double)dt_inv, e0.y, g0.y, j0.y, n0.y, n1.y, ey, gy);
    e1.x = (T)ex; e1.y = (T)ey;
    g1.x = (T)gx; g1.y = (T)gy;
  } else {
    T c1 = ((const T *)sv.c1[m])[id], c2 = ((const T *)sv.c2[m])[id], c3 = ((const T *)sv.c3[m])[id];
    psatd<T>(c1, c2, c3, dt_inv, e0.x, g0.x, j0.x, n0.x, n1.x, e1.x, g1.x);
    psatd<T>(c1, c2, c3, dt_inv, e0.y, g0.y, j0.y, n0.y, n1.y, e1.y, g1.y);
  }
}

template <typename T> __global__ void k_advance(SolvPlanes sv, T dt_inv, int64_t n) {
  typedef typename C2<T>::t T2;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (e >= n) return;
  for (int k = 0; k < 3; k++) {
    T2 e0 = ((T2 *)sv.e[k][m])[e], g0 = ((T2 *)sv.g[k][m])[e];
    T2 e1, g1;
    psatd_elem<T>(sv, m, k, e, dt_inv, e0, g0, e1, g1);
    ((T2 *)sv.e[k][m])[e] = e1;
    ((T2 *)sv.g[k][m])[e] = g1;
  }
}

static void fill_solv(SolvPlanes &sv, const chm_spectral_t *sp, const chm_fb_t *E, const chm_fb_t *G,
                      const chm_fb_t *J, const chm_fb_t *dN0, const chm_fb_t *dN1) {
  for (int m = 0; m <= sp->M; m++) {
    for (int k = 0; k < 3; k++) {
      sv.e[k][m] = E[k].m[m];
      sv.g[k][m] = G[k].m[m];
      if (J) {
        sv.j[k][m] = J[k].m[m];
        sv.n0[k][m] = dN0[k].m[m];
        sv.n1[k][m] = dN1[k].m[m];
      }
    }
    sv.c1[m] = sp->c1[m];
    sv.c2[m] = sp->c2[m];
    sv.c3[m] = sp->c3[m];
  }
  // the FP32 forms take the FP64 coefficient tables when all are given
  bool d = true;
  for (int m = 0; m <= sp->M; m++) d = d && sp->c1d[m] && sp->c2d[m] && sp->c3d[m];
  sv.c64 = 0;
  if (d) {
    for (int m = 0; m <= sp->M; m++) {
      sv.c1[m] = sp->c1d[m];
      sv.c2[m] = sp->c2d[m];
      sv.c3[m] = sp->c3d[m];
    }
    sv.c64 = 1;
  }
}

template <typename T>
static int advance(const chm_spectral_t *sp, const chm_fb_t *E, const chm_fb_t *G, const chm_fb_t *J,
                   const chm_fb_t *dN0, const chm_fb_t *dN1, void *stream) {
  if (!sp || !E || !G || !J || !dN0 || !dN1 || sp->M < 0 || sp->M > CHM_MAXM) return CHM_EINVAL;
  SolvPlanes sv = {};
  fill_solv(sv, sp, E, G, J, dN0, dN1);
  int64_t n = (int64_t)(sp->Nr - 1) * sp->Nx;
  dim3 g(chm_blocks(n, 256), sp->M + 1);
  k_advance<T><<<g, 256, 0, (hipStream_t)stream>>>(sv, (T)sp->dt_inv, n);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_advance_fields_f64(const chm_spectral_t *sp, const chm_fb_t *E, const chm_fb_t *G,
                                      const chm_fb_t *J, const chm_fb_t *dN0, const chm_fb_t *dN1,
                                      void *stream) {
  return advance<double>(sp, E, G, J, dN0, dN1, stream);
}
extern "C" int chm_advance_fields_f32(const chm_spectral_t *sp, const chm_fb_t *E, const chm_fb_t *G,
                                      const chm_fb_t *J, const chm_fb_t *dN0, const chm_fb_t *dN1,
                                      void *stream) {
  return advance<float>(sp, E, G, J, dN0, dN1, stream);
}

// ---------------------------------------------------------------- damping
// One block per (row kr, mode m, component c): optional PSATD of the row,
// then for E and G the damping round trip (Solver.damp_fields,
// solver.py:32-35): *phase_plus -> IFFT -> (m=0: real part) -> *profile(x)
// -> FFT -> *phase_minus.  The forward kernel computes it in correction
// form, row - FFT(v - f g(v)) phase_minus with v = IFFT(row phase_plus)/N,
// g = Re for m = 0 (identity otherwise) and f the profile: the same map in
// exact arithmetic, but the rounding of the transforms and of the phase
// factors then only touches the correction, which vanishes where f = 1
// (away from the edges), instead of the whole row in every step.
// Rows of 4096 run with 512 threads (one radix-8 butterfly per thread and
// pass, eight row elements per thread), shorter rows with 256.
template <int N> struct DampCfg {
  static constexpr int NT = N >= 4096 ? 512 : FFT_NT;
  static constexpr int PER = (N + NT - 1) / NT;
  // FP64 4096-point rows: registers held to 4 waves per SIMD (a few
  // spilled words), two 512-thread blocks per CU: 1080 -> 873 us per call on
  // the 4096 x 512, M = 2 step
  static constexpr int wpe(int tbytes) { return (N >= 4096 && tbytes == 8) ? 4 : 1; }
};

// The damping round trip of the row in buf (already multiplied by
// phase_plus), in place (direct form; used by the adjoint: the map is
// self-adjoint, so the adjoint kernel applies it unchanged).  Call after a
// barrier; returns after one.
template <typename T, int N, int NT>
__device__ __forceinline__ void damp_lds(typename C2<T>::t *buf, int m, const T *__restrict__ damp,
                                         int nd, const T *__restrict__ tw) {
  typedef typename C2<T>::t T2;
  const T inv = (T)1 / (T)N;
  fft_lds<T, N, 1, 1, NT>(buf, (const T2 *)tw);
  for (int x = threadIdx.x; x < N; x += NT) {
    T2 v = buf[x];
    v.x *= inv;
    v.y *= inv;
    if (m == 0) v.y = 0;
    if (x < nd) { v.x *= damp[x]; v.y *= damp[x]; }
    if (x > N - nd) { v.x *= damp[N - x]; v.y *= damp[N - x]; }
    buf[x] = v;
  }
  __syncthreads();
  fft_lds<T, N, -1, 1, NT>(buf, (const T2 *)tw);
}

// The correction of the round trip: buf (row * phase_plus) becomes
// FFT(v - f g(v)), v = IFFT(buf)/N.  Call after a barrier; returns after one.
template <typename T, int N, int NT>
__device__ __forceinline__ void damp_corr_lds(typename C2<T>::t *buf, int m,
                                              const T *__restrict__ damp, int nd,
                                              const T *__restrict__ tw) {
  typedef typename C2<T>::t T2;
  const T inv = (T)1 / (T)N;
  fft_lds<T, N, 1, 1, NT>(buf, (const T2 *)tw);
  for (int x = threadIdx.x; x < N; x += NT) {
    T2 v = buf[x];
    v.x *= inv;
    v.y *= inv;
    T f = (T)1;
    if (x < nd) f *= damp[x];
    if (x > N - nd) f *= damp[N - x];
    T2 r;
    r.x = v.x - f * v.x;
    r.y = m == 0 ? v.y : v.y - f * v.y;
    buf[x] = r;
  }
  __syncthreads();
  fft_lds<T, N, -1, 1, NT>(buf, (const T2 *)tw);
}

template <typename T, int N, bool ADV>
__global__ void __launch_bounds__(DampCfg<N>::NT) __attribute__((amdgpu_waves_per_eu(DampCfg<N>::wpe(sizeof(T))))) k_damp(SolvPlanes sv, const T *__restrict__ damp,
                                                         int nd, const T *__restrict__ php,
                                                         const T *__restrict__ phm,
                                                         const T *__restrict__ tw, T dt_inv, int Nk) {
  typedef typename C2<T>::t T2;
  constexpr int NT = DampCfg<N>::NT, PER = DampCfg<N>::PER;
  __shared__ T2 buf[N];
  const int kr = blockIdx.x, m = blockIdx.y, c = blockIdx.z;
  const int64_t rowoff = (int64_t)kr * N;
  const T2 *pp = (const T2 *)php, *pm = (const T2 *)phm;
  T2 *E = (T2 *)sv.e[c][m] + rowoff, *G = (T2 *)sv.g[c][m] + rowoff;
  if constexpr (sizeof(T) == 4) {
    // FP32: the new E and G of the thread's elements (x = threadIdx.x + q NT)
    // stay in registers until their corrected values are written (no staging
    // writes, no rereads: 560 -> 510 us per call on the 4096 x 512 step)
    T2 es[PER], gs[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int x = threadIdx.x + q * NT;
      if (x < N) {
        T2 e0 = E[x], g0 = G[x];
        if (ADV) {
          T2 e1, g1;
          psatd_elem<T>(sv, m, c, rowoff + x, dt_inv, e0, g0, e1, g1);
          e0 = e1;
          g0 = g1;
        }
        es[q] = e0;
        gs[q] = g0;
        buf[x] = c_mul(e0, pp[x]);
      }
    }
    __syncthreads();
    damp_corr_lds<T, N, NT>(buf, m, damp, nd, tw);
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int x = threadIdx.x + q * NT;
      if (x < N) E[x] = c_sub(es[q], c_mul(buf[x], pm[x]));
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int x = threadIdx.x + q * NT;
      if (x < N) buf[x] = c_mul(gs[q], pp[x]);
    }
    __syncthreads();
    damp_corr_lds<T, N, NT>(buf, m, damp, nd, tw);
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int x = threadIdx.x + q * NT;
      if (x < N) G[x] = c_sub(gs[q], c_mul(buf[x], pm[x]));
    }
    return;
  }
  // FP64: the new E and G go to memory right away (holding both rows in
  // registers leaves one block per CU: 1300 us per call instead of 873);
  // each thread later rereads only the elements it wrote itself
  // (x = threadIdx.x + q NT throughout)
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int x = threadIdx.x + q * NT;
    if (x < N) {
      T2 e0 = E[x];
      if (ADV) {
        T2 e1, g1;
        psatd_elem<T>(sv, m, c, rowoff + x, dt_inv, e0, G[x], e1, g1);
        E[x] = e1;
        G[x] = g1;
        e0 = e1;
      }
      buf[x] = c_mul(e0, pp[x]);
    }
  }
  __syncthreads();
  damp_corr_lds<T, N, NT>(buf, m, damp, nd, tw);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int x = threadIdx.x + q * NT;
    if (x < N) E[x] = c_sub(E[x], c_mul(buf[x], pm[x]));
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int x = threadIdx.x + q * NT;
    if (x < N) buf[x] = c_mul(G[x], pp[x]);
  }
  __syncthreads();
  damp_corr_lds<T, N, NT>(buf, m, damp, nd, tw);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int x = threadIdx.x + q * NT;
    if (x < N) G[x] = c_sub(G[x], c_mul(buf[x], pm[x]));
  }
}

template <typename T>
static int damp_fields(const chm_spectral_t *sp, const chm_fb_t *E, const chm_fb_t *G, int32_t adv,
                       const chm_fb_t *J, const chm_fb_t *dN0, const chm_fb_t *dN1, void *stream) {
  if (!sp || !E || !G || sp->M < 0 || sp->M > CHM_MAXM) return CHM_EINVAL;
  if (adv && (!J || !dN0 || !dN1)) return CHM_EINVAL;
  if (sp->ndamp <= 0 || !sp->damp) {
    if (adv) return advance<T>(sp, E, G, J, dN0, dN1, stream);
    return CHM_OK;
  }
  const int Nx = sp->Nx;
  if (!fft_size_ok(Nx)) return CHM_EUNSUPPORTED;
  SolvPlanes sv = {};
  fill_solv(sv, sp, E, G, adv ? J : nullptr, dN0, dN1);
  dim3 g(sp->Nr - 1, sp->M + 1, 3);
  hipStream_t st = (hipStream_t)stream;
  const T *dp = (const T *)sp->damp, *php = (const T *)sp->phase_plus, *phm = (const T *)sp->phase_minus,
          *tw = (const T *)sp->twiddle;
  T dti = (T)sp->dt_inv;
  switch (Nx) {
#define CASE(NN)                                                                                       \
  case NN:                                                                                             \
    if (adv) k_damp<T, NN, true><<<g, DampCfg<NN>::NT, 0, st>>>(sv, dp, sp->ndamp, php, phm, tw, dti, \
                                                                 sp->Nr - 1);                        \
    else k_damp<T, NN, false><<<g, DampCfg<NN>::NT, 0, st>>>(sv, dp, sp->ndamp, php, phm, tw, dti,   \
                                                              sp->Nr - 1);                           \
    break;
    CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048) CASE(4096)
#undef CASE
  }
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_damp_fields_f64(const chm_spectral_t *sp, const chm_fb_t *E, const chm_fb_t *G,
                                   int32_t advance_first, const chm_fb_t *J, const chm_fb_t *dN0,
                                   const chm_fb_t *dN1, void *stream) {
  return damp_fields<double>(sp, E, G, advance_first, J, dN0, dN1, stream);
}
extern "C" int chm_damp_fields_f32(const chm_spectral_t *sp, const chm_fb_t *E, const chm_fb_t *G,
                                   int32_t advance_first, const chm_fb_t *J, const chm_fb_t *dN0,
                                   const chm_fb_t *dN1, void *stream) {
  return damp_fields<float>(sp, E, G, advance_first, J, dN0, dN1, stream);
}

// ---------------------------------------------------------------- adjoints
// Adjoints of the spectral operators for the differentiable form.  Complex
// adjoints are dL/dRe + i dL/dIm of every stored element (the m = 0 planes'
// entries are treated as independent).  The DHT-type products use a
// chm_spectral_t whose dDp / dDm slots hold the TRANSPOSED tables
// (chm_pad_table of dDHT_plus^T, dDHT_minus^T), so that one product
// A^T (alpha X + beta Y) is a forward product with another table.
//
// For a forward product P = D (alpha X) read through the m = -1 map
// (X -> -conj(X(-kx))), the adjoint is Xbar += D^T (alpha remap(Pbar));
// without the map it is Xbar += D^T (conj(alpha) Pbar).

struct AdjTerm {
  const void *p;
  double cr, ci;
};
struct AdjTarget {
  void *out;
  const void *kxsrc;  // optional: adds (kr + i ki) kx * kxsrc
  double kr, ki;
  int nt;
  AdjTerm t[3];
};
#define ADJ_MAXT (3 * (CHM_MAXM + 1))
struct AdjAssemble {
  AdjTarget tg[ADJ_MAXT];
  int acc;
};

template <typename T>
__global__ void k_adj_assemble(AdjAssemble aa, const T *__restrict__ kx, int Nx, int64_t n) {
  typedef typename C2<T>::t T2;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const AdjTarget &tg = aa.tg[blockIdx.y];
  T2 v;
  v.x = 0;
  v.y = 0;
  if (aa.acc) v = ((const T2 *)tg.out)[e];
  for (int i = 0; i < tg.nt; i++) {
    T2 s = ((const T2 *)tg.t[i].p)[e];
    T cr = (T)tg.t[i].cr, ci = (T)tg.t[i].ci;
    v.x += cr * s.x - ci * s.y;
    v.y += cr * s.y + ci * s.x;
  }
  if (tg.kxsrc) {
    T2 s = ((const T2 *)tg.kxsrc)[e];
    T k = kx[e % Nx];
    T cr = (T)tg.kr * k, ci = (T)tg.ki * k;
    v.x += cr * s.x - ci * s.y;
    v.y += cr * s.y + ci * s.x;
  }
  ((T2 *)tg.out)[e] = v;
}

// OP 0: grad^H  (in = adjoint of (x, y, z) of grad, out = adjoint of the scalar)
// OP 2: rot^H   (in = adjoint of rot's x, y, z; out = adjoint of its input;
//                with poisson, in is first multiplied by 1/w^2 in place)
template <typename T>
static int spectral_op_adj(int OP, const chm_spectral_t *sp, const chm_fb_t *in,
                           const chm_fb_t *out, int32_t poisson, int32_t acc, void *ws,
                           size_t ws_bytes, void *stream) {
  if (!sp || !in || !out || sp->M < 0 || sp->M > CHM_MAXM) return CHM_EINVAL;
  const int Nx = sp->Nx, Nk = sp->Nr - 1, M = sp->M;
  if (chm_spectral_workspace_bytes(Nx, sp->Nr, M, sizeof(T)) > ws_bytes || !ws) return CHM_EWORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = (int64_t)Nk * Nx;
  int rc = CHM_OK;
  if (OP == 2 && poisson) {
    TabPlanes tp = {};
    int np = 0;
    for (int m = 0; m <= M; m++)
      for (int c = 0; c < 3; c++) {
        tp.p[np] = in[c].m[m];
        tp.t[np++] = sp->c3[m];
      }
    k_mult_table<T><<<dim3(chm_blocks(n, 256), np), 256, 0, st>>>(tp, n);
  }
  const size_t pe = (size_t)Nk * Nx * 2;
  T *scr = (T *)ws;
  GemmBatch gb = {};
  gb.M = Nk;
  gb.K = Nk;
  gb.N = 2 * Nx;
  gb.Nx = Nx;
  gb.lda = gemm_lda(Nk);
  gb.ldc = 2 * Nx;
  gb.bmode = 1;
  int ni = 0, ns = 0;
  const int kind = gemm_kind(sp, sizeof(T));
  auto item = [&](const void *A, const void *Alo, const void *X, double2 a, int rx, const void *Y,
                  double2 b, int ry) -> const void * {
    if (ni == GEMM_MAXB) {
      if ((rc = chm_gemm_launch(gb, ni, kind, st))) return nullptr;
      ni = 0;
    }
    GemmItem &it = gb.it[ni++];
    it.A = A;
    it.A_lo = Alo;
    it.X = X;
    it.Y = Y;
    it.alpha = a;
    it.beta = b;
    it.remapX = rx;
    it.remapY = ry;
    it.C = scr + (size_t)(ns++) * pe;
    return it.C;
  };
  const double2 one = {1, 0}, mone = {-1, 0}, I = {0, 1}, mI = {0, -1}, zero = {0, 0};
  AdjAssemble aa = {};
  aa.acc = acc;
  const int ncomp_out = OP == 0 ? 1 : 3;
  for (int s = 0; s <= M; s++)
    for (int c = 0; c < ncomp_out; c++) aa.tg[s * ncomp_out + c].out = out[c].m[s];
  auto add = [&](int s, int c, const void *p, double cr, double ci) {
    AdjTarget &t = aa.tg[s * ncomp_out + c];
    t.t[t.nt].p = p;
    t.t[t.nt].cr = cr;
    t.t[t.nt].ci = ci;
    t.nt++;
  };
  for (int m = 0; m <= M; m++) {
    const bool lower = m > 0 || M > 0;
    if (OP == 0) {
      // grad: y = -P0 + P1, z = -i P0 - i P1, P0 = D-_m S_{m-1}, P1 = D+_m S_{m+1}
      if (lower) {
        const void *q;
        if (m > 0) q = item(sp->dDm[m], sp->dDm_lo[m], in[1].m[m], mone, 0, in[2].m[m], I, 0);
        else q = item(sp->dDm[0], sp->dDm_lo[0], in[1].m[0], mone, 1, in[2].m[0], mI, 1);
        if (rc) return rc;
        add(m > 0 ? m - 1 : 1, 0, q, 1, 0);
      }
      if (m < M) {
        const void *q = item(sp->dDp[m], sp->dDp_lo[m], in[1].m[m], one, 0, in[2].m[m], I, 0);
        if (rc) return rc;
        add(m + 1, 0, q, 1, 0);
      }
    } else {
      // rot: bx = P0 + P2, by = -i (P1 + P3), bz = P1 - P3 (plus kx terms),
      // P0 = D-(-Gz + i Gy)_{m-1}, P1 = D- Gx_{m-1}, P2 = D+(Gz + i Gy)_{m+1}, P3 = D+ Gx_{m+1}
      if (lower) {
        if (m > 0) {
          const void *qa = item(sp->dDm[m], sp->dDm_lo[m], in[0].m[m], one, 0, nullptr, zero, 0);
          const void *qb = item(sp->dDm[m], sp->dDm_lo[m], in[1].m[m], I, 0, in[2].m[m], one, 0);
          if (rc) return rc;
          add(m - 1, 2, qa, -1, 0);
          add(m - 1, 1, qa, 0, -1);
          add(m - 1, 0, qb, 1, 0);
        } else {
          const void *qa = item(sp->dDm[0], sp->dDm_lo[0], in[0].m[0], one, 1, nullptr, zero, 0);
          const void *qb = item(sp->dDm[0], sp->dDm_lo[0], in[1].m[0], mI, 1, in[2].m[0], one, 1);
          if (rc) return rc;
          add(1, 2, qa, -1, 0);
          add(1, 1, qa, 0, 1);
          add(1, 0, qb, 1, 0);
        }
      }
      if (m < M) {
        const void *qc = item(sp->dDp[m], sp->dDp_lo[m], in[0].m[m], one, 0, nullptr, zero, 0);
        const void *qd = item(sp->dDp[m], sp->dDp_lo[m], in[1].m[m], I, 0, in[2].m[m], mone, 0);
        if (rc) return rc;
        add(m + 1, 2, qc, 1, 0);
        add(m + 1, 1, qc, 0, -1);
        add(m + 1, 0, qd, 1, 0);
      }
    }
  }
  if (ni > 0 && (rc = chm_gemm_launch(gb, ni, kind, st))) return rc;
  for (int s = 0; s <= M; s++) {
    if (OP == 0) {
      AdjTarget &t = aa.tg[s];
      t.kxsrc = in[0].m[s];
      t.kr = 0;
      t.ki = -1;
    } else {
      AdjTarget &tz = aa.tg[s * 3 + 2], &ty = aa.tg[s * 3 + 1];
      tz.kxsrc = in[1].m[s];
      tz.kr = 0;
      tz.ki = 1;
      ty.kxsrc = in[2].m[s];
      ty.kr = 0;
      ty.ki = -1;
    }
  }
  dim3 g(chm_blocks(n, 256), (M + 1) * ncomp_out);
  k_adj_assemble<T><<<g, 256, 0, st>>>(aa, (const T *)sp->kx, Nx, n);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_field_grad_adj_f64(const chm_spectral_t *sp, const chm_fb_t *vec_b,
                                      const chm_fb_t *scl_b, int32_t accumulate, void *ws,
                                      size_t ws_bytes, void *stream) {
  return spectral_op_adj<double>(0, sp, vec_b, scl_b, 0, accumulate, ws, ws_bytes, stream);
}
extern "C" int chm_field_grad_adj_f32(const chm_spectral_t *sp, const chm_fb_t *vec_b,
                                      const chm_fb_t *scl_b, int32_t accumulate, void *ws,
                                      size_t ws_bytes, void *stream) {
  return spectral_op_adj<float>(0, sp, vec_b, scl_b, 0, accumulate, ws, ws_bytes, stream);
}
extern "C" int chm_field_rot_adj_f64(const chm_spectral_t *sp, const chm_fb_t *in_b,
                                     const chm_fb_t *out_b, int32_t poisson, int32_t accumulate,
                                     void *ws, size_t ws_bytes, void *stream) {
  return spectral_op_adj<double>(2, sp, in_b, out_b, poisson, accumulate, ws, ws_bytes, stream);
}
extern "C" int chm_field_rot_adj_f32(const chm_spectral_t *sp, const chm_fb_t *in_b,
                                     const chm_fb_t *out_b, int32_t poisson, int32_t accumulate,
                                     void *ws, size_t ws_bytes, void *stream) {
  return spectral_op_adj<float>(2, sp, in_b, out_b, poisson, accumulate, ws, ws_bytes, stream);
}

// Transpose of the PSATD update of one element (real coefficients):
// (E', G') = L (E, G, J, n0, n1) -> (Ebar, Gbar, Jbar, n0bar, n1bar) = L^T (E'bar, G'bar).
template <typename T>
__device__ __forceinline__ void psatd_adj(T c1, T c2, T c3, T dt_inv, T eb1, T gb1, T &eb0, T &gb0,
                                          T &jb, T &n0b, T &n1b) {
  const T pi2 = (T)(2 * 3.14159265358979323846);
  const T c23 = c2 * c3;
  eb0 = c1 * eb1 - c2 * gb1;
  gb0 = c23 * eb1 + c1 * gb1;
  jb = pi2 * (-c23 * eb1 + ((T)1 - c1) * gb1);
  n0b = pi2 * (c3 * (c1 - dt_inv * c23) * eb1 + c3 * (dt_inv * ((T)1 - c1) - c2) * gb1);
  n1b = pi2 * (c3 * (dt_inv * c23 - (T)1) * eb1 - c3 * dt_inv * ((T)1 - c1) * gb1);
}

struct SolvAdj {
  void *e[3][CHM_MAXM + 1], *g[3][CHM_MAXM + 1];
  const void *eadd[3][CHM_MAXM + 1];
  void *j[3][CHM_MAXM + 1], *carry[3][CHM_MAXM + 1], *n1[3][CHM_MAXM + 1];
  const void *c1[CHM_MAXM + 1], *c2[CHM_MAXM + 1], *c3[CHM_MAXM + 1];
};

template <typename T>
__device__ __forceinline__ void psatd_adj_store(const SolvAdj &sa, int c, int m, int64_t id,
                                                typename C2<T>::t eb, typename C2<T>::t gb,
                                                T dt_inv) {
  typedef typename C2<T>::t T2;
  T c1 = ((const T *)sa.c1[m])[id], c2 = ((const T *)sa.c2[m])[id], c3 = ((const T *)sa.c3[m])[id];
  T2 e0, g0, jb, a, b;
  psatd_adj<T>(c1, c2, c3, dt_inv, eb.x, gb.x, e0.x, g0.x, jb.x, a.x, b.x);
  psatd_adj<T>(c1, c2, c3, dt_inv, eb.y, gb.y, e0.y, g0.y, jb.y, a.y, b.y);
  T2 cin = ((T2 *)sa.carry[c][m])[id];
  b.x += cin.x;
  b.y += cin.y;
  ((T2 *)sa.e[c][m])[id] = e0;
  ((T2 *)sa.g[c][m])[id] = g0;
  ((T2 *)sa.j[c][m])[id] = jb;
  ((T2 *)sa.n1[c][m])[id] = b;
  ((T2 *)sa.carry[c][m])[id] = a;
}

template <typename T, int N>
__global__ void __launch_bounds__(DampCfg<N>::NT) k_damp_adj(SolvAdj sa, const T *__restrict__ damp,
                                                             int nd, const T *__restrict__ php,
                                                             const T *__restrict__ phm,
                                                             const T *__restrict__ tw, T dt_inv) {
  typedef typename C2<T>::t T2;
  constexpr int NT = DampCfg<N>::NT, PER = DampCfg<N>::PER;
  __shared__ T2 buf[N];
  const int kr = blockIdx.x, m = blockIdx.y, c = blockIdx.z;
  const int64_t rowoff = (int64_t)kr * N;
  const T2 *pp = (const T2 *)php, *pm = (const T2 *)phm;
  T2 ev[PER], gv[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int x = threadIdx.x + q * NT;
    if (x < N) {
      T2 e = ((T2 *)sa.e[c][m])[rowoff + x];
      if (sa.eadd[c][m]) {
        T2 a = ((const T2 *)sa.eadd[c][m])[rowoff + x];
        e.x += a.x;
        e.y += a.y;
      }
      buf[x] = c_mul(e, pp[x]);
      gv[q] = ((T2 *)sa.g[c][m])[rowoff + x];
    }
  }
  __syncthreads();
  damp_lds<T, N, NT>(buf, m, damp, nd, tw);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int x = threadIdx.x + q * NT;
    if (x < N) ev[q] = c_mul(buf[x], pm[x]);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int x = threadIdx.x + q * NT;
    if (x < N) buf[x] = c_mul(gv[q], pp[x]);
  }
  __syncthreads();
  damp_lds<T, N, NT>(buf, m, damp, nd, tw);
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int x = threadIdx.x + q * NT;
    if (x < N) psatd_adj_store<T>(sa, c, m, rowoff + x, ev[q], c_mul(buf[x], pm[x]), dt_inv);
  }
}

template <typename T> __global__ void k_advance_adj(SolvAdj sa, T dt_inv, int64_t n) {
  typedef typename C2<T>::t T2;
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y, c = blockIdx.z;
  if (e >= n) return;
  T2 eb = ((T2 *)sa.e[c][m])[e];
  if (sa.eadd[c][m]) {
    T2 a = ((const T2 *)sa.eadd[c][m])[e];
    eb.x += a.x;
    eb.y += a.y;
  }
  T2 gb = ((T2 *)sa.g[c][m])[e];
  psatd_adj_store<T>(sa, c, m, e, eb, gb, dt_inv);
}

// Adjoint of damp_fields with the PSATD advance first (the loop's
// push_fields + damp_fields).  In: E_b, G_b = adjoints of the new E, G
// (plus E_add, if given, added to E_b), carry = adjoint of the dN1 this
// step produces as seen by the NEXT step's advance (as its dN0).  Out:
// E_b, G_b = adjoints of the old E, G; J_b; N1_b = total adjoint of this
// step's dN1; carry = adjoint of this step's dN0 (for the previous step).
template <typename T>
static int damp_fields_adj(const chm_spectral_t *sp, const chm_fb_t *E, const chm_fb_t *G,
                           const chm_fb_t *Eadd, const chm_fb_t *J, const chm_fb_t *carry,
                           const chm_fb_t *N1, void *stream) {
  if (!sp || !E || !G || !J || !carry || !N1 || sp->M < 0 || sp->M > CHM_MAXM) return CHM_EINVAL;
  SolvAdj sa = {};
  for (int m = 0; m <= sp->M; m++) {
    for (int k = 0; k < 3; k++) {
      sa.e[k][m] = E[k].m[m];
      sa.g[k][m] = G[k].m[m];
      sa.eadd[k][m] = Eadd ? Eadd[k].m[m] : nullptr;
      sa.j[k][m] = J[k].m[m];
      sa.carry[k][m] = carry[k].m[m];
      sa.n1[k][m] = N1[k].m[m];
    }
    sa.c1[m] = sp->c1[m];
    sa.c2[m] = sp->c2[m];
    sa.c3[m] = sp->c3[m];
  }
  hipStream_t st = (hipStream_t)stream;
  T dti = (T)sp->dt_inv;
  if (sp->ndamp <= 0 || !sp->damp) {
    int64_t n = (int64_t)(sp->Nr - 1) * sp->Nx;
    k_advance_adj<T><<<dim3(chm_blocks(n, 256), sp->M + 1, 3), 256, 0, st>>>(sa, dti, n);
    CHM_CHECK_LAUNCH();
    return CHM_OK;
  }
  const int Nx = sp->Nx;
  if (!fft_size_ok(Nx)) return CHM_EUNSUPPORTED;
  dim3 g(sp->Nr - 1, sp->M + 1, 3);
  const T *dp = (const T *)sp->damp, *php = (const T *)sp->phase_plus, *phm = (const T *)sp->phase_minus,
          *tw = (const T *)sp->twiddle;
  switch (Nx) {
#define CASE(NN)                                                                   \
  case NN:                                                                         \
    k_damp_adj<T, NN><<<g, DampCfg<NN>::NT, 0, st>>>(sa, dp, sp->ndamp, php, phm, tw, dti); \
    break;
    CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(512) CASE(1024) CASE(2048) CASE(4096)
#undef CASE
  }
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

extern "C" int chm_damp_fields_adj_f64(const chm_spectral_t *sp, const chm_fb_t *E_b,
                                       const chm_fb_t *G_b, const chm_fb_t *E_add,
                                       const chm_fb_t *J_b, const chm_fb_t *carry,
                                       const chm_fb_t *N1_b, void *stream) {
  return damp_fields_adj<double>(sp, E_b, G_b, E_add, J_b, carry, N1_b, stream);
}
extern "C" int chm_damp_fields_adj_f32(const chm_spectral_t *sp, const chm_fb_t *E_b,
                                       const chm_fb_t *G_b, const chm_fb_t *E_add,
                                       const chm_fb_t *J_b, const chm_fb_t *carry,
                                       const chm_fb_t *N1_b, void *stream) {
  return damp_fields_adj<float>(sp, E_b, G_b, E_add, J_b, carry, N1_b, stream);
}

// ---------------------------------------------------------------- laser
// laser.py:17-19
__global__ void k_laser_pulse(double *Ez, int Nx, int Nr, double xmin, double dx, double rmin,
                              double dr, double k0, double a0, double x0, double Lx, double R) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)(Nr - 1) * Nx) return;
  int ir = 1 + (int)(e / Nx), ix = (int)(e % Nx);
  double x = xmin + dx * (double)ix;
  double r = rmin + dr * (double)ir;
  double xs = x - x0;
  double v = a0 * sin(k0 * xs) * exp(-xs * xs / (Lx * Lx) - r * r / (R * R));
  if (!(fabs(r) < 3.5 * R) || !(fabs(xs) < 3.5 * Lx)) v = 0.0 * v;
  Ez[(int64_t)ir * Nx + ix] = v;
}

extern "C" int chm_laser_pulse_f64(double *Ez_m0, int32_t Nx, int32_t Nr, double xmin, double dx,
                                   double rmin, double dr, double k0, double a0, double x0,
                                   double Lx, double R, void *stream) {
  if (!Ez_m0 || Nx < 1 || Nr < 2) return CHM_EINVAL;
  int64_t n = (int64_t)(Nr - 1) * Nx;
  k_laser_pulse<<<chm_blocks(n, 256), 256, 0, (hipStream_t)stream>>>(Ez_m0, Nx, Nr, xmin, dx, rmin,
                                                                      dr, k0, a0, x0, Lx, R);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// laser.py:24-34 on mode 0
__global__ void k_laser_spectral(const double *kx, const double *w, double2 *E, double2 *G, double X,
                                 int Nx, int64_t n) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int ix = (int)(e % Nx);
  double k = kx[ix], ww = w[e];
  double sg = (k + (k == 0 ? 1.0 : 0.0)) > 0 ? 1.0 : -1.0;
  double2 EE = E[e];
  // DT = -i w sign(kx); GG = DT*EE
  double2 GG;
  GG.x = ww * sg * EE.y;
  GG.y = -ww * sg * EE.x;
  double cw = cos(ww * X), sw = sin(ww * X);
  double2 Et, Gt;
  Et.x = cw * EE.x + sw / ww * GG.x;
  Et.y = cw * EE.y + sw / ww * GG.y;
  Gt.x = -ww * sw * EE.x + cw * GG.x;
  Gt.y = -ww * sw * EE.y + cw * GG.y;
  double c = cos(k * X), s = sin(k * X);
  E[e].x = Et.x * c - Et.y * s;
  E[e].y = Et.x * s + Et.y * c;
  G[e].x = Gt.x * c - Gt.y * s;
  G[e].y = Gt.x * s + Gt.y * c;
}

extern "C" int chm_laser_spectral_f64(const chm_spectral_t *sp, const double *w_m0, void *Ez_fb_m0,
                                      void *Gz_fb_m0, double X_focus, void *stream) {
  if (!sp || !w_m0 || !Ez_fb_m0 || !Gz_fb_m0) return CHM_EINVAL;
  int64_t n = (int64_t)(sp->Nr - 1) * sp->Nx;
  k_laser_spectral<<<chm_blocks(n, 256), 256, 0, (hipStream_t)stream>>>(
      (const double *)sp->kx, w_m0, (double2 *)Ez_fb_m0, (double2 *)Gz_fb_m0, X_focus, sp->Nx, n);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}

// d/dR of the pulse above (R = waist): Ez * 2 r^2 / R^3 inside the cut-offs.
__global__ void k_laser_pulse_dR(double *Ez, int Nx, int Nr, double xmin, double dx, double rmin,
                                 double dr, double k0, double a0, double x0, double Lx, double R) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)(Nr - 1) * Nx) return;
  int ir = 1 + (int)(e / Nx), ix = (int)(e % Nx);
  double x = xmin + dx * (double)ix;
  double r = rmin + dr * (double)ir;
  double xs = x - x0;
  double v = a0 * sin(k0 * xs) * exp(-xs * xs / (Lx * Lx) - r * r / (R * R));
  v = v * (2.0 * r * r / (R * R * R));
  if (!(fabs(r) < 3.5 * R) || !(fabs(xs) < 3.5 * Lx)) v = 0.0 * v;
  Ez[(int64_t)ir * Nx + ix] = v;
}

extern "C" int chm_laser_pulse_dR_f64(double *Ez_m0, int32_t Nx, int32_t Nr, double xmin,
                                      double dx, double rmin, double dr, double k0, double a0,
                                      double x0, double Lx, double R, void *stream) {
  if (!Ez_m0 || Nx < 1 || Nr < 2) return CHM_EINVAL;
  int64_t n = (int64_t)(Nr - 1) * Nx;
  k_laser_pulse_dR<<<chm_blocks(n, 256), 256, 0, (hipStream_t)stream>>>(Ez_m0, Nx, Nr, xmin, dx,
                                                                         rmin, dr, k0, a0, x0, Lx, R);
  CHM_CHECK_LAUNCH();
  return CHM_OK;
}
