/*
 * chimera_hip.h -- C ABI of libchimera_hip.so, the MI355X (gfx950) kernels of
 * the chimeraCL quasi-cylindrical PSATD PIC step.
 *
 * Conventions
 *   - Every entry point is asynchronous on the HIP stream it is given
 *     (`stream`, a hipStream_t passed as void*; NULL = the null stream) and
 *     returns 0 on success, a negative CHM_E* code for a bad argument, or a
 *     positive hipError_t value when a launch failed.
 *   - All array arguments are device pointers.  Struct arguments are host
 *     pointers to plain structs holding device pointers and scalars.
 *   - No entry point allocates memory: scratch space comes in as `ws`
 *     (device pointer) with its size `ws_bytes`; each family has a
 *     *_workspace_bytes() query.
 *   - Suffix _f64: the plain FP64 form.  Suffix _f32: the mixed form's FP32
 *     kernels (particles, gathers, deposits, FFT, pointwise spectral ops).
 *     The mixed DHT product is selected in chm_spectral_t.dht_kind.
 *
 * Data layouts (the reference's, chimeraCL/grid.py:106-127,
 * transformer.py:100-132):
 *   particles     SoA [Np] (x, y, z, px, py, pz, w, g_inv)
 *   real planes   [Nr][Nx], x fastest; mode 0 real, modes m>=1 complex
 *                 interleaved (re, im); row 0 is the ghost row r = -dr/2
 *   spectral      [Nr-1][Nx] complex per mode, kr-major, kx in FFT order
 *   DHT tables    [Nr-1][Nr-1] real per mode
 *
 * Reference interfaces replaced (hightower8083/chimeraCL v0.0.1; the
 * reference binds its OpenCL kernels from Python with PyOpenCL, so these are
 * the methods of its *MethodsCL mixins and the kernels they launch):
 *   chm_push_x_*            ParticleMethodsCL.push_coords
 *                           (methods/particles_methods_cl.py:206-223,
 *                            kernels/particles_generic.cl:129-153)
 *   chm_index_sort_*        ParticleMethodsCL.index_sort
 *                           (methods/particles_methods_cl.py:225-261,
 *                            kernels/particles_generic.cl:88-126, 186-200)
 *   chm_dens_profile_*      ParticleMethodsCL.dens_profile
 *                           (methods/particles_methods_cl.py:179-204,
 *                            kernels/particles_generic.cl:6-30)
 *   chm_deposit_scalar_*    GridMethodsCL.depose_scalar
 *                           (methods/grid_methods_cl.py:44-69,
 *                            kernels/grid_deposit_m1.cl:20-152)
 *   chm_deposit_vector_*    GridMethodsCL.depose_vector
 *                           (methods/grid_methods_cl.py:71-95,
 *                            kernels/grid_deposit_m1.cl:171-327)
 *   chm_postproc_deposit_*  GridMethodsCL.postproc_depose_scalar/vector
 *                           (methods/grid_methods_cl.py:98-151,
 *                            kernels/grid_generic.cl:4-60)
 *   chm_warp_axis_*         GridMethodsCL.preproc_project_vec
 *                           (methods/grid_methods_cl.py:153-166,
 *                            kernels/grid_generic.cl:63-86)
 *   chm_gather_push_*       GridMethodsCL._gather_and_push
 *                           (methods/grid_methods_cl.py:168-192,
 *                            kernels/grid_deposit_m1.cl:330-511)
 *   chm_transform_*         TransformerMethodsCL.transform_field, full and
 *                           half modes (methods/transformer_methods_cl.py:
 *                           38-64, 290-455; Reikna FFT and MatrixMul)
 *   chm_mult_table_*        TransformerMethodsCL.fields_smooth and
 *                           field_poiss_vec (transformer_methods_cl.py:66-85)
 *   chm_field_grad_*        TransformerMethodsCL.field_grad (:87-133)
 *   chm_field_div_*         TransformerMethodsCL.field_div (:135-183)
 *   chm_field_rot_*         TransformerMethodsCL.field_rot (:185-263)
 *   chm_advance_fields_*    SolverMethodsCL.advance_fields
 *                           (methods/solver_methods_cl.py:43-62,
 *                            kernels/solver_ms_pic.cl:57-143)
 *   chm_damp_fields_*       Solver.damp_fields (solver.py:32-35:
 *                           half transforms + profile_edges,
 *                           solver_methods_cl.py:64-83)
 *   chm_build_tables_f64    Transformer._make_spectral_axes / _make_DHT
 *                           (transformer.py:28-97), Solver._make_ms_coefficients
 *                           (solver.py:41-52), _init_field_damping
 *                           (solver_methods_cl.py:32-41), get_phase_*
 *                           (kernels/transformer_generic.cl:28-55)
 *   chm_laser_*             add_gausian_pulse (laser.py:3-37)
 */
#ifndef CHIMERA_HIP_H
#define CHIMERA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHM_MAXM 4 /* highest azimuthal mode supported (modes 0..CHM_MAXM) */

/* Particle-grid geometry, from Grid.Args (grid.py:68-103). */
typedef struct chm_grid {
  int32_t Nx, Nr, M, pad_;
  double xmin, dx_inv, rmin, dr_inv;
} chm_grid_t;

/* Real-space planes of one field component: m0 is [Nr][Nx] real, mc[m-1] is
 * [Nr][Nx] complex for m = 1..M.  Element type: double (_f64) or float (_f32). */
typedef struct chm_planes {
  void *m0;
  void *mc[CHM_MAXM];
} chm_planes_t;

/* Spectral planes of one field component: m[m] is [Nr-1][Nx] complex. */
typedef struct chm_fb {
  void *m[CHM_MAXM + 1];
} chm_fb_t;

/* DHT product used by the _f32 (mixed) transforms. */
#define CHM_DHT_F32 0     /* FP32 MFMA product, FP32 accumulation            */
#define CHM_DHT_BF16X3 1  /* split-BF16 3-term MFMA product, FP32 accumulate */

/* Spectral tables (device pointers).  For _f64 every table is double; for
 * _f32 every table is float, and the DHT matrices are additionally given as
 * split-BF16 pairs (hi, lo) [Nk][Nk] each when dht_kind = CHM_DHT_BF16X3. */
typedef struct chm_spectral {
  int32_t Nx, Nr, M, ndamp; /* ndamp = 2*DampCells, 0 = no damping table   */
  int32_t dht_kind, pad_;
  double dt_inv;
  const void *kx;          /* [Nx]                                         */
  const void *phase_minus; /* [Nx] complex exp(-i kx Xmin)                 */
  const void *phase_plus;  /* [Nx] complex exp(+i kx Xmin)                 */
  const void *damp;        /* [ndamp]                                      */
  const void *twiddle;     /* [Nx] complex exp(-2 pi i t / Nx)             */
  const void *DHT[CHM_MAXM + 1];
  const void *DHT_inv[CHM_MAXM + 1];
  const void *dDp[CHM_MAXM + 1];
  const void *dDm[CHM_MAXM + 1];
  const void *DHT_lo[CHM_MAXM + 1];     /* bf16 low parts (BF16X3 only)  */
  const void *DHT_inv_lo[CHM_MAXM + 1];
  const void *dDp_lo[CHM_MAXM + 1];
  const void *dDm_lo[CHM_MAXM + 1];
  const void *c1[CHM_MAXM + 1];     /* cos(w dt)    [Nk][Nx]              */
  const void *c2[CHM_MAXM + 1];     /* w sin(w dt)  [Nk][Nx]              */
  const void *c3[CHM_MAXM + 1];     /* 1/w^2        [Nk][Nx] (= Poisson)  */
  const void *smooth[CHM_MAXM + 1]; /* smoothing filter [Nk][Nx]          */
  /* _f32 only, optional: the PSATD coefficients c1, c2, c3 in FP64; when
   * set, the FP32 forms' PSATD update (chm_advance_fields_f32,
   * chm_damp_fields_f32 with advance_first) runs in FP64 arithmetic from
   * them, so that the rounding of the coefficient tables does not act as
   * the same systematic rotation error in every step. */
  const double *c1d[CHM_MAXM + 1];
  const double *c2d[CHM_MAXM + 1];
  const double *c3d[CHM_MAXM + 1];
} chm_spectral_t;

/* Outputs of chm_build_tables_f64 (device pointers, all double). */
typedef struct chm_tables_out {
  double *kx;                /* [Nx]                                   */
  double *kr;                /* [M+2][Nk]: kr_m for m = 0..M+1         */
  double *phase_minus;       /* [Nx] complex                           */
  double *phase_plus;        /* [Nx] complex                           */
  double *damp;              /* [2*damp_cells] (may be NULL if 0)      */
  double *twiddle;           /* [Nx] complex exp(-2 pi i t / Nx)       */
  double *w[CHM_MAXM + 1];   /* [Nk][Nx]                               */
  double *c1[CHM_MAXM + 1];
  double *c2[CHM_MAXM + 1];
  double *c3[CHM_MAXM + 1];  /* 1/w^2, also Poiss_m                    */
  double *smooth[CHM_MAXM + 1];
  double *DHT[CHM_MAXM + 1]; /* [Nk][Nk]                               */
  double *DHT_inv[CHM_MAXM + 1];
  double *dDp[CHM_MAXM + 1];
  double *dDm[CHM_MAXM + 1];
} chm_tables_out_t;

/* ---------------------------------------------------------------- particles */
int chm_push_x_f64(double *x, double *y, double *z, const double *px,
                   const double *py, const double *pz, const double *g_inv,
                   double dt, int64_t Np, void *stream);
int chm_push_x_f32(float *x, float *y, float *z, const float *px,
                   const float *py, const float *pz, const float *g_inv,
                   float dt, int64_t Np, void *stream);

/* Cell keys, the stable sort and the offsets of index_sort.
 * nbins = (Nx-1)*(Nr-1)+1 (last bin = particles outside the valid cells).
 * Outputs: indx_in_cell [Np] (cell key), cell_offset [nbins+1]
 * (cell_offset[c] = first sorted slot of bin c, cell_offset[nbins] = Np),
 * sort_indx [Np] (particle indices ordered by key, ascending index within a
 * key: numpy.argsort(key, kind='stable')), np_stay [1] (= cell_offset[nbins-1],
 * the reference's Np_stay, kept on the device). */
size_t chm_sort_workspace_bytes(int64_t Np, int64_t nbins);
int chm_index_sort_f64(const double *x, const double *y, const double *z,
                       int64_t Np, const chm_grid_t *g, uint32_t *indx_in_cell,
                       uint32_t *cell_offset, uint32_t *sort_indx,
                       uint32_t *np_stay, void *ws, size_t ws_bytes,
                       void *stream);
int chm_index_sort_f32(const float *x, const float *y, const float *z,
                       int64_t Np, const chm_grid_t *g, uint32_t *indx_in_cell,
                       uint32_t *cell_offset, uint32_t *sort_indx,
                       uint32_t *np_stay, void *ws, size_t ws_bytes,
                       void *stream);
/* The two halves of index_sort, for the step's push-then-sort sequence
 * (pic_loop.py:70-91, push_coords followed by sort_parts):
 * chm_push_x_keys_* applies the coordinate push of chm_push_x_* and writes
 * the cell keys in the same pass (ws is checked for size only);
 * chm_sort_from_keys then produces cell_offset, sort_indx and np_stay
 * exactly as chm_index_sort_* does. */
int chm_push_x_keys_f64(double *x, double *y, double *z, const double *px,
                        const double *py, const double *pz,
                        const double *g_inv, double dt, int64_t Np,
                        const chm_grid_t *g, uint32_t *indx_in_cell, void *ws,
                        size_t ws_bytes, void *stream);
int chm_push_x_keys_f32(float *x, float *y, float *z, const float *px,
                        const float *py, const float *pz, const float *g_inv,
                        float dt, int64_t Np, const chm_grid_t *g,
                        uint32_t *indx_in_cell, void *ws, size_t ws_bytes,
                        void *stream);
int chm_sort_from_keys(int64_t Np, const chm_grid_t *g,
                       const uint32_t *indx_in_cell, uint32_t *cell_offset,
                       uint32_t *sort_indx, uint32_t *np_stay, void *ws,
                       size_t ws_bytes, void *stream);
/* Device address, inside a sort workspace, of the sorted keys that the last
 * chm_index_sort_* / chm_sort_from_keys call on it left there (key of each
 * sorted slot, [Np]; valid until the workspace's next sort).  Host-side
 * address arithmetic only, nothing is launched; NULL for bad arguments.  No
 * counterpart in the reference: the gather (chm_gather_push_*) takes it. */
const uint32_t *chm_sort_sorted_keys(const void *ws, int64_t Np,
                                     const chm_grid_t *g);
/* chm_index_sort_f64 without a workspace argument and without allocating:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The sort's scratch is a fixed device array of
 * CHM_SORT_SCRATCH_BYTES per GPU, held until the process ends; inputs larger
 * than it allows return CHM_ESCRATCH (-4) before anything is launched.  Any
 * grid of up to 2^22 cells with up to 2^21 particles fits
 * (chm_sort_workspace_bytes(Np, nbins) <= CHM_SORT_SCRATCH_BYTES is the
 * exact rule).  Every call on one GPU uses that GPU's one array: calls that
 * overlap in time on the same GPU (several streams or host threads) race and
 * must be serialised by the caller.  Outputs are those of
 * chm_index_sort_f64 bit for bit. */
#define CHM_SORT_SCRATCH_BYTES (48u << 20)
#define CHM_ESCRATCH (-4)
int chm_index_sort_f64_noalloc(const double *x, const double *y,
                               const double *z, int64_t Np,
                               const chm_grid_t *g, uint32_t *indx_in_cell,
                               uint32_t *cell_offset, uint32_t *sort_indx,
                               uint32_t *np_stay, void *stream);
/* data_align_dbl / data_align_int (kernels/particles_generic.cl:156-186,
 * ParticleMethodsCL.align_and_damp, methods/particles_methods_cl.py:263-286):
 * dst[i] = src[sort_indx[i]] for i < n; elem_bytes 8 (double) or 4 (float,
 * uint32).  dst must not be src. */
int chm_align_elems(const uint32_t *sort_indx, int64_t n, const void *src,
                    void *dst, int32_t elem_bytes, void *stream);
/* The same permutation of narr (1..CHM_ALIGN_MAXARR) arrays in one launch:
 * dst[a][i] = src[a][sort_indx[i]] for i < n, elem_bytes[a] 4 or 8; src,
 * dst and elem_bytes are host arrays of narr entries.  No dst may equal any
 * src, nor another dst.  The reference aligns one array per call
 * (particles_methods_cl.py:263-286); this groups the calls so that sort_indx
 * is read once (PIC_loop(align_every) reorders). */
#define CHM_ALIGN_MAXARR 16
int chm_align_multi(const uint32_t *sort_indx, int64_t n, int32_t narr,
                    const void *const *src, void *const *dst,
                    const int32_t *elem_bytes, void *stream);
int chm_dens_profile_f64(const double *x, double *w, int64_t Np,
                         const double *xx, const double *ff,
                         const double *dxm1, int32_t nknots, void *stream);
int chm_dens_profile_f32(const float *x, float *w, int64_t Np,
                         const double *xx, const double *ff,
                         const double *dxm1, int32_t nknots, void *stream);

/* ------------------------------------------------------------ grid <-> particles */
size_t chm_deposit_workspace_bytes(int32_t Nx, int32_t Nr, int32_t M,
                                   int32_t vector);
/* Adds the species' charge density (w*charge, CIC, modes 0..M) into rho. */
int chm_deposit_scalar_f64(const uint32_t *sort_indx,
                           const uint32_t *cell_offset, const double *x,
                           const double *y, const double *z, const double *w,
                           int32_t charge, const chm_grid_t *g,
                           const chm_planes_t *rho, void *ws, size_t ws_bytes,
                           void *stream);
int chm_deposit_scalar_f32(const uint32_t *sort_indx,
                           const uint32_t *cell_offset, const float *x,
                           const float *y, const float *z, const float *w,
                           int32_t charge, const chm_grid_t *g,
                           const chm_planes_t *rho, void *ws, size_t ws_bytes,
                           void *stream);
/* Adds the species' current (w*g_inv*charge*u) into J[0..2]. */
int chm_deposit_vector_f64(const uint32_t *sort_indx,
                           const uint32_t *cell_offset, const double *x,
                           const double *y, const double *z, const double *ux,
                           const double *uy, const double *uz,
                           const double *g_inv, const double *w, int32_t charge,
                           const chm_grid_t *g, const chm_planes_t *J,
                           void *ws, size_t ws_bytes, void *stream);
int chm_deposit_vector_f32(const uint32_t *sort_indx,
                           const uint32_t *cell_offset, const float *x,
                           const float *y, const float *z, const float *ux,
                           const float *uy, const float *uz,
                           const float *g_inv, const float *w, int32_t charge,
                           const chm_grid_t *g, const chm_planes_t *J,
                           void *ws, size_t ws_bytes, void *stream);
/* chm_deposit_vector_* followed by the step's next coordinate push and the
 * cell keys of its sort (pic_loop.py:82-91: depose_currents, then
 * push_coords('half') and the key pass of sort_parts), in one pass over the
 * particles: J from the positions xin, yin, zin, then every one of the Np
 * particles is pushed by dt (x = xin + u*dt*g_inv, chm_push_x_*'s
 * operations) into x, y, z (which may be xin, yin, zin), and indx_in_cell
 * receives the keys chm_push_x_keys_* would write, ready for
 * chm_sort_from_keys.  sort_indx / cell_offset must be the current sort of
 * these particles on grid g. */
int chm_deposit_vector_push_f64(const uint32_t *sort_indx,
                                const uint32_t *cell_offset, const double *xin,
                                const double *yin, const double *zin,
                                double *x, double *y, double *z,
                                const double *ux, const double *uy,
                                const double *uz, const double *g_inv,
                                const double *w, int32_t charge, double dt,
                                int64_t Np, uint32_t *indx_in_cell,
                                const chm_grid_t *g, const chm_planes_t *J,
                                void *ws, size_t ws_bytes, void *stream);
int chm_deposit_vector_push_f32(const uint32_t *sort_indx,
                                const uint32_t *cell_offset, const float *xin,
                                const float *yin, const float *zin, float *x,
                                float *y, float *z, const float *ux,
                                const float *uy, const float *uz,
                                const float *g_inv, const float *w,
                                int32_t charge, float dt, int64_t Np,
                                uint32_t *indx_in_cell, const chm_grid_t *g,
                                const chm_planes_t *J, void *ws,
                                size_t ws_bytes, void *stream);
/* The node pass of a deposit alone: adds to the planes the per-cell sums a
 * previous chm_deposit_scalar_* (vector = 0) or chm_deposit_vector_*
 * (vector = 1) call left in ws, in the same order.  For a species whose
 * particles have not changed since (immobile ions: their per-cell sums are
 * the same every step). */
int chm_deposit_add_staged_f64(const uint32_t *cell_offset,
                               const chm_grid_t *g, const chm_planes_t *flds,
                               int32_t vector, void *ws, size_t ws_bytes,
                               void *stream);
int chm_deposit_add_staged_f32(const uint32_t *cell_offset,
                               const chm_grid_t *g, const chm_planes_t *flds,
                               int32_t vector, void *ws, size_t ws_bytes,
                               void *stream);
/* treat_axis (row1 -= row0) then *= dV_inv[ir], for nflds fields. */
int chm_postproc_deposit_f64(const chm_planes_t *flds, int32_t nflds,
                             const double *dV_inv, int32_t Nx, int32_t Nr,
                             int32_t M, void *stream);
int chm_postproc_deposit_f32(const chm_planes_t *flds, int32_t nflds,
                             const double *dV_inv, int32_t Nx, int32_t Nr,
                             int32_t M, void *stream);
/* Ghost row: row0 = row1 (m=0), row0 = -row1 (m>=1), for nflds fields. */
/* Grid.set_to(plane, 0) of every plane (modes 0..M) of nflds fields in one
 * launch, as Grid.depose_charge / depose_currents do before depositing
 * (grid.py:23-54). */
int chm_zero_planes_f64(const chm_planes_t *flds, int32_t nflds, int32_t Nx,
                        int32_t Nr, int32_t M, void *stream);
int chm_zero_planes_f32(const chm_planes_t *flds, int32_t nflds, int32_t Nx,
                        int32_t Nr, int32_t M, void *stream);
int chm_warp_axis_f64(const chm_planes_t *flds, int32_t nflds, int32_t Nx,
                      int32_t Nr, int32_t M, void *stream);
int chm_warp_axis_f32(const chm_planes_t *flds, int32_t nflds, int32_t Nx,
                      int32_t Nr, int32_t M, void *stream);
/* Field gather and relativistic Boris push of every particle p with
 * p < *np_stay inside a valid cell (the set the reference's per-slot launch
 * visits).  factor_push = Args['FactorPush'] = 2*pi*dt*q/m.  sort_indx: the
 * current sorted order (chm_index_sort_* / chm_sort_from_keys of these
 * positions); the kernel then visits slots ip < *np_stay with
 * p = sort_indx[ip] as the reference does and stages the fields of each
 * slot block's cells in LDS.  NULL: one lane per particle index, fields
 * read directly.  Both give identical results.  sorted_keys (may be NULL):
 * the keys of that sort's slots (chm_sort_sorted_keys); with them each slot
 * block finds its cells from two keys and loads the fields it stages while
 * its particles are being loaded, instead of after (identical results). */
int chm_gather_push_f64(const double *x, const double *y, const double *z,
                        double *px, double *py, double *pz, double *g_inv,
                        int64_t Np, const uint32_t *sort_indx,
                        const uint32_t *np_stay, const uint32_t *sorted_keys,
                        double factor_push, const chm_grid_t *g,
                        const chm_planes_t *E, const chm_planes_t *B,
                        void *stream);
int chm_gather_push_f32(const float *x, const float *y, const float *z,
                        float *px, float *py, float *pz, float *g_inv,
                        int64_t Np, const uint32_t *sort_indx,
                        const uint32_t *np_stay, const uint32_t *sorted_keys,
                        float factor_push, const chm_grid_t *g,
                        const chm_planes_t *E, const chm_planes_t *B,
                        void *stream);
/* The same, for storage reordered by chm_align_elems (PIC_loop(align_every)):
 * pid[p] is particle p's index in the numbering without reordering, and the
 * reference's skip (grid_deposit_m1.cl:367, sorting_indx[ip] < Np_stay)
 * compares pid[p] with *np_stay instead of p, so the same particles are
 * skipped as without the reordering. */
int chm_gather_push_ids_f64(const double *x, const double *y, const double *z,
                            double *px, double *py, double *pz, double *g_inv,
                            int64_t Np, const uint32_t *sort_indx,
                            const uint32_t *np_stay, const uint32_t *pid,
                            const uint32_t *sorted_keys, double factor_push,
                            const chm_grid_t *g, const chm_planes_t *E,
                            const chm_planes_t *B, void *stream);
int chm_gather_push_ids_f32(const float *x, const float *y, const float *z,
                            float *px, float *py, float *pz, float *g_inv,
                            int64_t Np, const uint32_t *sort_indx,
                            const uint32_t *np_stay, const uint32_t *pid,
                            const uint32_t *sorted_keys, float factor_push,
                            const chm_grid_t *g, const chm_planes_t *E,
                            const chm_planes_t *B, void *stream);
/* chm_gather_push_ids_* (pid may be NULL) followed by the next step's first
 * coordinate push and the keys of its sort (pic_loop.py:134 gather_and_push,
 * then pic_loop.py:70-77 push_coords('half') + the key pass of sort_parts of
 * the next step), for consecutive steps with nothing in between: every one of
 * the Np particles (the ones the gather skips and those outside the domain
 * included) is pushed by dt from x, y, z with the momenta and g_inv it now
 * holds, into xh, yh, zh (x, y, z are left as they are), and indx_in_cell
 * receives the keys of those positions, ready for chm_sort_from_keys.
 * cell_offset: the current sort's offsets (as sort_indx). */
int chm_gather_push_pre_f64(const double *x, const double *y, const double *z,
                            double *px, double *py, double *pz, double *g_inv,
                            int64_t Np, const uint32_t *sort_indx,
                            const uint32_t *cell_offset, const uint32_t *np_stay,
                            const uint32_t *pid, const uint32_t *sorted_keys,
                            double factor_push, double dt, double *xh,
                            double *yh, double *zh, uint32_t *indx_in_cell,
                            const chm_grid_t *g, const chm_planes_t *E,
                            const chm_planes_t *B, void *stream);
int chm_gather_push_pre_f32(const float *x, const float *y, const float *z,
                            float *px, float *py, float *pz, float *g_inv,
                            int64_t Np, const uint32_t *sort_indx,
                            const uint32_t *cell_offset, const uint32_t *np_stay,
                            const uint32_t *pid, const uint32_t *sorted_keys,
                            float factor_push, float dt, float *xh, float *yh,
                            float *zh, uint32_t *indx_in_cell,
                            const chm_grid_t *g, const chm_planes_t *E,
                            const chm_planes_t *B, void *stream);

/* ------------------------------------------------------------- transforms */
size_t chm_transform_workspace_bytes(int32_t Nx, int32_t Nr, int32_t M,
                                     int32_t nflds, int32_t elem_bytes);
/* dir 0: real -> spectral, dir 1: spectral -> real (rows 1..Nr-1).
 * mode 0: full (DHT + FFT_x + phase); mode 1: half (FFT_x + phase only).
 * post_smooth (dir 0 only): if nonzero, multiply the result by sp->smooth. */
int chm_transform_f64(const chm_spectral_t *sp, const chm_planes_t *real,
                      const chm_fb_t *fb, int32_t nflds, int32_t dir,
                      int32_t mode, int32_t post_smooth, void *ws,
                      size_t ws_bytes, void *stream);
int chm_transform_f32(const chm_spectral_t *sp, const chm_planes_t *real,
                      const chm_fb_t *fb, int32_t nflds, int32_t dir,
                      int32_t mode, int32_t post_smooth, void *ws,
                      size_t ws_bytes, void *stream);
/* which 0: *= smooth (fields_smooth); which 1: *= 1/w^2 (field_poiss_*). */
int chm_mult_table_f64(const chm_spectral_t *sp, const chm_fb_t *flds,
                       int32_t nflds, int32_t which, void *stream);
int chm_mult_table_f32(const chm_spectral_t *sp, const chm_fb_t *flds,
                       int32_t nflds, int32_t which, void *stream);
/* out[3] = grad(scl).  If prev is not NULL, the old out[3] is first copied to
 * prev[3] (the dN0 <- dN1 copy of pic_loop.py:106-112). */
size_t chm_spectral_workspace_bytes(int32_t Nx, int32_t Nr, int32_t M,
                                    int32_t elem_bytes);
int chm_field_grad_f64(const chm_spectral_t *sp, const chm_fb_t *scl,
                       const chm_fb_t *out, const chm_fb_t *prev, void *ws,
                       size_t ws_bytes, void *stream);
int chm_field_grad_f32(const chm_spectral_t *sp, const chm_fb_t *scl,
                       const chm_fb_t *out, const chm_fb_t *prev, void *ws,
                       size_t ws_bytes, void *stream);
int chm_field_div_f64(const chm_spectral_t *sp, const chm_fb_t *vec,
                      const chm_fb_t *out, void *ws, size_t ws_bytes,
                      void *stream);
int chm_field_div_f32(const chm_spectral_t *sp, const chm_fb_t *vec,
                      const chm_fb_t *out, void *ws, size_t ws_bytes,
                      void *stream);
/* out[3] = rot(in[3]); poisson != 0 also multiplies by 1/w^2 (restore_B_fb). */
int chm_field_rot_f64(const chm_spectral_t *sp, const chm_fb_t *in,
                      const chm_fb_t *out, int32_t poisson, void *ws,
                      size_t ws_bytes, void *stream);
int chm_field_rot_f32(const chm_spectral_t *sp, const chm_fb_t *in,
                      const chm_fb_t *out, int32_t poisson, void *ws,
                      size_t ws_bytes, void *stream);

/* ----------------------------------------------------------------- solver */
int chm_advance_fields_f64(const chm_spectral_t *sp, const chm_fb_t *E,
                           const chm_fb_t *G, const chm_fb_t *J,
                           const chm_fb_t *dN0, const chm_fb_t *dN1,
                           void *stream);
int chm_advance_fields_f32(const chm_spectral_t *sp, const chm_fb_t *E,
                           const chm_fb_t *G, const chm_fb_t *J,
                           const chm_fb_t *dN0, const chm_fb_t *dN1,
                           void *stream);
/* Edge damping of E[3] and G[3] in (kr, x) space, in place in spectral
 * space; with advance != 0 the PSATD advance is applied first in the same
 * kernel (J, dN0, dN1 then required). */
int chm_damp_fields_f64(const chm_spectral_t *sp, const chm_fb_t *E,
                        const chm_fb_t *G, int32_t advance, const chm_fb_t *J,
                        const chm_fb_t *dN0, const chm_fb_t *dN1,
                        void *stream);
int chm_damp_fields_f32(const chm_spectral_t *sp, const chm_fb_t *E,
                        const chm_fb_t *G, int32_t advance, const chm_fb_t *J,
                        const chm_fb_t *dN0, const chm_fb_t *dN1,
                        void *stream);

/* ----------------------------------------------------------------- tables */
size_t chm_tables_workspace_bytes(int32_t Nx, int32_t Nr, int32_t M);
/* Builds every spectral table in FP64 on the GPU: Bessel roots (McMahon
 * start + Newton), J_m matrices, their inverse (Gauss-Jordan with partial
 * pivoting), dDHT+-, kx, w, PSATD coefficients, smoothing filter, phases and
 * the damping profile. */
int chm_build_tables_f64(int32_t Nx, int32_t Nr, int32_t M, double dx,
                         double dr, double dt, double xmin, int32_t damp_cells,
                         const chm_tables_out_t *out, void *ws,
                         size_t ws_bytes, void *stream);
/* The DHT-type matrices of chm_spectral_t (DHT, DHT_inv, dDp, dDm and their
 * _lo parts) are read in a zero-padded layout: [round_up(Nr-1, 64)] rows of
 * round_up(Nr-1, 64) elements.  chm_dht_table_elems gives its size in
 * elements; chm_pad_table writes it from a row-major [M][K] FP64 matrix as
 * double (d64), float (f32) and/or a split-BF16 pair (bf16_hi + bf16_lo ~=
 * src); any output may be NULL (hi and lo together). */
size_t chm_dht_table_elems(int32_t M, int32_t K);
int chm_pad_table(const double *src, int32_t M, int32_t K, double *d64,
                  float *f32, uint16_t *bf16_hi, uint16_t *bf16_lo,
                  void *stream);
/* Plain FP64 product C[M x N] = A[M x K] @ B[K x N] on the matrix cores,
 * A in the padded layout above, B and C row-major and contiguous; used for
 * the dDHT tables and by the tests. */
int chm_gemm_f64(const double *A, const double *B, double *C, int32_t M,
                 int32_t K, int32_t N, void *stream);
/* Measurement aid (no counterpart in the reference): chm_dht_timing(1)
 * clears the record and makes every DHT product launch on a stream that is
 * not being captured record a HIP event before and after it on that stream,
 * with the launch's algorithmic flops (2 M K N per product; the FP32-
 * equivalent work for split-BF16) and its form (0 FP64, 1 FP32, 2
 * split-BF16); chm_dht_timing(0) stops and clears.
 * chm_dht_timing_collect waits for the recorded launches, writes up to max
 * (duration ms, flops, form) triples (any output may be NULL), clears the
 * record and returns the count (-1 on a HIP error). */
int chm_dht_timing(int32_t enable);
int64_t chm_dht_timing_collect(double *ms, double *flops, int32_t *kind,
                               int64_t max);
/* Converts n doubles to float (round to nearest) and, if hi/lo are not NULL,
 * to a split-BF16 pair with hi + lo ~= src. */
int chm_convert_table(const double *src, int64_t n, float *f32,
                      uint16_t *bf16_hi, uint16_t *bf16_lo, void *stream);

/* ------------------------------------------------------------------ laser */
/* Ez_m0[1:] = a0 sin(k0 (x - x0)) exp(-(x-x0)^2/Lx^2 - r^2/R^2) inside the
 * 3.5 R, 3.5 Lx cut-offs (laser.py:17-19); row 0 untouched. */
int chm_laser_pulse_f64(double *Ez_m0, int32_t Nx, int32_t Nr, double xmin,
                        double dx, double rmin, double dr, double k0,
                        double a0, double x0, double Lx, double R,
                        void *stream);
/* Spectral part of laser.py:24-34 on mode 0: G = -i w sign(kx) E, then the
 * vacuum propagation over X_focus; writes E and G in place. */
int chm_laser_spectral_f64(const chm_spectral_t *sp, const double *w_m0,
                           void *Ez_fb_m0, void *Gz_fb_m0, double X_focus,
                           void *stream);

/* ------------------------------------------------------------ moving window */
/* ParticleMethodsCL.make_new_domain (methods/particles_methods_cl.py:66-147)
 * with fill_grid (kernels/particles_generic.cl:33-84): for every cell of the
 * local grids xgrid[Nx_loc] x rgrid (ncells = (Nx_loc-1)(Nr_loc-1)),
 * npt x npr x npx particles at the regular sub-positions, the theta set
 * rotated by a per-cell uniform offset in [0, 2 pi); w = r * w0; when px
 * is not NULL, momenta p_c + dp n per component (n a standard normal per
 * particle and component; a zero spread dp gives p_c exactly) and
 * g_inv = 1/sqrt(1 + p^2).  The offset of cell c is 2 pi u with u the top 53
 * bits of splitmix64 at (seed, counter0 + c) (the reference draws it from
 * PyOpenCL's generator, which is not reproducible here); with a nonzero
 * spread, particle p draws its normals by Box-Muller from counters
 * counter0 + ncells + 4p .. + 3, so a fill uses ncells + 4 Np counters (else
 * ncells).  sin/cos/log are evaluated from IEEE +, -, *, / only, so a CPU
 * restatement reproduces every bit. */
int chm_fill_domain_f64(double *x, double *y, double *z, double *w, double *px,
                        double *py, double *pz, double *g_inv,
                        const double *xgrid, const double *rgrid,
                        int32_t Nx_loc, int64_t ncells, int32_t npx,
                        int32_t npr, int32_t npt, double w0, double px_c,
                        double py_c, double pz_c, double dpx, double dpy,
                        double dpz, uint64_t seed, uint64_t counter0,
                        void *stream);
int chm_fill_domain_f32(float *x, float *y, float *z, float *w, float *px,
                        float *py, float *pz, float *g_inv,
                        const double *xgrid, const double *rgrid,
                        int32_t Nx_loc, int64_t ncells, int32_t npx,
                        int32_t npr, int32_t npt, double w0, double px_c,
                        double py_c, double pz_c, double dpx, double dpy,
                        double dpz, uint64_t seed, uint64_t counter0,
                        void *stream);
/* ParticleMethodsCL.make_new_beam (methods/particles_methods_cl.py:149-177):
 * Np particles with positions c + L n, momenta pc + dp n (n standard
 * normals), weights w, g_inv = 1/sqrt(1 + p^2).  Particle p's six normals
 * come from Box-Muller on splitmix64 at (seed, counter0 + 6p .. 6p + 5),
 * with log, sin and cos evaluated from IEEE operations only (bit-exact
 * against a CPU restatement). */
typedef struct chm_beam {
  double c[3];   /* x_c, y_c, z_c */
  double L[3];   /* Lx, Ly, Lz */
  double pc[3];  /* px_c, py_c, pz_c */
  double dp[3];  /* dpx, dpy, dpz */
  double w;      /* FullCharge / Np */
} chm_beam_t;
int chm_fill_beam_f64(double *x, double *y, double *z, double *px, double *py,
                      double *pz, double *w, double *g_inv, int64_t Np,
                      const chm_beam_t *b, uint64_t seed, uint64_t counter0,
                      void *stream);
int chm_fill_beam_f32(float *x, float *y, float *z, float *px, float *py,
                      float *pz, float *w, float *g_inv, int64_t Np,
                      const chm_beam_t *b, uint64_t seed, uint64_t counter0,
                      void *stream);
/* get_phase_minus / get_phase_plus (kernels/transformer_generic.cl:28-55)
 * for the window origin xmin after Frame.shift_grids (frame.py:22-30): the
 * transforms read these tables instead of recomputing them per call. */
int chm_phase_tables_f64(const double *kx, int32_t Nx, double xmin,
                         void *phase_minus, void *phase_plus, void *stream);

/* ------------------------------------------------------------------ adjoints
 * Reverse-mode kernels of the differentiable form (no counterpart in the
 * reference, whose step is not differentiable).  Adjoints of real values
 * are dL/dv; of complex values dL/dRe + i dL/dIm, every stored element
 * independent.  Particle adjoints are SoA arrays indexed like the
 * particles, and every entry point ADDS into them, except where stated.
 * Each is the exact transpose (vector-Jacobian product) of the forward
 * entry point named, at the forward inputs given. */

/* push_x (chm_push_x_*): x' = x + u g_inv dt.  u_b += x_b g_inv dt,
 * g_b += dt (x_b . u); x_b passes through unchanged. */
int chm_push_x_adj_f64(const double *xb, const double *yb, const double *zb,
                       const double *ux, const double *uy, const double *uz,
                       const double *g_inv, double *uxb, double *uyb,
                       double *uzb, double *gb, double dt, int64_t Np,
                       void *stream);
int chm_push_x_adj_f32(const float *xb, const float *yb, const float *zb,
                       const float *ux, const float *uy, const float *uz,
                       const float *g_inv, float *uxb, float *uyb, float *uzb,
                       float *gb, float dt, int64_t Np, void *stream);

/* deposit (chm_deposit_scalar_*, one species) given the adjoint planes of
 * rho after chm_postproc_deposit_adj_*: adds to x_b, y_b, z_b, w_b. */
int chm_deposit_scalar_adj_f64(const double *x, const double *y,
                               const double *z, const double *w,
                               int32_t charge, int64_t Np,
                               const chm_grid_t *g, const chm_planes_t *rho_b,
                               double *xb, double *yb, double *zb, double *wb,
                               void *stream);
int chm_deposit_scalar_adj_f32(const float *x, const float *y, const float *z,
                               const float *w, int32_t charge, int64_t Np,
                               const chm_grid_t *g, const chm_planes_t *rho_b,
                               float *xb, float *yb, float *zb, float *wb,
                               void *stream);
/* deposit (chm_deposit_vector_*) given J_b[3]: adds to x_b, y_b, z_b,
 * u_b (3), g_b, w_b. */
int chm_deposit_vector_adj_f64(const double *x, const double *y,
                               const double *z, const double *ux,
                               const double *uy, const double *uz,
                               const double *g_inv, const double *w,
                               int32_t charge, int64_t Np, const chm_grid_t *g,
                               const chm_planes_t *J_b, double *xb, double *yb,
                               double *zb, double *uxb, double *uyb,
                               double *uzb, double *gb, double *wb,
                               void *stream);
int chm_deposit_vector_adj_f32(const float *x, const float *y, const float *z,
                               const float *ux, const float *uy,
                               const float *uz, const float *g_inv,
                               const float *w, int32_t charge, int64_t Np,
                               const chm_grid_t *g, const chm_planes_t *J_b,
                               float *xb, float *yb, float *zb, float *uxb,
                               float *uyb, float *uzb, float *gb, float *wb,
                               void *stream);

/* gather + Boris push (chm_gather_push_*) at positions x, y, z and the
 * momenta BEFORE the push, with the fields E[3], B[3] of that push.
 * In: u_b, g_b = adjoints of the pushed momenta and g_inv.  Out (replaced):
 * u_b = adjoint of the momenta before the push, g_b = 0 for pushed
 * particles; added: x_b, y_b, z_b; written: fb[0..5] = per-particle
 * adjoints of the gathered E_p, B_p (0 for particles not pushed), to be
 * scattered with chm_gather_adj_scatter_*. */
int chm_gather_push_adj_f64(const double *x, const double *y, const double *z,
                            const double *ux, const double *uy,
                            const double *uz, int64_t Np,
                            const uint32_t *np_stay, double factor_push,
                            const chm_grid_t *g, const chm_planes_t *E,
                            const chm_planes_t *B, double *uxb, double *uyb,
                            double *uzb, double *gb, double *xb, double *yb,
                            double *zb, double *const *fb, void *stream);
int chm_gather_push_adj_f32(const float *x, const float *y, const float *z,
                            const float *ux, const float *uy, const float *uz,
                            int64_t Np, const uint32_t *np_stay,
                            float factor_push, const chm_grid_t *g,
                            const chm_planes_t *E, const chm_planes_t *B,
                            float *uxb, float *uyb, float *uzb, float *gb,
                            float *xb, float *yb, float *zb,
                            float *const *fb, void *stream);
/* out[k] += sum_p C(p) v_k[p] (mode 0) and C(p) v_k[p] 2 ((y+iz)/r)^m
 * (modes m >= 1): the transpose of the gather's interpolation, with the
 * sorted order of the positions (sort_indx, cell_offset) and the vector
 * deposit's two-pass scheme and workspace. */
int chm_gather_adj_scatter_f64(const uint32_t *sort_indx,
                               const uint32_t *cell_offset, const double *x,
                               const double *y, const double *z,
                               const double *v0, const double *v1,
                               const double *v2, const chm_grid_t *g,
                               const chm_planes_t *out, void *ws,
                               size_t ws_bytes, void *stream);
int chm_gather_adj_scatter_f32(const uint32_t *sort_indx,
                               const uint32_t *cell_offset, const float *x,
                               const float *y, const float *z, const float *v0,
                               const float *v1, const float *v2,
                               const chm_grid_t *g, const chm_planes_t *out,
                               void *ws, size_t ws_bytes, void *stream);

/* Transposes of chm_postproc_deposit_* and chm_warp_axis_*, in place. */
int chm_postproc_deposit_adj_f64(const chm_planes_t *flds, int32_t nflds,
                                 const double *dV_inv, int32_t Nx, int32_t Nr,
                                 int32_t M, void *stream);
int chm_postproc_deposit_adj_f32(const chm_planes_t *flds, int32_t nflds,
                                 const double *dV_inv, int32_t Nx, int32_t Nr,
                                 int32_t M, void *stream);
int chm_warp_axis_adj_f64(const chm_planes_t *flds, int32_t nflds, int32_t Nx,
                          int32_t Nr, int32_t M, void *stream);
int chm_warp_axis_adj_f32(const chm_planes_t *flds, int32_t nflds, int32_t Nx,
                          int32_t Nr, int32_t M, void *stream);

/* Density profile (chm_dens_profile_*): knot_grad[j] += sum_p w_b[p]
 * w_base[p] df(x_p)/dff_j, with x and w_base the positions and weights
 * the profile was applied to; at most 128 knots; deterministic. */
size_t chm_profile_adj_workspace_bytes(int32_t nknots);
int chm_dens_profile_adj_f64(const double *x, const double *w_base,
                             const double *w_bar, int64_t Np,
                             const double *xx, const double *dxm1,
                             int32_t nknots, double *knot_grad, void *ws,
                             size_t ws_bytes, void *stream);
int chm_dens_profile_adj_f32(const float *x, const float *w_base,
                             const float *w_bar, int64_t Np, const double *xx,
                             const double *dxm1, int32_t nknots,
                             double *knot_grad, void *ws, size_t ws_bytes,
                             void *stream);

/* Spectral operators.  sp must hold the TRANSPOSED dDHT tables in dDp/dDm
 * (and the transposed-and-scaled DHT tables, Nx DHT^T in DHT_inv and
 * DHT_inv^T / Nx in DHT, for chm_transform_* to act as the adjoint of the
 * opposite transform).  accumulate = 1 adds into the outputs. */
int chm_field_grad_adj_f64(const chm_spectral_t *sp, const chm_fb_t *vec_b,
                           const chm_fb_t *scl_b, int32_t accumulate, void *ws,
                           size_t ws_bytes, void *stream);
int chm_field_grad_adj_f32(const chm_spectral_t *sp, const chm_fb_t *vec_b,
                           const chm_fb_t *scl_b, int32_t accumulate, void *ws,
                           size_t ws_bytes, void *stream);
/* poisson = 1: the adjoint of rot followed by *1/w^2; in_b is multiplied by
 * 1/w^2 in place first. */
int chm_field_rot_adj_f64(const chm_spectral_t *sp, const chm_fb_t *in_b,
                          const chm_fb_t *out_b, int32_t poisson,
                          int32_t accumulate, void *ws, size_t ws_bytes,
                          void *stream);
int chm_field_rot_adj_f32(const chm_spectral_t *sp, const chm_fb_t *in_b,
                          const chm_fb_t *out_b, int32_t poisson,
                          int32_t accumulate, void *ws, size_t ws_bytes,
                          void *stream);
/* Transpose of chm_damp_fields_* with advance = 1 (PSATD, then damping):
 * in E_b, G_b (+ E_add if not NULL) = adjoints of the new E, G; carry =
 * adjoint of this step's dN1 as the next step's dN0.  Out: E_b, G_b of the
 * old E, G; J_b; N1_b = total adjoint of dN1; carry = adjoint of dN0.
 * Uses the forward tables (sp as for chm_damp_fields_*). */
int chm_damp_fields_adj_f64(const chm_spectral_t *sp, const chm_fb_t *E_b,
                            const chm_fb_t *G_b, const chm_fb_t *E_add,
                            const chm_fb_t *J_b, const chm_fb_t *carry,
                            const chm_fb_t *N1_b, void *stream);
int chm_damp_fields_adj_f32(const chm_spectral_t *sp, const chm_fb_t *E_b,
                            const chm_fb_t *G_b, const chm_fb_t *E_add,
                            const chm_fb_t *J_b, const chm_fb_t *carry,
                            const chm_fb_t *N1_b, void *stream);
/* d/dR of chm_laser_pulse_f64 (R = waist), same layout and cut-offs. */
int chm_laser_pulse_dR_f64(double *Ez_m0, int32_t Nx, int32_t Nr, double xmin,
                           double dx, double rmin, double dr, double k0,
                           double a0, double x0, double Lx, double R,
                           void *stream);

#ifdef __cplusplus
}
#endif
#endif
