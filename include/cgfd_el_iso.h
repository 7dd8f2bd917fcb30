/*
 * cgfd_el_iso.h -- C ABI of the MI355X (gfx950) elastic-isotropic RK4
 * time-stepping path of CGFD3D.
 *
 * The library owns every device buffer; callers pass plain host arrays in the
 * reference's host layout and get results copied back on request.  No torch
 * or HIP types appear in these signatures.
 *
 * Host array layout ("reference layout"): one component is nx*ny*nz values,
 * nx = ni+6, ny = nj+6, nz = nk+6 (3 ghost points per side), i fastest, then
 * j, then k; component c of a multi-component array starts at c*nx*ny*nz.
 * This is the layout of gd_t / gd_metric_t / md_t / wav_t in the reference
 * (src/forward/gd_t.cu:1985-2021, wav_t.cu:35-40).
 *
 * Wavefield component order: Vx,Vy,Vz,Txx,Tyy,Tzz,Tyz,Txz,Txy
 * (src/forward/wav_t.cu:100-140).  Metric order: jac, xi_x, xi_y, xi_z,
 * eta_x, eta_y, eta_z, zeta_x, zeta_y, zeta_z (src/forward/gd_t.cu:115-200).
 *
 * Error convention: every int-returning call returns CGFD_OK (0) or a
 * negative CGFD_E_* code; cgfd_last_error() gives a thread-local message.
 * The library never aborts the process.
 */
#ifndef CGFD_EL_ISO_H
#define CGFD_EL_ISO_H

#ifdef __cplusplus
extern "C" {
#endif

#define CGFD_OK 0
#define CGFD_E_ARG (-1)      /* invalid argument */
#define CGFD_E_HIP (-2)      /* HIP runtime error */
#define CGFD_E_NOMEM (-3)    /* insufficient HBM for the request */
#define CGFD_E_STATE (-4)    /* call out of order (e.g. run before upload) */
#define CGFD_E_NODEV (-5)    /* no usable gfx950 device */

/* wavefield / PML-memory storage type; arithmetic is FP32 except F64 */
#define CGFD_STORE_F32 0
#define CGFD_STORE_F16 1
#define CGFD_STORE_BF16 2
#define CGFD_STORE_F64 3

/* dtype of host arrays handed to upload/fetch calls */
#define CGFD_DT_F32 0
#define CGFD_DT_F64 1

/* wavefield levels of one shot */
/* Four wavefield levels per shot.  A step stores the stage values
 * U_{i+1} = W^n + a_i dt K_i and forms W^{n+1} = W^n + w1 (U1 - W^n) +
 * w2 (U2 - W^n) + w3 (U3 - W^n) + b3 dt K3 (w = 1/3, 2/3, 1/3), the classical
 * RK4 step of drv_rk_curv_col.cu:193-542 with 117 instead of 144 level
 * words moved per point and step. */
#define CGFD_LEVEL_PRE 0   /* W^n: the level receivers and snapshots read */
#define CGFD_LEVEL_TMPA 1  /* after a step: W^{n-1}; during it U1 (stage value of stage 0) */
#define CGFD_LEVEL_TMPB 2  /* U2 of the last step */
#define CGFD_LEVEL_END 3   /* U3 of the last step */
#define CGFD_LEVEL_RHS 4   /* right-hand side written by cgfd_el_iso_onestage */

/* PML faces */
#define CGFD_FACE_X1 0
#define CGFD_FACE_X2 1
#define CGFD_FACE_Y1 2
#define CGFD_FACE_Y2 3
#define CGFD_FACE_Z1 4
#define CGFD_FACE_Z2 5

typedef struct cgfd_session cgfd_session;

typedef struct {
  int ni, nj, nk;   /* physical points per axis, PML layers included (gd_t ni/nj/nk) */
  int storage;      /* CGFD_STORE_* */
  int nshots;       /* independent shots sharing the model; 1 for a plain run */
  int flags;        /* CGFD_FLAG_* */
} cgfd_desc;

/* use the plain per-point kernel everywhere (no z-marching interior kernel) */
#define CGFD_FLAG_PLAIN 1
/* PML face slabs by the per-point kernel, the z-marching kernel only on the
 * box inside them (by default the z-marching kernel sweeps whole planes,
 * PML included, and the per-point kernel takes the 4 free-surface rows) */
#define CGFD_FLAG_PML_POINT 2
/* keep the per-point metric arrays even when the grid is Cartesian (by
 * default a metric whose components are constant to 1e-10 of their largest
 * magnitude runs the constant-metric sweep with the first point's values) */
#define CGFD_FLAG_FULL_METRIC 4

/* Point-source tables of one shot: the device mirror of src_t
 * (src/forward/src_t.h:14-53, filled by src_read_locate_file src_t.cu:143-790). */
typedef struct {
  int nsrc;                  /* point sources of this shot */
  int force_on, moment_on;   /* which tables are present */
  int max_nt;                /* STF samples per source */
  int max_ext;               /* spreading points per source (<= 343) */
  int dtype;                 /* CGFD_DT_* of ext_coef/F/M */
  const int *it_begin;       /* [nsrc] first time step the source acts */
  const int *it_end;         /* [nsrc] last time step */
  const int *ext_num;        /* [nsrc] valid spreading points */
  const long long *ext_indx; /* [nsrc*max_ext] flat index in reference layout */
  const void *ext_coef;      /* [nsrc*max_ext] spreading weights */
  const void *F;             /* [3][nsrc][max_nt][4] Fx,Fy,Fz per RK stage, or NULL */
  const void *M;             /* [6][nsrc][max_nt][4] Mxx,Myy,Mzz,Myz,Mxz,Mxy, or NULL */
} cgfd_source_desc;

/* ---- session ---------------------------------------------------------- */

/* Allocate a session on the current HIP device (device memory for nshots
 * wavefields of 4 levels each plus the shared model).  Returns CGFD_E_NOMEM
 * when free HBM is insufficient.  Replaces the device setup of
 * drv_rk_curv_col_allstep (src/forward/drv_rk_curv_col.cu:98-133). */
int cgfd_create(const cgfd_desc *desc, cgfd_session **out);

/* Free every device buffer (src/forward/drv_rk_curv_col.cu:544-556). */
void cgfd_destroy(cgfd_session *s);

/* Bytes of device memory a session with this descriptor would allocate. */
long long cgfd_footprint_bytes(const cgfd_desc *desc, int nrecv, int nt_total);

/* Thread-local description of the last error. */
const char *cgfd_last_error(void);

/* ---- model upload (shared by all shots) -------------------------------- */

/* 10 metric components in reference layout (init_metric_device, alloc.cu). */
int cgfd_upload_metric(cgfd_session *s, int dtype, const void *metric);

/* lambda, mu and 1/rho (after md_rho_to_slow, src/forward/md_t.cu:622-637). */
int cgfd_upload_medium(cgfd_session *s, int dtype, const void *lam, const void *mu, const void *slw);

/* Free-surface velocity-gradient conversion matrices at the top (k = nk2):
 * matVx2Vz, matVy2Vz of nx*ny*9 values [j][i][row][col]
 * (sv_curv_col_el_iso_dvh2dvz, src/forward/sv_curv_col_el_iso_gpu.cu:1152-1261).
 * free_top = 0 disables the free surface. */
int cgfd_upload_free_surface(cgfd_session *s, int free_top, int dtype, const void *matVx2Vz, const void *matVy2Vz);

/* CFS-PML profile of one face: A, B, D of nlay+1 values ordered along the
 * axis from low to high index (bdry_pml_set, src/forward/bdry_t.cu:85-257).
 * nlay = 0 disables the face. */
int cgfd_upload_pml(cgfd_session *s, int face, int nlay, int dtype, const void *A, const void *B, const void *D);

/* Per-component power-of-two storage scales for F16/BF16 sessions:
 * stored = value * 2^-exp2[c] for wavefield component c (and its PML memory). */
int cgfd_set_storage_scale(cgfd_session *s, const int exp2[9]);

/* ---- per shot ----------------------------------------------------------- */

int cgfd_upload_source(cgfd_session *s, int shot, const cgfd_source_desc *src);

/* Receivers of one shot: 8 trilinear corner indices (flat, reference layout,
 * order of iorecv_one_t.indx1d, src/forward/io_funcs.cu:227-234) and the
 * fractions di,dj,dk; nt_total samples are kept per component. */
int cgfd_upload_receivers(cgfd_session *s, int shot, int nrecv, const long long *indx8, int dtype,
                          const void *frac3, int nt_total);

/* Set one level of one shot from a host array (9 components, reference
 * layout; ghosts must be zero).  Levels not set are zero. */
int cgfd_set_level(cgfd_session *s, int shot, int level, int dtype, const void *w9);

/* Zero every level and PML memory of a shot and rewind its step counter. */
int cgfd_reset_shot(cgfd_session *s, int shot);

/* ---- time stepping --------------------------------------------------------- */

/* Advance every shot by nsteps full RK4 steps starting at step it0, with
 * receivers sampled from W^n before each step, as drv_rk_curv_col_allstep
 * (src/forward/drv_rk_curv_col.cu:193-542): step it uses operator pair
 * it % 8, stage times {0,1/2,1/2,1}*dt, and W^{n+1} becomes the PRE level.
 * dt is rounded to float for FP32/F16/BF16 sessions, as the reference keeps it. */
int cgfd_run(cgfd_session *s, int it0, int nsteps, double dt);

/* One RK-stage right-hand side into the RHS level (and the PML memory RHS):
 * the operator sv_curv_col_el_iso_onestage
 * (src/forward/sv_curv_col_el_iso_gpu.cu:23-285) reading level in_level. */
int cgfd_el_iso_onestage(cgfd_session *s, int shot, int in_level, int ipair, int istage, int it);

/* ---- results ---------------------------------------------------------------- */

/* Seismograms of one shot: [nrecv][9][nt_total] (iorecv_one_t.seismo). */
int cgfd_read_seismo(cgfd_session *s, int shot, int dtype, void *out);

/* One level of one shot back to a host array in reference layout (ghosts zero). */
int cgfd_fetch_level(cgfd_session *s, int shot, int level, int dtype, void *w9);

/* Strided sub-volume of one level, gathered on the device, for snapshots and
 * slices (io_snap_pack_buff, src/forward/io_funcs.cu:1568-1596): the points
 * (i1 + a*di, j1 + b*dj, k1 + c*dk), a < ni, b < nj, c < nk, in physical
 * 0-based indices, of every component whose bit is set in cmp_mask (bit c =
 * component c in the order Vx,Vy,Vz,Txx,Tyy,Tzz,Tyz,Txz,Txy), written to out
 * as [selected component][nk][nj][ni] in dtype.  level: PRE/TMPA/TMPB/END. */
int cgfd_fetch_subvolume(cgfd_session *s, int shot, int level, unsigned cmp_mask, int i1, int j1, int k1, int ni,
                         int nj, int nk, int di, int dj, int dk, int dtype, void *out);

/* cgfd_fetch_subvolume for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it has no counterpart in the reference.  Same selection and
 * output layout, but the gather stages through a fixed 8 MiB device array
 * the library holds per GPU (never hipMalloc/hipFree): one component of the
 * sub-volume (ni*nj*nk values of the model type) must fit in it, and dtype
 * must be the session's model type (CGFD_DT_F64 for FP64 sessions,
 * CGFD_DT_F32 otherwise); larger requests fail with CGFD_E_ARG.  Calls
 * that overlap on one GPU share that array and race. */
int cgfd_fetch_subvolume_noalloc(cgfd_session *s, int shot, int level, unsigned cmp_mask, int i1, int j1, int k1,
                                 int ni, int nj, int nk, int di, int dj, int dk, int dtype, void *out);

/* PML memory of one face: level = PRE/TMPA/TMPB/END or RHS; [9][face box]
 * with the face box indexed iz*(nj*ni)+iy*ni+ix (sv_curv_col_el_iso_gpu.cu:775). */
int cgfd_fetch_pml_aux(cgfd_session *s, int shot, int face, int level, int dtype, void *out);

/* Tile of the z-marching kernel: shape id (0..3: 64x4, 32x8, 64x8, 32x16
 * threads, one per column) and z planes per workgroup.  By default the first
 * cgfd_run times every candidate tile under each work split (below) and keeps
 * the fastest pair (cached per grid, storage, shots and flags for the
 * process); shape < 0 returns to that automatic choice.  A pinned tile runs
 * with the whole-plane split unless cgfd_set_split pinned another or
 * CGFD_FLAG_PML_POINT is set. */
int cgfd_set_tile(cgfd_session *s, int shape, int zchunk);
int cgfd_get_tile(cgfd_session *s, int *shape, int *tx, int *ty, int *zchunk);

/* Work split of a stage between the z-marching and the per-point kernel. */
#define CGFD_SPLIT_SHELL 0  /* sweep inside the PML face boxes; per-point kernel on the slabs and 4 surface rows */
#define CGFD_SPLIT_PLANES 1 /* sweep whole planes with the PML terms; per-point kernel on the 4 surface rows */
#define CGFD_SPLIT_CORE 2   /* whole planes in two sweeps: the PML-free form on a tile-aligned core inside the
                               face boxes, the PML form on the boxes around it; per-point kernel on the 4
                               surface rows */
int cgfd_get_split(cgfd_session *s, int *split);
/* Pin the work split (split < 0: the autotuner chooses again).  With an
 * automatic tile the autotuner then tunes the tile under that split only.
 * A tuning control with no counterpart in the reference, like cgfd_set_tile;
 * every split gives the same results to rounding. */
int cgfd_set_split(cgfd_session *s, int split);

/* 1 when the uploaded metric runs the constant-metric sweep (a Cartesian
 * grid: one value per metric component, off-diagonal ones zero; the three
 * diagonal values are kernel arguments instead of nine streams). */
int cgfd_get_const_metric(cgfd_session *s, int *const_metric);

/* Enable or disable (default) HIP-event timing of every interior kernel
 * launch.  Timed runs are enqueued launch by launch, without graphs. */
int cgfd_set_timing(cgfd_session *s, int on);

/* Peak ground motion on a free surface (PG_calcu_gpu,
 * src/forward/wav_t.cu:281-360, called after every step at
 * drv_rk_curv_col.cu:519-527): when enabled (and the top is free), every step
 * updates per shot the running maxima PGV, PGVh, PGVx, PGVy, PGVz, PGA, PGAh,
 * PGAx, PGAy, PGAz, PGD, PGDh, PGDx, PGDy, PGDz on the top plane, the
 * acceleration being |V^{n+1}-V^n|/dt and the displacement the trapezoid sum
 * of V.  Maps start at zero on enabling and on cgfd_reset_shot.
 * cgfd_read_peak_motion returns them as [15][ny][nx] with ghosts (zero), the
 * layout the reference writes (PG_slice_output, io_funcs.cu:2065-2125). */
int cgfd_set_peak_motion(cgfd_session *s, int on);
int cgfd_read_peak_motion(cgfd_session *s, int shot, int dtype, void *out);

/* Run whole 8-step periods (starting at a step that is a multiple of 8) as
 * one captured hipGraph each (default on); no reference counterpart, the
 * reference launches and synchronises per kernel. */
int cgfd_set_graph(cgfd_session *s, int on);

/* Reserve seismogram memory for up to max_nrecv receivers over max_nt steps
 * per shot, so that later cgfd_upload_receivers within these bounds never
 * reallocate device memory (no reference counterpart; used when shots are
 * replaced while others run). */
int cgfd_reserve_receivers(cgfd_session *s, int max_nrecv, int max_nt);

/* Seismogram ring: ring_steps > 0 (a multiple of 16) keeps only ring_steps
 * samples per trace on the device; runs are cut into chunks of ring_steps/2
 * steps and each chunk's samples are copied, on a second stream and
 * overlapping the next chunk, into a pinned host record from which
 * cgfd_read_seismo reads.  Device memory for seismograms is then
 * nrecv*9*ring_steps per shot whatever nt_total is (dense receiver arrays
 * over long runs).  0 keeps the whole record on the device (default).
 * Changing the mode discards samples recorded so far.  Replaces the
 * per-step, per-receiver cudaMemcpy of io_recv_keep / io_line_keep
 * (src/forward/io_funcs.cu:1625-1698). */
int cgfd_set_seismo_ring(cgfd_session *s, int ring_steps);

/* Continuous batching of shots (no reference counterpart).  A shot with
 * active = 1 advances with every later run, its own step number being the
 * session step minus it_start; its source and receivers use that number, and
 * it_start must be a multiple of 8 so that its operator pairs are those of a
 * run started at 0.  Inactive shots are skipped by every kernel.  Pair with
 * cgfd_reset_shot and cgfd_upload_source / cgfd_upload_receivers (which keep
 * the other shots' seismograms) to admit a new shot between runs. */
int cgfd_set_shot(cgfd_session *s, int shot, int active, int it_start);

/* cgfd_run without waiting: enqueue on the session stream and return.
 * cgfd_poll returns 0 when all enqueued work is done, 1 while it runs;
 * cgfd_wait blocks until it is done.  Timing must be off. */
int cgfd_run_async(cgfd_session *s, int it0, int nsteps, double dt);
int cgfd_poll(cgfd_session *s);
int cgfd_wait(cgfd_session *s);

/* Milliseconds of device time spent in the last cgfd_run, and the average
 * duration of its z-marching interior kernel launches (events on the
 * session stream). */
int cgfd_last_run_timing(cgfd_session *s, double *run_ms, double *interior_avg_ms, int *interior_launches);

/* Achievable HBM bandwidth of the current device (no reference counterpart;
 * the ceiling the sweep kernels are compared with): a streaming read and a
 * streaming copy of `bytes` bytes (two buffers allocated and freed here),
 * 16-byte non-temporal accesses over a grid covering every XCD, each timed
 * over `reps` launches.  read_gbps counts the bytes read, copy_gbps the
 * bytes read plus written, in 1e9 bytes per second. */
int cgfd_hbm_stream(long long bytes, int reps, double *read_gbps, double *copy_gbps);

/* The same read stream with 4-byte cached loads per lane, the access size of
 * the sweep kernels (no reference counterpart): the read ceiling at that
 * granularity, and a known-size dispatch for scaling HBM byte counters. */
int cgfd_hbm_stream_dword(long long bytes, int reps, double *read_gbps);

#ifdef __cplusplus
}
#endif
#endif
