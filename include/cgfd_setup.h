/*
 * cgfd_setup.h -- host-side preparation of the inputs of the el_iso path.
 *
 * These functions replace the reference's setup code that feeds
 * drv_rk_curv_col_allstep: grid generation and metrics (gd_t), media
 * (md_t / src/media), CFS-PML profiles (bdry_t), free-surface conversion
 * matrices (sv_curv_col_el_iso_dvh2dvz), source tables (src_t) and receiver
 * location (io_funcs).  All arithmetic is FP64; arrays are double and use the
 * reference host layout described in cgfd_el_iso.h (nx = ni+6, i fastest).
 * Return values: CGFD_OK (0) or a negative CGFD_E_* code (cgfd_el_iso.h).
 */
#ifndef CGFD_SETUP_H
#define CGFD_SETUP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- grids --------------------------------------------------------------- */

/* Cartesian grid with ghosts, x0/y0/z0 = coordinate of physical point 0
 * (gd_curv_gen_cart, src/forward/gd_t.cu:599-630). */
int cgfd_grid_cartesian(int ni, int nj, int nk, double dx, double x0, double dy, double y0, double dz, double z0,
                        double *x, double *y, double *z);

/* Curvilinear grid under a topographic free surface (no reference
 * counterpart: the reference imports such grids).  Surface
 * z(x,y) = sum of nhill Gaussian hills + a Fourier perturbation of given rms
 * made of modes with wavelength >= min_wavelength; a flat bottom; vertical
 * spacing stretched linearly from 1 (top) to stretch (bottom) times dh.
 * Hill centres/amplitudes/widths and the Fourier modes come from seed. */
typedef struct {
  int ni, nj, nk;
  double dh;               /* horizontal spacing and top vertical spacing [m] */
  double stretch;          /* bottom/top vertical spacing ratio (e.g. 1.5) */
  int nhill;               /* Gaussian hills */
  double amp_min, amp_max; /* hill amplitudes [m] */
  double sig_min, sig_max; /* hill widths (sigma) [m] */
  double rough_rms;        /* rms of the Fourier perturbation [m] */
  double min_wavelength;   /* shortest perturbation wavelength [m] */
  unsigned long long seed;
} cgfd_topo_desc;

int cgfd_grid_topography(const cgfd_topo_desc *d, double *x, double *y, double *z);

/* Metrics and Jacobian with the centred 7-point operator and the
 * geometric ghost extrapolation, in FP64 (gd_curv_metric_cal +
 * geometric_symmetry, src/forward/gd_t.cu:209-314, 404-498).
 * metric: 10 components jac, xi_x..zeta_z. */
int cgfd_metric_calc(int ni, int nj, int nk, const double *x, const double *y, const double *z, double *metric);

/* Largest stable dt: min over cells of CFL * (distance to the 8 corner
 * planes) / Vp (blk_dt_esti_curv, src/forward/blk_t.cu:673-763). rho is the
 * density (not yet inverted).  Returns dtmax > 0, or a negative error code. */
double cgfd_dt_estimate(int ni, int nj, int nk, const double *x, const double *y, const double *z,
                        const double *lam, const double *mu, const double *rho, double cfl);

/* ---- media ----------------------------------------------------------------- */

/* Flat layers below interfaces at absolute depths z_top_of_layer[l] (layer 0
 * is the top layer, unbounded above); Vs = Vp/vpvs; rho from Gardner
 * rho = 310*Vp^0.25 (kg/m^3) when rho_l <= 0; each layer gets a Gaussian-
 * correlated relative Vp/Vs perturbation of std pert_std clipped at
 * +-pert_clip, correlation length corr_len [m], seeded per layer.  Values
 * are set at every point, ghosts included. */
typedef struct {
  int nlayer;
  const double *z_interface; /* nlayer-1 interface depths (z, decreasing) */
  const double *vp;          /* nlayer */
  const double *vpvs;        /* nlayer, Vp/Vs ratio */
  const double *rho;         /* nlayer, <= 0 selects Gardner */
  double pert_std, pert_clip, corr_len;
  unsigned long long seed;
} cgfd_layer_desc;

int cgfd_media_layered(int ni, int nj, int nk, const double *x, const double *y, const double *z,
                       const cgfd_layer_desc *d, double *lam, double *mu, double *rho);

/* Medium from a layer file in the reference's md3lay format, media_type
 * elastic_isotropic (read_interface_file, src/media/media_read_file.cpp:22-260):
 * after '#' comment lines, "elastic_isotropic", the number of interfaces NI,
 * "NX NY MINX MINY DX DY", then NI x NY x NX rows of
 * "elevation rho rho_grad rho_pow vp vp_grad vp_pow vs vs_grad vs_pow",
 * interfaces listed from the top.  Method "loc" takes each grid point's own
 * layer: v = v0 + dz^pow * grad with dz the depth below the layer's top
 * interface (media_layer2model.cpp:401-533, 1313-1362); "har" and "ari" then
 * replace, at points whose dual cell (corners at the 8 surrounding cell
 * centres) meets more than one layer, the values by the average over the 9^3
 * points subdividing that cell: harmonic for kappa and mu with arithmetic rho
 * ("har"), or arithmetic for lam, mu and rho ("ari")
 * (media_layer2model.cpp:1494-1580, 1963-2174).  All in float as the
 * reference.  Reference quirk kept: a point within 1 m (in z) of an interface
 * counts as on it (its isEqual binds C's int abs) and takes the mean of the
 * two layers.  Fills lam, mu, rho at every point, ghosts included.
 * CGFD_E_ARG (message in cgfd_last_error) for an unreadable file, a point
 * outside the interface mesh or vp <= sqrt(2) vs. */
int cgfd_media_layer_file(int ni, int nj, int nk, const double *x, const double *y, const double *z, const char *path,
                          const char *method, double *lam, double *mu, double *rho);

/* Medium from a grid media file in the reference's md3grd format, media_type
 * elastic_isotropic (read_grid_file, src/media/media_read_file.cpp:670-925,
 * 1243-1266): "elastic_isotropic", the number of layers NL, the number of
 * surfaces in each layer, "NX NY MINX MINY DX DY", then (sum of the surface
 * counts) x NY x NX rows of "elevation rho vp vs", surfaces listed from the
 * top; the last surface of a layer must equal the first of the next.  Only
 * the window of the mesh covering the grid's x, y extent is kept, as there.
 * Method "loc": rho, vp, vs linear in z between the two surfaces around each
 * point, the top values above the first surface and the last surface's below
 * it (media_grid2model.cpp:424-560, 1067-1110); "har" / "ari" average over the
 * 9^3 points subdividing each dual cell that meets more than one layer, as
 * cgfd_media_layer_file, with layers numbered by their top surfaces and the
 * sub-points on a layer boundary left out (media_grid2model.cpp:1269-1355,
 * 1819-2079); with one layer they give "loc".  Reference quirks kept: a point
 * within 1 m under the top surface of a lower layer takes the mean of that
 * surface and the one above; a point on the window's last mesh row or column
 * takes the value of the node at its truncated index, without interpolation
 * along the other axis (BilinearInterpolation, media_utility.cpp:145-182).
 * Unlike the reference, a mesh whose window would
 * reach past its last node is rejected rather than read uninitialised.
 * CGFD_E_ARG (message in cgfd_last_error) for an unreadable file, a mesh that
 * does not cover the grid or vp <= sqrt(2) vs. */
int cgfd_media_grid_file(int ni, int nj, int nk, const double *x, const double *y, const double *z, const char *path,
                         const char *method, double *lam, double *mu, double *rho);

/* Medium from three raw binary volumes of float32 rho, vp and vs
 * (media_bin2model_el_iso, src/media/media_bin2model.cpp:27-230, with
 * read_bin_file, media_read_file.cpp:1274-1320, and TrilinearInterpolation,
 * media_utility.cpp:189-284; the driver's "binfile" medium,
 * main_curv_col_el_3d.cu:439-470).  File dimension d (d = 0 fastest) has
 * size[d] nodes at origin[d] + n spacing[d] and holds axis order[d]
 * (0 = x, 1 = y, 2 = z), as the parameter file's dim1..dim3.  Only the window
 * of file nodes covering the grid's x, y extent is kept.  Each grid point
 * takes the trilinear interpolation of the 8 nodes around it, in float with
 * the reference's order; outside the file a coordinate clamps to the edge
 * node.  lam = vp^2 rho - 2 mu, mu = vs^2 rho.  CGFD_E_ARG (message in
 * cgfd_last_error) for a bad order, a model that does not meet the grid, a
 * covering window of one node on an axis (the reference reads past its
 * mesh there), or a missing or short file. */
int cgfd_media_bin_file(int ni, int nj, int nk, const double *x, const double *y, const double *z,
                        const int order[3], const int size[3], const double spacing[3], const double origin[3],
                        const char *rho_path, const char *vp_path, const char *vs_path, double *lam, double *mu,
                        double *rho);

/* Homogeneous medium at every point. */
int cgfd_media_homogeneous(int ni, int nj, int nk, double vp, double vs, double rho, double *lam, double *mu,
                           double *rhoa);

/* rho -> 1/rho (0 where rho <= 1e-10) (md_rho_to_slow, src/forward/md_t.cu:622-637). */
int cgfd_rho_to_slow(long long n, const double *rho, double *slw);

/* ---- boundaries -------------------------------------------------------------- */

/* Mean spacing dh and layer thickness L0 = dh*nlay across the box of PML
 * face `face` (0..5 = x1,x2,y1,y2,z1,z2) with nlay layers
 * (bdry_cal_abl_len_dh, src/forward/bdry_t.cu:334-451). */
int cgfd_pml_len_dh(int ni, int nj, int nk, const double *x, const double *y, const double *z, int face, int nlay,
                    double *L0, double *dh);

/* CFS-PML profile of one face: A = alpha + d/beta, B = 1/beta, D = d/beta
 * with d = d0 (r)^2, alpha = amax (1-r), beta = 1 + (bmax-1) r^2, r = L/L0,
 * d0 = -3 vel ln R / (2 L0), R = 10^-((log10 nlay - 1)/log10 2 + 4); ordered
 * from low to high index (bdry_pml_set, src/forward/bdry_t.cu:193-241, 297-330). */
int cgfd_pml_profile(int nlay, int face, double L0, double dh, double vel, double amax, double bmax, double *A,
                     double *B, double *D);

/* Zero-traction velocity-gradient conversion at the top surface:
 * matVx2Vz = A^-1 B, matVy2Vz = A^-1 C per (i,j), nx*ny*9 values
 * (sv_curv_col_el_iso_dvh2dvz, src/forward/sv_curv_col_el_iso_gpu.cu:1152-1261). */
int cgfd_free_surface_matrices(int ni, int nj, int nk, const double *metric, const double *lam, const double *mu,
                               double *matVx2Vz, double *matVy2Vz);

/* ---- sources --------------------------------------------------------------------- */

#define CGFD_WAVELET_RICKER 0
#define CGFD_WAVELET_GAUSSIAN 1
#define CGFD_WAVELET_RICKER_DERIV 2
#define CGFD_WAVELET_GAUSSIAN_DERIV 3
#define CGFD_WAVELET_KLAUDER 4 /* fixed 0.2 s, 1.5-41.5 Hz, 20 s sweep (fun_klauder, src_t.cu:869-876) */

/* Wavelet value (fun_ricker & co, src/forward/src_t.cu:834-878). */
double cgfd_wavelet(int kind, double t, double c0, double c1);

/* Moment tensor (Mxx,Myy,Mzz,Myz,Mxz,Mxy; x east, y north, z up) of a unit
 * double couple (angle2moment, src/forward/src_t.cu:910-950). */
int cgfd_angle2moment(double strike, double dip, double rake, double *m6);

/* One point source given by 0-based global grid index (sx,sy,sz); depth_flag
 * = 1 counts sz down from the surface (k = nk-1-sz).  Fills the spreading
 * (Gaussian of width 1.5 over 7^3 points, normalised before clipping to the
 * physical grid) and the per-(step, RK stage) wavelet table, as
 * src_read_locate_file does for index input and named wavelets
 * (src/forward/src_t.cu:296-330, 596-712, 796-827).
 * ext_indx/ext_coef hold 343 entries; stf holds max_nt*4 values (may be NULL). */
int cgfd_source_point(int ni, int nj, int nk, double sx, double sy, double sz, int depth_flag, double t0, double dt,
                      double wavelet_tstart, int wavelet, double c0, double c1, double stf_length, int *it_begin,
                      int *it_end, int *ext_num, long long *ext_indx, double *ext_coef, int *max_nt,
                      double *stf);

/* What a source file holds (src_read_locate_file, src/forward/src_t.cu:143-795). */
typedef struct {
  char event[256];     /* event name: the first line */
  int nsrc_file;       /* sources the file lists */
  int nsrc;            /* sources kept: located inside the physical grid */
  int stf_given;       /* time function by wavelet name (0) or by sampled values (1) */
  double stf_length;   /* by name: length in seconds */
  double stf_dt;       /* by values: sample interval */
  int stf_nt;          /* by values: samples */
  int cmp_type;        /* 1 force, 2 moment, 3 force and moment (the reference then activates the force only) */
  int mechanism;       /* moment as a tensor (0) or strike, dip, rake, mu, D, A (1) */
  int by_coord;        /* locations by grid index (0) or by coordinates (1) */
  int depth;           /* the third location value is a depth below the free surface (1) */
  int force_on, moment_on;
  int max_nt;          /* time-function samples per source (the tables' step count) */
} cgfd_source_file_info;

/* Source file (src_read_locate_file, src/forward/src_t.cu:143-795; lines read as
 * io_get_nextline does): event name, source count, time-function line ("0
 * length" or "1 dt nt"), component line ("cmp_type mechanism"), location
 * line ("by_coord depth"), one location per source, then per source either
 * "t_start wavelet_name coef..." and one component line, or "t_start" and
 * nt component lines.  Locations by index are rounded to the nearest node
 * (0-based, depth counted down from the surface); by coordinates they are
 * located as cgfd_locate_points does, keeping the nearest node and a signed
 * shift.  Component values missing from a line keep the previous source's
 * values (zero at first), as in the reference; strike/dip/rake/mu/D/A become
 * the tensor mu D A * cgfd_angle2moment (mu < 0: the model mu at the source
 * node, from `mu` in the reference layout).  Sampled time functions are
 * interpolated with the reference's piecewise cubic Lagrange rule
 * (LagInterp_Piecewise_1d, src/forward/interp.cu:220-298).
 * Call with all table pointers NULL to get *info (sizes); then with arrays
 * of info->nsrc entries (it_begin, it_end, ext_num), nsrc*343 (ext_indx,
 * ext_coef), 3*nsrc*max_nt*4 (F, when force_on) and 6*nsrc*max_nt*4 (M, when
 * moment_on), the layout of cgfd_source_desc.  node[nsrc][3] (0-based
 * physical node) and shift[nsrc][3] may be NULL.  x, y, z are needed for
 * locations by coordinates only. */
int cgfd_source_file(const char *path, int ni, int nj, int nk, const double *x, const double *y, const double *z,
                     const double *mu, double t0, double dt, cgfd_source_file_info *info, int *it_begin,
                     int *it_end, int *ext_num, long long *ext_indx, double *ext_coef, double *F, double *M,
                     int *node, double *shift);

/* ---- receivers ------------------------------------------------------------------- */

/* Points given by physical coordinates (sources and stations located by
 * coordinates: src_read_locate_file / io_recv_read_locate with
 * gd_curv_depth_to_axis and gd_curv_coord_to_glob_indx, gd_t.cu:1190-1486,
 * 1574-1650, 1767-1800; io_funcs.cu:100-170).  pts[n][3] = x, y and z, or
 * depth below the free surface where depth_flag[q] is 1 (depth_flag may be
 * NULL).  For each point found inside the physical grid: idx[q][3] = the
 * physical 0-based index of the lower cell corner and frac[q][3] in [0,1)
 * its fractional position (resolution 0.01, from the reference's 101^3
 * sampling of the cell), found[q] = 1; otherwise found[q] = 0 and idx =
 * -1000.  Computed in float on float copies of x, y, z, as the reference's
 * grid arrays are float. */
int cgfd_locate_points(int ni, int nj, int nk, const double *x, const double *y, const double *z, int n,
                       const double *pts, const int *depth_flag, int *idx, double *frac, int *found);

/* Receiver line (io_line_locate, io_funcs.cu:302-417): the points start +
 * ipt*incre, ipt < count, in 0-based global physical indices; those inside
 * the grid are kept in line order with their number ipt in seq, as on-node
 * receivers (8 corner offsets in indx8, fractions 0 in frac3: recording
 * returns the node value, as io_line_keep does) and, when the grid arrays
 * are given, their coordinates in xyz (3 per point; may be NULL).  Arrays
 * hold count entries at most.  Returns the number kept, or CGFD_E_ARG. */
int cgfd_line_points(int ni, int nj, int nk, const double *x, const double *y, const double *z, const int *start,
                     const int *incre, int count, int *seq, long long *indx8, double *frac3, double *xyz);

/* Receiver given by grid index as in a station file (io_funcs.cu:162-251):
 * x,y 1-based; z depth below the surface (depth_flag 1) or 1-based index.
 * Fills the 8 trilinear corner indices and the fractions; returns
 * CGFD_E_ARG when the point lies outside the grid. */
int cgfd_receiver_index(int ni, int nj, int nk, double rx, double ry, double rz, int depth_flag, long long *indx8,
                        double *frac3);

/* One station of a station file, as io_recv_read_locate reads and locates
 * it (src/forward/io_funcs.cu:24-299). */
typedef struct {
  char name[64];      /* first field of the line (longer names are cut) */
  int by_coord;       /* file flag: 1 physical coordinates, 0 grid index */
  int depth;          /* file flag: 1 the third value is a depth below the free surface */
  double in[3];       /* the three values as read (float precision, as %g into float) */
  int inside;         /* 1: located inside the physical grid (kept by the reference) */
  long long indx8[8]; /* trilinear corner offsets in the reference host layout */
  double frac[3];     /* fractional position in the lower cell */
  double xyz[3];      /* station coordinates: the located point (depth turned into z), or for
                         grid-index input the grid node at the lower corner; 0 without x, y, z */
} cgfd_station;

/* Station file (io_recv_read_locate, src/forward/io_funcs.cu:24-299, lines read
 * as io_get_nextline does, io_funcs.cu:2131-2154: lines starting with '#' and
 * empty lines skipped): the number of stations, then per station
 * "name by_coord depth v1 v2 v3".  Grid-index stations are located as
 * cgfd_receiver_index does, coordinate stations as cgfd_locate_points (x, y,
 * z required when the file has any).  Fills out[0 .. min(count, max_n)) in
 * file order; *n_file receives the count the file states.  Returns the
 * number filled, or a negative CGFD_E_* (missing file, bad line). */
int cgfd_station_file(const char *path, int ni, int nj, int nk, const double *x, const double *y, const double *z,
                      int max_n, cgfd_station *out, int *n_file);

/* ---- reduced-precision storage -------------------------------------------------- */

/* Power-of-two scale exponents (for cgfd_set_storage_scale) that keep the
 * wavefield of a reduced-precision session inside the format's range (no
 * reference counterpart).  The bound is physical: a stress component cannot
 * exceed the total stress the moment sources inject (sum over source points
 * of coef/J times the time integral of |M|) plus what the force sources'
 * velocity times the largest impedance rho*Vp gives; a velocity cannot exceed
 * that stress over the smallest shear impedance rho*Vs plus the injected
 * force velocity.  FP16 maps the bound to 2^12 (16x headroom below 65504);
 * BF16, FP32 and FP64 keep exponent 0.  src may hold several sources; pass
 * the shot with the strongest sources when shots share one session. */
int cgfd_storage_scales(int storage, int ni, int nj, int nk, const double *metric, const double *lam,
                        const double *mu, const double *slw, const cgfd_source_desc *src, double dt, int *exp2);

#ifdef __cplusplus
}
#endif
#endif
