/*
 * cgfd_io.h -- output formats of the el_iso path: SAC seismograms and
 * NetCDF-3 (classic) snapshots, with the reference's file names, variable
 * names and attributes.  Host-only; no device call.
 *
 * Return values: CGFD_OK or a negative CGFD_E_* code (cgfd_el_iso.h); the
 * message is in cgfd_last_error().
 */
#ifndef CGFD_IO_H
#define CGFD_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* One SAC (v6, binary, native byte order) trace: the header fields the
 * reference sets (sacHeaderSet, src/lib/sacLib.cu:251-341: delta, b, e =
 * b + delta*(npts-1), nvhdr 6, npts, iftype ITIME, leven, event and station
 * location, cmpaz = cmpinc = 0, kstnm, kcmpnm), every other field undefined
 * (-12345), followed by npts floats.  ev/st = {la, lo, el, dp}. */
int cgfd_sac_write(const char *path, const float *data, int npts, float delta, float b, const float *ev,
                   const float *st, const char *kstnm, const char *kcmpnm);

/* The SAC files of one receiver as io_recv_output_sac writes them
 * (src/forward/io_funcs.cu:1776-1818): {dir}/{name}.{Vx,Vy,Vz,Txx,Tyy,Tzz,
 * Tyz,Txz,Txy}.sac from seismo [9][nt], event location 0, station (x, y, z),
 * b = 0 and, as there, the station depth field carrying dt.  With
 * with_strain, also {name}.{Exx,Eyy,Ezz,Eyz,Exz,Exy}.sac converted with the
 * receiver's lambda and mu (io_recv_output_sac_el_iso_strain, :1822-1902). */
int cgfd_station_sac(const char *dir, const char *name, const float *seismo, int nt, float dt, float x, float y,
                     float z, int with_strain, float lam, float mu);

/* Snapshot file (io_snap_nc_create / io_snap_nc_put, io_funcs.cu:948-1250):
 * NetCDF classic, dimensions time (unlimited), k, j, i; float variables time
 * and, as selected, Vx Vy Vz / Txx Tyy Tzz Txz Tyz Txy / Exx Eyy Ezz Exz Eyz
 * Exy over [time][k][j][i]; global int attributes
 * first_index_to_snapshot_output, first_index_in_this_thread_with_ghosts,
 * index_stride_in_this_thread, coords_of_mpi_topo (3 ints each). */
typedef struct cgfd_snap cgfd_snap;

typedef struct {
  int ni, nj, nk;          /* points kept per axis */
  int out_vel, out_stress, out_strain;
  int first_global[3];     /* first_index_to_snapshot_output */
  int first_local[3];      /* first_index_in_this_thread_with_ghosts */
  int stride[3];           /* index_stride_in_this_thread */
  int topo[3];             /* coords_of_mpi_topo */
} cgfd_snap_desc;

int cgfd_snap_open(const char *path, const cgfd_snap_desc *d, cgfd_snap **out);

/* Append one time level: w holds the 9 components Vx,Vy,Vz,Txx,Tyy,Tzz,
 * Tyz,Txz,Txy, each [nk][nj][ni]; lam and mu ([nk][nj][ni]) are needed for
 * strain output only (else NULL). */
int cgfd_snap_append(cgfd_snap *sn, float time, const float *w, const float *lam, const float *mu);

/* Writes the record count and closes the file. */
int cgfd_snap_close(cgfd_snap *sn);

/* Slice file (io_slice_locate / io_slice_nc_create / io_slice_nc_put,
 * io_funcs.cu:418-784): one plane of the nine components per record, named
 * by the caller slicex_i{gi+1}_{part}.nc, slicey_j{gj+1}_{part}.nc or
 * slicez_k{gk+1}_{part}.nc (gi: 0-based global physical index).  NetCDF
 * classic; dimensions time (unlimited) and, by axis, (k, j) for an x slice,
 * (k, i) for y, (j, i) for z; float variables time and Vx Vy Vz Txx Tyy Tzz
 * Tyz Txz Txy over [time][slow][fast]; global int attributes
 * {i|j|k}_index_with_ghosts_in_this_thread (the plane's local index, ghosts
 * counted: gi + 3 on one subdomain) and coords_of_mpi_topo (3). */
typedef struct cgfd_slice cgfd_slice;

/* axis 0 (x), 1 (y) or 2 (z); n_slow / n_fast: the two in-plane sizes in
 * file order (x: nk, nj; y: nk, ni; z: nj, ni). */
int cgfd_slice_open(const char *path, int axis, int index_with_ghosts, int n_slow, int n_fast, const int *topo,
                    cgfd_slice **out);

/* Append one record: w holds the nine components in the order above, each
 * [n_slow][n_fast] (a plane from cgfd_fetch_subvolume with count 1 on the
 * slice axis). */
int cgfd_slice_append(cgfd_slice *sl, float time, const float *w);

/* Writes the record count and closes the file. */
int cgfd_slice_close(cgfd_slice *sl);

/* The SAC files of one point of a receiver line as io_line_output_sac writes
 * them (io_funcs.cu:1941-1980): {dir}/{line}.no{seq}.{Vx..Txy}.sac from
 * seismo [9][nt] (seq = the point's number along the line), event location
 * 0, station (x, y, z), the station depth field carrying dt as there. */
int cgfd_line_sac(const char *dir, const char *line, int seq, const float *seismo, int nt, float dt, float x,
                  float y, float z);

/* Peak ground motion file PG_V_A_D_{coords}.nc (PG_slice_output,
 * io_funcs.cu:2065-2125): dimensions j = ny, i = nx (with ghosts), float
 * variables PGV PGVh PGVx PGVy PGVz PGA PGAh PGAx PGAy PGAz PGD PGDh PGDx PGDy
 * PGDz from pg [15][ny][nx], global int attributes
 * global_index_of_first_physical_points (2), count_index_of_physical_points
 * (2), coords_of_mpi_topo (3). */
int cgfd_peak_motion_write(const char *path, int nx, int ny, const float *pg, const int *first_physical,
                           const int *count_physical, const int *topo);

#ifdef __cplusplus
}
#endif
#endif
