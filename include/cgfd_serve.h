/*
 * cgfd_serve.h -- shot-serving runtime over the el_iso sessions (no reference
 * counterpart: the reference runs one shot per process).
 *
 * A server keeps registered models resident on the device and answers shot
 * requests (a source, receivers and a step count on a model) as they arrive.
 * Requests on the same model share one batched session ("group") whose slots
 * are filled and emptied at 8-step boundaries while the others keep running
 * (continuous batching); groups of different models run on their own streams
 * and overlap on the device.  A group is created when its model's first
 * request is admitted, sized to what the HBM budget leaves, and idle groups
 * are released (least recently used first) when another model needs the
 * memory.  A shot's seismograms do not depend on what else is in flight.
 *
 * One worker thread drives all device work; the functions below are safe to
 * call from any thread.  Return values: CGFD_OK or a negative CGFD_E_* code
 * (cgfd_el_iso.h), message in cgfd_last_error() of the calling thread.
 */
#ifndef CGFD_SERVE_H
#define CGFD_SERVE_H

#include "cgfd_el_iso.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cgfd_server cgfd_server;

typedef struct {
  long long hbm_budget; /* device bytes the server may hold; 0: 90% of the free HBM at creation */
  int max_batch;        /* shots per group, 1..64 */
  int chunk_steps;      /* a positive multiple of 8: the longest a group with free slots runs between
                           admission points; a full group runs up to its first finishing shot, at
                           most 8 x chunk_steps */
  double group_share;   /* most of the budget one group may take (a group always gets one shot);
                           0: 1/4, so the groups of several models stay resident together */
} cgfd_server_desc;

/* A model: grid, medium, boundaries and time step, arrays in the reference
 * host layout of cgfd_el_iso.h (copied by cgfd_server_add_model). */
typedef struct {
  int ni, nj, nk;
  int storage;               /* CGFD_STORE_* of the wavefield */
  int dtype;                 /* CGFD_DT_* of every array below */
  const void *metric;        /* 10 components */
  const void *lam, *mu, *slw;
  int free_top;
  const void *matVx2Vz, *matVy2Vz; /* [ny][nx][9] each, or NULL without free surface */
  int pml_nlay[6];           /* 0: no PML on that face */
  const void *pml_A[6], *pml_B[6], *pml_D[6];
  const int *scale_exp2;     /* 9 storage scale exponents for F16/BF16, or NULL */
  double dt;
  int max_nrecv;             /* bounds of its requests, for memory accounting */
  int max_nt;
} cgfd_model_desc;

typedef struct {
  int model;                       /* id from cgfd_server_add_model */
  int nt_total;                    /* RK steps; receivers record before each */
  const cgfd_source_desc *src;     /* copied at submission */
  int nrecv;
  const long long *recv_indx8;     /* [nrecv][8] */
  int recv_dtype;
  const void *recv_frac3;          /* [nrecv][3] */
} cgfd_request;

typedef struct {
  long long submitted, admitted, completed, rejected;
  int groups;                      /* resident groups */
  int active_shots;                /* shots on the device now */
  int queued;                      /* waiting for a slot or memory */
  long long hbm_in_use;            /* bytes of the resident groups */
  long long hbm_budget;
} cgfd_server_stats;

int cgfd_server_create(const cgfd_server_desc *desc, cgfd_server **out);

/* Waits for the work in flight, then frees every group. */
void cgfd_server_destroy(cgfd_server *srv);

int cgfd_server_add_model(cgfd_server *srv, const cgfd_model_desc *model, int *model_id);

/* Queue a request; *ticket identifies it. */
int cgfd_server_submit(cgfd_server *srv, const cgfd_request *req, long long *ticket);

/* Block until the request is done or timeout_s passes (< 0: no limit).
 * Returns CGFD_OK when done, 1 on time-out, or an error (a rejected request
 * returns CGFD_E_NOMEM with the reason). */
int cgfd_server_wait(cgfd_server *srv, long long ticket, double timeout_s);

/* Seismograms of a finished request ([nrecv][9][nt_total] in dtype) and its
 * times in seconds since submission: admitted, finished.  Releases it. */
int cgfd_server_result(cgfd_server *srv, long long ticket, int dtype, void *seismo, double *t_admit,
                       double *t_done);

int cgfd_server_get_stats(cgfd_server *srv, cgfd_server_stats *st);

#ifdef __cplusplus
}
#endif
#endif
