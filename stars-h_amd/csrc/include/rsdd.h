/* Internal C interface between the host drivers and the HIP pipeline of the
 * randomized SVD (one batch of tiles). */
#ifndef STARSH_RSDD_H
#define STARSH_RSDD_H

#include <stdint.h>
#include "starsh-hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One problem of a batch: its kernel, points and accuracy. */
typedef struct rsdd_problem
{
    const starsh_hip_kernel_desc *kd;
    const double *d_xr, *d_xc;   /* points in cluster order (device, SoA) */
    int64_t nr, nc;
    int maxrank, oversample;
    double tol;
} RsddProblem;

typedef struct rsdd_input
{
    int ntiles;
    const int *m, *n;            /* tile sizes */
    /* Several problems (all of one kernel family and dimension): tile t
     * belongs to probs[prob[t]].  When probs is NULL the single problem is
     * given by the fields kd ... tol below. */
    int nprob;
    const RsddProblem *probs;
    const int *prob;
    /* tile contents: either generated ... */
    const starsh_hip_kernel_desc *kd;
    const double *d_xr, *d_xc;   /* points in cluster order (device, SoA) */
    int64_t nr, nc;
    const int64_t *r0, *c0;      /* first row / column point of each tile */
    /* ... or explicit (column-major, ld = m) */
    const double *d_a;
    const int64_t *a_off;
    int maxrank, oversample;
    double tol;
    const int *k;                /* optional per-tile sketch width (default
                                    min(maxrank+oversample, m, n)) */
    int sketch_int8;             /* 1: int8-digit sketch (mixed form) */
    double eta;                  /* truncation of B's rows: discarded
                                    mass at most (eta tol)^2 |B|_F^2
                                    (0: 1e-5) */
    double tau;                  /* deflation threshold of the sketch
                                    columns (0: 1e-13) */
    double *S_out;               /* optional: k singular values per tile */
    const int64_t *soff;
    int *rows_out;               /* optional: sketch width k per tile */
    /* Exact engines (SDD, RRQR): RSDD_RULE_HALF also declares a tile
     * false far-field when 2 rank >= min(m, n) (dsdd.c:53, dqp3.c:84);
     * RSDD_RULE_SELECT replaces the random sketch by the k columns
     * sel[sel_off[t] ...] of the tile (Y = A[:, sel]) and keeps
     * orthogonalising after a fully dropped panel. */
    int rule;
    const int *sel;
    const int64_t *sel_off;
    /* Optional: for every tile with more than piv_k[t] > 0 kept sketch
     * columns, the first piv_k[t] column pivots of a column-pivoted QR of
     * B = Q^T A (greedy largest residual norm, the pivot order of
     * dgeqp3), written to piv[piv_off[t] ...]; piv_done[t] = 1 when they
     * were computed. */
    const int *piv_k;
    int *piv;
    const int64_t *piv_off;
    int *piv_done;
} RsddInput;

#define RSDD_RULE_HALF 1
#define RSDD_RULE_SELECT 2

typedef struct rsdd_output
{
    int *rank;                   /* ntiles: rank, -1 for a false far tile */
    int *p, *kc, *flags;         /* ntiles: diagnostics */
    int64_t *uoff, *voff;        /* ntiles: offsets of U, V in d_uv */
    double *d_uv;                /* device buffer with U and V, owned by
                                    the pipeline and valid until its next
                                    call */
    double device_seconds;
    double sketch_seconds;       /* sketch kernel (HIP events) */
    double proj_seconds;         /* projection kernels (HIP events) */
} RsddOutput;

int rsdd_batch(const RsddInput *in, RsddOutput *out);

/* Copy-back of the factors of one API call, overlapped with the
 * compression of later batches (csrc/hip/copyout.hip).
 *   rsdd_out_begin   reserves the device arena for the call (at most
 *                    max_doubles doubles of factors in max_segments
 *                    segments) and starts the copy threads;
 *   rsdd_out_append  packs U_t then V_t of every tile t < ntiles of the
 *                    last rsdd_batch with take[t] != 0 and rank > 0 into
 *                    the arena and queues their copy to the host buffers
 *                    dst_u[t] (m x rank) and dst_v[t] (n x rank), which
 *                    must stay valid until rsdd_out_end; returns without
 *                    waiting;
 *   rsdd_out_end     waits for every queued copy and stops the threads;
 *                    adds the copy window (first queued batch to last
 *                    copy) and the wait inside this call to the two
 *                    counters; returns the first error.  Must follow every
 *                    successful rsdd_out_begin, also on error paths. */
int rsdd_out_begin(int64_t max_doubles, int64_t max_segments);
int rsdd_out_append(const RsddOutput *out, int ntiles, const int *m,
        const int *n, const int *take, double *const *dst_u,
        double *const *dst_v);
/*   rsdd_out_submit  queues the copy of len doubles of device memory d_src
 *                    (filled by work already queued on the library stream,
 *                    and kept alive until rsdd_out_end) to host: nseg
 *                    segments, in increasing order, d_src[off[q] ..
 *                    off[q] + seglen[q]) to dst[q]. */
int rsdd_out_submit(const double *d_src, int64_t len, int nseg,
        const int64_t *off, const int64_t *seglen, double *const *dst);
int rsdd_out_end(double *window_seconds, double *tail_seconds);
void rsdd_out_release(void);

/* Frees the workspace and staging buffers kept between calls. */
void rsdd_release_pools(void);

/* Per tile |A|_F^2 and |A - U V^T|_F^2 with A regenerated on the GPU. */
int starsh_hip_tile_errors(const starsh_hip_kernel_desc *kd,
        const double *d_xr, int64_t nr, const double *d_xc, int64_t nc,
        int ntiles, const int64_t *r0, const int64_t *c0, const int *m,
        const int *n, const int *rank, const double *d_U,
        const int64_t *uoff, const double *d_V, const int64_t *voff,
        double *norm2, double *diff2);

#ifdef __cplusplus
}
#endif

#endif
