/* Internal: pieces shared by the plain and mixed-precision HIP drivers. */
#ifndef STARSH_HIP_DRIVER_H
#define STARSH_HIP_DRIVER_H

#include "starsh.h"
#include "starsh-hip.h"

/* Largest tile side and sketch width the device pipeline takes. */
#define HIP_TILE_MAX 4096
#define HIP_SKETCH_MAX 1024

typedef struct hip_problem
{
    starsh_hip_kernel_desc kd;
    double *d_xr, *d_xc;   /* points in cluster order on the device */
    STARSH_int nr, nc;
    int oversample;
} hip_problem;

typedef struct hip_factors
{
    Array **far_U, **far_V;
    int *far_rank;
    double *alloc_U, *alloc_V;
    int64_t *uoff, *voff;  /* offsets of each tile's U and V in alloc_U /
                              alloc_V (-1: none) */
    int64_t ucap, vcap;    /* doubles of alloc_U / alloc_V */
    int64_t uused, vused;  /* doubles taken so far */
    STARSH_int nfar;       /* entries of far_U, far_V and far_rank */
} hip_factors;

int hip_driver_prepare(STARSH_blrf *F, int maxrank, double tol, int onfly,
        hip_problem *hp);
void hip_driver_release(hip_problem *hp);
int hip_driver_alloc_factors(STARSH_blrf *F, int maxrank, hip_factors *hf);
void hip_driver_free_factors(STARSH_blrf *F, hip_factors *hf);
/* Near-field tiles of a problem generated on the device (d, n doubles)
 * and their host buffer (h, owned by the returned matrix); the caller
 * queues the copy on the copy-back engine, waits for it and frees d. */
typedef struct hip_near
{
    double *d, *h;
    int64_t n;
} hip_near;

/* Steps 3-5 (format rewrite, near-field tiles, STARSH_blrm). */
int hip_driver_finish(STARSH_blrm **matrix, STARSH_blrf *format,
        const hip_problem *hp, hip_factors *hf, int onfly,
        starsh_hip_stats *st, hip_near *nf);
void hip_driver_set_stats(const starsh_hip_stats *stats);
/* Makes the driver step numbered `site` fail once with STARSH_MALLOC_ERROR
 * (1: false far-field lists, 2: near-field lists, 3: near-field device
 * buffer, 4: format rebuild); 0 clears.  For tests of the error paths. */
int hip_driver_inject_failure(int site);

/* Several problems compressed together (the single-problem drivers use
 * nprob = 1). */
typedef struct hip_batch
{
    int nprob;
    hip_problem *hp;        /* [nprob] */
    STARSH_blrf **F;        /* [nprob] */
    const int *maxrank;     /* [nprob] */
    const double *tol;      /* [nprob] */
    hip_factors *hf;        /* [nprob] */
} hip_batch;

/* Decides whether a tile's result is final: (rank, sketch width used, full
 * sketch width of the plain form, pipeline flags, oversample, kept sketch
 * columns, min(m, n)). */
typedef int (*hip_accept_fn)(int rank, int k, int kfull, int flags,
        int oversample, int kc, int mn);

/* Inputs of the exact engines for hip_driver_run (per tile of the call):
 * rule (RSDD_RULE_* bits), selected columns sel[sel_off[t] ...] (k[t] of
 * them, RSDD_RULE_SELECT), and pivots: the first piv_k[t] column pivots
 * of B for tiles with more kept columns go to piv[piv_off[t] ...] with
 * piv_done[t] = 1.  Offsets increase with t. */
typedef struct hip_run_extra
{
    int rule;
    const int *sel;
    const int64_t *sel_off;
    const int *piv_k;
    int *piv;
    const int64_t *piv_off;
    int *piv_done;
    int *kc_out;            /* kept sketch columns per tile, or NULL */
} hip_run_extra;

/* Compresses far-field tiles idx[t] of problems pidx[t] (t < ntiles, all of
 * one kernel family and dimension) with sketch widths k[] in
 * workspace-sized chunks; accepted results (all when accept is NULL) go to
 * B->hf[pidx[t]] (rank and U/V copied to the host buffers).  When S_out is
 * not NULL the k[t] singular values of tile t (those of B, then the norm
 * of the truncated rows, then zeros) go to S_out + t*sstride. */
int hip_driver_run(const hip_batch *B, STARSH_int ntiles, const int *pidx,
        const STARSH_int *idx, const int *k, int int8, double tau, double eta,
        int *rank_out, int *flags_out, int *accepted, hip_accept_fn accept,
        double *S_out, int sstride, starsh_hip_stats *st,
        const hip_run_extra *x);
void hip_driver_rank_stats(const hip_batch *B, starsh_hip_stats *st);
/* Bound on the factor doubles and segments of a batch of problems. */
void hip_driver_factor_bound(const hip_batch *B, int64_t *doubles,
        int64_t *segments);
/* U/V wrappers of every tile after the copy-back has finished. */
int hip_driver_wrap(const hip_batch *B);
/* Tiles per pipeline batch for a run of ntiles tiles. */
STARSH_int hip_driver_chunk_tiles(STARSH_int ntiles);

/* The plain and mixed forms over a list of tiles of a batch. */
int hip_plain_tiles(const hip_batch *B, STARSH_int ntiles, const int *pidx,
        const STARSH_int *idx, starsh_hip_stats *st);
int hip_mixed_tiles(const hip_batch *B, STARSH_int ntiles, const int *pidx,
        const STARSH_int *idx, starsh_hip_stats *st);
/* The exact engines (exact_hip.c): SVD (rrqr == 0) or rank-revealing QR. */
int hip_exact_tiles(const hip_batch *B, STARSH_int ntiles, const int *pidx,
        const STARSH_int *idx, int rrqr, starsh_hip_stats *st);

/* Engines of hip_compress. */
enum { HIP_ENGINE_RSVD = 0, HIP_ENGINE_MIXED = 1, HIP_ENGINE_SDD = 2,
    HIP_ENGINE_RRQR = 3 };

/* Whole compression of nprob problems (plain form when mixed == 0):
 * prepare, compress every far-field tile, finish each problem. */
int hip_compress(int mixed, int nprob, STARSH_blrm **matrix,
        STARSH_blrf **format, const int *maxrank, const double *tol,
        int onfly);

/* First rank bucket (index into 32, 64, 128, 256, 512) of the mixed form. */
int hip_mixed_set_first_bucket(int b);
/* Early false far-field factor of the mixed form (0: off). */
int hip_mixed_set_early_dense(double f);
/* Row truncation factor of the mixed form (B's rows dropped before the
 * core carry at most (eta tol)^2 |B|^2). */
int hip_mixed_set_eta(double e);

/* Column block of the generated-operand GEMMs (0 = by size, 64/128/256). */
int starsh_hip_gen_gemm_bn(int bn);
/* 16-row groups per wave of the generated-tile GEMMs (0 = by size, 1, 2). */
int starsh_hip_gen_gemm_rw(int rw);
/* Panel width of the block Gram-Schmidt (8 or 4). */
int starsh_hip_qr_panel(int width);
/* Gram-Schmidt panels as one captured hipGraph (1) or host launches (0). */
int starsh_hip_qr_graph(int on);
/* Values chosen by the autotuner for the current call (0 = by size); the
 * options above take precedence.  proj: 0 sketch GEMM, 1 projection. */
int hip_set_tuned_gemm(int proj, int bn, int rw);
int hip_set_tuned_panel(int width);
/* Applies the autotuner's knobs recorded for problem 0 of a batch (or the
 * trial knobs while tuning) before a compression; mixed: 0 plain, 1 mixed
 * form. */
void hip_sched_apply(int mixed, const hip_batch *B);
/* Bucket ladder of the mixed form: option (-1 = the autotuner's, 0, 1),
 * the autotuner's choice for the current call (a tuning trial's wins over
 * the option), and the bucket count of a ladder. */
int hip_mixed_set_ladder(int l);
void hip_mixed_tuned_ladder(int l, int trial);
int hip_mixed_nbuckets(int l);
/* Per-tile schedules of the mixed form on (1, default) or off (0). */
int hip_mixed_set_tile_schedule(int on);
/* First bucket and per-tile schedule (or NULL) of problem q of a mixed
 * compression on bucket ladder `ladder`. */
void hip_sched_plan(const hip_batch *B, int q, int ladder, int *first,
        const signed char **tiles);
/* Outcome of far-field tile `tile` of problem q (final bucket, or -2 for a
 * false far-field tile), kept while the autotuner records. */
void hip_sched_record(int q, STARSH_int tile, int outcome);
/* Empties the in-process schedule table (the file is not read again). */
void hip_sched_clear(void);
/* Tiles pending after the first bucket of the last mixed call. */
void hip_sched_pending_after_first(long long n);
/* Projection of the mixed form on int8 digit planes (1) or FP64 MFMA (0). */
int starsh_hip_proj_int8(int on);
/* Sweep limit of the Jacobi core (1..60; 0 restores the default 60). */
int starsh_hip_jacobi_sweeps(int sweeps);
/* Copy-back options: which 0 = slot size in MiB (power of two, 1..64),
 * 1 = transfers on the library stream (1) or a copy stream (0). */
int starsh_hip_copy_option(int which, int value);
size_t hip_driver_workspace_budget(void);
double hip_driver_now(void);

#endif
