/* HIP backend for tile low-rank (TLR) compression on one MI355X (gfx950).
 *
 * Two drivers have the reference approximation signature
 * (include/starsh.h:563-564 in ecrc/stars-h 0.3.1) and are selected by
 * starsh_init() through STARSH_BACKEND=HIP / HIP_MIXED:
 *
 *   starsh_blrm__drsdd_hip        plain form: FP64 randomized SVD of every
 *                                 far-field tile with sketch width
 *                                 k = min(maxrank+oversample, m, n), the same
 *                                 dlarnv test matrix and the same rank rule as
 *                                 src/backends/openmp/blrm/drsdd.c:16-307 and
 *                                 src/backends/sequential/dense/drsdd.c:16-93.
 *   starsh_blrm__drsdd_hip_mixed  batched mixed-precision form: sketch in
 *                                 exact int8-sliced MFMA arithmetic, rank
 *                                 buckets for the sketch width, FP64
 *                                 re-orthogonalisation, core SVD and checks.
 *
 * starsh_blrm__dsdd_hip and starsh_blrm__dqp3_hip (STARSH_LRENGINE=SVD,
 * DCSVD, RRQR) are the exact engines, below.
 *
 * starsh_blrm__dfe_hip replaces src/backends/openmp/blrm/dfe.c:16-148.
 *
 * Tile contents are generated on the GPU from the problem's point data.  The
 * driver recognises the kernel by the identity of F->problem->kernel (one of
 * this library's STARSH_kernel functions); any other kernel pointer is
 * rejected with STARSH_WRONG_PARAMETER.
 *
 * Memory contract (as the reference): U/V/D are written to malloc'd host
 * buffers owned by the returned STARSH_blrm and released by
 * starsh_blrm_free().  *format is rewritten in place when false far-field
 * tiles exist. */
#ifndef STARSH_HIP_H
#define STARSH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

struct starsh_blrm;
struct starsh_blrf;

/* Plain FP64 driver; replaces starsh_blrm__drsdd_omp
 * (src/backends/openmp/blrm/drsdd.c:16-307). */
int starsh_blrm__drsdd_hip(struct starsh_blrm **matrix,
        struct starsh_blrf *format, int maxrank, double tol, int onfly);

/* Batched mixed-precision driver; same signature and output contract. */
int starsh_blrm__drsdd_hip_mixed(struct starsh_blrm **matrix,
        struct starsh_blrf *format, int maxrank, double tol, int onfly);

/* Exact engines in FP64, same signature and output contract, selected by
 * STARSH_LRENGINE=SVD or DCSVD (dsdd) and RRQR (dqp3) under either HIP
 * backend:
 *   starsh_blrm__dsdd_hip  replaces starsh_blrm__dsdd_omp
 *                          (src/backends/openmp/blrm/dsdd.c) with
 *                          starsh_dense_dlrsdd (sequential/dense/dsdd.c:
 *                          16-66): the SVD of each far-field tile, rank by
 *                          dsvfr over all singular values, low-rank when
 *                          rank < min(m, n)/2 and rank <= maxrank;
 *   starsh_blrm__dqp3_hip  replaces starsh_blrm__dqp3_omp with
 *                          starsh_dense_dlrqp3 (sequential/dense/dqp3.c:
 *                          16-93): column-pivoted QR, SVD of the leading
 *                          min(maxrank+oversample, m, n) rows of R, the
 *                          same rank rule over those singular values.
 * A tile with more than 1024 singular values above 1e-5 tol |A| keeps the
 * 1024 leading rows of B = Q^T A for its SVD core and counts the others in
 * the rank rule's tail (a conservative rank). */
int starsh_blrm__dsdd_hip(struct starsh_blrm **matrix,
        struct starsh_blrf *format, int maxrank, double tol, int onfly);
int starsh_blrm__dqp3_hip(struct starsh_blrm **matrix,
        struct starsh_blrf *format, int maxrank, double tol, int onfly);

/* Batch of independent problems in one call (no reference counterpart: the
 * reference compresses one problem per starsh_blrm_approximate call,
 * include/starsh.h:563-567).  Equivalent to
 *   for q < nprob: starsh_blrm_approximate(&matrix[q], format[q],
 *                                          maxrank[q], tol[q], onfly)
 * with the engine selected by starsh_init(); under the HIP backends the
 * far-field tiles of all problems go through the GPU pipeline together
 * (problems are grouped by kernel family and dimension).  On failure no
 * matrix is returned (all matrix[q] are NULL or freed). */
int starsh_blrm_approximate_batch(struct starsh_blrm **matrix,
        struct starsh_blrf **format, int nprob, const int *maxrank,
        const double *tol, int onfly);

/* Global relative Frobenius error on the GPU; replaces starsh_blrm__dfe_omp
 * (src/backends/openmp/blrm/dfe.c:16-148).  Returns -1 on failure. */
double starsh_blrm__dfe_hip(struct starsh_blrm *matrix);

/* Per-tile variant of starsh_blrm__dfe_hip (no reference counterpart): fills
 * far_err[bi] = |A_bi - U_bi V_bi^T|_F / |A_bi|_F (0 when |A_bi|_F == 0) and
 * far_norm[bi] = |A_bi|_F for every far-field tile, and returns the global
 * error.  Either array may be NULL. */
double starsh_blrm__dfe_hip_tiles(struct starsh_blrm *matrix,
        double *far_err, double *far_norm);

/* Run statistics of the last HIP compression on this thread (no reference
 * counterpart; the reference keeps only internal timers,
 * src/backends/openmp/blrm/drsdd.c:128-140). */
typedef struct starsh_hip_stats
{
    double device_seconds;   /* first kernel to last kernel, HIP events */
    double wall_seconds;     /* whole API call incl. copy-back */
    double copy_seconds;     /* device-to-host copy of U, V and D */
    double dominant_kernel_seconds; /* summed duration of the sketch kernel */
    long long dominant_kernel_launches;
    long long ntiles;        /* far-field tiles compressed */
    long long nfalse_far;    /* tiles that turned out to be dense */
    int rank_min, rank_max;
    double rank_avg;
    double flops_algorithmic;   /* work of the reference algorithm */
    double sketch_bytes_or_flops; /* algorithmic flops of the sketch kernel */
    /* host-side timeline of the call (seconds; their sum with the device
     * batches is the wall time) */
    double prepare_seconds;  /* argument checks, point upload, U/V malloc */
    double batch_seconds;    /* wall time inside the device batches */
    double d2h_seconds;      /* pack + device-to-host copy of U and V */
    double scatter_seconds;  /* host copy into the caller's buffers */
    double finish_seconds;   /* false far-field rewrite, near tiles, blrm */
    double proj_kernel_seconds; /* summed duration of the projection
                                   kernels B^T = A^T Q (HIP events) */
    long long early_dense;   /* mixed form: tiles declared false far-field
                                from their extrapolated rank */
    long long tile_rounds;   /* mixed form: tiles summed over rounds */
} starsh_hip_stats;

int starsh_hip_get_stats(starsh_hip_stats *stats);

/* Product with a compressed matrix on the GPU, B = alpha*M*A + beta*B with
 * host arrays (vectors in cluster order); replaces starsh_blrm__dmml
 * (src/backends/sequential/blrm/dmml.c:16-145) with the same signature and
 * semantics: far-field tiles add U (V^T A_j) and, symmetric with i != j,
 * V (U^T A_i); near-field tiles add D A_j (and D^T A_i), D from near_D
 * (onfly == 0) or generated by the kernel (onfly == 1).  The factors are
 * uploaded for the call; use starsh_hip_tlr_* for repeated products. */
int starsh_blrm__dmml_hip(struct starsh_blrm *matrix, int nrhs, double alpha,
        double *A, int lda, double beta, double *B, int ldb);

/* Conjugate gradients with GPU products; replaces
 * starsh_itersolvers__dcg_omp (src/itersolvers/cg.c:24-95): same arguments,
 * iteration and stopping rule (column q stops once |r_q| < tol |r_q(0)|),
 * X updated in place; returns the iteration at which every column
 * converged, or -1.  work is not used (the vectors live on the device). */
int starsh_itersolvers__dcg_hip(struct starsh_blrm *matrix, int nrhs,
        double *B, int ldb, double *X, int ldx, double tol, double *work);

/* Device-resident copy of a compressed matrix for repeated products (no
 * reference counterpart): starsh_hip_tlr_new uploads the factors and near
 * tiles once; starsh_hip_tlr_dmml computes B = alpha*M*A + beta*B with A, B
 * in device memory (ld >= columns / rows of M); starsh_hip_tlr_bench times
 * nrhs-column products (seconds per product, mean over reps) and reports
 * the bytes a product must move at least (factors, near tiles, A, B). */
typedef struct starsh_hip_tlr starsh_hip_tlr;
int starsh_hip_tlr_new(starsh_hip_tlr **T, struct starsh_blrm *matrix);
int starsh_hip_tlr_dmml(starsh_hip_tlr *T, int nrhs, double alpha,
        const double *d_A, int lda, double beta, double *d_B, int ldb);
int starsh_hip_tlr_bench(starsh_hip_tlr *T, int nrhs, int reps,
        double *seconds, double *bytes);
void starsh_hip_tlr_free(starsh_hip_tlr *T);

/* Allocation-free variant of starsh_hip_tlr_dmml.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * It computes the same product with the same kernels and allocates nothing:
 * the r x nrhs coefficient blocks of the far-field tiles and the n x nrhs
 * transposed contributions of the symmetric format's off-diagonal tiles,
 * which starsh_hip_tlr_dmml keeps in a workspace of T grown on demand, go
 * to a static device array of 2^22 doubles (32 MiB) that each GPU holds
 * until the process ends.  Products that need more than that,
 * (2 x sum of far-field ranks + sum of n over those tiles) x nrhs doubles,
 * return STARSH_WRONG_PARAMETER (2D sqexp N = 64000, tile 1600: up to 3
 * right-hand sides).  All calls on one GPU share the array: calls that
 * overlap on one GPU (from several host threads) race. */
int starsh_hip_tlr_dmml_noalloc(starsh_hip_tlr *T, int nrhs, double alpha,
        const double *d_A, int lda, double beta, double *d_B, int ldb);

/* Measured FP64 rates of the current GPU in TFLOP/s (no reference
 * counterpart): independent v_mfma_f64_16x16x4f64 chains on every SIMD,
 * and FP64 fused multiply-adds on the vector ALUs.  For reading the
 * roofline fractions of the FP64 kernels against what the chip sustains. */
int starsh_hip_fp64_probe(double *mfma_tflops, double *valu_tflops);

/* Tuning options of the HIP drivers (no reference counterpart):
 *   "workspace_gb"        device workspace budget per batch (default 96);
 *   "mixed_first_bucket"  first rank bucket of the mixed form, an index
 *                         0..4 into the ladder's widths ({32, 64, 128, 256,
 *                         512}; on ladder 1 index 4 is 384), or -1
 *                         (default) for the autotuner's schedule table (1,
 *                         width 64, without an entry);
 *   "gemm_bn"             column block of the generated-tile GEMMs: 0 (by
 *                         size, default), 64, 128 or 256;
 *   "gemm_rw"             16-row groups per wave of those GEMMs: 0 (by
 *                         size, default), 1 or 2;
 *   "qr_panel"            panel width of the block Gram-Schmidt: 0 (by
 *                         batch, default: 8 columns for batches of more
 *                         than 1536 tiles), 4 or 8 (8 only for tiles of
 *                         at most 2048 rows);
 *   "proj_int8"           projection B^T = A^T Q of the mixed form on
 *                         int8 digit planes of both operands (1, default)
 *                         or on FP64 MFMA (0);
 *   "jacobi_max_sweeps"   sweep limit of the Jacobi core SVD, 1..60 (0:
 *                         the default 60); a tile whose core has not
 *                         converged within it is a false far-field tile
 *                         (rank -1), as for a failed dgesdd in
 *                         src/backends/sequential/dense/drsdd.c:78;
 *   "mixed_trunc"         truncation factor eta of the mixed form in (0,
 *                         0.5] (default 1e-5): rows of B whose tail is at
 *                         most (eta tol)^2 |B|_F^2 are dropped before the
 *                         core SVD and counted in the rank rule's tail;
 *   "mixed_early_dense"   the mixed form declares a tile false far-field
 *                         when its rank, extrapolated from the decay of
 *                         its sketch's singular values, exceeds this
 *                         factor times maxrank (default 1.5; 0 disables,
 *                         values in (0, 1) are refused);
 *   "qr_graph"            Gram-Schmidt panels as one captured hipGraph,
 *                         kept and replayed while its arguments are
 *                         unchanged (1), or host launches (0, default);
 *   "mixed_ladder"        bucket ladder of the mixed form: 0 (widths 32,
 *                         64, 128, 256, 512; the next bucket from the
 *                         extrapolated rank), 1 (a 384 bucket added; the
 *                         next bucket from the computed rank when the
 *                         sketch was not saturated), or -1 (default) for
 *                         the autotuner's schedule table (0 without an
 *                         entry);
 *   "tile_schedule"       per-tile schedules recorded by the autotuner
 *                         used by the mixed form (1, default) or not (0);
 *   "schedule_clear"      (value ignored) empties the autotuner's
 *                         in-process schedule table;
 *   "pipeline_batches"    batches per run of tiles (0 or 1, default: all
 *                         tiles of a run together; n > 1 splits them so
 *                         that the copy-back of one batch overlaps the
 *                         next, which pays only with "copy_stream" 0);
 *   "copy_slot_mb"        slot size of the pinned copy-back ring, MiB
 *                         (power of two, 1..64; default 32);
 *   "copy_stream"         device-to-host transfers of the factors on the
 *                         library stream (1, default) or on a stream of
 *                         their own (0);
 *   "inject_failure"      test hook: the driver step with this number
 *                         (1..4) fails once with STARSH_MALLOC_ERROR
 *                         (0 clears);
 *   "release_workspace"   (value ignored) frees the device workspace and
 *                         the pinned staging buffer, which are otherwise
 *                         kept and reused between calls.
 * Returns STARSH_WRONG_PARAMETER for an unknown name or value. */
int starsh_hip_set_option(const char *name, double value);

/* Autotuner of the mixed form (no reference counterpart): compresses
 * *format once per candidate first rank bucket, keeps the one with the
 * shortest device time and records it for this problem signature (kernel
 * and its parameters, point count, tile size, maxrank, oversample, tol) in
 * the in-process schedule table, which starsh_blrm__drsdd_hip_mixed
 * consults, and appends it to the file named by STARSH_HIP_SCHEDULE when
 * that variable is set (the file is read on first use).  *first_out gets
 * the chosen bucket; seconds_out (5 entries, may be NULL) the device time
 * of each candidate tried.  Every candidate is timed on *format as passed
 * (its tile lists are restored after each trial, which rewrites them when
 * it finds false far-field tiles), and *format is left as passed. */
int starsh_hip_autotune(struct starsh_blrf *format, int maxrank, double tol,
        int onfly, int *first_out, double *seconds_out);

/* Result of starsh_hip_autotune_batch. */
typedef struct starsh_hip_tuning
{
    int form;               /* 0 plain FP64, 1 mixed */
    int first_bucket;       /* mixed: chosen first bucket (0..4 on ladder
                             * 0, 0..5 on ladder 1), else -1 */
    int tile_schedule;      /* mixed: per-tile schedule kept (1) or not */
    int qr_panel;           /* chosen Gram-Schmidt panel width (0: by batch,
                             * 4, 8) */
    int sketch_bn, sketch_rw;   /* plain: sketch GEMM column block and
                                 * 16-row groups per wave (0: by size) */
    int proj_bn, proj_rw;       /* plain: projection GEMM, the same */
    int ntrials;            /* compressions run, warm-up included */
    double first_seconds[5];    /* mixed: device time per first bucket
                                 * tried, -1 if not tried */
    double schedule_seconds;    /* mixed: device time with the per-tile
                                 * schedule, -1 if none was recorded */
    double untuned_seconds;     /* device time with the built-in choices */
    double best_seconds;        /* device time with the chosen knobs */
    int ladder;                 /* mixed: chosen bucket ladder (0, 1) */
    double ladder_seconds;      /* mixed: device time on ladder 1 at the
                                 * chosen first width, -1 if not tried */
} starsh_hip_tuning;

/* Autotuner of both HIP forms for one problem or a batch (no reference
 * counterpart); form 0 tunes starsh_blrm__drsdd_hip, 1 the mixed form.
 * The nprob problems are compressed together (as by
 * starsh_blrm_approximate_batch) once per trial, each trial timed on the
 * device and run on the formats as passed (restored after each trial):
 *   plain: sketch GEMM (column block 256 or 128, one or two 16-row groups
 *          per wave), then projection GEMM (128 or 64, the same), then the
 *          Gram-Schmidt panel width (4, 8), each kept if faster;
 *   mixed: the first bucket as starsh_hip_autotune; then the bucket
 *          ladder: 0 (widths 32, 64, 128, 256, 512; a tile that did not
 *          finish moves to the width its extrapolated rank needs) or 1
 *          (a 384 bucket added; the rank a round computed steers the next
 *          bucket when that round's sketch was not saturated); then one
 *          compression at that bucket recording every far-field tile's
 *          outcome (the
 *          bucket it finished in, or false far-field) and one with that
 *          per-tile schedule (every tile starts in its bucket, false
 *          far-field tiles are stored densely at once), kept if faster;
 *          then the panel width.
 * The chosen knobs go to the schedule table (and to STARSH_HIP_SCHEDULE's
 * file) under each problem's signature; later compressions of that form
 * use them, options set by starsh_hip_set_option taking precedence.  A
 * per-tile schedule also applies to a problem with the same point
 * coordinates (hashed), kernel family, tile size, maxrank, oversample,
 * tolerance and number of far-field tiles but other kernel parameters (a
 * parameter sweep over one point set): it only moves work between buckets
 * and stores tiles densely, so it can cost time, never accuracy.  Option
 * "tile_schedule" 0 turns per-tile schedules off. */
int starsh_hip_autotune_batch(struct starsh_blrf **format, int nprob,
        const int *maxrank, const double *tol, int onfly, int form,
        starsh_hip_tuning *out);

/* ---- Thin C ABI over the HIP kernels.  Plain pointers and sizes; device
 * pointers are marked d_.  Every function returns STARSH_SUCCESS or a
 * STARSH_ERRNO code and never aborts. */

/* Kernel families the device generators implement. */
enum STARSH_HIP_KERNEL_FAMILY
{
    STARSH_HIP_EXP = 1,       /* sigma*exp(-r/beta); r==0 -> sigma+noise */
    STARSH_HIP_SQREXP = 2,    /* sigma*exp(-r^2/(2 beta^2)) */
    STARSH_HIP_MATERN = 3,    /* x = sqrt(2 nu) r/beta, nu in {.5,1.5,2.5} */
    STARSH_HIP_MATERN2 = 4,   /* x = r/beta */
    STARSH_HIP_COULOMB = 5,   /* 1/r; r==0 -> 0 */
    STARSH_HIP_SIN = 6,       /* sin(k r)/r; r==0 -> diag */
    STARSH_HIP_COS = 7        /* cos(k r)/r; r==0 -> diag */
};

typedef struct starsh_hip_kernel_desc
{
    int family;     /* enum STARSH_HIP_KERNEL_FAMILY */
    int ndim;       /* 1..4 */
    double beta;
    double nu;
    double noise;
    double sigma;
    double wave_k;
    double diag;
} starsh_hip_kernel_desc;

/* Selects the device and creates the library stream.  Called lazily by the
 * drivers; exposed for callers that want to pick a device. */
int starsh_hip_init(int device);

/* Fills d_omega (ncols x k, column-major, leading dimension ncols) with the
 * first ncols*k values of LAPACK dlarnv(idist=3, iseed={0,0,0,1}) — the
 * test matrix of src/backends/sequential/dense/drsdd.c:57-59. */
int starsh_hip_omega(double *d_omega, int ncols, int k);

/* Generates ntiles dense tiles.  Tile t covers rows [r0[t], r0[t]+m[t]) of
 * d_rowpts and columns [c0[t], c0[t]+n[t]) of d_colpts (points already in
 * cluster order, SoA with strides nrowpts / ncolpts) and is written
 * column-major with leading dimension m[t] at d_out + off[t].  The host
 * arrays r0, c0, m, n, off have ntiles entries. */
int starsh_hip_generate_tiles(const starsh_hip_kernel_desc *kd,
        const double *d_rowpts, int64_t nrowpts,
        const double *d_colpts, int64_t ncolpts, int ntiles,
        const int64_t *r0, const int64_t *c0, const int *m, const int *n,
        const int64_t *off, double *d_out);

/* Morton (Z-order) sort of n points on the GPU (no counterpart in the
 * reference's device code; the host sort is starsh_particles_zsort_inplace,
 * src/applications/particles.c:1090-1221): d_x holds ndim x n coordinates
 * (SoA, ndim <= 4, n < 2^31), d_y receives them in the reference's Morton
 * order and d_order (optional, int64) the permutation, d_y[., i] =
 * d_x[., d_order[i]].  The order equals the reference's, ties included:
 * points of equal key keep their input order when the key has an even
 * number of 1 bits and reverse it when odd, as the reference's partition
 * (1-bit points moved to the back in reverse) does. */
int starsh_hip_zsort(const double *d_x, int64_t n, int ndim, double *d_y,
        int64_t *d_order);

/* Raw device memory helpers used by the drivers and by tests. */
int starsh_hip_malloc(void **d_ptr, size_t bytes);
int starsh_hip_free(void *d_ptr);
int starsh_hip_memcpy_h2d(void *d_dst, const void *src, size_t bytes);
int starsh_hip_memcpy_d2h(void *dst, const void *d_src, size_t bytes);
int starsh_hip_synchronize(void);

/* Dense FP64 randomized-SVD of ntiles explicit tiles already in device
 * memory (tile t: m x n, column-major, leading dimension m, at d_tiles +
 * off[t]).  Same algorithm and rank rule as starsh_dense_dlrrsdd
 * (src/backends/sequential/dense/drsdd.c:16-93) with sketch width
 * k = min(maxrank+oversample, m, n); rank[t] = -1 marks a tile whose rank
 * exceeds maxrank.  U (m x rank) and V (n x rank, V = S*Vs^T rows) are
 * written to host arrays U + uoff[t], V + voff[t]; singular values (k per
 * tile) to S + soff[t] when S is not NULL.  This is the per-tile
 * entry point the parity tests use for arbitrary matrices. */
int starsh_hip_dense_rsdd(int ntiles, const int *m, const int *n,
        const double *d_tiles, const int64_t *off, int maxrank,
        int oversample, double tol, int *rank, double *U,
        const int64_t *uoff, double *V, const int64_t *voff, double *S,
        const int64_t *soff);

/* The same for the three engines: engine 0 randomized SVD (as above), 1
 * SVD (starsh_dense_dlrsdd: all singular values, rank < min(m, n)/2 and
 * rank <= maxrank), 2 rank-revealing QR (starsh_dense_dlrqp3); S + soff[t]
 * receives the singular values of the tile's last round (up to min(m, n)
 * for engine 1, min(maxrank+oversample, m, n) otherwise). */
int starsh_hip_dense_lowrank(int engine, int ntiles, const int *m,
        const int *n, const double *d_tiles, const int64_t *off, int maxrank,
        int oversample, double tol, int *rank, double *U,
        const int64_t *uoff, double *V, const int64_t *voff, double *S,
        const int64_t *soff);

#ifdef __cplusplus
}
#endif

#endif
