/* Batched mixed-precision HIP driver of tile low-rank compression.
 *
 * Same contract and output as starsh_blrm__drsdd_hip (and the reference's
 * src/backends/openmp/blrm/drsdd.c:16-307), but the sketch width adapts to
 * each tile's rank instead of being fixed at maxrank+oversample:
 *
 *   - tiles go through rank buckets of sketch width 32, 64, 128, 256, 512
 *     (capped at the plain width min(maxrank+oversample, m, n));
 *   - the sketch Y = A Omega runs on int8 MFMA with exact integer
 *     accumulation of 7-bit digit planes of A (sketch_i8.hip);
 *   - orthogonalisation, the projection B = Q^T A, the core SVD and the rank
 *     rule stay in FP64;
 *   - a tile's result is final when its rank leaves at least `oversample`
 *     columns of the sketch unused (the margin the reference keeps with its
 *     fixed width), when the sketch exhausted the tile's numerical range, or
 *     at the plain width; otherwise the tile moves to the next bucket.
 * All pending tiles of a bucket go through the pipeline as one batch. */
#include "common.h"
#include "rsdd.h"
#include "hip_driver.h"

static const int bucket_widths[] = {32, 64, 128, 256, 512};
#define NBUCKETS ((int)(sizeof(bucket_widths)/sizeof(bucket_widths[0])))

static int first_bucket = 0;

int hip_mixed_set_first_bucket(int b)
{
    if(b < 0 || b >= NBUCKETS)
        return STARSH_WRONG_PARAMETER;
    first_bucket = b;
    return STARSH_SUCCESS;
}

/* A randomized range finder with k = r + p samples leaves an expected
 * error factor sqrt(1 + r/(p-1)) over the best rank-r approximation; the
 * reference keeps p = maxrank + oversample - r.  A bucket result is final
 * when p >= max(oversample, r), i.e. the factor stays below sqrt(2). */
static int accept_mixed(int rank, int k, int kfull, int flags, int os)
{
    if(k >= kfull)
        return 1;
    if(rank < 0)
        return 0;
    if(rank+(rank > os ? rank : os) <= k)
        return 1;
    /* flag 1: a whole panel of sketch columns was numerically dependent */
    return (flags & 1) != 0;
}

int starsh_blrm__drsdd_hip_mixed(STARSH_blrm **matrix, STARSH_blrf *format,
        int maxrank, double tol, int onfly)
{
    double t_start = hip_driver_now();
    if(matrix == NULL)
    {
        STARSH_ERROR("Invalid value of `matrix`");
        return STARSH_WRONG_PARAMETER;
    }
    hip_problem hp;
    int info = hip_driver_prepare(format, maxrank, tol, onfly, &hp);
    if(info != STARSH_SUCCESS)
        return info;
    STARSH_blrf *F = format;
    STARSH_cluster *RC = F->row_cluster, *CC = F->col_cluster;
    STARSH_int nfar = F->nblocks_far;
    hip_factors hf;
    info = hip_driver_alloc_factors(F, maxrank, &hf);
    starsh_hip_stats st;
    memset(&st, 0, sizeof(st));
    STARSH_int *pending = malloc(sizeof(STARSH_int)*(nfar+1));
    STARSH_int *next = malloc(sizeof(STARSH_int)*(nfar+1));
    int *k = malloc(sizeof(int)*(nfar+1)), *rk = malloc(sizeof(int)*(nfar+1));
    int *fl = malloc(sizeof(int)*(nfar+1));
    int *acc = malloc(sizeof(int)*(nfar+1));
    if(info == STARSH_SUCCESS && (!pending || !next || !k || !rk || !fl ||
                !acc))
        info = STARSH_MALLOC_ERROR;
    STARSH_int npending = nfar;
    for(STARSH_int bi = 0; bi < nfar && info == STARSH_SUCCESS; bi++)
        pending[bi] = bi;
    /* Deflation threshold of the sketch columns: above the int8 sketch's
     * representation noise (2^-28 of a row's largest entry) and well below
     * the requested accuracy. */
    double tau = 0.1*tol;
    tau = tau > 1e-8 ? 1e-8 : tau;
    tau = tau < 1e-13 ? 1e-13 : tau;
    /* Four 7-bit digits represent A to 2^-28 of each row's largest entry;
     * tighter tolerances keep the FP64 sketch (still rank-bucketed). */
    int int8 = tol >= 3e-8;
    if(!int8)
        tau = 1e-13;
    for(int b = first_bucket; b < NBUCKETS && npending > 0 &&
            info == STARSH_SUCCESS; b++)
    {
        int last = 1;
        for(STARSH_int t = 0; t < npending; t++)
        {
            STARSH_int bi = pending[t];
            int m = (int)RC->size[F->block_far[2*bi]];
            int n = (int)CC->size[F->block_far[2*bi+1]];
            int mn = m < n ? m : n;
            int kfull = maxrank+hp.oversample < mn ? maxrank+hp.oversample :
                mn;
            k[t] = bucket_widths[b] < kfull ? bucket_widths[b] : kfull;
            if(b == NBUCKETS-1)
                k[t] = kfull;
            if(k[t] < kfull)
                last = 0;
        }
        (void)last;
        info = hip_driver_run(&hp, F, maxrank, tol, npending, pending, k,
                int8, tau, &hf, rk, fl, acc, accept_mixed, &st);
        STARSH_int nn = 0;
        for(STARSH_int t = 0; t < npending && info == STARSH_SUCCESS; t++)
            if(!acc[t])
                next[nn++] = pending[t];
        STARSH_int *tmp = pending;
        pending = next;
        next = tmp;
        npending = nn;
    }
    if(info == STARSH_SUCCESS)
        hip_driver_rank_stats(&hf, nfar, &st);
    free(pending); free(next); free(k); free(rk); free(fl); free(acc);
    if(info == STARSH_SUCCESS)
        info = hip_driver_finish(matrix, F, &hp, &hf, onfly, &st);
    if(info != STARSH_SUCCESS)
        hip_driver_free_factors(F, &hf);
    hip_driver_release(&hp);
    st.wall_seconds = hip_driver_now()-t_start;
    hip_driver_set_stats(&st);
    return info;
}
