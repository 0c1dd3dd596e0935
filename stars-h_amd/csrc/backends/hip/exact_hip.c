/* Exact low-rank engines on the GPU: SVD (STARSH_LRENGINE SVD and DCSVD,
 * the reference's starsh_blrm__dsdd_omp with starsh_dense_dlrsdd,
 * src/backends/openmp/blrm/dsdd.c and sequential/dense/dsdd.c:16-66) and
 * rank-revealing QR (RRQR, starsh_blrm__dqp3_omp with starsh_dense_dlrqp3,
 * sequential/dense/dqp3.c:16-93).  Same far-field driver, output and
 * false far-field rewrite as the randomized engine (drsdd_hip.c); the
 * per-tile computation is the HIP pipeline run so that it yields what
 * dgesdd and dgeqp3 yield:
 *
 * SVD:  the singular values and vectors of A are those of B = Q^T A once
 *       the range of Q holds the range of A.  Round 1 sketches with width
 *       min(maxrank+oversample, m, n) (deflation at 1e-13 of each column);
 *       a tile whose sketch exhausted its range (a whole panel of sketch
 *       columns dropped) is final.  The others go to round 2 with the
 *       sketch min(m, n) wide, whose Gaussian test matrix makes range(Y) =
 *       range(A).  The rank is dsvfr over all singular values; the tile is
 *       low-rank when rank < min(m, n)/2 and rank <= maxrank (dsdd.c:53).
 * RRQR: dgeqp3 pivots the columns of A by largest residual norm; with
 *       mn2 = min(maxrank+oversample, m, n) the reference takes the SVD of
 *       R[:mn2, :] (dqp3.c:70-82), whose singular values are those of
 *       Q1^T A for Q1 an orthonormal basis of the first mn2 pivot columns.
 *       A tile whose range has at most mn2 dimensions has R[:mn2, :] = the
 *       whole R, and the SVD result; it is final after round 1 or 2 as for
 *       SVD.  For the others round 2 also computes the first mn2 pivots on
 *       B (whose columns are A's in the basis Q of A's range, so with the
 *       same residual norms), and round 3 runs the pipeline with the sketch
 *       Y = A[:, pivots] (rank by dsvfr over mn2 values, dqp3.c:83-84).
 *
 * A tile whose truncated core would exceed 1024 rows (more than 1024
 * singular values above 1e-5 tol |A|, only possible in round 2) keeps
 * 1024 rows of B and counts the others in the rank rule's tail, like the
 * truncation itself: the rank can then exceed dgesdd's by the margin that
 * extra mass allows, never fall below it. */
#include "common.h"
#include "rsdd.h"
#include "hip_driver.h"

/* The sketch holds the whole range of the tile. */
static int accept_range(int rank, int k, int kfull, int flags, int os,
        int kc, int mn)
{
    (void)rank; (void)kfull; (void)os; (void)kc;
    return (flags & 1) != 0 || k >= mn;
}

/* ... and the range has at most mn2 = kfull dimensions (RRQR). */
static int accept_range_rrqr(int rank, int k, int kfull, int flags, int os,
        int kc, int mn)
{
    return accept_range(rank, k, kfull, flags, os, kc, mn) && kc <= kfull;
}

int hip_exact_tiles(const hip_batch *B, STARSH_int ntiles, const int *pidx,
        const STARSH_int *idx, int rrqr, starsh_hip_stats *st)
{
    if(ntiles == 0)
        return STARSH_SUCCESS;
    int info = STARSH_SUCCESS;
    int *k = calloc(ntiles, sizeof(int)), *rk = calloc(ntiles, sizeof(int));
    int *fl = malloc(sizeof(int)*ntiles), *acc = malloc(sizeof(int)*ntiles);
    int *kc = malloc(sizeof(int)*ntiles), *mn = malloc(sizeof(int)*ntiles);
    int *kfull = malloc(sizeof(int)*ntiles);
    int *sp = malloc(sizeof(int)*ntiles);
    STARSH_int *si = malloc(sizeof(STARSH_int)*ntiles);
    STARSH_int *pos = malloc(sizeof(STARSH_int)*ntiles);
    int *pk = malloc(sizeof(int)*ntiles), *pdone = malloc(sizeof(int)*ntiles);
    int64_t *poff = malloc(sizeof(int64_t)*(ntiles+1));
    int *piv = NULL;
    if(!k || !rk || !fl || !acc || !kc || !mn || !kfull || !sp || !si ||
            !pos || !pk || !pdone || !poff)
        info = STARSH_MALLOC_ERROR;
    for(STARSH_int t = 0; info == STARSH_SUCCESS && t < ntiles; t++)
    {
        const int q = pidx[t];
        const STARSH_blrf *F = B->F[q];
        const STARSH_int bi = idx[t];
        const int m = (int)F->row_cluster->size[F->block_far[2*bi]];
        const int n = (int)F->col_cluster->size[F->block_far[2*bi+1]];
        mn[t] = m < n ? m : n;
        const int w = B->maxrank[q]+B->hp[q].oversample;
        kfull[t] = w < mn[t] ? w : mn[t];
        k[t] = kfull[t];
    }
    hip_run_extra x;
    memset(&x, 0, sizeof(x));
    x.rule = RSDD_RULE_HALF;
    x.kc_out = kc;
    /* round 1: sketch of the plain width */
    st->tile_rounds += ntiles;
    if(info == STARSH_SUCCESS)
        info = hip_driver_run(B, ntiles, pidx, idx, k, 0, 0.0, 0.0, rk, fl,
                acc, rrqr ? accept_range_rrqr : accept_range, NULL, 0, st,
                &x);
    /* round 2: the tiles whose range was not exhausted, full width; RRQR
     * also takes the pivots of those whose range exceeds mn2 */
    STARSH_int n2 = 0;
    for(STARSH_int t = 0; info == STARSH_SUCCESS && t < ntiles; t++)
        if(!acc[t])
        {
            pos[n2] = t;
            sp[n2] = pidx[t];
            si[n2] = idx[t];
            k[n2] = mn[t];
            n2++;
        }
    if(info == STARSH_SUCCESS && n2 > 0)
    {
        if(rrqr)
        {
            poff[0] = 0;
            for(STARSH_int u = 0; u < n2; u++)
            {
                pk[u] = kfull[pos[u]];
                poff[u+1] = poff[u]+pk[u];
                pdone[u] = 0;
            }
            piv = malloc(sizeof(int)*(size_t)(poff[n2] > 0 ? poff[n2] : 1));
            if(piv == NULL)
                info = STARSH_MALLOC_ERROR;
            x.piv_k = pk;
            x.piv = piv;
            x.piv_off = poff;
            x.piv_done = pdone;
        }
        /* kfull of the compacted tiles for the acceptance test comes from
         * the driver (min(maxrank+oversample, m, n)) */
        st->tile_rounds += n2;
        if(info == STARSH_SUCCESS)
            info = hip_driver_run(B, n2, sp, si, k, 0, 0.0, 0.0, rk, fl, acc,
                    rrqr ? accept_range_rrqr : NULL, NULL, 0, st, &x);
    }
    /* round 3 (RRQR): the pivot columns as the sketch */
    if(info == STARSH_SUCCESS && rrqr && n2 > 0)
    {
        STARSH_int n3 = 0;
        for(STARSH_int u = 0; u < n2; u++)
        {
            if(acc[u])
                continue;
            if(!pdone[u])
            {
                info = STARSH_UNKNOWN_ERROR;
                break;
            }
            /* pivots are columns of the tile */
            const STARSH_blrf *F = B->F[sp[u]];
            const int n = (int)F->col_cluster->size[F->block_far[2*si[u]+1]];
            for(int j = 0; j < pk[u]; j++)
                if(piv[poff[u]+j] < 0 || piv[poff[u]+j] >= n)
                    info = STARSH_UNKNOWN_ERROR;
            if(info != STARSH_SUCCESS)
                break;
            /* compact in place: entries u >= n3 are not read again */
            sp[n3] = sp[u];
            si[n3] = si[u];
            k[n3] = pk[u];
            if(n3 != u)
                memmove(piv+poff[n3], piv+poff[u], sizeof(int)*pk[u]);
            poff[n3+1] = poff[n3]+pk[u];
            n3++;
        }
        if(info == STARSH_SUCCESS && n3 > 0)
        {
            hip_run_extra x3;
            memset(&x3, 0, sizeof(x3));
            x3.rule = RSDD_RULE_HALF | RSDD_RULE_SELECT;
            x3.sel = piv;
            x3.sel_off = poff;
            st->tile_rounds += n3;
            info = hip_driver_run(B, n3, sp, si, k, 0, 0.0, 0.0, rk, fl,
                    acc, NULL, NULL, 0, st, &x3);
        }
    }
    free(k); free(rk); free(fl); free(acc); free(kc); free(mn); free(kfull);
    free(sp); free(si); free(pos); free(pk); free(pdone); free(poff);
    free(piv);
    return info;
}
