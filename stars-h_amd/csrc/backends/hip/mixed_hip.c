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
#include <math.h>
#include "common.h"
#include "kernels_host.h"
#include "rsdd.h"
#include "hip_driver.h"

/* Sketch widths of the rank buckets (the ladder).  Ladder 0: 32, 64, 128,
 * 256, 512.  Ladder 1 adds 384, and a tile that did not finish moves to
 * the bucket that fits the rank its round computed when that round's
 * sketch was not saturated (rank + 2 oversample <= width), instead of the
 * rank extrapolated from the decay of its singular values, which
 * overestimates the ranks of slowly decaying tiles (3D Matern: tiles of
 * rank 145-192 extrapolate to 222-833 from a 256-wide round).  Ladder 1
 * takes 5% off the 3D Matern workload and adds 15-19% to others (more,
 * smaller batches per round), so the autotuner chooses it per problem.
 * The last bucket is always the plain width.  The ladder of a call: the
 * option "mixed_ladder" when set (0/1), else the autotuner's. */
static const int ladders[2][6] = {
    {32, 64, 128, 256, 512},
    {32, 64, 128, 256, 384, 512}};
static const int ladder_len[2] = {5, 6};
static int ladder_forced = -1, ladder_tuned = 0, ladder_trial = 0;
static int ladder = 0;                  /* of the current call */
#define bucket_widths (ladders[ladder])
#define NBUCKETS (ladder_len[ladder])
#define NBUCKETS_MAX 6

int hip_mixed_set_ladder(int l)
{
    if(l < -1 || l > 1)
        return STARSH_WRONG_PARAMETER;
    ladder_forced = l;
    return STARSH_SUCCESS;
}

void hip_mixed_tuned_ladder(int l, int trial)
{
    ladder_tuned = l == 1 ? 1 : 0;
    ladder_trial = trial;
}

int hip_mixed_nbuckets(int l)
{
    return ladder_len[l == 1 ? 1 : 0];
}

/* Early false far-field decision: a tile that has not finished and whose
 * rank, extrapolated from the singular values of its sketch, exceeds
 * early_dense times maxrank is declared a false far-field tile (a dense
 * near-field tile, exact) without the full-width round that the reference
 * rule would take to reach the same verdict.  0 disables it.  On the 3D
 * Matern workload (N=131072, tile 2048, maxrank 512, first width 256) the
 * 70 false far-field tiles extrapolate to ranks 470-850 and the tiles that
 * finish with ranks up to 512 to at most 463; 1.5 x maxrank = 768 separates
 * them with a wide margin; on the sin(kr)/r workload (tile 1536, maxrank
 * 384) every false far-field tile above 576 is caught and no other tile.
 * Tiles no wider than maxrank are never declared this way. */
static double early_dense = 1.5;

/* Truncation of B's rows in the mixed form: the rows dropped before the
 * core SVD carry at most (mixed_eta tol)^2 |B|_F^2, counted in full in the
 * rank rule's tail, so the rank can only grow and the error bound holds.
 * The plain form keeps 1e-5 (the reference's ranks up to near-ties). */
static double mixed_eta = 1e-5;

int hip_mixed_set_eta(double e)
{
    if(!(e > 0.0 && e <= 0.5))
        return STARSH_WRONG_PARAMETER;
    mixed_eta = e;
    return STARSH_SUCCESS;
}

int hip_mixed_set_early_dense(double f)
{
    if(!(f >= 0.0) || (f > 0.0 && f < 1.0))
        return STARSH_WRONG_PARAMETER;
    early_dense = f;
    return STARSH_SUCCESS;
}


/* A randomized range finder with k = r + p samples leaves an expected
 * error factor sqrt(1 + r/(p-1)) over the best rank-r approximation; the
 * reference keeps p = maxrank + oversample - r.  A bucket result is final
 * when p >= max(oversample, r), i.e. the factor stays below sqrt(2). */
static int accept_mixed(int rank, int k, int kfull, int flags, int os,
        int kc, int mn)
{
    (void)kc;
    (void)mn;
    /* flag 4: the core SVD did not converge; the tile is a false far-field
     * tile (rank -1), as when dgesdd fails in drsdd.c:78 */
    if(k >= kfull || (flags & 4) != 0)
        return 1;
    if(rank < 0)
        return 0;
    if(rank+(rank > os ? rank : os) <= k)
        return 1;
    /* flag 1: a whole panel of sketch columns was numerically dependent */
    return (flags & 1) != 0;
}

/* Predicted rank of a tile from the singular values found with sketch
 * width k (s[0..k): descending values, then the norm of the truncated rows,
 * then zeros): the values away from the last `os` ones (which the sketch
 * resolves least well) fix a geometric decay rate, and the tail is
 * extrapolated until it drops below tol^2 times the energy, as in the rank
 * rule of dsvfr.c:16-47.  Returns a large value for a decay too slow to
 * extrapolate. */
static double predicted_rank(const double *s, int k, int os, double tol)
{
    int p = 0;
    double e = 0.0;
    while(p < k && s[p] > 0.0)
    {
        e += s[p]*s[p];
        p++;
    }
    const int hi = p-os-1, lo = hi/2;
    if(hi < 4 || !(s[lo] > s[hi]))
        return 1e9;
    const double rho = pow(s[hi]/s[lo], 1.0/(hi-lo));
    const double target = tol*sqrt(e*(1.0-rho*rho));
    if(s[hi] <= target)
        return hi;
    return hi+log(target/s[hi])/log(rho);
}

int hip_mixed_tiles(const hip_batch *B, STARSH_int ntiles, const int *pidx,
        const STARSH_int *idx, starsh_hip_stats *st)
{
    if(ntiles == 0)
        return STARSH_SUCCESS;
    ladder = ladder_forced >= 0 && !ladder_trial ? ladder_forced :
        ladder_tuned;
    int info = STARSH_SUCCESS;
    /* pending tiles (positions into pidx/idx) and their current bucket */
    STARSH_int *pend = malloc(sizeof(STARSH_int)*ntiles);
    STARSH_int *next = malloc(sizeof(STARSH_int)*ntiles);
    int *bucket = malloc(sizeof(int)*ntiles);
    int *pq = malloc(sizeof(int)*ntiles);
    STARSH_int *pi = malloc(sizeof(STARSH_int)*ntiles);
    int *k = malloc(sizeof(int)*ntiles), *rk = malloc(sizeof(int)*ntiles);
    int *fl = malloc(sizeof(int)*ntiles), *acc = malloc(sizeof(int)*ntiles);
    int *first = malloc(sizeof(int)*B->nprob);
    /* singular values of each tile of a round, NBMAX apart (the largest
     * sketch width of any tile) */
    int NBMAX = 1;
    for(STARSH_int t = 0; t < ntiles; t++)
    {
        const int q = pidx[t];
        const STARSH_blrf *F = B->F[q];
        const int m = (int)F->row_cluster->size[F->block_far[2*idx[t]]];
        const int n = (int)F->col_cluster->size[F->block_far[2*idx[t]+1]];
        int kf = B->maxrank[q]+B->hp[q].oversample;
        kf = kf < m ? kf : m;
        kf = kf < n ? kf : n;
        NBMAX = kf > NBMAX ? kf : NBMAX;
    }
    double *sv = malloc(sizeof(double)*(size_t)ntiles*NBMAX);
    if(!pend || !next || !bucket || !pq || !pi || !k || !rk || !fl || !acc ||
            !first || !sv)
        info = STARSH_MALLOC_ERROR;
    /* Deflation threshold of the sketch columns: above the int8 sketch's
     * representation noise (2^-28 of a chunk's largest entry) and well
     * below the requested accuracy (the tightest one of the batch). */
    double tmin = 1e300;
    for(int q = 0; info == STARSH_SUCCESS && q < B->nprob; q++)
        tmin = B->tol[q] < tmin ? B->tol[q] : tmin;
    double tau = 0.1*tmin;
    tau = tau > 1e-8 ? 1e-8 : tau;
    tau = tau < 1e-13 ? 1e-13 : tau;
    /* Four byte digits represent A to 2^-31 of a chunk's largest entry;
     * tighter tolerances keep the FP64 sketch and projection (still
     * rank-bucketed). */
    const int int8 = tmin >= 3e-8;
    if(!int8)
        tau = 1e-13;
    /* first bucket and per-tile schedule of each problem (autotuner) */
    const signed char **tsched = calloc(B->nprob, sizeof(*tsched));
    if(tsched == NULL)
        info = STARSH_MALLOC_ERROR;
    for(int q = 0; info == STARSH_SUCCESS && q < B->nprob; q++)
    {
        hip_sched_plan(B, q, ladder, first+q, tsched+q);
        if(first[q] < 0 || first[q] >= NBUCKETS)
            first[q] = NBUCKETS-1;
    }
    STARSH_int npend = 0;
    for(STARSH_int t = 0; info == STARSH_SUCCESS && t < ntiles; t++)
    {
        const int q = pidx[t];
        int b = tsched[q] ? tsched[q][idx[t]] : first[q];
        if(b != -2 && (b < 0 || b >= NBUCKETS))
            b = NBUCKETS-1;
        if(b == -2)
        {
            /* a false far-field tile in the tuning run: dense at once */
            B->hf[q].far_rank[idx[t]] = -1;
            st->early_dense++;
            hip_sched_record(q, idx[t], -2);
            continue;
        }
        pend[npend] = t;
        bucket[t] = b;
        npend++;
    }
    hip_sched_pending_after_first(npend);
    for(int round = 0; npend > 0 && info == STARSH_SUCCESS; round++)
    {
        st->tile_rounds += npend;
        /* pending tiles grouped by bucket (stable), one batch per group so
         * that every batch has one sketch width class */
        STARSH_int cnt[NBUCKETS_MAX+1];
        memset(cnt, 0, sizeof(cnt));
        for(STARSH_int u = 0; u < npend; u++)
            cnt[bucket[pend[u]]+1]++;
        for(int b = 0; b < NBUCKETS; b++)
            cnt[b+1] += cnt[b];
        for(STARSH_int u = 0; u < npend; u++)
            next[cnt[bucket[pend[u]]]++] = pend[u];
        memcpy(pend, next, sizeof(STARSH_int)*npend);
        for(STARSH_int u = 0; u < npend; u++)
        {
            const STARSH_int t = pend[u];
            const int q = pidx[t];
            const STARSH_blrf *F = B->F[q];
            const STARSH_int bi = idx[t];
            const int m = (int)F->row_cluster->size[F->block_far[2*bi]];
            const int n = (int)F->col_cluster->size[F->block_far[2*bi+1]];
            const int mn = m < n ? m : n;
            const int os = B->hp[q].oversample;
            const int kfull = B->maxrank[q]+os < mn ? B->maxrank[q]+os : mn;
            const int b = bucket[t];
            k[u] = b >= NBUCKETS-1 || bucket_widths[b] >= kfull ? kfull :
                bucket_widths[b];
            pq[u] = q;
            pi[u] = bi;
        }
        for(STARSH_int g0 = 0; g0 < npend && info == STARSH_SUCCESS; )
        {
            STARSH_int g1 = g0+1;
            while(g1 < npend && bucket[pend[g1]] == bucket[pend[g0]])
                g1++;
            info = hip_driver_run(B, g1-g0, pq+g0, pi+g0, k+g0, int8, tau,
                    mixed_eta,
                    rk+g0, fl+g0, acc+g0, accept_mixed,
                    sv+(size_t)g0*NBMAX, NBMAX, st, NULL);
            g0 = g1;
        }
        /* optional per-tile record of the rounds (STARSH_HIP_MIXED_TRACE):
         * round, problem, tile, sketch width, rank, accepted, predicted
         * rank, first and last singular values, Jacobi sweeps and size p
         * of the core */
        FILE *tf = getenv("STARSH_HIP_MIXED_TRACE") ?
            fopen(getenv("STARSH_HIP_MIXED_TRACE"), round ? "a" : "w") : NULL;
        for(STARSH_int u = 0; tf && u < npend; u++)
        {
            const int q = pq[u], os = B->hp[q].oversample;
            const double pr = predicted_rank(sv+(size_t)u*NBMAX, k[u], os,
                    B->tol[q]);
            fprintf(tf, "%d %d %lld %d %d %d %.3g %.3e %.3e %d %d\n", round,
                    q,
                    (long long)pi[u], k[u], rk[u], acc[u], pr,
                    sv[(size_t)u*NBMAX], sv[(size_t)u*NBMAX+k[u]-1],
                    (fl[u] >> 8) & 255, fl[u] >> 16);
        }
        if(tf)
            fclose(tf);
        for(STARSH_int u = 0; u < npend; u++)
            if(acc[u])
                hip_sched_record(pq[u], pi[u], rk[u] < 0 ? -2 :
                        bucket[pend[u]]);
        STARSH_int nn = 0;
        for(STARSH_int u = 0; u < npend && info == STARSH_SUCCESS; u++)
            if(!acc[u])
            {
                /* next bucket: at least one up, and wide enough for the
                 * extrapolated rank r with max(oversample, r) spare
                 * columns */
                const STARSH_int t = pend[u];
                const int q = pq[u], os = B->hp[q].oversample;
                double r = predicted_rank(sv+(size_t)u*NBMAX, k[u], os,
                        B->tol[q]);
                /* 1e9 means "decay too slow to extrapolate": no verdict;
                 * a tile with min(m, n) <= maxrank cannot exceed maxrank
                 * (the ragged 1536 x 160 tiles of the sin(kr)/r workload
                 * extrapolate far beyond their 160 columns) */
                const STARSH_blrf *Fq = B->F[q];
                const STARSH_int mt = Fq->row_cluster->size[
                    Fq->block_far[2*pi[u]]];
                const STARSH_int nt = Fq->col_cluster->size[
                    Fq->block_far[2*pi[u]+1]];
                const STARSH_int mnt = mt < nt ? mt : nt;
                if(early_dense > 0.0 && r < 1e9 && mnt > B->maxrank[q] &&
                        r > early_dense*B->maxrank[q])
                {
                    B->hf[q].far_rank[pi[u]] = -1;
                    st->early_dense++;
                    hip_sched_record(q, pi[u], -2);
                    continue;
                }
                if(ladder == 1 && rk[u] >= 0 && rk[u]+2*os <= k[u])
                    r = rk[u];
                else if(r < rk[u])
                    r = rk[u];
                const double want = r+(r > os ? r : os);
                int b = bucket[t]+1;
                while(b < NBUCKETS-1 && bucket_widths[b] < want)
                    b++;
                bucket[t] = b;
                next[nn++] = t;
            }
        STARSH_int *tmp = pend;
        pend = next;
        next = tmp;
        npend = nn;
        if(round == 0)
            hip_sched_pending_after_first(nn);
    }
    free(pend); free(next); free(bucket); free(pq); free(pi); free(k);
    free(rk); free(fl); free(acc); free(first); free(sv); free(tsched);
    return info;
}
