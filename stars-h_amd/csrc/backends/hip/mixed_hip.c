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
#include "rsdd.h"
#include "hip_driver.h"

static const int bucket_widths[] = {32, 64, 128, 256, 512};
#define NBUCKETS ((int)(sizeof(bucket_widths)/sizeof(bucket_widths[0])))

/* First bucket forced by starsh_hip_set_option("mixed_first_bucket"), or -1
 * for the schedule table, then DEFAULT_FIRST (width 64) when the problem has
 * no entry: on every workload measured (2D sqexp and exp, 3D Matern,
 * Coulomb and sin(kr)/r) width 64 beat width 32 as the first bucket. */
#define DEFAULT_FIRST 1
static int forced_bucket = -1;
/* Tiles still pending after the first bucket of the last call. */
static long long last_pending_after_first = 0;

int hip_mixed_set_first_bucket(int b)
{
    if(b < -1 || b >= NBUCKETS)
        return STARSH_WRONG_PARAMETER;
    forced_bucket = b;
    return STARSH_SUCCESS;
}

/* Schedule table of the autotuner: the first bucket that gave the shortest
 * compression time for a problem signature (kernel, its parameters, point
 * count, largest cluster, maxrank, oversample, tolerance).  Entries come
 * from starsh_hip_autotune() and from the text file named by the
 * environment variable STARSH_HIP_SCHEDULE, one entry per line:
 *   family ndim beta nu noise sigma wave_k diag n nb maxrank os tol first */
typedef struct sched_entry
{
    starsh_hip_kernel_desc kd;
    long long n;
    int nb, maxrank, os;
    double tol;
    int first;
} sched_entry;

#define SCHED_MAX 256
static sched_entry sched[SCHED_MAX];
static int nsched = 0, sched_loaded = 0;

static void sched_key(const hip_problem *hp, const STARSH_blrf *F,
        int maxrank, double tol, sched_entry *e)
{
    memset(e, 0, sizeof(*e));
    e->kd = hp->kd;
    e->n = hp->nr;
    int nb = 0;
    for(STARSH_int i = 0; i < F->row_cluster->nblocks; i++)
        if(F->row_cluster->size[i] > nb)
            nb = (int)F->row_cluster->size[i];
    e->nb = nb;
    e->maxrank = maxrank;
    e->os = hp->oversample;
    e->tol = tol;
    e->first = -1;
}

static int sched_same(const sched_entry *a, const sched_entry *b)
{
    const starsh_hip_kernel_desc *x = &a->kd, *y = &b->kd;
    return x->family == y->family && x->ndim == y->ndim &&
        x->beta == y->beta && x->nu == y->nu && x->noise == y->noise &&
        x->sigma == y->sigma && x->wave_k == y->wave_k &&
        x->diag == y->diag && a->n == b->n && a->nb == b->nb &&
        a->maxrank == b->maxrank && a->os == b->os && a->tol == b->tol;
}

static void sched_put(const sched_entry *e)
{
    for(int i = 0; i < nsched; i++)
        if(sched_same(&sched[i], e))
        {
            sched[i].first = e->first;
            return;
        }
    if(nsched < SCHED_MAX)
        sched[nsched++] = *e;
}

static void sched_load(void)
{
    if(sched_loaded)
        return;
    sched_loaded = 1;
    const char *path = getenv("STARSH_HIP_SCHEDULE");
    FILE *f = path ? fopen(path, "r") : NULL;
    if(f == NULL)
        return;
    sched_entry e;
    memset(&e, 0, sizeof(e));
    while(fscanf(f, "%d %d %lg %lg %lg %lg %lg %lg %lld %d %d %d %lg %d",
                &e.kd.family, &e.kd.ndim, &e.kd.beta, &e.kd.nu, &e.kd.noise,
                &e.kd.sigma, &e.kd.wave_k, &e.kd.diag, &e.n, &e.nb,
                &e.maxrank, &e.os, &e.tol, &e.first) == 14)
        if(e.first >= 0 && e.first < NBUCKETS)
            sched_put(&e);
    fclose(f);
}

static void sched_save(const sched_entry *e)
{
    const char *path = getenv("STARSH_HIP_SCHEDULE");
    FILE *f = path ? fopen(path, "a") : NULL;
    if(f == NULL)
        return;
    fprintf(f, "%d %d %.17g %.17g %.17g %.17g %.17g %.17g %lld %d %d %d "
            "%.17g %d\n", e->kd.family, e->kd.ndim, e->kd.beta, e->kd.nu,
            e->kd.noise, e->kd.sigma, e->kd.wave_k, e->kd.diag, e->n, e->nb,
            e->maxrank, e->os, e->tol, e->first);
    fclose(f);
}

static int sched_first(const sched_entry *key)
{
    sched_load();
    for(int i = 0; i < nsched; i++)
        if(sched_same(&sched[i], key))
            return sched[i].first;
    return -1;
}

/* A randomized range finder with k = r + p samples leaves an expected
 * error factor sqrt(1 + r/(p-1)) over the best rank-r approximation; the
 * reference keeps p = maxrank + oversample - r.  A bucket result is final
 * when p >= max(oversample, r), i.e. the factor stays below sqrt(2). */
static int accept_mixed(int rank, int k, int kfull, int flags, int os)
{
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
    for(int q = 0; info == STARSH_SUCCESS && q < B->nprob; q++)
    {
        first[q] = forced_bucket;
        if(first[q] < 0)
        {
            sched_entry key;
            sched_key(B->hp+q, B->F[q], B->maxrank[q], B->tol[q], &key);
            first[q] = sched_first(&key);
            if(first[q] < 0)
                first[q] = DEFAULT_FIRST;
        }
    }
    STARSH_int npend = info == STARSH_SUCCESS ? ntiles : 0;
    for(STARSH_int t = 0; t < npend; t++)
    {
        pend[t] = t;
        bucket[t] = first[pidx[t]];
    }
    last_pending_after_first = npend;
    for(int round = 0; npend > 0 && info == STARSH_SUCCESS; round++)
    {
        /* pending tiles grouped by bucket (stable), one batch per group so
         * that every batch has one sketch width class */
        STARSH_int cnt[NBUCKETS+1];
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
                    rk+g0, fl+g0, acc+g0, accept_mixed,
                    sv+(size_t)g0*NBMAX, NBMAX, st);
            g0 = g1;
        }
        /* optional per-tile record of the rounds (STARSH_HIP_MIXED_TRACE):
         * round, problem, tile, sketch width, rank, accepted, predicted
         * rank, first singular values */
        FILE *tf = getenv("STARSH_HIP_MIXED_TRACE") ?
            fopen(getenv("STARSH_HIP_MIXED_TRACE"), round ? "a" : "w") : NULL;
        for(STARSH_int u = 0; tf && u < npend; u++)
        {
            const int q = pq[u], os = B->hp[q].oversample;
            const double pr = predicted_rank(sv+(size_t)u*NBMAX, k[u], os,
                    B->tol[q]);
            fprintf(tf, "%d %d %lld %d %d %d %.3g %.3e %.3e\n", round, q,
                    (long long)pi[u], k[u], rk[u], acc[u], pr,
                    sv[(size_t)u*NBMAX], sv[(size_t)u*NBMAX+k[u]-1]);
        }
        if(tf)
            fclose(tf);
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
                if(r < rk[u])
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
            last_pending_after_first = nn;
    }
    free(pend); free(next); free(bucket); free(pq); free(pi); free(k);
    free(rk); free(fl); free(acc); free(first); free(sv);
    return info;
}

/* Runs the mixed form once per candidate first bucket (after one untimed
 * run), keeps the fastest by device time, records it in the schedule table
 * (and in the STARSH_HIP_SCHEDULE file) and returns it in *first_out.  The
 * candidates stop at the first bucket that finishes every tile. */
/* The tile lists of a format, kept so that every trial of the autotuner
 * starts from the format the caller passed (a compression rewrites it when
 * it finds false far-field tiles). */
typedef struct
{
    STARSH_int nfar, nnear;
    STARSH_int *far, *near;
} format_lists;

static int lists_save(const STARSH_blrf *F, format_lists *L)
{
    L->nfar = F->nblocks_far;
    L->nnear = F->nblocks_near;
    L->far = malloc(sizeof(STARSH_int)*2*(L->nfar > 0 ? L->nfar : 1));
    L->near = malloc(sizeof(STARSH_int)*2*(L->nnear > 0 ? L->nnear : 1));
    if(!L->far || !L->near)
        return STARSH_MALLOC_ERROR;
    if(L->nfar > 0)
        memcpy(L->far, F->block_far, sizeof(STARSH_int)*2*L->nfar);
    if(L->nnear > 0)
        memcpy(L->near, F->block_near, sizeof(STARSH_int)*2*L->nnear);
    return STARSH_SUCCESS;
}

/* Rebuilds *F from the saved lists if a trial changed them. */
static int lists_restore(STARSH_blrf *F, const format_lists *L)
{
    if(F->nblocks_far == L->nfar && F->nblocks_near == L->nnear)
        return STARSH_SUCCESS;
    STARSH_int *far = malloc(sizeof(STARSH_int)*2*(L->nfar > 0 ? L->nfar :
                1));
    STARSH_int *near = malloc(sizeof(STARSH_int)*2*(L->nnear > 0 ?
                L->nnear : 1));
    if(!far || !near)
    {
        free(far);
        free(near);
        return STARSH_MALLOC_ERROR;
    }
    memcpy(far, L->far, sizeof(STARSH_int)*2*L->nfar);
    memcpy(near, L->near, sizeof(STARSH_int)*2*L->nnear);
    STARSH_blrf *F2;
    int info = starsh_blrf_new_from_coo(&F2, F->problem, F->symm,
            F->row_cluster, F->col_cluster, L->nfar, far, L->nnear, near,
            F->type);
    if(info != STARSH_SUCCESS)
        return info;
    if(L->nfar == 0)
        free(far);
    if(L->nnear == 0)
        free(near);
    STARSH_blrf tmp = *F;
    *F = *F2;
    *F2 = tmp;
    starsh_blrf_free(F2);
    return STARSH_SUCCESS;
}

int starsh_hip_autotune(STARSH_blrf *format, int maxrank, double tol,
        int onfly, int *first_out, double *seconds_out)
{
    if(format == NULL)
        return STARSH_WRONG_PARAMETER;
    format_lists lists = {0, 0, NULL, NULL};
    int info = lists_save(format, &lists);
    if(info != STARSH_SUCCESS)
    {
        free(lists.far);
        free(lists.near);
        return info;
    }
    const int saved = forced_bucket;
    STARSH_blrm *M = NULL;
    forced_bucket = 0;
    info = starsh_blrm__drsdd_hip_mixed(&M, format, maxrank, tol, onfly);
    if(info == STARSH_SUCCESS)
    {
        starsh_blrm_free(M);
        info = lists_restore(format, &lists);
    }
    if(info != STARSH_SUCCESS)
    {
        forced_bucket = saved;
        free(lists.far);
        free(lists.near);
        return info;
    }
    int best = 0;
    double best_t = 1e300;
    for(int b = 0; b < NBUCKETS; b++)
    {
        forced_bucket = b;
        M = NULL;
        info = starsh_blrm__drsdd_hip_mixed(&M, format, maxrank, tol, onfly);
        if(info != STARSH_SUCCESS)
            break;
        starsh_hip_stats st;
        starsh_hip_get_stats(&st);
        starsh_blrm_free(M);
        info = lists_restore(format, &lists);
        if(info != STARSH_SUCCESS)
            break;
        if(seconds_out)
            seconds_out[b] = st.device_seconds;
        if(st.device_seconds < best_t)
        {
            best_t = st.device_seconds;
            best = b;
        }
        if(last_pending_after_first == 0)
            break;
    }
    forced_bucket = saved;
    free(lists.far);
    free(lists.near);
    if(info != STARSH_SUCCESS)
        return info;
    hip_problem hp;
    info = hip_driver_prepare(format, maxrank, tol, onfly, &hp);
    if(info != STARSH_SUCCESS)
        return info;
    sched_entry e;
    sched_key(&hp, format, maxrank, tol, &e);
    hip_driver_release(&hp);
    e.first = best;
    sched_load();
    sched_put(&e);
    sched_save(&e);
    if(first_out)
        *first_out = best;
    return STARSH_SUCCESS;
}
