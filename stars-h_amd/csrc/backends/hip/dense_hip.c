/* Randomized SVD of explicit dense tiles already in device memory: the
 * pipeline of the plain driver with A read from memory instead of being
 * generated.  Same algorithm and rank rule as starsh_dense_dlrrsdd
 * (src/backends/sequential/dense/drsdd.c:16-93). */
#include "common.h"
#include "rsdd.h"

int starsh_hip_dense_rsdd(int ntiles, const int *m, const int *n,
        const double *d_tiles, const int64_t *off, int maxrank,
        int oversample, double tol, int *rank, double *U,
        const int64_t *uoff, double *V, const int64_t *voff, double *S,
        const int64_t *soff)
{
    if(ntiles < 0 || (ntiles > 0 && (m == NULL || n == NULL ||
                    d_tiles == NULL || off == NULL || rank == NULL)) ||
            maxrank <= 0 || oversample <= 0 || !(tol >= 0))
        return STARSH_WRONG_PARAMETER;
    if(ntiles == 0)
        return STARSH_SUCCESS;
    for(int t = 0; t < ntiles; t++)
        if(m[t] <= 0 || n[t] <= 0)
            return STARSH_WRONG_PARAMETER;
    int info = starsh_hip_init(0);
    if(info != STARSH_SUCCESS)
        return info;
    int *pp = malloc(sizeof(int)*ntiles), *kc = malloc(sizeof(int)*ntiles);
    int *fl = malloc(sizeof(int)*ntiles);
    int64_t *uo = malloc(sizeof(int64_t)*ntiles);
    int64_t *vo = malloc(sizeof(int64_t)*ntiles);
    if(!pp || !kc || !fl || !uo || !vo)
    {
        free(pp); free(kc); free(fl); free(uo); free(vo);
        return STARSH_MALLOC_ERROR;
    }
    RsddInput in;
    memset(&in, 0, sizeof(in));
    in.ntiles = ntiles;
    in.m = m;
    in.n = n;
    in.d_a = d_tiles;
    in.a_off = off;
    in.maxrank = maxrank;
    in.oversample = oversample;
    in.tol = tol;
    in.S_out = S;
    in.soff = soff;
    RsddOutput out;
    memset(&out, 0, sizeof(out));
    out.rank = rank;
    out.p = pp;
    out.kc = kc;
    out.flags = fl;
    out.uoff = uo;
    out.voff = vo;
    info = rsdd_batch(&in, &out);
    for(int t = 0; info == STARSH_SUCCESS && t < ntiles; t++)
    {
        if(rank[t] <= 0)
            continue;
        if(U)
            info = starsh_hip_memcpy_d2h(U+uoff[t], out.d_uv+uo[t],
                    sizeof(double)*(size_t)m[t]*rank[t]);
        if(info == STARSH_SUCCESS && V)
            info = starsh_hip_memcpy_d2h(V+voff[t], out.d_uv+vo[t],
                    sizeof(double)*(size_t)n[t]*rank[t]);
    }
    free(pp); free(kc); free(fl); free(uo); free(vo);
    return info;
}

/* ---- Exact engines on explicit tiles: the rounds of exact_hip.c with the
 * tiles read from device memory. */

typedef struct
{
    int cnt;                /* tiles of the round */
    int *t;                 /* their indices */
    int *m, *n, *k, *rank, *p, *kc, *fl;
    int64_t *off, *soff, *uo, *vo;
} dense_sub;

static int sub_alloc(dense_sub *d, int ntiles)
{
    memset(d, 0, sizeof(*d));
    d->t = malloc(sizeof(int)*ntiles);
    d->m = malloc(sizeof(int)*ntiles);
    d->n = malloc(sizeof(int)*ntiles);
    d->k = malloc(sizeof(int)*ntiles);
    d->rank = malloc(sizeof(int)*ntiles);
    d->p = malloc(sizeof(int)*ntiles);
    d->kc = malloc(sizeof(int)*ntiles);
    d->fl = malloc(sizeof(int)*ntiles);
    d->off = malloc(sizeof(int64_t)*ntiles);
    d->soff = malloc(sizeof(int64_t)*ntiles);
    d->uo = malloc(sizeof(int64_t)*ntiles);
    d->vo = malloc(sizeof(int64_t)*ntiles);
    return d->t && d->m && d->n && d->k && d->rank && d->p && d->kc &&
        d->fl && d->off && d->soff && d->uo && d->vo ? STARSH_SUCCESS :
        STARSH_MALLOC_ERROR;
}

static void sub_free(dense_sub *d)
{
    free(d->t); free(d->m); free(d->n); free(d->k); free(d->rank);
    free(d->p); free(d->kc); free(d->fl); free(d->off); free(d->soff);
    free(d->uo); free(d->vo);
}

/* Runs the pipeline on the round's tiles; the accepted ones (acc[u] after
 * the call, all when accept_all) get rank, U, V and S. */
static int dense_round(dense_sub *d, const double *d_tiles, int maxrank,
        int oversample, double tol, int rule, const int *sel,
        const int64_t *sel_off, const int *piv_k, int *piv,
        const int64_t *piv_off, int *piv_done, int rrqr, int last,
        int *acc, int *rank, double *U, const int64_t *uoff, double *V,
        const int64_t *voff, double *S)
{
    RsddInput in;
    memset(&in, 0, sizeof(in));
    in.ntiles = d->cnt;
    in.m = d->m;
    in.n = d->n;
    in.k = d->k;
    in.d_a = d_tiles;
    in.a_off = d->off;
    in.maxrank = maxrank;
    in.oversample = oversample;
    in.tol = tol;
    in.rule = rule;
    in.sel = sel;
    in.sel_off = sel_off;
    in.piv_k = piv_k;
    in.piv = piv;
    in.piv_off = piv_off;
    in.piv_done = piv_done;
    if(S)
    {
        in.S_out = S;
        in.soff = d->soff;
    }
    RsddOutput out;
    memset(&out, 0, sizeof(out));
    out.rank = d->rank;
    out.p = d->p;
    out.kc = d->kc;
    out.flags = d->fl;
    out.uoff = d->uo;
    out.voff = d->vo;
    int info = rsdd_batch(&in, &out);
    for(int u = 0; info == STARSH_SUCCESS && u < d->cnt; u++)
    {
        const int mn = d->m[u] < d->n[u] ? d->m[u] : d->n[u];
        int kfull = maxrank+oversample;
        kfull = kfull < mn ? kfull : mn;
        const int range = (d->fl[u] & 1) != 0 || d->k[u] >= mn;
        acc[u] = last || (range && (!rrqr || d->kc[u] <= kfull));
        if(!acc[u])
            continue;
        const int t = d->t[u], r = d->rank[u];
        rank[t] = r;
        if(r <= 0)
            continue;
        if(U)
            info = starsh_hip_memcpy_d2h(U+uoff[t], out.d_uv+d->uo[u],
                    sizeof(double)*(size_t)d->m[u]*r);
        if(info == STARSH_SUCCESS && V)
            info = starsh_hip_memcpy_d2h(V+voff[t], out.d_uv+d->vo[u],
                    sizeof(double)*(size_t)d->n[u]*r);
    }
    return info;
}

int starsh_hip_dense_lowrank(int engine, int ntiles, const int *m,
        const int *n, const double *d_tiles, const int64_t *off, int maxrank,
        int oversample, double tol, int *rank, double *U,
        const int64_t *uoff, double *V, const int64_t *voff, double *S,
        const int64_t *soff)
{
    if(engine == 0)
        return starsh_hip_dense_rsdd(ntiles, m, n, d_tiles, off, maxrank,
                oversample, tol, rank, U, uoff, V, voff, S, soff);
    if((engine != 1 && engine != 2) || ntiles < 0 || (ntiles > 0 &&
                (m == NULL || n == NULL || d_tiles == NULL || off == NULL ||
                 rank == NULL || (U && !uoff) || (V && !voff) ||
                 (S && !soff))) || maxrank <= 0 || oversample <= 0 ||
            !(tol >= 0))
        return STARSH_WRONG_PARAMETER;
    if(ntiles == 0)
        return STARSH_SUCCESS;
    for(int t = 0; t < ntiles; t++)
        if(m[t] <= 0 || n[t] <= 0)
            return STARSH_WRONG_PARAMETER;
    int info = starsh_hip_init(0);
    if(info != STARSH_SUCCESS)
        return info;
    const int rrqr = engine == 2;
    dense_sub d;
    int *acc = malloc(sizeof(int)*ntiles);
    int *pk = malloc(sizeof(int)*ntiles), *pdone = malloc(sizeof(int)*ntiles);
    int64_t *poff = malloc(sizeof(int64_t)*(ntiles+1));
    int *piv = NULL;
    info = sub_alloc(&d, ntiles);
    if(!acc || !pk || !pdone || !poff)
        info = STARSH_MALLOC_ERROR;
    /* round 1: sketch of the plain width */
    for(int t = 0; info == STARSH_SUCCESS && t < ntiles; t++)
    {
        const int mn = m[t] < n[t] ? m[t] : n[t];
        const int w = maxrank+oversample;
        d.t[t] = t;
        d.m[t] = m[t];
        d.n[t] = n[t];
        d.k[t] = w < mn ? w : mn;
        d.off[t] = off[t];
        d.soff[t] = S ? soff[t] : 0;
    }
    d.cnt = ntiles;
    if(info == STARSH_SUCCESS)
        info = dense_round(&d, d_tiles, maxrank, oversample, tol,
                RSDD_RULE_HALF, NULL, NULL, NULL, NULL, NULL, NULL, rrqr, 0,
                acc, rank, U, uoff, V, voff, S);
    /* round 2: full width; RRQR also takes pivots */
    int c = 0;
    for(int u = 0; info == STARSH_SUCCESS && u < d.cnt; u++)
        if(!acc[u])
        {
            const int t = d.t[u];
            d.t[c] = t;
            d.m[c] = m[t];
            d.n[c] = n[t];
            d.k[c] = m[t] < n[t] ? m[t] : n[t];
            d.off[c] = off[t];
            d.soff[c] = S ? soff[t] : 0;
            c++;
        }
    d.cnt = c;
    if(info == STARSH_SUCCESS && c > 0)
    {
        if(rrqr)
        {
            poff[0] = 0;
            for(int u = 0; u < c; u++)
            {
                const int mn = d.m[u] < d.n[u] ? d.m[u] : d.n[u];
                const int w = maxrank+oversample;
                pk[u] = w < mn ? w : mn;
                poff[u+1] = poff[u]+pk[u];
                pdone[u] = 0;
            }
            piv = malloc(sizeof(int)*(size_t)(poff[c] > 0 ? poff[c] : 1));
            if(piv == NULL)
                info = STARSH_MALLOC_ERROR;
        }
        if(info == STARSH_SUCCESS)
            info = dense_round(&d, d_tiles, maxrank, oversample, tol,
                    RSDD_RULE_HALF, NULL, NULL, rrqr ? pk : NULL, piv,
                    rrqr ? poff : NULL, rrqr ? pdone : NULL, rrqr, !rrqr,
                    acc, rank, U, uoff, V, voff, S);
    }
    /* round 3 (RRQR): the pivot columns as the sketch */
    if(info == STARSH_SUCCESS && rrqr && c > 0)
    {
        int c3 = 0;
        for(int u = 0; info == STARSH_SUCCESS && u < c; u++)
        {
            if(acc[u])
                continue;
            if(!pdone[u])
            {
                info = STARSH_UNKNOWN_ERROR;
                break;
            }
            for(int j = 0; j < pk[u]; j++)
                if(piv[poff[u]+j] < 0 || piv[poff[u]+j] >= d.n[u])
                    info = STARSH_UNKNOWN_ERROR;
            d.t[c3] = d.t[u];
            d.m[c3] = d.m[u];
            d.n[c3] = d.n[u];
            d.k[c3] = pk[u];
            d.off[c3] = d.off[u];
            d.soff[c3] = d.soff[u];
            if(c3 != u)
                memmove(piv+poff[c3], piv+poff[u], sizeof(int)*pk[u]);
            poff[c3+1] = poff[c3]+pk[u];
            c3++;
        }
        d.cnt = c3;
        if(info == STARSH_SUCCESS && c3 > 0)
            info = dense_round(&d, d_tiles, maxrank, oversample, tol,
                    RSDD_RULE_HALF | RSDD_RULE_SELECT, piv, poff, NULL, NULL,
                    NULL, NULL, 1, 1, acc, rank, U, uoff, V, voff, S);
    }
    sub_free(&d);
    free(acc); free(pk); free(pdone); free(poff); free(piv);
    return info;
}

/* ---- The reference's dense engines on one host tile, run on the GPU
 * (include/starsh.h:642-652; work, lwork and iwork are not used). */
static int dense_host(int engine, int nrows, int ncols, const double *D,
        int ldD, double *U, int ldU, double *V, int ldV, int *rank,
        int maxrank, int oversample, double tol)
{
    if(nrows <= 0 || ncols <= 0 || ldD < nrows || ldU < nrows ||
            ldV < ncols || D == NULL || rank == NULL)
        return STARSH_WRONG_PARAMETER;
    const size_t mn = (size_t)nrows*ncols;
    double *d = NULL;
    int info = starsh_hip_init(0);
    if(info == STARSH_SUCCESS)
        info = starsh_hip_malloc((void **)&d, sizeof(double)*mn);
    for(int j = 0; info == STARSH_SUCCESS && j < ncols; j++)
        info = starsh_hip_memcpy_h2d(d+(size_t)j*nrows, D+(size_t)j*ldD,
                sizeof(double)*nrows);
    double *u = malloc(sizeof(double)*(size_t)nrows*(maxrank > 0 ?
                maxrank : 1));
    double *v = malloc(sizeof(double)*(size_t)ncols*(maxrank > 0 ?
                maxrank : 1));
    if(info == STARSH_SUCCESS && (!u || !v))
        info = STARSH_MALLOC_ERROR;
    const int64_t zero = 0;
    if(info == STARSH_SUCCESS)
        info = starsh_hip_dense_lowrank(engine, 1, &nrows, &ncols, d, &zero,
                maxrank, oversample, tol, rank, u, &zero, v, &zero, NULL,
                NULL);
    for(int i = 0; info == STARSH_SUCCESS && i < *rank; i++)
    {
        memcpy(U+(size_t)i*ldU, u+(size_t)i*nrows, sizeof(double)*nrows);
        memcpy(V+(size_t)i*ldV, v+(size_t)i*ncols, sizeof(double)*ncols);
    }
    if(d)
        starsh_hip_free(d);
    free(u);
    free(v);
    return info;
}

void starsh_dense_dlrrsdd(int nrows, int ncols, double *D, int ldD,
        double *U, int ldU, double *V, int ldV, int *rank, int maxrank,
        int oversample, double tol, double *work, int lwork, int *iwork)
{
    (void)work; (void)lwork; (void)iwork;
    if(dense_host(0, nrows, ncols, D, ldD, U, ldU, V, ldV, rank, maxrank,
                oversample, tol) != STARSH_SUCCESS)
    {
        STARSH_ERROR("starsh_dense_dlrrsdd failed");
        *rank = -1;
    }
}

void starsh_dense_dlrsdd(int nrows, int ncols, double *D, int ldD,
        double *U, int ldU, double *V, int ldV, int *rank, int maxrank,
        double tol, double *work, int lwork, int *iwork)
{
    (void)work; (void)lwork; (void)iwork;
    /* the SVD engine's first round uses the plain width with oversample
     * 10; its result does not depend on it */
    if(dense_host(1, nrows, ncols, D, ldD, U, ldU, V, ldV, rank, maxrank,
                10, tol) != STARSH_SUCCESS)
    {
        STARSH_ERROR("starsh_dense_dlrsdd failed");
        *rank = -1;
    }
}

void starsh_dense_dlrqp3(int nrows, int ncols, double *D, int ldD,
        double *U, int ldU, double *V, int ldV, int *rank, int maxrank,
        int oversample, double tol, double *work, int lwork, int *iwork)
{
    (void)work; (void)lwork; (void)iwork;
    if(dense_host(2, nrows, ncols, D, ldD, U, ldU, V, ldV, rank, maxrank,
                oversample, tol) != STARSH_SUCCESS)
    {
        STARSH_ERROR("starsh_dense_dlrqp3 failed");
        *rank = -1;
    }
}

