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
