/* Plain FP64 HIP driver of tile low-rank compression.
 *
 * Follows src/backends/openmp/blrm/drsdd.c:16-307 of ecrc/stars-h 0.3.1:
 *   1. allocate U/V host buffers maxrank columns wide per far-field tile
 *      (:53-90);
 *   2. compress every far-field tile (:94-145) -- here all tiles of a chunk
 *      go through the batched HIP pipeline at once (rsdd_batch), tile
 *      contents generated on the GPU from the points;
 *   3. collect false far-field tiles (rank -1) in ascending order, rebuild
 *      the format and swap it into *format (:146-212);
 *   4. generate near-field tiles when onfly == 0 (:214-256), on the GPU;
 *   5. compact the far-field arrays (:259-282) and build the STARSH_blrm
 *      (:305-306).
 * The mixed-precision driver shares steps 1 and 3-5 (hip_driver_finish). */
#include <time.h>
#include "common.h"
#include "kernels_host.h"
#include "rsdd.h"
#include "hip_driver.h"

static __thread starsh_hip_stats last_stats;
static size_t workspace_budget = (size_t)96 << 30;

int starsh_hip_get_stats(starsh_hip_stats *stats)
{
    if(stats == NULL)
        return STARSH_WRONG_PARAMETER;
    *stats = last_stats;
    return STARSH_SUCCESS;
}

void hip_driver_set_stats(const starsh_hip_stats *stats)
{
    last_stats = *stats;
}

int starsh_hip_set_option(const char *name, double value)
{
    if(name == NULL)
        return STARSH_WRONG_PARAMETER;
    if(!strcmp(name, "workspace_gb"))
    {
        if(value <= 0)
            return STARSH_WRONG_PARAMETER;
        workspace_budget = (size_t)(value*(double)(1 << 30));
        return STARSH_SUCCESS;
    }
    if(!strcmp(name, "mixed_first_bucket"))
        return hip_mixed_set_first_bucket((int)value);
    if(!strcmp(name, "gemm_bn"))
        return starsh_hip_gen_gemm_bn((int)value);
    if(!strcmp(name, "gemm_rw"))
        return starsh_hip_gen_gemm_rw((int)value);
    if(!strcmp(name, "qr_panel"))
        return starsh_hip_qr_panel((int)value);
    if(!strcmp(name, "proj_int8"))
        return starsh_hip_proj_int8((int)value);
    if(!strcmp(name, "release_workspace"))
    {
        rsdd_release_pools();
        return STARSH_SUCCESS;
    }
    return STARSH_WRONG_PARAMETER;
}

size_t hip_driver_workspace_budget(void)
{
    return workspace_budget;
}

double hip_driver_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec+1e-9*ts.tv_nsec;
}

/* Copies the points of a cluster into cluster (pivot) order, SoA, and
 * uploads them. */
static int upload_points(const double *x, STARSH_int count, int ndim,
        const STARSH_cluster *C, double **d_x)
{
    STARSH_int n = C->ndata;
    double *h = malloc(sizeof(double)*(size_t)n*ndim + 1);
    if(h == NULL)
        return STARSH_MALLOC_ERROR;
    for(int d = 0; d < ndim; d++)
        for(STARSH_int i = 0; i < n; i++)
        {
            STARSH_int p = C->pivot[i];
            if(p < 0 || p >= count)
            {
                free(h);
                STARSH_ERROR("pivot entry out of range");
                return STARSH_WRONG_PARAMETER;
            }
            h[(size_t)d*n+i] = x[(size_t)d*count+p];
        }
    int info = starsh_hip_malloc((void **)d_x, sizeof(double)*(size_t)n*ndim);
    if(info == STARSH_SUCCESS)
        info = starsh_hip_memcpy_h2d(*d_x, h, sizeof(double)*(size_t)n*ndim);
    free(h);
    return info;
}

int hip_driver_prepare(STARSH_blrf *F, int maxrank, double tol, int onfly,
        hip_problem *hp)
{
    memset(hp, 0, sizeof(*hp));
    if(F == NULL || F->problem == NULL)
    {
        STARSH_ERROR("Invalid value of `format`");
        return STARSH_WRONG_PARAMETER;
    }
    if(maxrank <= 0 || !(tol >= 0) || (onfly != 0 && onfly != 1))
    {
        STARSH_ERROR("Invalid value of `maxrank`, `tol` or `onfly`");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_problem *P = F->problem;
    if(P->ndim != 2 || P->dtype != 'd')
    {
        STARSH_ERROR("Only scalar double precision problems are supported");
        return STARSH_WRONG_PARAMETER;
    }
    const double *xr, *xc;
    STARSH_int cr, cc;
    starsh_hip_kernel_desc kd2;
    if(starsh_kernel_describe(P->kernel, P->row_data, &hp->kd, &xr, &cr) != 0
            || starsh_kernel_describe(P->kernel, P->col_data, &kd2, &xc,
                &cc) != 0)
    {
        STARSH_ERROR("The problem kernel has no HIP generator (only this "
                "library's spatial exp/sqrexp/matern/matern2, "
                "electrostatics and electrodynamics kernels with ndim <= 4 "
                "are supported)");
        return STARSH_WRONG_PARAMETER;
    }
    int info = starsh_hip_init(0);
    if(info != STARSH_SUCCESS)
        return info;
    hp->oversample = starsh_params.oversample > 0 ? starsh_params.oversample
        : 10;
    STARSH_cluster *RC = F->row_cluster, *CC = F->col_cluster;
    info = upload_points(xr, cr, hp->kd.ndim, RC, &hp->d_xr);
    if(info != STARSH_SUCCESS)
        return info;
    hp->nr = RC->ndata;
    if(CC == RC && P->col_data == P->row_data)
    {
        hp->d_xc = hp->d_xr;
        hp->nc = hp->nr;
    }
    else
    {
        info = upload_points(xc, cc, hp->kd.ndim, CC, &hp->d_xc);
        if(info != STARSH_SUCCESS)
        {
            hip_driver_release(hp);
            return info;
        }
        hp->nc = CC->ndata;
    }
    return STARSH_SUCCESS;
}

void hip_driver_release(hip_problem *hp)
{
    if(hp->d_xc != NULL && hp->d_xc != hp->d_xr)
        starsh_hip_free(hp->d_xc);
    if(hp->d_xr != NULL)
        starsh_hip_free(hp->d_xr);
    hp->d_xr = hp->d_xc = NULL;
}

/* Steps 1: host buffers and wrappers (drsdd.c:53-90). */
int hip_driver_alloc_factors(STARSH_blrf *F, int maxrank, hip_factors *hf)
{
    memset(hf, 0, sizeof(*hf));
    STARSH_int nfar = F->nblocks_far;
    STARSH_cluster *RC = F->row_cluster, *CC = F->col_cluster;
    if(nfar == 0)
        return STARSH_SUCCESS;
    STARSH_MALLOC(hf->far_U, nfar);
    STARSH_MALLOC(hf->far_V, nfar);
    STARSH_MALLOC(hf->far_rank, nfar);
    STARSH_MALLOC(hf->uoff, nfar);
    STARSH_MALLOC(hf->voff, nfar);
    size_t su = 0, sv = 0;
    for(STARSH_int bi = 0; bi < nfar; bi++)
    {
        hf->uoff[bi] = su;
        hf->voff[bi] = sv;
        su += (size_t)RC->size[F->block_far[2*bi]]*maxrank;
        sv += (size_t)CC->size[F->block_far[2*bi+1]]*maxrank;
    }
    STARSH_MALLOC(hf->alloc_U, su);
    STARSH_MALLOC(hf->alloc_V, sv);
    for(STARSH_int bi = 0; bi < nfar; bi++)
    {
        int shape_U[2] = {(int)RC->size[F->block_far[2*bi]], maxrank};
        int shape_V[2] = {(int)CC->size[F->block_far[2*bi+1]], maxrank};
        int info = array_from_buffer(hf->far_U+bi, 2, shape_U, 'd', 'F',
                hf->alloc_U+hf->uoff[bi]);
        if(info == STARSH_SUCCESS)
            info = array_from_buffer(hf->far_V+bi, 2, shape_V, 'd', 'F',
                    hf->alloc_V+hf->voff[bi]);
        if(info != STARSH_SUCCESS)
            return info;
    }
    return STARSH_SUCCESS;
}

void hip_driver_free_factors(STARSH_blrf *F, hip_factors *hf)
{
    if(hf->far_U)
        for(STARSH_int bi = 0; bi < F->nblocks_far; bi++)
            if(hf->far_U[bi])
            {
                hf->far_U[bi]->data = NULL;
                array_free(hf->far_U[bi]);
            }
    if(hf->far_V)
        for(STARSH_int bi = 0; bi < F->nblocks_far; bi++)
            if(hf->far_V[bi])
            {
                hf->far_V[bi]->data = NULL;
                array_free(hf->far_V[bi]);
            }
    free(hf->far_U);
    free(hf->far_V);
    free(hf->far_rank);
    free(hf->alloc_U);
    free(hf->alloc_V);
    free(hf->uoff);
    free(hf->voff);
    memset(hf, 0, sizeof(*hf));
}

/* Steps 3-5 (drsdd.c:146-306). */
int hip_driver_finish(STARSH_blrm **matrix, STARSH_blrf *format,
        const hip_problem *hp, hip_factors *hf, int onfly,
        starsh_hip_stats *st)
{
    STARSH_blrf *F = format;
    STARSH_problem *P = F->problem;
    STARSH_cluster *RC = F->row_cluster, *CC = F->col_cluster;
    STARSH_int nblocks_far = F->nblocks_far, nblocks_near = F->nblocks_near;
    STARSH_int new_far = nblocks_far, new_near = nblocks_near;
    STARSH_int *block_far = F->block_far, *block_near = F->block_near;
    int *far_rank = hf->far_rank;
    Array **far_U = hf->far_U, **far_V = hf->far_V, **near_D = NULL;
    double *alloc_U = hf->alloc_U, *alloc_V = hf->alloc_V, *alloc_D = NULL;
    STARSH_int nfalse = 0, bi, bj;
    for(bi = 0; bi < nblocks_far; bi++)
        if(far_rank[bi] == -1)
            nfalse++;
    st->nfalse_far += nfalse;
    if(nfalse > 0)
    {
        STARSH_int *false_far;
        STARSH_MALLOC(false_far, nfalse);
        for(bi = 0, bj = 0; bi < nblocks_far; bi++)
            if(far_rank[bi] == -1)
                false_far[bj++] = bi;
        new_near = nblocks_near+nfalse;
        STARSH_MALLOC(block_near, 2*new_near);
        for(bi = 0; bi < 2*nblocks_near; bi++)
            block_near[bi] = F->block_near[bi];
        for(bi = 0; bi < nfalse; bi++)
        {
            block_near[2*(bi+nblocks_near)] = F->block_far[2*false_far[bi]];
            block_near[2*(bi+nblocks_near)+1] =
                F->block_far[2*false_far[bi]+1];
        }
        new_far = nblocks_far-nfalse;
        block_far = NULL;
        if(new_far > 0)
        {
            STARSH_MALLOC(block_far, 2*new_far);
            for(bi = 0, bj = 0; bi < nblocks_far; bi++)
            {
                if(bj < nfalse && false_far[bj] == bi)
                    bj++;
                else
                {
                    block_far[2*(bi-bj)] = F->block_far[2*bi];
                    block_far[2*(bi-bj)+1] = F->block_far[2*bi+1];
                }
            }
        }
        STARSH_blrf *F2;
        int info = starsh_blrf_new_from_coo(&F2, P, F->symm, RC, CC, new_far,
                block_far, new_near, block_near, F->type);
        if(info != STARSH_SUCCESS)
        {
            free(false_far);
            return info;
        }
        STARSH_blrf tmp = *F;
        *F = *F2;
        *F2 = tmp;
        STARSH_WARNING("`F` was modified due to false far-field blocks");
        starsh_blrf_free(F2);
        /* Compaction of the far-field arrays (drsdd.c:259-282); the old
         * wrappers of false far-field tiles are released. */
        for(bi = 0, bj = 0; bi < nblocks_far; bi++)
        {
            if(far_rank[bi] == -1)
            {
                far_U[bi]->data = NULL;
                array_free(far_U[bi]);
                far_V[bi]->data = NULL;
                array_free(far_V[bi]);
                bj++;
                continue;
            }
            int shape_U[2] = {far_U[bi]->shape[0], far_rank[bi]};
            int shape_V[2] = {far_V[bi]->shape[0], far_rank[bi]};
            void *du = far_U[bi]->data, *dv = far_V[bi]->data;
            far_U[bi]->data = NULL;
            array_free(far_U[bi]);
            far_V[bi]->data = NULL;
            array_free(far_V[bi]);
            array_from_buffer(far_U+bi-bj, 2, shape_U, 'd', 'F', du);
            array_from_buffer(far_V+bi-bj, 2, shape_V, 'd', 'F', dv);
            far_rank[bi-bj] = far_rank[bi];
        }
        free(false_far);
        if(new_far > 0)
        {
            STARSH_REALLOC(far_rank, new_far);
            STARSH_REALLOC(far_U, new_far);
            STARSH_REALLOC(far_V, new_far);
        }
        else
        {
            free(far_rank);
            free(far_U);
            free(far_V);
            free(alloc_U);
            free(alloc_V);
            far_rank = NULL;
            far_U = far_V = NULL;
            alloc_U = alloc_V = NULL;
        }
    }
    /* Near-field tiles (drsdd.c:214-256), generated on the GPU. */
    if(onfly == 0 && new_near > 0)
    {
        STARSH_MALLOC(near_D, new_near);
        int64_t *r0 = malloc(sizeof(int64_t)*new_near);
        int64_t *c0 = malloc(sizeof(int64_t)*new_near);
        int64_t *off = malloc(sizeof(int64_t)*new_near);
        int *mm = malloc(sizeof(int)*new_near), *nn = malloc(sizeof(int)*
                new_near);
        if(!r0 || !c0 || !off || !mm || !nn)
        {
            free(r0); free(c0); free(off); free(mm); free(nn);
            return STARSH_MALLOC_ERROR;
        }
        size_t size_D = 0;
        for(bi = 0; bi < new_near; bi++)
        {
            STARSH_int i = block_near[2*bi], j = block_near[2*bi+1];
            r0[bi] = RC->start[i];
            c0[bi] = CC->start[j];
            mm[bi] = (int)RC->size[i];
            nn[bi] = (int)CC->size[j];
            off[bi] = size_D;
            size_D += (size_t)mm[bi]*nn[bi];
        }
        int info = STARSH_SUCCESS;
        alloc_D = malloc(sizeof(double)*size_D + 1);
        double *d_D = NULL;
        if(alloc_D == NULL)
            info = STARSH_MALLOC_ERROR;
        if(info == STARSH_SUCCESS)
            info = starsh_hip_malloc((void **)&d_D, sizeof(double)*size_D);
        if(info == STARSH_SUCCESS)
            info = starsh_hip_generate_tiles(&hp->kd, hp->d_xr, hp->nr,
                    hp->d_xc, hp->nc, (int)new_near, r0, c0, mm, nn, off,
                    d_D);
        if(info == STARSH_SUCCESS)
            info = starsh_hip_memcpy_d2h(alloc_D, d_D, sizeof(double)*size_D);
        starsh_hip_free(d_D);
        for(bi = 0; bi < new_near && info == STARSH_SUCCESS; bi++)
        {
            int shape[2] = {mm[bi], nn[bi]};
            info = array_from_buffer(near_D+bi, 2, shape, 'd', 'F',
                    alloc_D+off[bi]);
        }
        free(r0); free(c0); free(off); free(mm); free(nn);
        if(info != STARSH_SUCCESS)
            return info;
    }
    hf->far_U = NULL;
    hf->far_V = NULL;
    hf->far_rank = NULL;
    hf->alloc_U = NULL;
    hf->alloc_V = NULL;
    return starsh_blrm_new(matrix, F, far_rank, far_U, far_V, onfly, near_D,
            alloc_U, alloc_V, alloc_D, '1');
}

/* Bytes of the pipeline's workspace for one tile. */
static size_t tile_bytes(int m, int n, int k, int maxrank)
{
    size_t r = (size_t)(maxrank < k ? maxrank : k);
    /* FP64 buffers, factors, int8 Omega, int8 digit planes of Q */
    return sizeof(double)*((size_t)m*k+2*(size_t)n*k+4*(size_t)k*k+
            (size_t)k+(size_t)(m+n)*r)+(size_t)(n+63)*k+
        4*(size_t)(m+63)*(size_t)(k+15)+8*(size_t)(k+15);
}

int hip_driver_run(const hip_batch *B, STARSH_int ntiles, const int *pidx,
        const STARSH_int *idx, const int *k, int int8, double tau,
        int *rank_out, int *flags_out, int *accepted, hip_accept_fn accept,
        double *S_out, int sstride, starsh_hip_stats *st)
{
    if(ntiles == 0)
        return STARSH_SUCCESS;
    int info = STARSH_SUCCESS;
    int *m = malloc(sizeof(int)*ntiles), *n = malloc(sizeof(int)*ntiles);
    int *pp = malloc(sizeof(int)*ntiles), *kc = malloc(sizeof(int)*ntiles);
    int64_t *r0 = malloc(sizeof(int64_t)*ntiles);
    int64_t *c0 = malloc(sizeof(int64_t)*ntiles);
    int64_t *uo = malloc(sizeof(int64_t)*ntiles);
    int64_t *vo = malloc(sizeof(int64_t)*ntiles);
    RsddProblem *probs = malloc(sizeof(RsddProblem)*B->nprob);
    double **du = malloc(sizeof(double *)*ntiles);
    double **dv = malloc(sizeof(double *)*ntiles);
    int64_t *so = malloc(sizeof(int64_t)*ntiles);
    if(!m || !n || !pp || !kc || !r0 || !c0 || !uo || !vo || !probs || !du ||
            !dv || !so)
        info = STARSH_MALLOC_ERROR;
    for(STARSH_int t = 0; info == STARSH_SUCCESS && t < ntiles; t++)
        so[t] = (int64_t)t*sstride;
    for(int q = 0; info == STARSH_SUCCESS && q < B->nprob; q++)
    {
        const hip_problem *hp = B->hp+q;
        probs[q].kd = &hp->kd;
        probs[q].d_xr = hp->d_xr;
        probs[q].d_xc = hp->d_xc;
        probs[q].nr = hp->nr;
        probs[q].nc = hp->nc;
        probs[q].maxrank = B->maxrank[q];
        probs[q].oversample = hp->oversample;
        probs[q].tol = B->tol[q];
    }
    for(STARSH_int t = 0; info == STARSH_SUCCESS && t < ntiles; t++)
    {
        const STARSH_blrf *F = B->F[pidx[t]];
        const STARSH_cluster *RC = F->row_cluster, *CC = F->col_cluster;
        STARSH_int bi = idx[t];
        STARSH_int i = F->block_far[2*bi], j = F->block_far[2*bi+1];
        m[t] = (int)RC->size[i];
        n[t] = (int)CC->size[j];
        r0[t] = RC->start[i];
        c0[t] = CC->start[j];
        if(m[t] != n[t])
            STARSH_WARNING("This was only tested on square tiles, error of "
                    "approximation may be much higher, than demanded");
    }
    STARSH_int b0 = 0;
    while(info == STARSH_SUCCESS && b0 < ntiles)
    {
        size_t bytes = 0;
        STARSH_int b1 = b0;
        while(b1 < ntiles)
        {
            const int q = pidx[b1];
            const int mr = B->maxrank[q], os = B->hp[q].oversample;
            int mn = m[b1] < n[b1] ? m[b1] : n[b1];
            int kk = k[b1] < mn ? k[b1] : mn;
            size_t tb = tile_bytes(m[b1], n[b1], kk, mr);
            if(b1 > b0 && bytes+tb > hip_driver_workspace_budget())
                break;
            bytes += tb;
            st->flops_algorithmic += 4.0*m[b1]*(double)n[b1]*
                (mr+os < mn ? mr+os : mn);
            st->sketch_bytes_or_flops += 2.0*m[b1]*(double)n[b1]*kk;
            b1++;
        }
        RsddInput in;
        memset(&in, 0, sizeof(in));
        in.ntiles = (int)(b1-b0);
        in.m = m+b0;
        in.n = n+b0;
        in.k = k+b0;
        in.nprob = B->nprob;
        in.probs = probs;
        in.prob = pidx+b0;
        in.r0 = r0+b0;
        in.c0 = c0+b0;
        in.sketch_int8 = int8;
        in.tau = tau;
        if(S_out)
        {
            in.S_out = S_out;
            in.soff = so+b0;
        }
        RsddOutput out;
        memset(&out, 0, sizeof(out));
        out.rank = rank_out+b0;
        out.p = pp+b0;
        out.kc = kc+b0;
        out.flags = flags_out+b0;
        out.uoff = uo+b0;
        out.voff = vo+b0;
        info = rsdd_batch(&in, &out);
        if(info != STARSH_SUCCESS)
            break;
        st->device_seconds += out.device_seconds;
        st->dominant_kernel_seconds += out.sketch_seconds;
        st->dominant_kernel_launches += 1;
        double tc = hip_driver_now();
        for(STARSH_int t = b0; t < b1; t++)
        {
            const int q = pidx[t];
            const int mn = m[t] < n[t] ? m[t] : n[t];
            const int kk = k[t] < mn ? k[t] : mn;
            int kfull = B->maxrank[q]+B->hp[q].oversample;
            kfull = kfull < mn ? kfull : mn;
            accepted[t] = accept ? accept(rank_out[t], kk, kfull,
                    flags_out[t], B->hp[q].oversample) : 1;
            du[t] = dv[t] = NULL;
            if(!accepted[t])
                continue;
            hip_factors *hf = B->hf+q;
            STARSH_int bi = idx[t];
            hf->far_rank[bi] = rank_out[t];
            if(rank_out[t] <= 0)
                continue;
            du[t] = hf->alloc_U+hf->uoff[bi];
            dv[t] = hf->alloc_V+hf->voff[bi];
        }
        info = rsdd_copy_factors(&out, (int)(b1-b0), m+b0, n+b0, du+b0,
                dv+b0);
        st->copy_seconds += hip_driver_now()-tc;
        b0 = b1;
    }
    free(m); free(n); free(pp); free(kc); free(r0); free(c0); free(uo);
    free(vo); free(probs); free(du); free(dv); free(so);
    return info;
}

void hip_driver_rank_stats(const hip_batch *B, starsh_hip_stats *st)
{
    long long cnt = 0, sum = 0, nfar = 0;
    st->rank_min = 1 << 30;
    st->rank_max = -1;
    for(int q = 0; q < B->nprob; q++)
        for(STARSH_int bi = 0; bi < B->F[q]->nblocks_far; bi++)
        {
            nfar++;
            const int r = B->hf[q].far_rank[bi];
            if(r < 0)
                continue;
            cnt++;
            sum += r;
            st->rank_min = r < st->rank_min ? r : st->rank_min;
            st->rank_max = r > st->rank_max ? r : st->rank_max;
        }
    st->rank_avg = cnt ? (double)sum/cnt : 0;
    if(cnt == 0)
        st->rank_min = 0;
    st->ntiles = nfar;
}

/* Plain form over a batch: every far-field tile at the full sketch width
 * k = min(maxrank+oversample, m, n). */
int hip_plain_tiles(const hip_batch *B, STARSH_int ntiles, const int *pidx,
        const STARSH_int *idx, starsh_hip_stats *st)
{
    int *k = malloc(sizeof(int)*(ntiles+1));
    int *rk = malloc(sizeof(int)*(ntiles+1));
    int *fl = malloc(sizeof(int)*(ntiles+1));
    int *acc = malloc(sizeof(int)*(ntiles+1));
    int info = STARSH_SUCCESS;
    if(!k || !rk || !fl || !acc)
        info = STARSH_MALLOC_ERROR;
    for(STARSH_int t = 0; info == STARSH_SUCCESS && t < ntiles; t++)
        k[t] = B->maxrank[pidx[t]]+B->hp[pidx[t]].oversample;
    if(info == STARSH_SUCCESS)
        info = hip_driver_run(B, ntiles, pidx, idx, k, 0, 0.0, rk, fl, acc,
                NULL, NULL, 0, st);
    free(k); free(rk); free(fl); free(acc);
    return info;
}

/* Kernel family and dimension of a prepared problem, as one sort key. */
static int group_key(const hip_problem *hp)
{
    return hp->kd.family*16+hp->kd.ndim;
}

int hip_compress(int mixed, int nprob, STARSH_blrm **matrix,
        STARSH_blrf **format, const int *maxrank, const double *tol,
        int onfly)
{
    double t_start = hip_driver_now();
    if(nprob <= 0 || matrix == NULL || format == NULL || maxrank == NULL ||
            tol == NULL)
    {
        STARSH_ERROR("Invalid arguments");
        return STARSH_WRONG_PARAMETER;
    }
    for(int q = 0; q < nprob; q++)
        matrix[q] = NULL;
    hip_batch B;
    memset(&B, 0, sizeof(B));
    B.nprob = nprob;
    B.F = format;
    B.maxrank = maxrank;
    B.tol = tol;
    B.hp = calloc(nprob, sizeof(hip_problem));
    B.hf = calloc(nprob, sizeof(hip_factors));
    starsh_hip_stats st;
    memset(&st, 0, sizeof(st));
    int info = B.hp && B.hf ? STARSH_SUCCESS : STARSH_MALLOC_ERROR;
    int nprep = 0;
    for(int q = 0; info == STARSH_SUCCESS && q < nprob; q++)
    {
        info = hip_driver_prepare(format[q], maxrank[q], tol[q], onfly,
                B.hp+q);
        if(info != STARSH_SUCCESS)
            break;
        nprep++;
        info = hip_driver_alloc_factors(format[q], maxrank[q], B.hf+q);
    }
    /* all far-field tiles, grouped by kernel family and dimension */
    STARSH_int ntot = 0;
    for(int q = 0; info == STARSH_SUCCESS && q < nprob; q++)
        ntot += format[q]->nblocks_far;
    int *pidx = malloc(sizeof(int)*(ntot+1));
    STARSH_int *idx = malloc(sizeof(STARSH_int)*(ntot+1));
    int *done = calloc(nprob, sizeof(int));
    if(info == STARSH_SUCCESS && (!pidx || !idx || !done))
        info = STARSH_MALLOC_ERROR;
    for(int q0 = 0; info == STARSH_SUCCESS && q0 < nprob; q0++)
    {
        if(done[q0])
            continue;
        const int key = group_key(B.hp+q0);
        STARSH_int nt = 0;
        for(int q = q0; q < nprob; q++)
            if(!done[q] && group_key(B.hp+q) == key)
            {
                done[q] = 1;
                for(STARSH_int bi = 0; bi < format[q]->nblocks_far; bi++)
                {
                    pidx[nt] = q;
                    idx[nt++] = bi;
                }
            }
        info = mixed ? hip_mixed_tiles(&B, nt, pidx, idx, &st) :
            hip_plain_tiles(&B, nt, pidx, idx, &st);
    }
    free(pidx); free(idx); free(done);
    if(info == STARSH_SUCCESS)
        hip_driver_rank_stats(&B, &st);
    for(int q = 0; info == STARSH_SUCCESS && q < nprob; q++)
        info = hip_driver_finish(matrix+q, format[q], B.hp+q, B.hf+q, onfly,
                &st);
    if(info != STARSH_SUCCESS)
        for(int q = 0; q < nprob; q++)
        {
            if(matrix[q])
            {
                starsh_blrm_free(matrix[q]);
                matrix[q] = NULL;
            }
            else if(B.hf && q < nprep)
                hip_driver_free_factors(format[q], B.hf+q);
        }
    for(int q = 0; B.hp && q < nprep; q++)
        hip_driver_release(B.hp+q);
    free(B.hp);
    free(B.hf);
    st.wall_seconds = hip_driver_now()-t_start;
    hip_driver_set_stats(&st);
    return info;
}

int starsh_blrm__drsdd_hip(STARSH_blrm **matrix, STARSH_blrf *format,
        int maxrank, double tol, int onfly)
{
    if(matrix == NULL || format == NULL)
    {
        STARSH_ERROR("Invalid value of `matrix` or `format`");
        return STARSH_WRONG_PARAMETER;
    }
    return hip_compress(0, 1, matrix, &format, &maxrank, &tol, onfly);
}

int starsh_blrm__drsdd_hip_mixed(STARSH_blrm **matrix, STARSH_blrf *format,
        int maxrank, double tol, int onfly)
{
    if(matrix == NULL || format == NULL)
    {
        STARSH_ERROR("Invalid value of `matrix` or `format`");
        return STARSH_WRONG_PARAMETER;
    }
    return hip_compress(1, 1, matrix, &format, &maxrank, &tol, onfly);
}

int starsh_blrm_approximate_batch(STARSH_blrm **matrix, STARSH_blrf **format,
        int nprob, const int *maxrank, const double *tol, int onfly)
{
    if(nprob < 0 || (nprob > 0 && (matrix == NULL || format == NULL ||
                    maxrank == NULL || tol == NULL)))
    {
        STARSH_ERROR("Invalid arguments");
        return STARSH_WRONG_PARAMETER;
    }
    if(nprob == 0)
        return STARSH_SUCCESS;
    for(int q = 0; q < nprob; q++)
        if(format[q] == NULL)
        {
            STARSH_ERROR("Invalid value of `format[%d]`", q);
            return STARSH_WRONG_PARAMETER;
        }
    if(starsh_blrm_approximate == starsh_blrm__drsdd_hip)
        return hip_compress(0, nprob, matrix, format, maxrank, tol, onfly);
    if(starsh_blrm_approximate == starsh_blrm__drsdd_hip_mixed)
        return hip_compress(1, nprob, matrix, format, maxrank, tol, onfly);
    if(starsh_blrm_approximate == NULL)
    {
        STARSH_ERROR("starsh_init() was not called");
        return STARSH_WRONG_PARAMETER;
    }
    /* any other engine: one problem after the other */
    for(int q = 0; q < nprob; q++)
    {
        int info = starsh_blrm_approximate(matrix+q, format[q], maxrank[q],
                tol[q], onfly);
        if(info != STARSH_SUCCESS)
        {
            for(int r = 0; r < q; r++)
            {
                starsh_blrm_free(matrix[r]);
                matrix[r] = NULL;
            }
            return info;
        }
    }
    return STARSH_SUCCESS;
}
