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
#include <sys/mman.h>
#include "common.h"
#include "kernels_host.h"
#include "rsdd.h"
#include "hip_driver.h"

static __thread starsh_hip_stats last_stats;
/* Batches per pipeline run (besides the split by the workspace budget):
 * 0 or 1 runs all tiles of a run together (default), n > 1 splits them into
 * n batches, so that the copy-back of one overlaps the compression of the
 * next.  At C2 (820 tiles) the split costs more device time (9.9 ms in one
 * batch, 14.1 in two, 15.4 in three) than the overlap saves. */
static int pipeline_batches = 0;
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
    if(!strcmp(name, "qr_graph"))
        return starsh_hip_qr_graph((int)value);
    if(!strcmp(name, "proj_int8"))
        return starsh_hip_proj_int8((int)value);
    if(!strcmp(name, "jacobi_max_sweeps"))
        return starsh_hip_jacobi_sweeps((int)value);
    if(!strcmp(name, "mixed_trunc"))
        return hip_mixed_set_eta(value);
    if(!strcmp(name, "mixed_early_dense"))
        return hip_mixed_set_early_dense(value);
    if(!strcmp(name, "mixed_ladder"))
        return hip_mixed_set_ladder((int)value);
    if(!strcmp(name, "tile_schedule"))
        return hip_mixed_set_tile_schedule((int)value);
    if(!strcmp(name, "schedule_clear"))
    {
        hip_sched_clear();
        return STARSH_SUCCESS;
    }
    if(!strcmp(name, "pipeline_batches"))
    {
        if(value < 0 || value > 64)
            return STARSH_WRONG_PARAMETER;
        pipeline_batches = (int)value;
        return STARSH_SUCCESS;
    }
    if(!strcmp(name, "copy_slot_mb"))
        return starsh_hip_copy_option(0, (int)value);
    if(!strcmp(name, "copy_stream"))
        return starsh_hip_copy_option(1, (int)value);
    if(!strcmp(name, "inject_failure"))
        return hip_driver_inject_failure((int)value);
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

STARSH_int hip_driver_chunk_tiles(STARSH_int ntiles)
{
    const STARSH_int nb = pipeline_batches > 1 ? pipeline_batches : 1;
    return (ntiles+nb-1)/nb;
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
    hp->oversample = starsh_params.oversample > 0 ? starsh_params.oversample
        : 10;
    STARSH_cluster *RC = F->row_cluster, *CC = F->col_cluster;
    /* Limits of the device pipeline, checked before any GPU work: tiles of
     * at most HIP_TILE_MAX rows and columns (Gram-Schmidt panels keep a
     * column in registers), sketch widths of at most HIP_SKETCH_MAX (core
     * SVD in one workgroup). */
    for(STARSH_int bi = 0; bi < F->nblocks_far; bi++)
    {
        const STARSH_int m = RC->size[F->block_far[2*bi]];
        const STARSH_int n = CC->size[F->block_far[2*bi+1]];
        const STARSH_int mn = m < n ? m : n;
        const STARSH_int k = (STARSH_int)maxrank+hp->oversample < mn ?
            (STARSH_int)maxrank+hp->oversample : mn;
        if(m > HIP_TILE_MAX || n > HIP_TILE_MAX || k > HIP_SKETCH_MAX)
        {
            STARSH_ERROR("far-field tile %lld is %lld x %lld with sketch "
                    "width %lld; the HIP backend takes tiles of at most %d "
                    "rows and columns and sketch widths "
                    "min(maxrank+oversample, m, n) of at most %d",
                    (long long)bi, (long long)m, (long long)n, (long long)k,
                    HIP_TILE_MAX, HIP_SKETCH_MAX);
            return STARSH_WRONG_PARAMETER;
        }
    }
    int info = starsh_hip_init(0);
    if(info != STARSH_SUCCESS)
        return info;
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

/* Host buffer for factors: 2 MiB aligned and, when large, backed by huge
 * pages where the kernel allows it, so that first touches fault 2 MiB at a
 * time.  Released with free(), as the reference's malloc'd buffers.  Only
 * the pages written are ever backed by memory. */
static double *alloc_host(size_t n)
{
    void *p = NULL;
    const size_t bytes = sizeof(double)*(n > 0 ? n : 1);
    if(posix_memalign(&p, (size_t)2 << 20, bytes) != 0)
        return NULL;
    if(bytes >= ((size_t)4 << 20))
        (void)madvise(p, bytes, MADV_HUGEPAGE);
    return p;
}

/* Sketch width of the plain form for a far-field tile. */
static int full_width(const STARSH_blrf *F, STARSH_int bi, int maxrank,
        int oversample)
{
    const STARSH_int m = F->row_cluster->size[F->block_far[2*bi]];
    const STARSH_int n = F->col_cluster->size[F->block_far[2*bi+1]];
    STARSH_int k = (STARSH_int)maxrank+oversample;
    k = k < m ? k : m;
    return (int)(k < n ? k : n);
}

/* Step 1 (drsdd.c:53-90): host buffers for U and V of every far-field
 * tile.  The reference gives each tile maxrank columns at a fixed offset;
 * here each finished tile takes the next m x r (n x r) block of alloc_U
 * (alloc_V) in the order tiles finish, so the factors are contiguous and
 * the copy from the device writes only what it needs.  The buffers are
 * sized for the largest possible ranks, min(maxrank, k) per tile; pages
 * never written cost nothing.  hf owns what it points to; on failure
 * hip_driver_free_factors releases it. */
int hip_driver_alloc_factors(STARSH_blrf *F, int maxrank, hip_factors *hf)
{
    memset(hf, 0, sizeof(*hf));
    STARSH_int nfar = F->nblocks_far;
    if(nfar == 0)
        return STARSH_SUCCESS;
    hf->far_U = calloc(nfar, sizeof(Array *));
    hf->far_V = calloc(nfar, sizeof(Array *));
    hf->nfar = nfar;
    STARSH_MALLOC(hf->far_rank, nfar);
    STARSH_MALLOC(hf->uoff, nfar);
    STARSH_MALLOC(hf->voff, nfar);
    if(hf->far_U == NULL || hf->far_V == NULL)
        return STARSH_MALLOC_ERROR;
    const int os = starsh_params.oversample > 0 ? starsh_params.oversample :
        10;
    size_t su = 0, sv = 0;
    for(STARSH_int bi = 0; bi < nfar; bi++)
    {
        hf->far_rank[bi] = 0;
        hf->uoff[bi] = hf->voff[bi] = -1;
        int r = full_width(F, bi, maxrank, os);
        r = r < maxrank ? r : maxrank;
        su += (size_t)F->row_cluster->size[F->block_far[2*bi]]*r;
        sv += (size_t)F->col_cluster->size[F->block_far[2*bi+1]]*r;
    }
    hf->ucap = (int64_t)su;
    hf->vcap = (int64_t)sv;
    hf->alloc_U = alloc_host(su);
    hf->alloc_V = alloc_host(sv);
    if(hf->alloc_U == NULL || hf->alloc_V == NULL)
    {
        STARSH_ERROR("malloc() failed");
        return STARSH_MALLOC_ERROR;
    }
    return STARSH_SUCCESS;
}

/* Doubles and segments the factors of a batch of problems can take at
 * most (the device arena of the copy-back). */
void hip_driver_factor_bound(const hip_batch *B, int64_t *doubles,
        int64_t *segments)
{
    *doubles = 0;
    *segments = 0;
    for(int q = 0; q < B->nprob; q++)
    {
        *doubles += B->hf[q].ucap+B->hf[q].vcap;
        *segments += 2*B->hf[q].nfar;
    }
}

/* The U/V wrappers (shape m x r, n x r) of every tile once all copies have
 * landed; false far-field tiles (rank -1) get none. */
int hip_driver_wrap(const hip_batch *B)
{
    for(int q = 0; q < B->nprob; q++)
    {
        hip_factors *hf = B->hf+q;
        const STARSH_blrf *F = B->F[q];
        for(STARSH_int bi = 0; bi < hf->nfar; bi++)
        {
            const int r = hf->far_rank[bi];
            if(r < 0)
                continue;
            const int m = (int)F->row_cluster->size[F->block_far[2*bi]];
            const int n = (int)F->col_cluster->size[F->block_far[2*bi+1]];
            int shape_U[2] = {m, r}, shape_V[2] = {n, r};
            /* rank-0 tiles point at the start of the buffers */
            double *u = hf->alloc_U+(hf->uoff[bi] >= 0 ? hf->uoff[bi] : 0);
            double *v = hf->alloc_V+(hf->voff[bi] >= 0 ? hf->voff[bi] : 0);
            int info = array_from_buffer(hf->far_U+bi, 2, shape_U, 'd', 'F',
                    u);
            if(info == STARSH_SUCCESS)
                info = array_from_buffer(hf->far_V+bi, 2, shape_V, 'd', 'F',
                        v);
            if(info != STARSH_SUCCESS)
                return info;
        }
    }
    return STARSH_SUCCESS;
}

/* Releases what hf holds.  The wrappers only view alloc_U/alloc_V, so
 * their data pointers are detached before array_free. */
void hip_driver_free_factors(STARSH_blrf *F, hip_factors *hf)
{
    (void)F;
    for(STARSH_int bi = 0; hf->far_U && bi < hf->nfar; bi++)
        if(hf->far_U[bi])
        {
            hf->far_U[bi]->data = NULL;
            array_free(hf->far_U[bi]);
        }
    for(STARSH_int bi = 0; hf->far_V && bi < hf->nfar; bi++)
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

/* Releases near-field wrappers and their buffer after a failure. */
static void free_near(Array **near_D, STARSH_int n, double *alloc_D)
{
    for(STARSH_int bi = 0; near_D && bi < n; bi++)
        if(near_D[bi])
        {
            near_D[bi]->data = NULL;
            array_free(near_D[bi]);
        }
    free(near_D);
    free(alloc_D);
}

/* Fault injection for tests of the error paths: the next call of
 * inject_failure() with this site returns nonzero once. */
static int injected_site = 0;

int hip_driver_inject_failure(int site)
{
    if(site < 0)
        return STARSH_WRONG_PARAMETER;
    injected_site = site;
    return STARSH_SUCCESS;
}

static int inject_failure(int site)
{
    if(injected_site != 0 && injected_site == site)
    {
        injected_site = 0;
        return 1;
    }
    return 0;
}

/* Steps 3-5 (drsdd.c:146-306).  After each step hf describes exactly the
 * far-field arrays still owned by the driver, so that a failure at any
 * point can be cleaned up by hip_driver_free_factors. */
int hip_driver_finish(STARSH_blrm **matrix, STARSH_blrf *format,
        const hip_problem *hp, hip_factors *hf, int onfly,
        starsh_hip_stats *st, hip_near *nf)
{
    STARSH_blrf *F = format;
    STARSH_problem *P = F->problem;
    STARSH_cluster *RC = F->row_cluster, *CC = F->col_cluster;
    STARSH_int nblocks_far = F->nblocks_far, nblocks_near = F->nblocks_near;
    STARSH_int new_far = nblocks_far, new_near = nblocks_near;
    STARSH_int *block_near = F->block_near;
    STARSH_int nfalse = 0, bi, bj;
    for(bi = 0; bi < nblocks_far; bi++)
        if(hf->far_rank[bi] == -1)
            nfalse++;
    st->nfalse_far += nfalse;
    if(nfalse > 0)
    {
        STARSH_int *false_far = malloc(sizeof(STARSH_int)*nfalse);
        STARSH_int *bnear = NULL, *bfar = NULL;
        new_near = nblocks_near+nfalse;
        new_far = nblocks_far-nfalse;
        bnear = malloc(sizeof(STARSH_int)*2*new_near);
        if(new_far > 0)
            bfar = malloc(sizeof(STARSH_int)*2*new_far);
        if(false_far == NULL || bnear == NULL || (new_far > 0 &&
                    bfar == NULL) || inject_failure(1))
        {
            free(false_far);
            free(bnear);
            free(bfar);
            STARSH_ERROR("malloc() failed");
            return STARSH_MALLOC_ERROR;
        }
        for(bi = 0, bj = 0; bi < nblocks_far; bi++)
            if(hf->far_rank[bi] == -1)
                false_far[bj++] = bi;
        for(bi = 0; bi < 2*nblocks_near; bi++)
            bnear[bi] = F->block_near[bi];
        for(bi = 0; bi < nfalse; bi++)
        {
            bnear[2*(bi+nblocks_near)] = F->block_far[2*false_far[bi]];
            bnear[2*(bi+nblocks_near)+1] = F->block_far[2*false_far[bi]+1];
        }
        for(bi = 0, bj = 0; bi < nblocks_far; bi++)
        {
            if(bj < nfalse && false_far[bj] == bi)
                bj++;
            else
            {
                bfar[2*(bi-bj)] = F->block_far[2*bi];
                bfar[2*(bi-bj)+1] = F->block_far[2*bi+1];
            }
        }
        free(false_far);
        /* on success starsh_blrf_new_from_coo owns both lists */
        STARSH_blrf *F2;
        int info = inject_failure(4) ? STARSH_MALLOC_ERROR :
            starsh_blrf_new_from_coo(&F2, P, F->symm, RC, CC, new_far, bfar,
                    new_near, bnear, F->type);
        if(info != STARSH_SUCCESS)
        {
            free(bnear);
            free(bfar);
            return info;
        }
        STARSH_blrf tmp = *F;
        *F = *F2;
        *F2 = tmp;
        STARSH_WARNING("`F` was modified due to false far-field blocks");
        starsh_blrf_free(F2);
        block_near = F->block_near;
        /* Compaction of the far-field arrays (drsdd.c:259-282): false
         * far-field tiles have no wrappers (rank -1, no factors). */
        for(bi = 0, bj = 0; bi < nblocks_far; bi++)
        {
            if(hf->far_rank[bi] == -1)
            {
                bj++;
                continue;
            }
            hf->far_U[bi-bj] = hf->far_U[bi];
            hf->far_V[bi-bj] = hf->far_V[bi];
            hf->far_rank[bi-bj] = hf->far_rank[bi];
            if(bj > 0)
                hf->far_U[bi] = hf->far_V[bi] = NULL;
        }
        hf->nfar = new_far;
        if(new_far == 0)
        {
            free(hf->alloc_U);
            free(hf->alloc_V);
            hf->alloc_U = hf->alloc_V = NULL;
        }
    }
    /* Near-field tiles (drsdd.c:214-256), generated on the GPU. */
    Array **near_D = NULL;
    double *alloc_D = NULL;
    if(onfly == 0 && new_near > 0)
    {
        near_D = calloc(new_near, sizeof(Array *));
        int64_t *r0 = malloc(sizeof(int64_t)*new_near);
        int64_t *c0 = malloc(sizeof(int64_t)*new_near);
        int64_t *off = malloc(sizeof(int64_t)*new_near);
        int *mm = malloc(sizeof(int)*new_near), *nn = malloc(sizeof(int)*
                new_near);
        int info = STARSH_SUCCESS;
        size_t size_D = 0;
        if(!near_D || !r0 || !c0 || !off || !mm || !nn || inject_failure(2))
            info = STARSH_MALLOC_ERROR;
        for(bi = 0; info == STARSH_SUCCESS && bi < new_near; bi++)
        {
            STARSH_int i = block_near[2*bi], j = block_near[2*bi+1];
            r0[bi] = RC->start[i];
            c0[bi] = CC->start[j];
            mm[bi] = (int)RC->size[i];
            nn[bi] = (int)CC->size[j];
            off[bi] = size_D;
            size_D += (size_t)mm[bi]*nn[bi];
        }
        double *d_D = NULL;
        if(info == STARSH_SUCCESS)
        {
            alloc_D = alloc_host(size_D);
            if(alloc_D == NULL)
                info = STARSH_MALLOC_ERROR;
        }
        if(info == STARSH_SUCCESS)
            info = inject_failure(3) ? STARSH_MALLOC_ERROR :
                starsh_hip_malloc((void **)&d_D, sizeof(double)*size_D);
        if(info == STARSH_SUCCESS)
            info = starsh_hip_generate_tiles(&hp->kd, hp->d_xr, hp->nr,
                    hp->d_xc, hp->nc, (int)new_near, r0, c0, mm, nn, off,
                    d_D);
        /* the tiles go to alloc_D through the copy-back engine, which
         * the caller starts once every problem is finished */
        nf->d = d_D;
        nf->h = alloc_D;
        nf->n = (int64_t)size_D;
        for(bi = 0; bi < new_near && info == STARSH_SUCCESS; bi++)
        {
            int shape[2] = {mm[bi], nn[bi]};
            info = array_from_buffer(near_D+bi, 2, shape, 'd', 'F',
                    alloc_D+off[bi]);
        }
        free(r0); free(c0); free(off); free(mm); free(nn);
        if(info != STARSH_SUCCESS)
        {
            STARSH_ERROR("near-field tiles failed");
            free_near(near_D, new_near, alloc_D);
            return info;
        }
    }
    int *far_rank = hf->far_rank;
    Array **far_U = hf->far_U, **far_V = hf->far_V;
    double *alloc_U = hf->alloc_U, *alloc_V = hf->alloc_V;
    if(new_far == 0)
    {
        free(far_rank);
        free(far_U);
        free(far_V);
        far_rank = NULL;
        far_U = far_V = NULL;
    }
    else
    {
        /* shrink the pointer arrays; a failed shrink keeps the old ones */
        void *q;
        if((q = realloc(far_rank, sizeof(int)*new_far)) != NULL)
            far_rank = q;
        if((q = realloc(far_U, sizeof(Array *)*new_far)) != NULL)
            far_U = q;
        if((q = realloc(far_V, sizeof(Array *)*new_far)) != NULL)
            far_V = q;
    }
    free(hf->uoff);
    free(hf->voff);
    memset(hf, 0, sizeof(*hf));
    int info = starsh_blrm_new(matrix, F, far_rank, far_U, far_V, onfly,
            near_D, alloc_U, alloc_V, alloc_D, '1');
    if(info != STARSH_SUCCESS)
    {
        /* give the buffers back to hf so the caller's cleanup frees them */
        hf->far_rank = far_rank;
        hf->far_U = far_U;
        hf->far_V = far_V;
        hf->alloc_U = alloc_U;
        hf->alloc_V = alloc_V;
        hf->nfar = new_far;
        free_near(near_D, onfly == 0 ? new_near : 0, alloc_D);
    }
    return info;
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
        const STARSH_int *idx, const int *k, int int8, double tau, double eta,
        int *rank_out, int *flags_out, int *accepted, hip_accept_fn accept,
        double *S_out, int sstride, starsh_hip_stats *st,
        const hip_run_extra *x)
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
    int64_t *xso = malloc(sizeof(int64_t)*ntiles);
    int64_t *xpo = malloc(sizeof(int64_t)*ntiles);
    if(!m || !n || !pp || !kc || !r0 || !c0 || !uo || !vo || !probs || !du ||
            !dv || !so || !xso || !xpo)
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
    /* batches of at most chunk_tiles tiles, so that the copy-back of one
     * overlaps the compression of the next */
    const STARSH_int chunk_tiles = hip_driver_chunk_tiles(ntiles);
    STARSH_int b0 = 0;
    while(info == STARSH_SUCCESS && b0 < ntiles)
    {
        size_t bytes = 0;
        STARSH_int b1 = b0;
        while(b1 < ntiles && (chunk_tiles <= 0 || b1-b0 < chunk_tiles))
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
        in.eta = eta;
        if(S_out)
        {
            in.S_out = S_out;
            in.soff = so+b0;
        }
        if(x)
        {
            /* offsets relative to the chunk's first tile */
            in.rule = x->rule;
            if(x->sel)
            {
                for(STARSH_int t = b0; t < b1; t++)
                    xso[t] = x->sel_off[t]-x->sel_off[b0];
                in.sel = x->sel+x->sel_off[b0];
                in.sel_off = xso+b0;
            }
            if(x->piv_k)
            {
                for(STARSH_int t = b0; t < b1; t++)
                    xpo[t] = x->piv_off[t]-x->piv_off[b0];
                in.piv_k = x->piv_k+b0;
                in.piv = x->piv+x->piv_off[b0];
                in.piv_off = xpo+b0;
                in.piv_done = x->piv_done+b0;
            }
        }
        RsddOutput out;
        memset(&out, 0, sizeof(out));
        out.rank = rank_out+b0;
        out.p = pp+b0;
        out.kc = kc+b0;
        out.flags = flags_out+b0;
        out.uoff = uo+b0;
        out.voff = vo+b0;
        double tb0 = hip_driver_now();
        info = rsdd_batch(&in, &out);
        st->batch_seconds += hip_driver_now()-tb0;
        if(info != STARSH_SUCCESS)
            break;
        st->device_seconds += out.device_seconds;
        st->dominant_kernel_seconds += out.sketch_seconds;
        st->proj_kernel_seconds += out.proj_seconds;
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
                    flags_out[t], B->hp[q].oversample, kc[t], mn) : 1;
            if(x && x->kc_out)
                x->kc_out[t] = kc[t];
        }
        /* finished tiles take the next blocks of their problem's host
         * buffers; their factors are queued for the copy-back, which runs
         * while the next batch computes */
        for(STARSH_int t = b0; t < b1; t++)
        {
            du[t] = dv[t] = NULL;
            if(!accepted[t])
                continue;
            hip_factors *hf = B->hf+pidx[t];
            const STARSH_int bi = idx[t];
            const int r = rank_out[t];
            hf->far_rank[bi] = r;
            if(r <= 0)
                continue;
            if(hf->uused+(int64_t)m[t]*r > hf->ucap ||
                    hf->vused+(int64_t)n[t]*r > hf->vcap)
            {
                info = STARSH_UNKNOWN_ERROR;
                break;
            }
            hf->uoff[bi] = hf->uused;
            hf->voff[bi] = hf->vused;
            hf->uused += (int64_t)m[t]*r;
            hf->vused += (int64_t)n[t]*r;
            du[t] = hf->alloc_U+hf->uoff[bi];
            dv[t] = hf->alloc_V+hf->voff[bi];
        }
        if(info == STARSH_SUCCESS)
            info = rsdd_out_append(&out, (int)(b1-b0), m+b0, n+b0,
                    accepted+b0, du+b0, dv+b0);
        st->copy_seconds += hip_driver_now()-tc;
        b0 = b1;
    }
    free(m); free(n); free(pp); free(kc); free(r0); free(c0); free(uo);
    free(vo); free(probs); free(du); free(dv); free(so); free(xso);
    free(xpo);
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
        info = hip_driver_run(B, ntiles, pidx, idx, k, 0, 0.0, 0.0, rk, fl, acc,
                NULL, NULL, 0, st, NULL);
    free(k); free(rk); free(fl); free(acc);
    return info;
}

/* Kernel family and dimension of a prepared problem, as one sort key. */
static int group_key(const hip_problem *hp)
{
    return hp->kd.family*16+hp->kd.ndim;
}

int hip_compress(int engine, int nprob, STARSH_blrm **matrix,
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
    int out_running = 0;
    if(info == STARSH_SUCCESS)
    {
        int64_t nd, ns;
        hip_driver_factor_bound(&B, &nd, &ns);
        info = rsdd_out_begin(nd, ns);
        out_running = info == STARSH_SUCCESS;
    }
    /* knobs recorded by the autotuner for this problem, if any */
    if(info == STARSH_SUCCESS)
        hip_sched_apply(engine == HIP_ENGINE_MIXED, &B);
    st.prepare_seconds = hip_driver_now()-t_start;
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
        if(engine == HIP_ENGINE_MIXED)
            info = hip_mixed_tiles(&B, nt, pidx, idx, &st);
        else if(engine == HIP_ENGINE_SDD || engine == HIP_ENGINE_RRQR)
            info = hip_exact_tiles(&B, nt, pidx, idx,
                    engine == HIP_ENGINE_RRQR, &st);
        else
            info = hip_plain_tiles(&B, nt, pidx, idx, &st);
    }
    free(pidx); free(idx); free(done);
    if(info == STARSH_SUCCESS)
        info = hip_driver_wrap(&B);
    if(info == STARSH_SUCCESS)
        hip_driver_rank_stats(&B, &st);
    /* the format rewrite and the near-field tiles, whose copy to the host
     * joins the factors' in the copy-back engine */
    hip_near *nf = calloc(nprob, sizeof(hip_near));
    if(info == STARSH_SUCCESS && nf == NULL)
        info = STARSH_MALLOC_ERROR;
    double tf = hip_driver_now();
    for(int q = 0; info == STARSH_SUCCESS && q < nprob; q++)
        info = hip_driver_finish(matrix+q, format[q], B.hp+q, B.hf+q, onfly,
                &st, nf+q);
    for(int q = 0; out_running && info == STARSH_SUCCESS && q < nprob; q++)
        if(nf[q].d != NULL)
        {
            const int64_t zero = 0;
            info = rsdd_out_submit(nf[q].d, nf[q].n, 1, &zero, &nf[q].n,
                    &nf[q].h);
        }
    st.finish_seconds = hip_driver_now()-tf;
    if(out_running)
    {
        /* the copy threads must be done with the host buffers before
         * anything else touches or frees them, also after an error */
        double td = hip_driver_now();
        int e = rsdd_out_end(&st.d2h_seconds, &st.scatter_seconds);
        st.copy_seconds += hip_driver_now()-td;
        if(info == STARSH_SUCCESS)
            info = e;
    }
    for(int q = 0; nf && q < nprob; q++)
        if(nf[q].d)
            starsh_hip_free(nf[q].d);
    free(nf);
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

int starsh_blrm__dsdd_hip(STARSH_blrm **matrix, STARSH_blrf *format,
        int maxrank, double tol, int onfly)
{
    if(matrix == NULL || format == NULL)
    {
        STARSH_ERROR("Invalid value of `matrix` or `format`");
        return STARSH_WRONG_PARAMETER;
    }
    return hip_compress(HIP_ENGINE_SDD, 1, matrix, &format, &maxrank, &tol,
            onfly);
}

int starsh_blrm__dqp3_hip(STARSH_blrm **matrix, STARSH_blrf *format,
        int maxrank, double tol, int onfly)
{
    if(matrix == NULL || format == NULL)
    {
        STARSH_ERROR("Invalid value of `matrix` or `format`");
        return STARSH_WRONG_PARAMETER;
    }
    return hip_compress(HIP_ENGINE_RRQR, 1, matrix, &format, &maxrank, &tol,
            onfly);
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
    if(starsh_blrm_approximate == starsh_blrm__dsdd_hip)
        return hip_compress(HIP_ENGINE_SDD, nprob, matrix, format, maxrank,
                tol, onfly);
    if(starsh_blrm_approximate == starsh_blrm__dqp3_hip)
        return hip_compress(HIP_ENGINE_RRQR, nprob, matrix, format, maxrank,
                tol, onfly);
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
