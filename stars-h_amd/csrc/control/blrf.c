/* Tile low-rank (TLR) block format: COO lists of far-field and near-field
 * tiles plus CSR/CSC views.  Follows src/control/blrf.c:90-526 of
 * ecrc/stars-h 0.3.1.  For symm == 'S' only tiles (i, j) with j <= i are
 * stored and the column views alias the row views. */
#include "common.h"

/* Builds start/list arrays grouping the `n` pairs of `coo` by coo[2*b+k]. */
static int group_by(STARSH_int nb, STARSH_int n, const STARSH_int *coo, int k,
        STARSH_int **start_out, STARSH_int **list_out)
{
    STARSH_int *start = malloc(sizeof(*start)*(nb+1));
    STARSH_int *list = malloc(sizeof(*list)*n+1);
    STARSH_int *fill = calloc(nb+1, sizeof(*fill));
    if(start == NULL || list == NULL || fill == NULL)
    {
        free(start);
        free(list);
        free(fill);
        STARSH_ERROR("malloc() failed");
        return STARSH_MALLOC_ERROR;
    }
    for(STARSH_int b = 0; b < n; b++)
        fill[coo[2*b+k]]++;
    start[0] = 0;
    for(STARSH_int i = 0; i < nb; i++)
        start[i+1] = start[i]+fill[i];
    memset(fill, 0, sizeof(*fill)*(nb+1));
    for(STARSH_int b = 0; b < n; b++)
    {
        STARSH_int i = coo[2*b+k];
        list[start[i]+fill[i]] = b;
        fill[i]++;
    }
    free(fill);
    *start_out = start;
    *list_out = list;
    return STARSH_SUCCESS;
}

int starsh_blrf_new_from_coo(STARSH_blrf **format, STARSH_problem *problem,
        char symm, STARSH_cluster *row_cluster, STARSH_cluster *col_cluster,
        STARSH_int nblocks_far, STARSH_int *block_far,
        STARSH_int nblocks_near, STARSH_int *block_near,
        enum STARSH_BLRF_TYPE type)
{
    if(format == NULL || problem == NULL || row_cluster == NULL ||
            col_cluster == NULL)
    {
        STARSH_ERROR("Invalid value of `format`, `problem` or clusters");
        return STARSH_WRONG_PARAMETER;
    }
    if(symm != 'S' && symm != 'N')
    {
        STARSH_ERROR("Invalid value of `symm`");
        return STARSH_WRONG_PARAMETER;
    }
    if((nblocks_far > 0 && block_far == NULL) ||
            (nblocks_near > 0 && block_near == NULL))
    {
        STARSH_ERROR("Invalid value of `block_far` or `block_near`");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_blrf *F;
    STARSH_MALLOC(F, 1);
    memset(F, 0, sizeof(*F));
    F->problem = problem;
    F->symm = symm;
    F->row_cluster = row_cluster;
    F->nbrows = row_cluster->nblocks;
    F->nblocks_far = nblocks_far;
    F->block_far = nblocks_far > 0 ? block_far : NULL;
    F->nblocks_near = nblocks_near;
    F->block_near = nblocks_near > 0 ? block_near : NULL;
    F->nblocks_far_local = nblocks_far;
    F->block_far_local = F->block_far;
    F->nblocks_near_local = nblocks_near;
    F->block_near_local = F->block_near;
    F->type = type;
    int info = STARSH_SUCCESS;
    if(nblocks_far > 0)
        info = group_by(F->nbrows, nblocks_far, block_far, 0,
                &F->brow_far_start, &F->brow_far);
    if(info == STARSH_SUCCESS && nblocks_near > 0)
        info = group_by(F->nbrows, nblocks_near, block_near, 0,
                &F->brow_near_start, &F->brow_near);
    if(info == STARSH_SUCCESS && symm == 'N')
    {
        F->col_cluster = col_cluster;
        F->nbcols = col_cluster->nblocks;
        if(nblocks_far > 0)
            info = group_by(F->nbcols, nblocks_far, block_far, 1,
                    &F->bcol_far_start, &F->bcol_far);
        if(info == STARSH_SUCCESS && nblocks_near > 0)
            info = group_by(F->nbcols, nblocks_near, block_near, 1,
                    &F->bcol_near_start, &F->bcol_near);
    }
    if(info != STARSH_SUCCESS)
    {
        /* the caller keeps the COO lists on failure; everything built
         * here is released */
        free(F->brow_far_start);
        free(F->brow_far);
        free(F->brow_near_start);
        free(F->brow_near);
        free(F->bcol_far_start);
        free(F->bcol_far);
        free(F->bcol_near_start);
        free(F->bcol_near);
        free(F);
        return info;
    }
    if(symm == 'S')
    {
        F->col_cluster = row_cluster;
        F->nbcols = row_cluster->nblocks;
        F->bcol_far_start = F->brow_far_start;
        F->bcol_far = F->brow_far;
        F->bcol_near_start = F->brow_near_start;
        F->bcol_near = F->brow_near;
    }
    *format = F;
    return STARSH_SUCCESS;
}

/* All tiles start as far-field, the diagonal included (blrf.c:291-371). */
int starsh_blrf_new_tlr(STARSH_blrf **format, STARSH_problem *problem,
        char symm, STARSH_cluster *row_cluster, STARSH_cluster *col_cluster)
{
    if(format == NULL || problem == NULL || row_cluster == NULL ||
            col_cluster == NULL)
    {
        STARSH_ERROR("Invalid value of `format`, `problem` or clusters");
        return STARSH_WRONG_PARAMETER;
    }
    if(symm != 'S' && symm != 'N')
    {
        STARSH_ERROR("Invalid value of `symm`");
        return STARSH_WRONG_PARAMETER;
    }
    if(symm == 'S' && problem->symm == 'N')
    {
        STARSH_ERROR("Invalid value of `symm`");
        return STARSH_WRONG_PARAMETER;
    }
    if(symm == 'S' && row_cluster != col_cluster)
    {
        STARSH_ERROR("`row_cluster` and `col_cluster` should be equal");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_int nbrows = row_cluster->nblocks, nbcols = col_cluster->nblocks;
    STARSH_int nfar = symm == 'N' ? nbrows*nbcols : nbrows*(nbrows+1)/2;
    STARSH_int *block_far, k = 0;
    STARSH_MALLOC(block_far, 2*nfar);
    for(STARSH_int i = 0; i < nbrows; i++)
    {
        STARSH_int jend = symm == 'N' ? nbcols : i+1;
        for(STARSH_int j = 0; j < jend; j++)
        {
            block_far[2*k] = i;
            block_far[2*k+1] = j;
            k++;
        }
    }
    int info = starsh_blrf_new_from_coo(format, problem, symm, row_cluster,
            col_cluster, nfar, block_far, 0, NULL, STARSH_TLR);
    if(info != STARSH_SUCCESS)
        free(block_far);
    return info;
}

void starsh_blrf_free(STARSH_blrf *format)
{
    if(format == NULL)
        return;
    if(format->nblocks_far > 0)
    {
        free(format->block_far);
        free(format->brow_far_start);
        free(format->brow_far);
    }
    if(format->nblocks_near > 0)
    {
        free(format->block_near);
        free(format->brow_near_start);
        free(format->brow_near);
    }
    if(format->symm == 'N')
    {
        if(format->nblocks_far > 0)
        {
            free(format->bcol_far_start);
            free(format->bcol_far);
        }
        if(format->nblocks_near > 0)
        {
            free(format->bcol_near_start);
            free(format->bcol_near);
        }
    }
    free(format);
}

void starsh_blrf_info(STARSH_blrf *format)
{
    STARSH_blrf *F = format;
    if(F == NULL)
    {
        STARSH_ERROR("Invalid value of `format`");
        return;
    }
    printf("<STARSH_blrf at %p, '%c' symmetric, %zd block rows, %zd "
            "block columns, %zd far-field blocks, %zd near-field blocks>\n",
            (void *)F, F->symm, (ssize_t)F->nbrows, (ssize_t)F->nbcols,
            (ssize_t)F->nblocks_far, (ssize_t)F->nblocks_near);
}

void starsh_blrf_print(STARSH_blrf *format)
{
    STARSH_blrf *F = format;
    if(F == NULL)
    {
        STARSH_ERROR("Invalid value of `format`");
        return;
    }
    starsh_blrf_info(F);
    for(int pass = 0; pass < 2; pass++)
    {
        STARSH_int n = pass ? F->nblocks_near : F->nblocks_far;
        STARSH_int *st = pass ? F->brow_near_start : F->brow_far_start;
        STARSH_int *li = pass ? F->brow_near : F->brow_far;
        if(n == 0)
            continue;
        for(STARSH_int i = 0; i < F->nbrows; i++)
        {
            if(st[i+1] == st[i])
                continue;
            printf("Admissible %s-field blocks for block row %zd:",
                    pass ? "near" : "far", (ssize_t)i);
            for(STARSH_int j = st[i]; j < st[i+1]; j++)
                printf(" %zd", (ssize_t)li[j]);
            printf("\n");
        }
    }
}

/* Dense tile (i, j) computed by the problem kernel into a new malloc'd
 * buffer (blrf.c:471-526). */
int starsh_blrf_get_block(STARSH_blrf *format, STARSH_int i, STARSH_int j,
        int *shape, void **D)
{
    STARSH_blrf *F = format;
    if(F == NULL || shape == NULL || D == NULL)
    {
        STARSH_ERROR("Invalid argument");
        return STARSH_WRONG_PARAMETER;
    }
    if(i < 0 || i >= F->nbrows || j < 0 || j >= F->nbcols)
    {
        STARSH_ERROR("Invalid value of `i` or `j`");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_problem *P = F->problem;
    if(P->ndim != 2)
    {
        STARSH_ERROR("Only scalar kernels are supported");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_cluster *R = F->row_cluster, *C = F->col_cluster;
    int nrows = R->size[i], ncols = C->size[j];
    shape[0] = nrows;
    shape[1] = ncols;
    double *buf = malloc(sizeof(double)*(size_t)nrows*ncols + 1);
    if(buf == NULL)
        return STARSH_MALLOC_ERROR;
    P->kernel(nrows, ncols, R->pivot+R->start[i], C->pivot+C->start[j],
            R->data, C->data, buf, nrows);
    *D = buf;
    return STARSH_SUCCESS;
}
