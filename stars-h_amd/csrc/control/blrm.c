/* Block low-rank matrix container.  Follows src/control/blrm.c:17-322 of
 * ecrc/stars-h 0.3.1: far_U/far_V wrap slices of alloc_U/alloc_V (alloc_type
 * '1') and near_D wrap slices of alloc_D; starsh_blrm_free releases the three
 * buffers with free(). */
#include "common.h"

int starsh_blrm_new(STARSH_blrm **matrix, STARSH_blrf *format, int *far_rank,
        Array **far_U, Array **far_V, int onfly, Array **near_D, void *alloc_U,
        void *alloc_V, void *alloc_D, char alloc_type)
{
    if(matrix == NULL)
    {
        STARSH_ERROR("Invalid value of `matrix`");
        return STARSH_WRONG_PARAMETER;
    }
    if(format == NULL)
    {
        STARSH_ERROR("Invalid value of `format`");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_blrf *F = format;
    if(F->nblocks_far > 0 && (far_rank == NULL || far_U == NULL ||
                far_V == NULL))
    {
        STARSH_ERROR("Invalid value of `far_rank`, `far_U` or `far_V`");
        return STARSH_WRONG_PARAMETER;
    }
    if(onfly != 0 && onfly != 1)
    {
        STARSH_ERROR("Invalid value of `onfly`");
        return STARSH_WRONG_PARAMETER;
    }
    if(near_D == NULL && F->nblocks_near > 0 && onfly == 0)
    {
        STARSH_ERROR("Invalid value of `near_D`");
        return STARSH_WRONG_PARAMETER;
    }
    if(alloc_type != '1' && alloc_type != '2')
    {
        STARSH_ERROR("Invalid value of `alloc_type`");
        return STARSH_WRONG_PARAMETER;
    }
    if(alloc_type == '1' && ((F->nblocks_far > 0 && (alloc_U == NULL ||
                        alloc_V == NULL)) || (F->nblocks_near > 0 &&
                        onfly == 0 && alloc_D == NULL)))
    {
        STARSH_ERROR("Invalid value of `alloc_U`, `alloc_V` or `alloc_D`");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_blrm *M;
    STARSH_MALLOC(M, 1);
    M->format = F;
    M->far_rank = far_rank;
    M->far_U = far_U;
    M->far_V = far_V;
    M->onfly = onfly;
    M->near_D = near_D;
    M->alloc_U = alloc_U;
    M->alloc_V = alloc_V;
    M->alloc_D = alloc_D;
    M->alloc_type = alloc_type;
    size_t size = sizeof(*M), data_size = 0;
    size += F->nblocks_far*(sizeof(*far_rank)+sizeof(*far_U)+sizeof(*far_V));
    for(STARSH_int bi = 0; bi < F->nblocks_far; bi++)
    {
        size += far_U[bi]->nbytes+far_V[bi]->nbytes;
        data_size += far_U[bi]->data_nbytes+far_V[bi]->data_nbytes;
    }
    if(onfly == 0)
    {
        size += F->nblocks_near*sizeof(*near_D);
        for(STARSH_int bi = 0; bi < F->nblocks_near; bi++)
        {
            size += near_D[bi]->nbytes;
            data_size += near_D[bi]->data_nbytes;
        }
    }
    M->nbytes = size;
    M->data_nbytes = data_size;
    *matrix = M;
    return STARSH_SUCCESS;
}

void starsh_blrm_free(STARSH_blrm *matrix)
{
    STARSH_blrm *M = matrix;
    if(M == NULL)
        return;
    STARSH_blrf *F = M->format;
    if(F->nblocks_far > 0)
    {
        for(STARSH_int bi = 0; bi < F->nblocks_far; bi++)
        {
            if(M->alloc_type == '1')
            {
                M->far_U[bi]->data = NULL;
                M->far_V[bi]->data = NULL;
            }
            array_free(M->far_U[bi]);
            array_free(M->far_V[bi]);
        }
        if(M->alloc_type == '1')
        {
            free(M->alloc_U);
            free(M->alloc_V);
        }
        free(M->far_rank);
        free(M->far_U);
        free(M->far_V);
    }
    if(F->nblocks_near > 0 && M->onfly == 0)
    {
        for(STARSH_int bi = 0; bi < F->nblocks_near; bi++)
        {
            if(M->alloc_type == '1')
                M->near_D[bi]->data = NULL;
            array_free(M->near_D[bi]);
        }
        if(M->alloc_type == '1')
            free(M->alloc_D);
        free(M->near_D);
    }
    free(M);
}

void starsh_blrm_info(STARSH_blrm *matrix)
{
    STARSH_blrm *M = matrix;
    if(M == NULL)
        return;
    printf("<STARSH_blrm at %p, %d onfly, allocation type '%c', %f MB memory "
            "footprint>\n", (void *)M, M->onfly, M->alloc_type,
            M->nbytes/1024./1024.);
}

/* Factors of tile (i, j) if it is far-field, otherwise its dense values
 * (stored, or computed into a new buffer when onfly == 1), as in
 * blrm.c:200-322.  The far-field lookup uses the tile index brow_far[k]. */
int starsh_blrm_get_block(STARSH_blrm *matrix, STARSH_int i, STARSH_int j,
        int *shape, int *rank, void **U, void **V, void **D)
{
    STARSH_blrm *M = matrix;
    if(M == NULL || shape == NULL || rank == NULL || U == NULL || V == NULL ||
            D == NULL)
    {
        STARSH_ERROR("Invalid argument");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_blrf *F = M->format;
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
    *rank = nrows < ncols ? nrows : ncols;
    *U = NULL;
    *V = NULL;
    *D = NULL;
    if(F->nblocks_far > 0)
        for(STARSH_int k = F->brow_far_start[i]; k < F->brow_far_start[i+1];
                k++)
        {
            STARSH_int bi = F->brow_far[k];
            if(F->block_far[2*bi+1] == j)
            {
                *rank = M->far_rank[bi];
                *U = M->far_U[bi]->data;
                *V = M->far_V[bi]->data;
                return STARSH_SUCCESS;
            }
        }
    if(F->nblocks_near > 0)
        for(STARSH_int k = F->brow_near_start[i]; k < F->brow_near_start[i+1];
                k++)
        {
            STARSH_int bi = F->brow_near[k];
            if(F->block_near[2*bi+1] == j)
            {
                if(M->onfly == 0)
                    *D = M->near_D[bi]->data;
                else
                {
                    double *buf = malloc(sizeof(double)*(size_t)nrows*ncols
                            + 1);
                    if(buf == NULL)
                        return STARSH_MALLOC_ERROR;
                    P->kernel(nrows, ncols, R->pivot+R->start[i],
                            C->pivot+C->start[j], R->data, C->data, buf,
                            nrows);
                    *D = buf;
                }
                return STARSH_SUCCESS;
            }
        }
    STARSH_ERROR("Requested block is not admissible in the format");
    return STARSH_WRONG_PARAMETER;
}

/* Rank from singular values (src/backends/sequential/dense/dsvfr.c:16-47):
 * the total sum of squares is accumulated in index order, then ranks size,
 * size-1, ... are tried while the discarded tail stays within tol^2 of it.
 * The summation order is part of the contract: the GPU core kernel repeats
 * it to make identical decisions on identical singular values. */
int starsh_dense_dsvfr(int size, double *S, double tol)
{
    double total = 0;
    for(int i = 0; i < size; i++)
        total += S[i]*S[i];
    if(total == 0)
        return 0;
    int r = size;
    double limit = total*(tol*tol);
    double tail = S[size-1]*S[size-1];
    while(r > 1 && limit >= tail)
    {
        r--;
        tail += S[r-1]*S[r-1];
    }
    return r;
}
