/* Matrix-free operator: kernel pointer plus row/column data.
 * Follows src/control/problem.c:16-316 of ecrc/stars-h 0.3.1
 * (starsh_problem_new, _free, _info, _get_block, _to_array). */
#include "common.h"

int starsh_problem_new(STARSH_problem **problem, int ndim, STARSH_int *shape,
        char symm, char dtype, void *row_data, void *col_data,
        STARSH_kernel *kernel, char *name)
{
    if(problem == NULL)
    {
        STARSH_ERROR("Invalid value of `problem`");
        return STARSH_WRONG_PARAMETER;
    }
    if(ndim < 2)
    {
        STARSH_ERROR("Invalid value of `ndim`");
        return STARSH_WRONG_PARAMETER;
    }
    if(shape == NULL)
    {
        STARSH_ERROR("Invalid value of `shape`");
        return STARSH_WRONG_PARAMETER;
    }
    if(kernel == NULL)
    {
        STARSH_ERROR("Invalid value of `kernel`");
        return STARSH_WRONG_PARAMETER;
    }
    size_t dtype_size;
    switch(dtype)
    {
        case 's': dtype_size = 4; break;
        case 'd': dtype_size = 8; break;
        case 'c': dtype_size = 8; break;
        case 'z': dtype_size = 16; break;
        default:
            STARSH_ERROR("Invalid value of `dtype`");
            return STARSH_WRONG_PARAMETER;
    }
    size_t entry_size = dtype_size;
    for(int i = 1; i < ndim-1; i++)
        entry_size *= shape[i];
    STARSH_problem *P;
    STARSH_MALLOC(P, 1);
    P->shape = malloc(sizeof(*P->shape)*ndim);
    if(P->shape == NULL)
    {
        free(P);
        STARSH_ERROR("malloc() failed");
        return STARSH_MALLOC_ERROR;
    }
    memcpy(P->shape, shape, sizeof(*P->shape)*ndim);
    P->ndim = ndim;
    P->symm = symm;
    P->dtype = dtype;
    P->dtype_size = dtype_size;
    P->entry_size = entry_size;
    P->row_data = row_data;
    P->col_data = col_data;
    P->kernel = kernel;
    P->name = NULL;
    if(name != NULL)
    {
        P->name = malloc(strlen(name)+1);
        if(P->name == NULL)
        {
            free(P->shape);
            free(P);
            STARSH_ERROR("malloc() failed");
            return STARSH_MALLOC_ERROR;
        }
        strcpy(P->name, name);
    }
    *problem = P;
    return STARSH_SUCCESS;
}

void starsh_problem_free(STARSH_problem *problem)
{
    if(problem == NULL)
        return;
    free(problem->shape);
    free(problem->name);
    free(problem);
}

void starsh_problem_info(STARSH_problem *problem)
{
    STARSH_problem *P = problem;
    if(P == NULL)
        return;
    printf("<STARSH_problem at %p, name \"%s\", shape (%zd", (void *)P,
            P->name ? P->name : "", (ssize_t)P->shape[0]);
    for(int i = 1; i < P->ndim; i++)
        printf(", %zd", (ssize_t)P->shape[i]);
    printf("), '%c' symmetric, '%c' precision>\n", P->symm, P->dtype);
}

int starsh_problem_get_block(STARSH_problem *problem, int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, Array **A)
{
    if(problem == NULL || irow == NULL || icol == NULL || A == NULL ||
            nrows < 0 || ncols < 0)
    {
        STARSH_ERROR("Invalid argument");
        return STARSH_WRONG_PARAMETER;
    }
    int ndim = problem->ndim;
    int *shape;
    STARSH_MALLOC(shape, ndim);
    shape[0] = nrows;
    shape[ndim-1] = ncols;
    for(int i = 1; i < ndim-1; i++)
        shape[i] = problem->shape[i];
    int info = array_new(A, ndim, shape, problem->dtype, 'F');
    free(shape);
    if(info != STARSH_SUCCESS)
        return info;
    problem->kernel(nrows, ncols, irow, icol, problem->row_data,
            problem->col_data, (*A)->data, nrows);
    return STARSH_SUCCESS;
}

/* Dense assembly of the whole operator (problem.c:279-316); meant for small
 * problems and checks. */
int starsh_problem_to_array(STARSH_problem *problem, Array **A)
{
    if(problem == NULL || A == NULL)
    {
        STARSH_ERROR("Invalid argument");
        return STARSH_WRONG_PARAMETER;
    }
    if(problem->ndim != 2)
    {
        STARSH_ERROR("Only scalar kernels are supported");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_int nrows = problem->shape[0], ncols = problem->shape[1];
    STARSH_int *irow, *icol;
    STARSH_MALLOC(irow, nrows);
    icol = malloc(sizeof(*icol)*(ncols+1));
    if(icol == NULL)
    {
        free(irow);
        return STARSH_MALLOC_ERROR;
    }
    for(STARSH_int i = 0; i < nrows; i++)
        irow[i] = i;
    for(STARSH_int i = 0; i < ncols; i++)
        icol[i] = i;
    int info = starsh_problem_get_block(problem, nrows, ncols, irow, icol, A);
    free(irow);
    free(icol);
    return info;
}
