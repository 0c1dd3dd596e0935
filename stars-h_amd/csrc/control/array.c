/* Minimal N-d array wrapper used for U, V and dense tiles.
 * Same fields and byte accounting as src/control/array.c:16-160 of
 * ecrc/stars-h 0.3.1 (array_from_buffer, array_new, array_free). */
#include "common.h"

static size_t dtype_bytes(char dtype)
{
    switch(dtype)
    {
        case 's': return 4;
        case 'd': return 8;
        case 'c': return 8;
        case 'z': return 16;
        default: return 0;
    }
}

int array_from_buffer(Array **A, int ndim, int *shape, char dtype, char order,
        void *buffer)
{
    if(A == NULL)
    {
        STARSH_ERROR("Invalid value of `A`");
        return STARSH_WRONG_PARAMETER;
    }
    if(ndim < 0)
    {
        STARSH_ERROR("Invalid value of `ndim`");
        return STARSH_WRONG_PARAMETER;
    }
    if(order != 'F' && order != 'C')
    {
        STARSH_ERROR("Invalid value of `order`");
        return STARSH_WRONG_PARAMETER;
    }
    size_t esize = dtype_bytes(dtype);
    if(esize == 0)
    {
        STARSH_ERROR("Invalid value of `dtype`");
        return STARSH_WRONG_PARAMETER;
    }
    Array *a;
    STARSH_MALLOC(a, 1);
    a->ndim = ndim;
    a->dtype = dtype;
    a->dtype_size = esize;
    a->data = buffer;
    if(ndim == 0)
    {
        a->shape = NULL;
        a->stride = NULL;
        a->order = 'F';
        a->size = 0;
        a->nbytes = 0;
        a->data_nbytes = sizeof(*a);
        *A = a;
        return STARSH_SUCCESS;
    }
    int *sh;
    ssize_t *st;
    sh = malloc(sizeof(*sh)*ndim);
    st = malloc(sizeof(*st)*ndim);
    if(sh == NULL || st == NULL)
    {
        free(sh);
        free(st);
        free(a);
        STARSH_ERROR("malloc() failed");
        return STARSH_MALLOC_ERROR;
    }
    memcpy(sh, shape, sizeof(*sh)*ndim);
    if(order == 'F')
    {
        st[0] = 1;
        for(int i = 1; i < ndim; i++)
            st[i] = st[i-1]*sh[i-1];
        a->size = (size_t)st[ndim-1]*sh[ndim-1];
    }
    else
    {
        st[ndim-1] = 1;
        for(int i = ndim-1; i > 0; i--)
            st[i-1] = st[i]*sh[i];
        a->size = (size_t)st[0]*sh[0];
    }
    a->shape = sh;
    a->stride = st;
    a->order = order;
    a->data_nbytes = a->size*esize;
    a->nbytes = a->data_nbytes + ndim*(sizeof(*sh)+sizeof(*st)) + sizeof(*a);
    *A = a;
    return STARSH_SUCCESS;
}

int array_new(Array **A, int ndim, int *shape, char dtype, char order)
{
    if(A == NULL)
    {
        STARSH_ERROR("Invalid value of `A`");
        return STARSH_WRONG_PARAMETER;
    }
    int info = array_from_buffer(A, ndim, shape, dtype, order, NULL);
    if(info != STARSH_SUCCESS)
        return info;
    (*A)->data = malloc((*A)->data_nbytes + ((*A)->data_nbytes == 0));
    if((*A)->data == NULL)
    {
        array_free(*A);
        *A = NULL;
        STARSH_ERROR("malloc() failed");
        return STARSH_MALLOC_ERROR;
    }
    return STARSH_SUCCESS;
}

/* Frees the data buffer and the wrapper (array.c:248-260).  Wrappers over
 * a shared buffer have their data pointer cleared first, as
 * starsh_blrm_free does. */
void array_free(Array *A)
{
    if(A == NULL)
        return;
    free(A->data);
    free(A->shape);
    free(A->stride);
    free(A);
}

void array_info(Array *A)
{
    if(A == NULL)
        return;
    printf("<Array at %p of shape (", (void *)A);
    for(int i = 0; i < A->ndim; i++)
        printf(i ? ",%d" : "%d", A->shape[i]);
    printf("), '%c' order, %zu elements, '%c' dtype, %zu bytes per element, "
            "%zu total bytes, data at %p>\n", A->order, A->size, A->dtype,
            A->dtype_size, A->nbytes, A->data);
}
