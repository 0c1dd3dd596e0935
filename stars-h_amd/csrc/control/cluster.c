/* Plain (tiled) clusterization of points into contiguous blocks.
 * Follows src/control/cluster.c:16-146 of ecrc/stars-h 0.3.1: block b covers
 * [start[b], start[b]+size[b]) of the identity pivot, the last block holds
 * the remainder (ragged edge). */
#include "common.h"

int starsh_cluster_new(STARSH_cluster **cluster, void *data, STARSH_int ndata,
        STARSH_int *pivot, STARSH_int nblocks, STARSH_int nlevels,
        STARSH_int *level, STARSH_int *start, STARSH_int *size,
        STARSH_int *parent, STARSH_int *child_start, STARSH_int *child,
        enum STARSH_CLUSTER_TYPE type)
{
    if(cluster == NULL)
    {
        STARSH_ERROR("Invalid value of `cluster`");
        return STARSH_WRONG_PARAMETER;
    }
    if(pivot == NULL || start == NULL || size == NULL)
    {
        STARSH_ERROR("Invalid value of `pivot`, `start` or `size`");
        return STARSH_WRONG_PARAMETER;
    }
    if(type != STARSH_PLAIN)
    {
        STARSH_ERROR("Only plain clusterization is supported");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_cluster *C;
    STARSH_MALLOC(C, 1);
    C->data = data;
    C->ndata = ndata;
    C->pivot = pivot;
    C->nblocks = nblocks;
    C->nlevels = nlevels;
    C->level = level;
    C->start = start;
    C->size = size;
    C->parent = parent;
    C->child_start = child_start;
    C->child = child;
    C->type = type;
    *cluster = C;
    return STARSH_SUCCESS;
}

void starsh_cluster_free(STARSH_cluster *cluster)
{
    if(cluster == NULL)
        return;
    free(cluster->pivot);
    free(cluster->level);
    free(cluster->start);
    free(cluster->size);
    free(cluster->parent);
    free(cluster->child_start);
    free(cluster->child);
    free(cluster);
}

void starsh_cluster_info(STARSH_cluster *cluster)
{
    if(cluster == NULL)
        return;
    printf("<STARSH_cluster at %p, tiled, %zd blocks>\n", (void *)cluster,
            (ssize_t)cluster->nblocks);
}

int starsh_cluster_new_plain(STARSH_cluster **cluster, void *data,
        STARSH_int ndata, STARSH_int block_size)
{
    if(cluster == NULL)
    {
        STARSH_ERROR("Invalid value of `cluster`");
        return STARSH_WRONG_PARAMETER;
    }
    if(ndata <= 0 || block_size <= 0)
    {
        STARSH_ERROR("Invalid value of `ndata` or `block_size`");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_int nblocks = (ndata-1)/block_size+1;
    STARSH_int *start, *size, *pivot;
    STARSH_MALLOC(start, nblocks);
    STARSH_MALLOC(size, nblocks);
    STARSH_MALLOC(pivot, ndata);
    for(STARSH_int b = 0; b < nblocks; b++)
    {
        start[b] = b*block_size;
        size[b] = ndata-start[b] < block_size ? ndata-start[b] : block_size;
    }
    for(STARSH_int i = 0; i < ndata; i++)
        pivot[i] = i;
    return starsh_cluster_new(cluster, data, ndata, pivot, nblocks, 0, NULL,
            start, size, NULL, NULL, NULL, STARSH_PLAIN);
}
