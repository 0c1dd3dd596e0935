/* Point sets: generators, Morton sort and file I/O.
 * Follows src/applications/particles.c of ecrc/stars-h 0.3.1.  Generators
 * draw from rand() in the same order as the reference, so srand(seed)
 * reproduces the reference's points on the same libc. */
#include <stdint.h>
#include "common.h"
#include "starsh-particles.h"

int starsh_particles_new(STARSH_particles **data, STARSH_int count, int ndim)
{
    if(data == NULL || ndim <= 0 || count < 0)
    {
        STARSH_ERROR("Invalid argument");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_particles *p;
    STARSH_MALLOC(p, 1);
    p->point = malloc(sizeof(double)*(size_t)count*ndim + 1);
    if(p->point == NULL)
    {
        free(p);
        return STARSH_MALLOC_ERROR;
    }
    p->count = count;
    p->ndim = ndim;
    *data = p;
    return STARSH_SUCCESS;
}

int starsh_particles_init(STARSH_particles **data, STARSH_int count, int ndim,
        double *point)
{
    if(data == NULL || ndim <= 0 || count < 0 || point == NULL)
    {
        STARSH_ERROR("Invalid argument");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_particles *p;
    STARSH_MALLOC(p, 1);
    p->count = count;
    p->ndim = ndim;
    p->point = point;
    *data = p;
    return STARSH_SUCCESS;
}

void starsh_particles_free(STARSH_particles *data)
{
    if(data == NULL)
        return;
    free(data->point);
    free(data);
}

static int wrap(STARSH_particles **data, STARSH_int count, int ndim,
        double *point)
{
    STARSH_particles *p = malloc(sizeof(*p));
    if(p == NULL)
    {
        free(point);
        return STARSH_MALLOC_ERROR;
    }
    p->count = count;
    p->ndim = ndim;
    p->point = point;
    int info = starsh_particles_zsort_inplace(p);
    if(info != STARSH_SUCCESS)
    {
        starsh_particles_free(p);
        return info;
    }
    *data = p;
    return STARSH_SUCCESS;
}

/* Smallest side with side^ndim >= count, computed as in particles.c:357-365. */
static STARSH_int grid_side(STARSH_int count, int ndim)
{
    STARSH_int side = (STARSH_int)ceil(pow((double)count, 1.0/ndim));
    STARSH_int total = 1;
    for(int d = 0; d < ndim; d++)
        total *= side;
    return total < count ? -1 : side;
}

/* Walks the grid with the last dimension fastest (particles.c:390-401) and
 * fills point[i + d*count] = coord[d*side + index_d] (+ jitter). */
static void fill_grid(double *point, STARSH_int count, int ndim,
        STARSH_int side, const double *coord, int jitter)
{
    STARSH_int idx[64] = {0};
    double mult = 0.8/side/RAND_MAX, add = 0.1/side;
    for(STARSH_int i = 0; i < count; i++)
    {
        for(int d = 0; d < ndim; d++)
            point[i+d*count] = jitter ? coord[idx[d]+d*side]+add+mult*rand()
                : coord[idx[d]+d*side];
        int d = ndim-1;
        idx[d]++;
        while(idx[d] == side)
        {
            idx[d] = 0;
            if(d > 0)
            {
                d--;
                idx[d]++;
            }
        }
    }
}

int starsh_particles_generate_rand(STARSH_particles **data, STARSH_int count,
        int ndim)
{
    if(data == NULL || count <= 0 || ndim <= 0)
        return STARSH_WRONG_PARAMETER;
    double *point;
    STARSH_MALLOC(point, (size_t)count*ndim);
    for(size_t i = 0; i < (size_t)count*ndim; i++)
        point[i] = (double)rand()/RAND_MAX;
    return wrap(data, count, ndim, point);
}

int starsh_particles_generate_randgrid(STARSH_particles **data,
        STARSH_int count, int ndim)
{
    if(data == NULL || count <= 0 || ndim <= 0 || ndim > 64)
        return STARSH_WRONG_PARAMETER;
    STARSH_int side = grid_side(count, ndim);
    if(side < 0)
        return STARSH_UNKNOWN_ERROR;
    double *point, *coord;
    STARSH_MALLOC(point, (size_t)count*ndim);
    coord = malloc(sizeof(double)*ndim*side);
    if(coord == NULL)
    {
        free(point);
        return STARSH_MALLOC_ERROR;
    }
    for(STARSH_int i = 0; i < ndim*side; i++)
        coord[i] = (double)rand()/RAND_MAX;
    fill_grid(point, count, ndim, side, coord, 0);
    free(coord);
    return wrap(data, count, ndim, point);
}

static int regular_grid(STARSH_particles **data, STARSH_int count, int ndim,
        int jitter_points, int jitter_coords)
{
    if(data == NULL || count <= 0 || ndim <= 0 || ndim > 64)
        return STARSH_WRONG_PARAMETER;
    STARSH_int side = grid_side(count, ndim);
    if(side < 0)
        return STARSH_UNKNOWN_ERROR;
    double *point, *coord;
    STARSH_MALLOC(point, (size_t)count*ndim);
    coord = malloc(sizeof(double)*ndim*side);
    if(coord == NULL)
    {
        free(point);
        return STARSH_MALLOC_ERROR;
    }
    double mult = 0.8/side/RAND_MAX, add = 0.1/side;
    for(STARSH_int i = 0; i < side; i++)
    {
        double val = (double)i/side;
        for(int d = 0; d < ndim; d++)
            coord[d*side+i] = jitter_coords ? val+add+mult*rand() : val;
    }
    fill_grid(point, count, ndim, side, coord, jitter_points);
    free(coord);
    return wrap(data, count, ndim, point);
}

/* particles.c:269-339: regular grid i/side. */
int starsh_particles_generate_uniform(STARSH_particles **data,
        STARSH_int count, int ndim)
{
    return regular_grid(data, count, ndim, 0, 0);
}

/* particles.c:341-411: grid cell corner + 0.1/side + 0.8/side*U[0,1]. */
int starsh_particles_generate_quasiuniform1(STARSH_particles **data,
        STARSH_int count, int ndim)
{
    return regular_grid(data, count, ndim, 1, 0);
}

/* particles.c:413-480: each grid coordinate jittered once. */
int starsh_particles_generate_quasiuniform2(STARSH_particles **data,
        STARSH_int count, int ndim)
{
    return regular_grid(data, count, ndim, 0, 1);
}

int starsh_particles_generate(STARSH_particles **data, STARSH_int count,
        int ndim, enum STARSH_PARTICLES_PLACEMENT ptype)
{
    switch(ptype)
    {
        case STARSH_PARTICLES_RAND:
            return starsh_particles_generate_rand(data, count, ndim);
        case STARSH_PARTICLES_UNIFORM:
            return starsh_particles_generate_uniform(data, count, ndim);
        case STARSH_PARTICLES_RANDGRID:
            return starsh_particles_generate_randgrid(data, count, ndim);
        case STARSH_PARTICLES_QUASIUNIFORM1:
            return starsh_particles_generate_quasiuniform1(data, count, ndim);
        case STARSH_PARTICLES_QUASIUNIFORM2:
            return starsh_particles_generate_quasiuniform2(data, count, ndim);
        default:
            STARSH_ERROR("Placement %d is not supported", (int)ptype);
            return STARSH_WRONG_PARAMETER;
    }
}

/* Standard normal variate by the polar Box-Muller method on rand(). */
static double gauss_rand(void)
{
    double u, v, s;
    do
    {
        u = 2.0*rand()/RAND_MAX-1.0;
        v = 2.0*rand()/RAND_MAX-1.0;
        s = u*u+v*v;
    } while(s >= 1.0 || s == 0.0);
    return u*sqrt(-2.0*log(s)/s);
}

int starsh_particles_generate_spheres(STARSH_particles **data,
        STARSH_int nspheres, STARSH_int npoints, double radius)
{
    if(data == NULL || nspheres <= 0 || npoints <= 0 || radius < 0)
        return STARSH_WRONG_PARAMETER;
    STARSH_int count = nspheres*npoints;
    double *point;
    STARSH_MALLOC(point, (size_t)count*3);
    for(STARSH_int s = 0; s < nspheres; s++)
    {
        double c[3];
        for(int d = 0; d < 3; d++)
            c[d] = (double)rand()/RAND_MAX;
        for(STARSH_int q = 0; q < npoints; q++)
        {
            double g[3], nrm;
            do
            {
                for(int d = 0; d < 3; d++)
                    g[d] = gauss_rand();
                nrm = sqrt(g[0]*g[0]+g[1]*g[1]+g[2]*g[2]);
            } while(nrm == 0.0);
            for(int d = 0; d < 3; d++)
                point[s*npoints+q+d*count] = c[d]+radius*g[d]/nrm;
        }
    }
    return wrap(data, count, 3, point);
}

/* ---- Morton order (particles.c:1090-1221).  Coordinates are scaled per
 * dimension to [0, UINT32_MAX] and truncated; a most-significant-bit-first
 * binary radix sort then visits dimensions ndim-1 .. 0 at every bit.  Items
 * with the bit clear keep their order at the front, items with the bit set
 * are placed from the back (their order reverses), exactly as the
 * reference's partition, so equal keys land in the same positions. */

typedef struct
{
    const uint32_t *key;
    STARSH_int count;
    int ndim;
    STARSH_int *order;
    STARSH_int *tmp;
} zsort_ctx;

static void radix_partition(zsort_ctx *z, int sdim, int sbit, STARSH_int lo,
        STARSH_int hi)
{
    /* Explicit stack instead of recursion: depth can reach 32*ndim. */
    typedef struct { int sdim, sbit; STARSH_int lo, hi; } frame;
    frame stack[32*64*2+8];
    int top = 0;
    stack[top++] = (frame){sdim, sbit, lo, hi};
    while(top > 0)
    {
        frame f = stack[--top];
        const uint32_t *k = z->key+(size_t)f.sdim*z->count;
        uint32_t mask = (uint32_t)1 << f.sbit;
        STARSH_int a = f.lo, b = f.hi;
        for(STARSH_int i = f.lo; i <= f.hi; i++)
        {
            STARSH_int o = z->order[i];
            if((k[o] & mask) == 0)
                z->tmp[a++] = o;
            else
                z->tmp[b--] = o;
        }
        memcpy(z->order+f.lo, z->tmp+f.lo, sizeof(STARSH_int)*(f.hi-f.lo+1));
        int nd, nb;
        if(f.sdim > 0)
        {
            nd = f.sdim-1;
            nb = f.sbit;
        }
        else if(f.sbit > 0)
        {
            nd = z->ndim-1;
            nb = f.sbit-1;
        }
        else
            continue;
        /* Push the upper range first so the lower one is processed first;
         * the ranges are disjoint, so the order does not change results. */
        if(f.hi-b > 1)
            stack[top++] = (frame){nd, nb, b+1, f.hi};
        if(a-f.lo > 1)
            stack[top++] = (frame){nd, nb, f.lo, a-1};
    }
}

int starsh_particles_zsort_inplace(STARSH_particles *data)
{
    if(data == NULL)
        return STARSH_WRONG_PARAMETER;
    STARSH_int count = data->count;
    int ndim = data->ndim;
    if(ndim <= 0 || ndim > 64)
        return STARSH_WRONG_PARAMETER;
    if(count <= 1)
        return STARSH_SUCCESS;
    double *point = data->point;
    uint32_t *key;
    STARSH_MALLOC(key, (size_t)count*ndim);
    for(int d = 0; d < ndim; d++)
    {
        const double *x = point+(size_t)d*count;
        double mn = x[0], mx = x[0];
        for(STARSH_int j = 1; j < count; j++)
        {
            if(mn > x[j])
                mn = x[j];
            else if(mx < x[j])
                mx = x[j];
        }
        double range = mx-mn;
        for(STARSH_int j = 0; j < count; j++)
            key[(size_t)d*count+j] = (uint32_t)((x[j]-mn)/range*UINT32_MAX);
    }
    zsort_ctx z = {key, count, ndim, NULL, NULL};
    z.order = malloc(sizeof(STARSH_int)*count);
    z.tmp = malloc(sizeof(STARSH_int)*count);
    double *np = malloc(sizeof(double)*(size_t)count*ndim);
    if(z.order == NULL || z.tmp == NULL || np == NULL)
    {
        free(z.order);
        free(z.tmp);
        free(np);
        free(key);
        return STARSH_MALLOC_ERROR;
    }
    for(STARSH_int j = 0; j < count; j++)
        z.order[j] = j;
    radix_partition(&z, ndim-1, 31, 0, count-1);
    for(STARSH_int j = 0; j < count; j++)
        for(int d = 0; d < ndim; d++)
            np[(size_t)d*count+j] = point[(size_t)d*count+z.order[j]];
    data->point = np;
    free(point);
    free(key);
    free(z.order);
    free(z.tmp);
    return STARSH_SUCCESS;
}

/* ---- File I/O (particles.c:789-1088).  ASCII: "count=%zu, ndim=%d" then one
 * line per particle; binary: [STARSH_int count][int ndim][count*ndim
 * doubles, SoA]. */

int starsh_particles_read_from_file_pointer_ascii(STARSH_particles **data,
        FILE *fp)
{
    size_t count;
    int ndim;
    if(fscanf(fp, "count=%zu, ndim=%d", &count, &ndim) != 2 || ndim <= 0)
        return STARSH_FILE_WRONG_INPUT;
    double *point;
    STARSH_MALLOC(point, count*ndim);
    for(size_t i = 0; i < count; i++)
    {
        if(fscanf(fp, "%lf", point+i) != 1)
        {
            free(point);
            return STARSH_FILE_WRONG_INPUT;
        }
        for(int d = 1; d < ndim; d++)
            if(fscanf(fp, ",%lf", point+i+d*count) != 1)
            {
                free(point);
                return STARSH_FILE_WRONG_INPUT;
            }
    }
    int info = starsh_particles_init(data, count, ndim, point);
    if(info != STARSH_SUCCESS)
        free(point);
    return info;
}

int starsh_particles_read_from_file_pointer_binary(STARSH_particles **data,
        FILE *fp)
{
    STARSH_int count;
    int ndim;
    if(fread(&count, sizeof(count), 1, fp) != 1 || count < 0)
        return STARSH_FILE_WRONG_INPUT;
    if(fread(&ndim, sizeof(ndim), 1, fp) != 1 || ndim <= 0)
        return STARSH_FILE_WRONG_INPUT;
    double *point;
    STARSH_MALLOC(point, (size_t)count*ndim);
    if(fread(point, sizeof(double), (size_t)count*ndim, fp) !=
            (size_t)count*ndim)
    {
        free(point);
        return STARSH_FILE_WRONG_INPUT;
    }
    int info = starsh_particles_init(data, count, ndim, point);
    if(info != STARSH_SUCCESS)
        free(point);
    return info;
}

int starsh_particles_read_from_file_pointer(STARSH_particles **data,
        FILE *fp, const enum STARSH_FILE_TYPE ftype)
{
    if(data == NULL || fp == NULL)
        return STARSH_WRONG_PARAMETER;
    if(ftype == STARSH_ASCII)
        return starsh_particles_read_from_file_pointer_ascii(data, fp);
    if(ftype == STARSH_BINARY)
        return starsh_particles_read_from_file_pointer_binary(data, fp);
    return STARSH_WRONG_PARAMETER;
}

int starsh_particles_read_from_file(STARSH_particles **data, const char *fname,
        const enum STARSH_FILE_TYPE ftype)
{
    if(fname == NULL)
        return STARSH_WRONG_PARAMETER;
    FILE *fp = fopen(fname, ftype == STARSH_BINARY ? "rb" : "r");
    if(fp == NULL)
    {
        STARSH_ERROR("Cannot open file %s", fname);
        return STARSH_FILE_NOT_EXIST;
    }
    int info = starsh_particles_read_from_file_pointer(data, fp, ftype);
    fclose(fp);
    return info;
}

int starsh_particles_write_to_file_pointer_ascii(const STARSH_particles *data,
        FILE *fp)
{
    size_t count = data->count;
    int ndim = data->ndim;
    if(fprintf(fp, "count=%zu, ndim=%d\n", count, ndim) < 0)
        return STARSH_FPRINTF_ERROR;
    for(size_t i = 0; i < count; i++)
    {
        if(fprintf(fp, "%lf", data->point[i]) < 0)
            return STARSH_FPRINTF_ERROR;
        for(int d = 1; d < ndim; d++)
            if(fprintf(fp, ", %lf", data->point[i+d*count]) < 0)
                return STARSH_FPRINTF_ERROR;
        if(fprintf(fp, "\n") < 0)
            return STARSH_FPRINTF_ERROR;
    }
    return STARSH_SUCCESS;
}

int starsh_particles_write_to_file_pointer_binary(const STARSH_particles *data,
        FILE *fp)
{
    STARSH_int count = data->count;
    int ndim = data->ndim;
    if(fwrite(&count, sizeof(count), 1, fp) != 1)
        return STARSH_FWRITE_ERROR;
    if(fwrite(&ndim, sizeof(ndim), 1, fp) != 1)
        return STARSH_FWRITE_ERROR;
    if(fwrite(data->point, sizeof(double), (size_t)count*ndim, fp) !=
            (size_t)count*ndim)
        return STARSH_FWRITE_ERROR;
    return STARSH_SUCCESS;
}

int starsh_particles_write_to_file_pointer(const STARSH_particles *data,
        FILE *fp, const enum STARSH_FILE_TYPE ftype)
{
    if(data == NULL || fp == NULL)
        return STARSH_WRONG_PARAMETER;
    if(ftype == STARSH_ASCII)
        return starsh_particles_write_to_file_pointer_ascii(data, fp);
    if(ftype == STARSH_BINARY)
        return starsh_particles_write_to_file_pointer_binary(data, fp);
    return STARSH_WRONG_PARAMETER;
}

int starsh_particles_write_to_file(const STARSH_particles *data,
        const char *fname, const enum STARSH_FILE_TYPE ftype)
{
    if(fname == NULL)
        return STARSH_WRONG_PARAMETER;
    FILE *fp = fopen(fname, ftype == STARSH_BINARY ? "wb" : "w");
    if(fp == NULL)
    {
        STARSH_ERROR("Cannot open file %s", fname);
        return STARSH_FILE_NOT_EXIST;
    }
    int info = starsh_particles_write_to_file_pointer(data, fp, ftype);
    fclose(fp);
    return info;
}

/* starsh_particles_zsort_inplace on the GPU: the same permutation (the
 * reference's partition order, particles.c:1090-1221), computed by
 * starsh_hip_zsort on the device.  data->point is replaced by a new
 * malloc'd array in Morton order, as in the reference (:1150-1160). */
int starsh_particles_zsort_inplace_hip(STARSH_particles *data)
{
    if(data == NULL || data->ndim < 1 || data->count < 0)
        return STARSH_WRONG_PARAMETER;
    const STARSH_int n = data->count;
    const int ndim = data->ndim;
    if(n == 0)
        return STARSH_SUCCESS;
    const size_t bytes = sizeof(double)*(size_t)n*ndim;
    double *d_x = NULL, *d_y = NULL;
    double *y = malloc(bytes);
    if(y == NULL)
        return STARSH_MALLOC_ERROR;
    int info = starsh_hip_malloc((void **)&d_x, bytes);
    if(info == STARSH_SUCCESS)
        info = starsh_hip_malloc((void **)&d_y, bytes);
    if(info == STARSH_SUCCESS)
        info = starsh_hip_memcpy_h2d(d_x, data->point, bytes);
    if(info == STARSH_SUCCESS)
        info = starsh_hip_zsort(d_x, n, ndim, d_y, NULL);
    if(info == STARSH_SUCCESS)
        info = starsh_hip_memcpy_d2h(y, d_y, bytes);
    starsh_hip_free(d_x);
    starsh_hip_free(d_y);
    if(info != STARSH_SUCCESS)
    {
        free(y);
        return info;
    }
    free(data->point);
    data->point = y;
    return STARSH_SUCCESS;
}
