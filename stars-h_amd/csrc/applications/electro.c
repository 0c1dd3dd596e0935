/* Electrostatics (Coulomb potential) and electrodynamics (sin/cos Helmholtz)
 * data and kernel selection.  Follow src/applications/electrostatics.c:17-359
 * and src/applications/electrodynamics.c:17-417 of ecrc/stars-h 0.3.1.  The
 * reference leaves *data unset in the _new and _init constructors; here
 * they set it. */
#include "common.h"
#include "starsh-electrostatics.h"
#include "starsh-electrodynamics.h"

int starsh_esdata_new(STARSH_esdata **data, STARSH_int count, int ndim)
{
    return starsh_particles_new(data, count, ndim);
}

int starsh_esdata_init(STARSH_esdata **data, STARSH_int count, int ndim,
        double *point)
{
    return starsh_particles_init(data, count, ndim, point);
}

int starsh_esdata_generate(STARSH_esdata **data, STARSH_int count, int ndim,
        enum STARSH_PARTICLES_PLACEMENT place)
{
    if(data == NULL)
        return STARSH_WRONG_PARAMETER;
    return starsh_particles_generate(data, count, ndim, place);
}

int starsh_esdata_generate_va(STARSH_esdata **data, STARSH_int count,
        va_list args)
{
    int ndim = 2, arg;
    enum STARSH_PARTICLES_PLACEMENT place = STARSH_PARTICLES_UNIFORM;
    while((arg = va_arg(args, int)) != 0)
    {
        switch(arg)
        {
            case STARSH_ELECTROSTATICS_NDIM: ndim = va_arg(args, int); break;
            case STARSH_ELECTROSTATICS_PLACE:
                place = (enum STARSH_PARTICLES_PLACEMENT)va_arg(args, int);
                break;
            default:
                STARSH_ERROR("Wrong parameter type");
                return STARSH_WRONG_PARAMETER;
        }
    }
    return starsh_esdata_generate(data, count, ndim, place);
}

int starsh_esdata_generate_el(STARSH_esdata **data, STARSH_int count, ...)
{
    va_list args;
    va_start(args, count);
    int info = starsh_esdata_generate_va(data, count, args);
    va_end(args);
    return info;
}

void starsh_esdata_free(STARSH_esdata *data)
{
    starsh_particles_free(data);
}

int starsh_esdata_get_kernel(STARSH_kernel **kernel, STARSH_esdata *data,
        enum STARSH_ELECTROSTATICS_KERNEL type)
{
    if(kernel == NULL || data == NULL)
        return STARSH_WRONG_PARAMETER;
    int simd = type == STARSH_ELECTROSTATICS_COULOMB_POTENTIAL_SIMD;
    if(type != STARSH_ELECTROSTATICS_COULOMB_POTENTIAL && !simd)
    {
        STARSH_ERROR("Wrong type of kernel");
        return STARSH_WRONG_PARAMETER;
    }
    switch(data->ndim)
    {
        case 1: *kernel = simd ? starsh_esdata_block_coulomb_potential_kernel_1d_simd
                : starsh_esdata_block_coulomb_potential_kernel_1d; break;
        case 2: *kernel = simd ? starsh_esdata_block_coulomb_potential_kernel_2d_simd
                : starsh_esdata_block_coulomb_potential_kernel_2d; break;
        case 3: *kernel = simd ? starsh_esdata_block_coulomb_potential_kernel_3d_simd
                : starsh_esdata_block_coulomb_potential_kernel_3d; break;
        case 4: *kernel = simd ? starsh_esdata_block_coulomb_potential_kernel_4d_simd
                : starsh_esdata_block_coulomb_potential_kernel_4d; break;
        default: *kernel = simd ? starsh_esdata_block_coulomb_potential_kernel_nd_simd
                : starsh_esdata_block_coulomb_potential_kernel_nd; break;
    }
    return STARSH_SUCCESS;
}

int starsh_eddata_new(STARSH_eddata **data, STARSH_int count, int ndim)
{
    if(data == NULL || ndim <= 0 || count < 0)
        return STARSH_WRONG_PARAMETER;
    STARSH_eddata *d;
    STARSH_MALLOC(d, 1);
    d->particles.count = count;
    d->particles.ndim = ndim;
    d->particles.point = malloc(sizeof(double)*(size_t)count*ndim + 1);
    if(d->particles.point == NULL)
    {
        free(d);
        return STARSH_MALLOC_ERROR;
    }
    d->k = 1.0;
    d->diag = 1.0;
    *data = d;
    return STARSH_SUCCESS;
}

int starsh_eddata_init(STARSH_eddata **data, STARSH_int count, int ndim,
        double *point, double k, double diag)
{
    if(data == NULL || ndim <= 0 || count < 0 || point == NULL)
        return STARSH_WRONG_PARAMETER;
    STARSH_eddata *d;
    STARSH_MALLOC(d, 1);
    d->particles.count = count;
    d->particles.ndim = ndim;
    d->particles.point = point;
    d->k = k;
    d->diag = diag;
    *data = d;
    return STARSH_SUCCESS;
}

int starsh_eddata_generate(STARSH_eddata **data, STARSH_int count, int ndim,
        double k, double diag, enum STARSH_PARTICLES_PLACEMENT place)
{
    if(data == NULL)
        return STARSH_WRONG_PARAMETER;
    STARSH_particles *p;
    int info = starsh_particles_generate(&p, count, ndim, place);
    if(info != STARSH_SUCCESS)
        return info;
    info = starsh_eddata_init(data, p->count, p->ndim, p->point, k, diag);
    if(info != STARSH_SUCCESS)
    {
        starsh_particles_free(p);
        return info;
    }
    free(p);
    return STARSH_SUCCESS;
}

int starsh_eddata_generate_va(STARSH_eddata **data, STARSH_int count,
        va_list args)
{
    int ndim = 2, arg;
    double k = 1.0, diag = 1.0;
    enum STARSH_PARTICLES_PLACEMENT place = STARSH_PARTICLES_UNIFORM;
    while((arg = va_arg(args, int)) != 0)
    {
        switch(arg)
        {
            case STARSH_ELECTRODYNAMICS_NDIM: ndim = va_arg(args, int); break;
            case STARSH_ELECTRODYNAMICS_K: k = va_arg(args, double); break;
            case STARSH_ELECTRODYNAMICS_DIAG: diag = va_arg(args, double); break;
            case STARSH_ELECTRODYNAMICS_PLACE:
                place = (enum STARSH_PARTICLES_PLACEMENT)va_arg(args, int);
                break;
            default:
                STARSH_ERROR("Wrong parameter type");
                return STARSH_WRONG_PARAMETER;
        }
    }
    return starsh_eddata_generate(data, count, ndim, k, diag, place);
}

int starsh_eddata_generate_el(STARSH_eddata **data, STARSH_int count, ...)
{
    va_list args;
    va_start(args, count);
    int info = starsh_eddata_generate_va(data, count, args);
    va_end(args);
    return info;
}

void starsh_eddata_free(STARSH_eddata *data)
{
    if(data == NULL)
        return;
    free(data->particles.point);
    free(data);
}

#define ED_PICK(nd) \
    *kernel = cosine ? (simd ? starsh_eddata_block_cos_kernel_##nd##_simd : \
            starsh_eddata_block_cos_kernel_##nd) : \
        (simd ? starsh_eddata_block_sin_kernel_##nd##_simd : \
            starsh_eddata_block_sin_kernel_##nd)

int starsh_eddata_get_kernel(STARSH_kernel **kernel, STARSH_eddata *data,
        enum STARSH_ELECTRODYNAMICS_KERNEL type)
{
    if(kernel == NULL || data == NULL)
        return STARSH_WRONG_PARAMETER;
    int cosine, simd;
    switch(type)
    {
        case STARSH_ELECTRODYNAMICS_SIN: cosine = 0; simd = 0; break;
        case STARSH_ELECTRODYNAMICS_COS: cosine = 1; simd = 0; break;
        case STARSH_ELECTRODYNAMICS_SIN_SIMD: cosine = 0; simd = 1; break;
        case STARSH_ELECTRODYNAMICS_COS_SIMD: cosine = 1; simd = 1; break;
        default:
            STARSH_ERROR("Wrong type of kernel");
            return STARSH_WRONG_PARAMETER;
    }
    switch(data->particles.ndim)
    {
        case 1: ED_PICK(1d); break;
        case 2: ED_PICK(2d); break;
        case 3: ED_PICK(3d); break;
        case 4: ED_PICK(4d); break;
        default: ED_PICK(nd); break;
    }
    return STARSH_SUCCESS;
}
