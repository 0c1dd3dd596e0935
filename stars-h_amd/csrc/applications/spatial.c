/* Spatial statistics data and kernel selection.
 * Follows src/applications/spatial.c:20-893 of ecrc/stars-h 0.3.1 for the
 * exponential, square exponential and Matern families. */
#include "common.h"
#include "starsh-spatial.h"

int starsh_ssdata_new(STARSH_ssdata **data, STARSH_int count, int ndim)
{
    if(data == NULL || ndim <= 0 || count < 0)
    {
        STARSH_ERROR("Invalid argument");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_ssdata *d;
    STARSH_MALLOC(d, 1);
    memset(d, 0, sizeof(*d));
    d->particles.count = count;
    d->particles.ndim = ndim;
    d->particles.point = malloc(sizeof(double)*(size_t)count*ndim + 1);
    if(d->particles.point == NULL)
    {
        free(d);
        return STARSH_MALLOC_ERROR;
    }
    *data = d;
    return STARSH_SUCCESS;
}

int starsh_ssdata_init(STARSH_ssdata **data, STARSH_int count, int ndim,
        double *point, double beta, double nu, double noise, double sigma)
{
    if(data == NULL || ndim <= 0 || count < 0 || point == NULL)
    {
        STARSH_ERROR("Invalid argument");
        return STARSH_WRONG_PARAMETER;
    }
    if(beta <= 0 || nu < 0 || noise < 0 || sigma < 0)
    {
        STARSH_ERROR("Invalid value of `beta`, `nu`, `noise` or `sigma`");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_ssdata *d;
    STARSH_MALLOC(d, 1);
    memset(d, 0, sizeof(*d));
    d->particles.count = count;
    d->particles.ndim = ndim;
    d->particles.point = point;
    d->beta = beta;
    d->nu = nu;
    d->noise = noise;
    d->sigma = sigma;
    *data = d;
    return STARSH_SUCCESS;
}

int starsh_ssdata_generate(STARSH_ssdata **data, STARSH_int count, int ndim,
        double beta, double nu, double noise,
        enum STARSH_PARTICLES_PLACEMENT place, double sigma)
{
    if(data == NULL)
    {
        STARSH_ERROR("Invalid value of `data`");
        return STARSH_WRONG_PARAMETER;
    }
    if(beta <= 0 || nu < 0 || noise < 0 || sigma < 0)
    {
        STARSH_ERROR("Invalid value of `beta`, `nu`, `noise` or `sigma`");
        return STARSH_WRONG_PARAMETER;
    }
    STARSH_particles *p;
    int info = starsh_particles_generate(&p, count, ndim, place);
    if(info != STARSH_SUCCESS)
        return info;
    info = starsh_ssdata_init(data, p->count, p->ndim, p->point, beta, nu,
            noise, sigma);
    if(info != STARSH_SUCCESS)
    {
        starsh_particles_free(p);
        return info;
    }
    free(p);
    return STARSH_SUCCESS;
}

int starsh_ssdata_generate_va(STARSH_ssdata **data, STARSH_int count,
        va_list args)
{
    int ndim = 2, arg;
    double beta = 0.1, nu = 0.5, noise = 0, sigma = 1.0;
    enum STARSH_PARTICLES_PLACEMENT place = STARSH_PARTICLES_UNIFORM;
    while((arg = va_arg(args, int)) != 0)
    {
        switch(arg)
        {
            case STARSH_SPATIAL_NDIM: ndim = va_arg(args, int); break;
            case STARSH_SPATIAL_BETA: beta = va_arg(args, double); break;
            case STARSH_SPATIAL_NU: nu = va_arg(args, double); break;
            case STARSH_SPATIAL_NOISE: noise = va_arg(args, double); break;
            case STARSH_SPATIAL_PLACE:
                place = (enum STARSH_PARTICLES_PLACEMENT)va_arg(args, int);
                break;
            case STARSH_SPATIAL_SIGMA: sigma = va_arg(args, double); break;
            default:
                STARSH_ERROR("Wrong parameter type");
                return STARSH_WRONG_PARAMETER;
        }
    }
    return starsh_ssdata_generate(data, count, ndim, beta, nu, noise, place,
            sigma);
}

int starsh_ssdata_generate_el(STARSH_ssdata **data, STARSH_int count, ...)
{
    va_list args;
    va_start(args, count);
    int info = starsh_ssdata_generate_va(data, count, args);
    va_end(args);
    return info;
}

void starsh_ssdata_free(STARSH_ssdata *data)
{
    if(data == NULL)
        return;
    free(data->particles.point);
    free(data);
}

#define PICK(fam, nd) \
    do { switch(type) { \
        case STARSH_SPATIAL_EXP: *kernel = starsh_ssdata_block_exp_kernel_##nd; break; \
        case STARSH_SPATIAL_EXP_SIMD: *kernel = starsh_ssdata_block_exp_kernel_##nd##_simd; break; \
        case STARSH_SPATIAL_SQREXP: *kernel = starsh_ssdata_block_sqrexp_kernel_##nd; break; \
        case STARSH_SPATIAL_SQREXP_SIMD: *kernel = starsh_ssdata_block_sqrexp_kernel_##nd##_simd; break; \
        case STARSH_SPATIAL_MATERN: *kernel = starsh_ssdata_block_matern_kernel_##nd; break; \
        case STARSH_SPATIAL_MATERN_SIMD: *kernel = starsh_ssdata_block_matern_kernel_##nd##_simd; break; \
        case STARSH_SPATIAL_MATERN2: *kernel = starsh_ssdata_block_matern2_kernel_##nd; break; \
        case STARSH_SPATIAL_MATERN2_SIMD: *kernel = starsh_ssdata_block_matern2_kernel_##nd##_simd; break; \
        default: STARSH_ERROR("Wrong type of kernel"); return STARSH_WRONG_PARAMETER; \
    } } while(0)

int starsh_ssdata_get_kernel(STARSH_kernel **kernel, STARSH_ssdata *data,
        enum STARSH_SPATIAL_KERNEL type)
{
    if(kernel == NULL || data == NULL)
        return STARSH_WRONG_PARAMETER;
    if(type == STARSH_SPATIAL_MATERN || type == STARSH_SPATIAL_MATERN_SIMD ||
            type == STARSH_SPATIAL_MATERN2 ||
            type == STARSH_SPATIAL_MATERN2_SIMD)
    {
        /* nu = 0 makes 2^(1-nu)/Gamma(nu) vanish against an infinite
         * K_0 x^0 at every distance; the reference fails there too */
        if(!(data->nu > 0))
        {
            STARSH_ERROR("Matern kernels need nu > 0");
            return STARSH_WRONG_PARAMETER;
        }
    }
    switch(data->particles.ndim)
    {
        case 1: PICK(fam, 1d); break;
        case 2: PICK(fam, 2d); break;
        case 3: PICK(fam, 3d); break;
        case 4: PICK(fam, 4d); break;
        default: PICK(fam, nd); break;
    }
    return STARSH_SUCCESS;
}
