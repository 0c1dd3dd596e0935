/* Host implementations of the STARSH_kernel tile generators.
 *
 * Values follow src/applications/spatial/kernel_{exp,sqrexp,matern,matern2}.c,
 * electrostatics/kernel_coulomb_potential.c and electrodynamics/kernel_{sin,cos}.c
 * of ecrc/stars-h 0.3.1 operation by operation: the squared distance is
 * accumulated over dimensions from 0.0, the exponent is formed with the same
 * divisions, and r == 0 takes the special value.  The Matern kernels use the
 * closed forms of 2^(1-nu)/Gamma(nu) x^nu K_nu(x) for nu = 1/2, 3/2, 5/2:
 * exp(-x), (1+x) exp(-x) and (1+x+x^2/3) exp(-x); any other nu > 0 takes
 * the reference's formula sigma 2^(1-nu)/Gamma(nu) x^nu K_nu(x)
 * (kernel_matern.c:115-117) with the K_nu of besselk.h.
 *
 * These host functions serve the API (starsh_problem_get_block, user code);
 * the HIP drivers never call them: they map the pointer to a device
 * generator with starsh_kernel_identify. */
#include "common.h"
#include "kernels_host.h"
#include "starsh-spatial.h"
#include "starsh-electrostatics.h"
#include "starsh-electrodynamics.h"
#include "besselk.h"

#pragma GCC optimize ("no-fast-math")

static double matern_closed(double nu, double x)
{
    if(nu == 0.5)
        return exp(-x);
    if(nu == 1.5)
        return (1.0+x)*exp(-x);
    return (1.0+x+x*x/3.0)*exp(-x);
}

static void eval_block(int family, int ndim, const double *x1,
        STARSH_int count1, const double *x2, STARSH_int count2,
        const starsh_hip_kernel_desc *kd, int nrows, int ncols,
        const STARSH_int *irow, const STARSH_int *icol, double *out, int ld)
{
    double beta_exp = -kd->beta;
    double beta_sqr = -2*kd->beta*kd->beta;
    double theta = sqrt(2*kd->nu)/kd->beta;
    const int closed = starsh_matern_closed_nu(kd->nu);
    const double mcoef = starsh_matern_coef(kd->sigma, kd->nu);
    for(int j = 0; j < ncols; j++)
        for(int i = 0; i < nrows; i++)
        {
            double dist = 0.0;
            for(int k = 0; k < ndim; k++)
            {
                double tmp = x1[k*count1+irow[i]]-x2[k*count2+icol[j]];
                dist += tmp*tmp;
            }
            double v;
            switch(family)
            {
                case STARSH_HIP_EXP:
                    dist = sqrt(dist)/beta_exp;
                    v = dist == 0 ? kd->sigma+kd->noise : kd->sigma*exp(dist);
                    break;
                case STARSH_HIP_SQREXP:
                    dist = dist/beta_sqr;
                    v = dist == 0 ? kd->sigma+kd->noise : kd->sigma*exp(dist);
                    break;
                case STARSH_HIP_MATERN:
                    dist = sqrt(dist)*theta;
                    v = dist == 0 ? kd->sigma+kd->noise : closed ?
                        kd->sigma*matern_closed(kd->nu, dist) :
                        starsh_matern_general(mcoef, kd->nu, dist);
                    break;
                case STARSH_HIP_MATERN2:
                    dist = sqrt(dist)/kd->beta;
                    v = dist == 0 ? kd->sigma+kd->noise : closed ?
                        kd->sigma*matern_closed(kd->nu, dist) :
                        starsh_matern_general(mcoef, kd->nu, dist);
                    break;
                case STARSH_HIP_COULOMB:
                    v = dist == 0 ? 0.0 : 1.0/sqrt(dist);
                    break;
                case STARSH_HIP_SIN:
                    if(dist == 0)
                        v = kd->diag;
                    else
                    {
                        dist = sqrt(dist);
                        v = sin(kd->wave_k*dist)/dist;
                    }
                    break;
                default: /* STARSH_HIP_COS */
                    if(dist == 0)
                        v = kd->diag;
                    else
                    {
                        dist = sqrt(dist);
                        v = cos(kd->wave_k*dist)/dist;
                    }
                    break;
            }
            out[(size_t)j*ld+i] = v;
        }
}

/* Descriptor of the row data for a family. */
static void describe(int family, void *data, starsh_hip_kernel_desc *kd,
        const double **points, STARSH_int *count, int *ndim)
{
    memset(kd, 0, sizeof(*kd));
    kd->family = family;
    if(family <= STARSH_HIP_MATERN2)
    {
        STARSH_ssdata *d = data;
        kd->beta = d->beta;
        kd->nu = d->nu;
        kd->noise = d->noise;
        kd->sigma = d->sigma;
        *points = d->particles.point;
        *count = d->particles.count;
        *ndim = d->particles.ndim;
    }
    else if(family == STARSH_HIP_COULOMB)
    {
        STARSH_esdata *d = data;
        *points = d->point;
        *count = d->count;
        *ndim = d->ndim;
    }
    else
    {
        STARSH_eddata *d = data;
        kd->wave_k = d->k;
        kd->diag = d->diag;
        *points = d->particles.point;
        *count = d->particles.count;
        *ndim = d->particles.ndim;
    }
}

static void run(int family, int fixed_ndim, int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld)
{
    starsh_hip_kernel_desc kd, kd2;
    const double *x1, *x2;
    STARSH_int c1, c2;
    int nd1, nd2;
    describe(family, row_data, &kd, &x1, &c1, &nd1);
    describe(family, col_data, &kd2, &x2, &c2, &nd2);
    int ndim = fixed_ndim ? fixed_ndim : nd1;
    eval_block(family, ndim, x1, c1, x2, c2, &kd, nrows, ncols, irow, icol,
            result, ld);
}

#define KERNEL_ONE(prefix, fam, famid, nd, ndval, suffix) \
void prefix##_block_##fam##_kernel_##nd##suffix(int nrows, int ncols, \
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data, \
        void *result, int ld) \
{ \
    run(famid, ndval, nrows, ncols, irow, icol, row_data, col_data, result, \
            ld); \
}

#define KERNEL_SET(prefix, fam, famid) \
    KERNEL_ONE(prefix, fam, famid, 1d, 1, ) \
    KERNEL_ONE(prefix, fam, famid, 2d, 2, ) \
    KERNEL_ONE(prefix, fam, famid, 3d, 3, ) \
    KERNEL_ONE(prefix, fam, famid, 4d, 4, ) \
    KERNEL_ONE(prefix, fam, famid, nd, 0, ) \
    KERNEL_ONE(prefix, fam, famid, 1d, 1, _simd) \
    KERNEL_ONE(prefix, fam, famid, 2d, 2, _simd) \
    KERNEL_ONE(prefix, fam, famid, 3d, 3, _simd) \
    KERNEL_ONE(prefix, fam, famid, 4d, 4, _simd) \
    KERNEL_ONE(prefix, fam, famid, nd, 0, _simd)

KERNEL_SET(starsh_ssdata, exp, STARSH_HIP_EXP)
KERNEL_SET(starsh_ssdata, sqrexp, STARSH_HIP_SQREXP)
KERNEL_SET(starsh_ssdata, matern, STARSH_HIP_MATERN)
KERNEL_SET(starsh_ssdata, matern2, STARSH_HIP_MATERN2)
KERNEL_SET(starsh_esdata, coulomb_potential, STARSH_HIP_COULOMB)
KERNEL_SET(starsh_eddata, sin, STARSH_HIP_SIN)
KERNEL_SET(starsh_eddata, cos, STARSH_HIP_COS)

#define ENTRY_ONE(prefix, fam, famid, nd, ndval, suffix) \
    {prefix##_block_##fam##_kernel_##nd##suffix, famid, ndval},
#define ENTRY_SET(prefix, fam, famid) \
    ENTRY_ONE(prefix, fam, famid, 1d, 1, ) \
    ENTRY_ONE(prefix, fam, famid, 2d, 2, ) \
    ENTRY_ONE(prefix, fam, famid, 3d, 3, ) \
    ENTRY_ONE(prefix, fam, famid, 4d, 4, ) \
    ENTRY_ONE(prefix, fam, famid, nd, 0, ) \
    ENTRY_ONE(prefix, fam, famid, 1d, 1, _simd) \
    ENTRY_ONE(prefix, fam, famid, 2d, 2, _simd) \
    ENTRY_ONE(prefix, fam, famid, 3d, 3, _simd) \
    ENTRY_ONE(prefix, fam, famid, 4d, 4, _simd) \
    ENTRY_ONE(prefix, fam, famid, nd, 0, _simd)

static const struct { STARSH_kernel *fn; int family; int ndim; } table[] =
{
    ENTRY_SET(starsh_ssdata, exp, STARSH_HIP_EXP)
    ENTRY_SET(starsh_ssdata, sqrexp, STARSH_HIP_SQREXP)
    ENTRY_SET(starsh_ssdata, matern, STARSH_HIP_MATERN)
    ENTRY_SET(starsh_ssdata, matern2, STARSH_HIP_MATERN2)
    ENTRY_SET(starsh_esdata, coulomb_potential, STARSH_HIP_COULOMB)
    ENTRY_SET(starsh_eddata, sin, STARSH_HIP_SIN)
    ENTRY_SET(starsh_eddata, cos, STARSH_HIP_COS)
};

int starsh_kernel_identify(STARSH_kernel *kernel, int *family, int *ndim)
{
    for(size_t i = 0; i < sizeof(table)/sizeof(table[0]); i++)
        if(table[i].fn == kernel)
        {
            *family = table[i].family;
            *ndim = table[i].ndim;
            return 0;
        }
    return -1;
}

int starsh_kernel_describe(STARSH_kernel *kernel, void *data,
        starsh_hip_kernel_desc *kd, const double **points, STARSH_int *count)
{
    int family, ndim, data_ndim;
    if(starsh_kernel_identify(kernel, &family, &ndim) != 0)
        return -1;
    describe(family, data, kd, points, count, &data_ndim);
    kd->ndim = ndim ? ndim : data_ndim;
    if(kd->ndim < 1 || kd->ndim > 4 || kd->ndim > data_ndim)
        return -1;
    if((family == STARSH_HIP_MATERN || family == STARSH_HIP_MATERN2) &&
            !(kd->nu > 0))
        return -1;
    return 0;
}
