// Shared definitions of the HIP backend: per-tile descriptor of a compression
// batch, kernel-value functions, launch helpers.  gfx950 (MI355X) only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "starsh.h"
#include "starsh-hip.h"

#define HIP_TRY(expr) do { hipError_t e_ = (expr); if(e_ != hipSuccess) { \
    fprintf(stderr, "STARSH HIP ERROR: %s at %s:%d\n", hipGetErrorString(e_), \
            __FILE__, __LINE__); \
    return e_ == hipErrorOutOfMemory ? STARSH_MALLOC_ERROR : \
        STARSH_UNKNOWN_ERROR; } } while(0)

// Kernel parameters in the form the device evaluates them.  Derived on the
// host with the same operations as the reference kernels.
struct KernelParams
{
    int family;
    int ndim;
    double bexp;   // -beta                 (kernel_exp.c:74)
    double bsqr;   // -2*beta*beta          (kernel_sqrexp.c:75)
    double theta;  // sqrt(2*nu)/beta       (kernel_matern.c:82)
    double beta;
    double nu;
    double sigma;
    double noise;
    double wave_k;
    double diag;
    double ibexp;  // 1/bexp, 1/bsqr, 1/beta (kval)
    double ibsqr;
    double ibeta;
};

static inline KernelParams make_params(const starsh_hip_kernel_desc *kd)
{
    KernelParams p;
    p.family = kd->family;
    p.ndim = kd->ndim;
    p.bexp = -kd->beta;
    p.bsqr = -2*kd->beta*kd->beta;
    p.theta = sqrt(2*kd->nu)/kd->beta;
    p.beta = kd->beta;
    p.nu = kd->nu;
    p.sigma = kd->sigma;
    p.noise = kd->noise;
    p.wave_k = kd->wave_k;
    p.diag = kd->diag;
    p.ibexp = 1.0/p.bexp;
    p.ibsqr = 1.0/p.bsqr;
    p.ibeta = 1.0/p.beta;
    return p;
}

// One far-field tile of a compression batch.  Offsets count doubles from
// the base of the batch workspace; the device writes kc, p, rank, flags,
// bnorm2, tail2.
struct TileDesc
{
    int m, n, k;       // tile rows, columns, sketch width
    int gen;           // 1: A generated from points, 0: explicit A
    int64_t r0, c0;    // first row / column point (cluster order)
    int64_t a_off;     // explicit A (m x n, ld m) when gen == 0
    int64_t om_off;    // test matrix Omega (n x k, ld n)
    int64_t q_off;     // Y, then Q (m x k, ld m)
    int64_t bt_off;    // Bt = A^T Q (n x k, ld n)
    int64_t q2_off;    // Q2 (n x k, ld n)
    int64_t r_off;     // R2 (k x k, ld k)
    int64_t w_off;     // X_r (p x r, ld p): left singular vectors of B
    int64_t s_off;     // singular values, k entries
    int64_t u_off;     // U (m x maxrank, ld m)
    int64_t v_off;     // V (n x maxrank, ld n)
    int64_t om8_off;   // integer test matrix (bytes, n x k, ld n up to 64)
    int kc;           // kept (numerically independent) sketch columns
    int p;             // size of the core problem after truncation
    int rank;          // chosen rank, -1 for a false far-field tile
    int flags;
    double bnorm2;     // |B|_F^2
    double tail2;      // squared norm of the truncated rows of B
    // the tile's problem (a batch may mix problems of one kernel family
    // and dimension)
    const double *xr;  // row points in cluster order, SoA with stride nr
    const double *xc;  // column points, SoA with stride nc
    int64_t nr, nc;
    KernelParams kp;
    double tol;        // accuracy of the rank rule
    int maxrank;
};

__device__ inline double matern_closed(double nu, double x)
{
    double e = exp(-x);
    if(nu == 0.5)
        return e;
    if(nu == 1.5)
        return (1.0+x)*e;
    return (1.0+x+x*x/3.0)*e;
}

// Entry A(i, j) = f(|x_i - y_j|) in FP64 with the reference's operation
// order (no contraction into fused multiply-adds).
template <int NDIM>
__device__ inline double kernel_value(const KernelParams &kp,
        const double *__restrict__ xr, int64_t nr, int64_t i,
        const double *__restrict__ xc, int64_t nc, int64_t j)
{
#pragma clang fp contract(off)
    double dist = 0.0;
#pragma unroll
    for(int d = 0; d < NDIM; d++)
    {
        double t = xr[d*nr+i]-xc[d*nc+j];
        dist += t*t;
    }
    double v;
    switch(kp.family)
    {
        case STARSH_HIP_EXP:
            dist = sqrt(dist)/kp.bexp;
            v = dist == 0 ? kp.sigma+kp.noise : kp.sigma*exp(dist);
            break;
        case STARSH_HIP_SQREXP:
            dist = dist/kp.bsqr;
            v = dist == 0 ? kp.sigma+kp.noise : kp.sigma*exp(dist);
            break;
        case STARSH_HIP_MATERN:
            dist = sqrt(dist)*kp.theta;
            v = dist == 0 ? kp.sigma+kp.noise :
                kp.sigma*matern_closed(kp.nu, dist);
            break;
        case STARSH_HIP_MATERN2:
            dist = sqrt(dist)/kp.beta;
            v = dist == 0 ? kp.sigma+kp.noise :
                kp.sigma*matern_closed(kp.nu, dist);
            break;
        case STARSH_HIP_COULOMB:
            v = dist == 0 ? 0.0 : 1.0/sqrt(dist);
            break;
        case STARSH_HIP_SIN:
            if(dist == 0)
                v = kp.diag;
            else
            {
                dist = sqrt(dist);
                v = sin(kp.wave_k*dist)/dist;
            }
            break;
        default:
            if(dist == 0)
                v = kp.diag;
            else
            {
                dist = sqrt(dist);
                v = cos(kp.wave_k*dist)/dist;
            }
            break;
    }
    return v;
}

// Entry A(i, j) for the compression GEMMs: the family fixed at compile
// time, the row point x_i held in registers, reciprocals of the kernel
// parameters instead of divisions and contraction into fused multiply-adds
// allowed.  It agrees with kernel_value to a few units in the last place;
// tiles handed back to the caller (near field, error check) use
// kernel_value.
template <int NDIM, int FAM>
__device__ inline double kval_pt(const KernelParams &kp,
        const double (&xi)[NDIM], const double (&yj)[NDIM])
{
    double dist = 0.0;
#pragma unroll
    for(int d = 0; d < NDIM; d++)
    {
        const double t = xi[d]-yj[d];
        dist = fma(t, t, dist);
    }
    double v;
    switch(FAM)
    {
        case STARSH_HIP_EXP:
            v = dist == 0 ? kp.sigma+kp.noise :
                kp.sigma*exp(sqrt(dist)*kp.ibexp);
            break;
        case STARSH_HIP_SQREXP:
            v = dist == 0 ? kp.sigma+kp.noise : kp.sigma*exp(dist*kp.ibsqr);
            break;
        case STARSH_HIP_MATERN:
            v = dist == 0 ? kp.sigma+kp.noise :
                kp.sigma*matern_closed(kp.nu, sqrt(dist)*kp.theta);
            break;
        case STARSH_HIP_MATERN2:
            v = dist == 0 ? kp.sigma+kp.noise :
                kp.sigma*matern_closed(kp.nu, sqrt(dist)*kp.ibeta);
            break;
        case STARSH_HIP_COULOMB:
            v = dist == 0 ? 0.0 : rsqrt(dist);
            break;
        case STARSH_HIP_SIN:
        {
            const double ri = rsqrt(dist);
            v = dist == 0 ? kp.diag : sin(kp.wave_k*(dist*ri))*ri;
            break;
        }
        default:
        {
            const double ri = rsqrt(dist);
            v = dist == 0 ? kp.diag : cos(kp.wave_k*(dist*ri))*ri;
            break;
        }
    }
    return v;
}

template <int NDIM, int FAM>
__device__ inline double kval(const KernelParams &kp, const double (&xi)[NDIM],
        const double *__restrict__ yc, int64_t ny, int64_t j)
{
    double yj[NDIM];
#pragma unroll
    for(int d = 0; d < NDIM; d++)
        yj[d] = yc[d*ny+j];
    return kval_pt<NDIM, FAM>(kp, xi, yj);
}

// Points of a batch: row and column point sets in cluster order, SoA.
struct PointSets
{
    const double *xr;
    int64_t nr;
    const double *xc;
    int64_t nc;
};

// Library stream (created by starsh_hip_init).
extern "C" hipStream_t starsh_hip_stream(void);

// Calls F.template operator()<NDIM, FAM>() for the runtime (ndim, family).
template <typename F>
inline int dispatch_kernel(int ndim, int family, F &&f)
{
#define STARSH_FAM_CASES(D) \
    switch(family) \
    { \
        case STARSH_HIP_EXP: return f.template operator()<D, \
            STARSH_HIP_EXP>(); \
        case STARSH_HIP_SQREXP: return f.template operator()<D, \
            STARSH_HIP_SQREXP>(); \
        case STARSH_HIP_MATERN: return f.template operator()<D, \
            STARSH_HIP_MATERN>(); \
        case STARSH_HIP_MATERN2: return f.template operator()<D, \
            STARSH_HIP_MATERN2>(); \
        case STARSH_HIP_COULOMB: return f.template operator()<D, \
            STARSH_HIP_COULOMB>(); \
        case STARSH_HIP_SIN: return f.template operator()<D, \
            STARSH_HIP_SIN>(); \
        default: return f.template operator()<D, STARSH_HIP_COS>(); \
    }
    switch(ndim)
    {
        case 1: STARSH_FAM_CASES(1)
        case 2: STARSH_FAM_CASES(2)
        case 3: STARSH_FAM_CASES(3)
        default: STARSH_FAM_CASES(4)
    }
#undef STARSH_FAM_CASES
}
