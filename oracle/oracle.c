/* ORACLE — test infrastructure only.
 *
 * A CPU restatement of the reference algorithm for tile low-rank compression
 * with the randomized-SVD engine, used to check the HIP backend.  Only the
 * tests, the smoke check and the CPU-baseline leg of bench.py load it.  The
 * product library never links or calls it.
 *
 * Functions and the reference code they restate (ecrc/stars-h 0.3.1):
 *   oracle_omega          LAPACK dlarnv(idist=3, iseed={0,0,0,1}) as called by
 *                         src/backends/sequential/dense/drsdd.c:57-59; closed
 *                         form of the 48-bit multiplicative congruential
 *                         generator of dlaruv and the Box-Muller transform of
 *                         dlarnv.  Pinned by tests/golden/omega_*.
 *   oracle_kernel_block   src/applications/spatial/kernel_exp.c:31-110,
 *                         kernel_sqrexp.c:31-111, kernel_matern.c:33-120 (nu in
 *                         {1/2,3/2,5/2} in closed form; other nu > 0 by the
 *                         reference's sigma 2^(1-nu)/Gamma(nu) x^nu K_nu(x)
 *                         with K_nu from its integral representation,
 *                         oracle_besselk), kernel_matern2.c,
 *                         electrostatics/kernel_coulomb_potential.c:31-99,
 *                         electrodynamics/kernel_sin.c:31-105, kernel_cos.c.
 *   oracle_dsvfr          src/backends/sequential/dense/dsvfr.c:16-47.
 *   oracle_dlrrsdd        src/backends/sequential/dense/drsdd.c:16-93 (calls
 *                         dgemm, dgeqrf, dorgqr, dgesdd of the host LAPACK).
 *   oracle_dlrsdd         src/backends/sequential/dense/dsdd.c:16-66 (dgesdd
 *                         of the whole tile).
 *   oracle_dlrqp3         src/backends/sequential/dense/dqp3.c:16-93 (dgeqp3,
 *                         dorgqr, dgesdd of the leading rows of R).
 *   oracle_compress       src/backends/openmp/blrm/drsdd.c:94-145 (tile loop,
 *                         schedule(dynamic,1)) on a plain tiling; engine 1
 *                         and 2 take the tile loops of openmp/blrm/dsdd.c
 *                         and dqp3.c with the two functions above.
 *   oracle_dfe            src/backends/openmp/blrm/dfe.c:16-148.
 *   oracle_gen_*          src/applications/particles.c:177-201 and 341-411.
 *   oracle_zsort          src/applications/particles.c:1090-1221.
 *
 * LAPACK/BLAS come from the host (MKL libmkl_rt or scipy's OpenBLAS), loaded
 * with dlopen and called through the Fortran interface with 32-bit
 * integers, in sequential mode. */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef void dgemm_t(const char *, const char *, const int *, const int *,
        const int *, const double *, const double *, const int *,
        const double *, const int *, const double *, double *, const int *);
typedef void dgeqrf_t(const int *, const int *, double *, const int *,
        double *, double *, const int *, int *);
typedef void dorgqr_t(const int *, const int *, const int *, double *,
        const int *, const double *, double *, const int *, int *);
typedef void dgesdd_t(const char *, const int *, const int *, double *,
        const int *, double *, double *, const int *, double *, const int *,
        double *, const int *, int *, int *);

typedef void dgeqp3_t(const int *, const int *, double *, const int *,
        int *, double *, double *, const int *, int *);

static dgemm_t *p_dgemm;
static dgeqp3_t *p_dgeqp3;
static dgeqrf_t *p_dgeqrf;
static dorgqr_t *p_dorgqr;
static dgesdd_t *p_dgesdd;
static char lapack_name[512];

/* Returns 0 when a LAPACK was found.  `path` may name a library to try
 * first. */
int oracle_lapack_load(const char *path)
{
    if(p_dgemm)
        return 0;
    const char *cands[] = {path, "/opt/conda/lib/libmkl_rt.so",
        "libmkl_rt.so", NULL};
    const char *prefixes[] = {"", "scipy_"};
    for(int c = 0; c < 3; c++)
    {
        if(cands[c] == NULL)
            continue;
        void *h = dlopen(cands[c], RTLD_NOW | RTLD_LOCAL);
        if(h == NULL)
            continue;
        for(int p = 0; p < 2; p++)
        {
            char nm[64];
            snprintf(nm, sizeof nm, "%sdgemm_", prefixes[p]);
            p_dgemm = (dgemm_t *)dlsym(h, nm);
            snprintf(nm, sizeof nm, "%sdgeqrf_", prefixes[p]);
            p_dgeqrf = (dgeqrf_t *)dlsym(h, nm);
            snprintf(nm, sizeof nm, "%sdorgqr_", prefixes[p]);
            p_dorgqr = (dorgqr_t *)dlsym(h, nm);
            snprintf(nm, sizeof nm, "%sdgesdd_", prefixes[p]);
            p_dgesdd = (dgesdd_t *)dlsym(h, nm);
            snprintf(nm, sizeof nm, "%sdgeqp3_", prefixes[p]);
            p_dgeqp3 = (dgeqp3_t *)dlsym(h, nm);
            if(p_dgemm && p_dgeqrf && p_dorgqr && p_dgesdd)
            {
                int (*layer)(int) = (int (*)(int))dlsym(h,
                        "MKL_Set_Threading_Layer");
                if(layer)
                    layer(1); /* MKL_THREADING_SEQUENTIAL */
                void (*nt)(int) = (void (*)(int))dlsym(h,
                        "scipy_openblas_set_num_threads");
                if(nt)
                    nt(1);
                snprintf(lapack_name, sizeof lapack_name, "%s", cands[c]);
                return 0;
            }
        }
        p_dgemm = NULL;
    }
    return -1;
}

const char *oracle_lapack_name(void)
{
    return lapack_name;
}

/* ---- dlarnv(3, {0,0,0,1}, n, x).  dlaruv advances the 48-bit state by the
 * multiplier a = 33952834046453 (mod 2^48) per uniform and returns
 * state * 2^-48; dlarnv pairs consecutive uniforms (u1, u2) into
 * sqrt(-2 log u1) cos(2 pi u2).  The state starts at 1 (iseed limbs
 * {0,0,0,1}, 12 bits each). */
void oracle_dlarnv3(int64_t n, double *x)
{
    const uint64_t a = 33952834046453ULL, mask = (1ULL << 48)-1;
    const double twopi = 6.2831853071795864769252867663;
    uint64_t s = 1;
    for(int64_t i = 0; i < n; i++)
    {
        s = (s*a) & mask;
        double u1 = (double)s*0x1p-48;
        s = (s*a) & mask;
        double u2 = (double)s*0x1p-48;
        x[i] = sqrt(-2.0*log(u1))*cos(twopi*u2);
    }
}

void oracle_omega(int ncols, int k, double *out)
{
    oracle_dlarnv3((int64_t)ncols*k, out);
}

/* ---- Kernels.  family: 1 exp, 2 sqrexp, 3 matern, 4 matern2, 5 coulomb,
 * 6 sin, 7 cos.  Points are SoA: x[d*count + i]. */
typedef struct
{
    int family, ndim;
    double beta, nu, noise, sigma, wave_k, diag;
} oracle_kernel;

static double matern_closed(double nu, double x)
{
    if(nu == 0.5)
        return exp(-x);
    if(nu == 1.5)
        return (1.0+x)*exp(-x);
    return (1.0+x+x*x/3.0)*exp(-x);
}

/* K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt (Watson 6.22 (5)) by
 * the trapezoidal rule with step min(1/20, 1/(2 sqrt x)) -- fine enough
 * for the Gaussian-like peak of width 1/sqrt(x) at large x -- and
 * compensated summation, until x cosh t exceeds 800.  A method of its own,
 * independent of the product's (continued fraction for x >= 2, a coarser
 * rule below); pinned against scipy.special.kv values in
 * tests/golden/kernel_golden.json. */
double oracle_besselk(double nu, double x)
{
    const double h = fmin(0.05, 0.5/sqrt(x));
    double s = 0.5*exp(-x), comp = 0.0;
    for(int64_t j = 1; ; j++)
    {
        const double t = j*h;
        const double xc = x*cosh(t);
        if(xc > 800.0)
            break;
        const double y = exp(-xc)*cosh(nu*t)-comp;
        const double u = s+y;
        comp = (u-s)-y;
        s = u;
    }
    return h*s;
}

static double matern_any(double sigma, double nu, double x)
{
    if(nu == 0.5 || nu == 1.5 || nu == 2.5)
        return sigma*matern_closed(nu, x);
    return sigma*pow(2.0, 1.0-nu)/tgamma(nu)*pow(x, nu)*
        oracle_besselk(nu, x);
}

void oracle_kernel_block(const oracle_kernel *kd, const double *x1,
        int64_t c1, const double *x2, int64_t c2, int nrows, int ncols,
        const int64_t *irow, const int64_t *icol, double *out, int ld)
{
    double bexp = -kd->beta;
    double bsqr = -2*kd->beta*kd->beta;
    double theta = sqrt(2*kd->nu)/kd->beta;
    for(int j = 0; j < ncols; j++)
        for(int i = 0; i < nrows; i++)
        {
            double dist = 0.0, t, v;
            for(int d = 0; d < kd->ndim; d++)
            {
                t = x1[d*c1+irow[i]]-x2[d*c2+icol[j]];
                dist += t*t;
            }
            switch(kd->family)
            {
                case 1:
                    dist = sqrt(dist)/bexp;
                    v = dist == 0 ? kd->sigma+kd->noise : kd->sigma*exp(dist);
                    break;
                case 2:
                    dist = dist/bsqr;
                    v = dist == 0 ? kd->sigma+kd->noise : kd->sigma*exp(dist);
                    break;
                case 3:
                    dist = sqrt(dist)*theta;
                    v = dist == 0 ? kd->sigma+kd->noise :
                        matern_any(kd->sigma, kd->nu, dist);
                    break;
                case 4:
                    dist = sqrt(dist)/kd->beta;
                    v = dist == 0 ? kd->sigma+kd->noise :
                        matern_any(kd->sigma, kd->nu, dist);
                    break;
                case 5:
                    v = dist == 0 ? 0.0 : 1.0/sqrt(dist);
                    break;
                case 6:
                    v = dist == 0 ? kd->diag :
                        sin(kd->wave_k*sqrt(dist))/sqrt(dist);
                    break;
                default:
                    v = dist == 0 ? kd->diag :
                        cos(kd->wave_k*sqrt(dist))/sqrt(dist);
                    break;
            }
            out[(size_t)j*ld+i] = v;
        }
}

/* ---- dsvfr.c:16-47 */
int oracle_dsvfr(int size, const double *S, double tol)
{
    int i;
    double err_tol = 0;
    for(i = 0; i < size; i++)
        err_tol += S[i]*S[i];
    if(err_tol == 0)
        return 0;
    i = size;
    err_tol *= tol*tol;
    double tmp_norm = S[size-1]*S[size-1];
    while(i > 1 && err_tol >= tmp_norm)
    {
        i--;
        tmp_norm += S[i-1]*S[i-1];
    }
    return i;
}

/* ---- drsdd.c:16-93.  Returns the rank (-1 for a false far-field tile).
 * U (nrows x maxrank, ld ldU) and V (ncols x maxrank, ld ldV) receive the
 * factors; S (mn2 values) the singular values of B when not NULL. */
int oracle_dlrrsdd(int nrows, int ncols, const double *D, int ldD, double *U,
        int ldU, double *V, int ldV, int maxrank, int oversample, double tol,
        double *S)
{
    int mn = nrows < ncols ? nrows : ncols;
    int mn2 = maxrank+oversample;
    if(mn2 > mn)
        mn2 = mn;
    size_t nX = (size_t)ncols*mn2, nQ = (size_t)nrows*mn2;
    double *X = malloc(sizeof(double)*nX);
    double *Q = malloc(sizeof(double)*nQ);
    double *sU = malloc(sizeof(double)*mn2*mn2);
    double *sS = malloc(sizeof(double)*mn2);
    double *sV = malloc(sizeof(double)*nX);
    double *tau = malloc(sizeof(double)*mn2);
    int *iwork = malloc(sizeof(int)*8*mn2);
    int lwork = -1, info;
    double wq;
    p_dgesdd("S", &mn2, &ncols, X, &mn2, sS, sU, &mn2, sV, &mn2, &wq, &lwork,
            iwork, &info);
    int lw1 = (int)wq;
    p_dgeqrf(&nrows, &mn2, Q, &nrows, tau, &wq, &lwork, &info);
    if((int)wq > lw1)
        lw1 = (int)wq;
    p_dorgqr(&nrows, &mn2, &mn2, Q, &nrows, tau, &wq, &lwork, &info);
    if((int)wq > lw1)
        lw1 = (int)wq;
    double *work = malloc(sizeof(double)*(lw1+1));
    const double one = 1.0, zero = 0.0;
    /* Omega: ncols x mn2, leading dimension ncols. */
    oracle_dlarnv3((int64_t)ncols*mn2, X);
    /* Q = D * X */
    p_dgemm("N", "N", &nrows, &mn2, &ncols, &one, D, &ldD, X, &ncols, &zero,
            Q, &nrows);
    p_dgeqrf(&nrows, &mn2, Q, &nrows, tau, work, &lw1, &info);
    p_dorgqr(&nrows, &mn2, &mn2, Q, &nrows, tau, work, &lw1, &info);
    /* X = Q^T * D (mn2 x ncols) */
    p_dgemm("T", "N", &mn2, &ncols, &nrows, &one, Q, &nrows, D, &ldD, &zero,
            X, &mn2);
    p_dgesdd("S", &mn2, &ncols, X, &mn2, sS, sU, &mn2, sV, &mn2, work, &lw1,
            iwork, &info);
    int rank = oracle_dsvfr(mn2, sS, tol);
    if(S)
        memcpy(S, sS, sizeof(double)*mn2);
    if(info == 0 && rank <= maxrank)
    {
        if(rank > 0 && U)
            p_dgemm("N", "N", &nrows, &rank, &mn2, &one, Q, &nrows, sU, &mn2,
                    &zero, U, &ldU);
        if(V)
            for(int i = 0; i < rank; i++)
                for(int j = 0; j < ncols; j++)
                    V[(size_t)i*ldV+j] = sV[(size_t)j*mn2+i]*sS[i];
    }
    else
        rank = -1;
    free(X); free(Q); free(sU); free(sS); free(sV); free(tau); free(iwork);
    free(work);
    return rank;
}

/* ---- dsdd.c:16-66.  Rank (-1 for a false far-field tile); U, V as for
 * oracle_dlrrsdd; S (min(m, n) values) when not NULL.  D is overwritten. */
int oracle_dlrsdd(int nrows, int ncols, double *D, int ldD, double *U,
        int ldU, double *V, int ldV, int maxrank, double tol, double *S)
{
    int mn = nrows < ncols ? nrows : ncols;
    double *sU = malloc(sizeof(double)*(size_t)nrows*mn);
    double *sS = malloc(sizeof(double)*mn);
    double *sV = malloc(sizeof(double)*(size_t)ncols*mn);
    int *iwork = malloc(sizeof(int)*8*mn);
    int lwork = -1, info;
    double wq;
    p_dgesdd("S", &nrows, &ncols, D, &ldD, sS, sU, &nrows, sV, &mn, &wq,
            &lwork, iwork, &info);
    lwork = (int)wq+1;
    double *work = malloc(sizeof(double)*lwork);
    p_dgesdd("S", &nrows, &ncols, D, &ldD, sS, sU, &nrows, sV, &mn, work,
            &lwork, iwork, &info);
    int rank = oracle_dsvfr(mn, sS, tol);
    if(S)
        memcpy(S, sS, sizeof(double)*mn);
    if(info == 0 && rank < mn/2 && rank <= maxrank)
    {
        for(int i = 0; i < rank; i++)
        {
            if(U)
                memcpy(U+(size_t)i*ldU, sU+(size_t)i*nrows,
                        sizeof(double)*nrows);
            if(V)
                for(int j = 0; j < ncols; j++)
                    V[(size_t)i*ldV+j] = sV[(size_t)j*mn+i]*sS[i];
        }
    }
    else
        rank = -1;
    free(sU); free(sS); free(sV); free(iwork); free(work);
    return rank;
}

/* ---- dqp3.c:16-93.  Rank (-1 for a false far-field tile); U, V as for
 * oracle_dlrrsdd; S (mn2 values) and the column pivots (ncols, 0-based)
 * when not NULL.  D is overwritten. */
int oracle_dlrqp3(int nrows, int ncols, double *D, int ldD, double *U,
        int ldU, double *V, int ldV, int maxrank, int oversample, double tol,
        double *S, int *piv)
{
    int mn = nrows < ncols ? nrows : ncols;
    int mn2 = maxrank+oversample;
    if(mn2 > mn)
        mn2 = mn;
    int *jpvt = calloc(ncols, sizeof(int));
    double *tau = malloc(sizeof(double)*mn);
    double *R = calloc((size_t)ncols*mn2, sizeof(double));
    double *sU = malloc(sizeof(double)*(size_t)mn2*mn2);
    double *sS = malloc(sizeof(double)*mn2);
    double *sV = malloc(sizeof(double)*(size_t)ncols*mn2);
    int *iwork = malloc(sizeof(int)*8*mn2);
    int lwork = -1, info;
    double wq;
    p_dgeqp3(&nrows, &ncols, D, &ldD, jpvt, tau, &wq, &lwork, &info);
    int lw = (int)wq;
    p_dorgqr(&nrows, &mn2, &mn2, D, &ldD, tau, &wq, &lwork, &info);
    lw = (int)wq > lw ? (int)wq : lw;
    p_dgesdd("S", &mn2, &ncols, R, &mn2, sS, sU, &mn2, sV, &mn2, &wq,
            &lwork, iwork, &info);
    lw = (int)wq > lw ? (int)wq : lw;
    double *work = malloc(sizeof(double)*(lw+1));
    p_dgeqp3(&nrows, &ncols, D, &ldD, jpvt, tau, work, &lw, &info);
    /* R (mn2 x ncols) with the columns put back in their original order:
     * column i of the factor is column jpvt[i]-1 of the tile */
    for(int i = 0; i < ncols; i++)
    {
        const int j = jpvt[i]-1;
        const int kk = i < mn2 ? i+1 : mn2;
        for(int l = 0; l < kk; l++)
            R[(size_t)j*mn2+l] = D[(size_t)i*ldD+l];
    }
    if(piv)
        for(int i = 0; i < ncols; i++)
            piv[i] = jpvt[i]-1;
    p_dorgqr(&nrows, &mn2, &mn2, D, &ldD, tau, work, &lw, &info);
    p_dgesdd("S", &mn2, &ncols, R, &mn2, sS, sU, &mn2, sV, &mn2, work, &lw,
            iwork, &info);
    int rank = oracle_dsvfr(mn2, sS, tol);
    if(S)
        memcpy(S, sS, sizeof(double)*mn2);
    if(info == 0 && rank < mn/2 && rank <= maxrank)
    {
        const double one = 1.0, zero = 0.0;
        if(rank > 0 && U)
            p_dgemm("N", "N", &nrows, &rank, &mn2, &one, D, &ldD, sU, &mn2,
                    &zero, U, &ldU);
        if(V)
            for(int i = 0; i < rank; i++)
                for(int j = 0; j < ncols; j++)
                    V[(size_t)i*ldV+j] = sV[(size_t)j*mn2+i]*sS[i];
    }
    else
        rank = -1;
    free(jpvt); free(tau); free(R); free(sU); free(sS); free(sV);
    free(iwork); free(work);
    return rank;
}

/* ---- openmp/blrm/drsdd.c:94-145 on a plain tiling with identity pivot.
 * Tile bi covers rows start[brow[bi]] .. +size and columns start[bcol[bi]]
 * .. +size of the point set(s).  U/V for tile bi go to U + uoff[bi] and
 * V + voff[bi] (maxrank columns each, ld = rows / cols) when U and V are
 * not NULL; S to S + bi*smax when S is not NULL. */
int oracle_compress_engine(int engine, const oracle_kernel *kd,
        const double *xr, int64_t cr, const double *xc, int64_t cc,
        int64_t nfar, const int64_t *brow, const int64_t *bcol,
        const int64_t *rstart, const int64_t *rsize, const int64_t *cstart,
        const int64_t *csize, int maxrank, int oversample, double tol,
        int nthreads, int *rank, double *U, const int64_t *uoff, double *V,
        const int64_t *voff, double *S, int smax)
{
    if(oracle_lapack_load(NULL) != 0 || (engine == 2 && p_dgeqp3 == NULL))
        return -1;
    if(nthreads > 0)
        omp_set_num_threads(nthreads);
    #pragma omp parallel for schedule(dynamic,1)
    for(int64_t bi = 0; bi < nfar; bi++)
    {
        int m = (int)rsize[brow[bi]], n = (int)csize[bcol[bi]];
        int64_t *ir = malloc(sizeof(int64_t)*m);
        int64_t *ic = malloc(sizeof(int64_t)*n);
        double *D = malloc(sizeof(double)*(size_t)m*n);
        for(int i = 0; i < m; i++)
            ir[i] = rstart[brow[bi]]+i;
        for(int j = 0; j < n; j++)
            ic[j] = cstart[bcol[bi]]+j;
        oracle_kernel_block(kd, xr, cr, xc, cc, m, n, ir, ic, D, m);
        double *Ub = U ? U+uoff[bi] : NULL, *Vb = V ? V+voff[bi] : NULL;
        double *Sb = S ? S+bi*(int64_t)smax : NULL;
        if(engine == 1)
            rank[bi] = oracle_dlrsdd(m, n, D, m, Ub, m, Vb, n, maxrank, tol,
                    Sb);
        else if(engine == 2)
            rank[bi] = oracle_dlrqp3(m, n, D, m, Ub, m, Vb, n, maxrank,
                    oversample, tol, Sb, NULL);
        else
            rank[bi] = oracle_dlrrsdd(m, n, D, m, Ub, m, Vb, n, maxrank,
                    oversample, tol, Sb);
        free(ir);
        free(ic);
        free(D);
    }
    return 0;
}

int oracle_compress(const oracle_kernel *kd, const double *xr, int64_t cr,
        const double *xc, int64_t cc, int64_t nfar, const int64_t *brow,
        const int64_t *bcol, const int64_t *rstart, const int64_t *rsize,
        const int64_t *cstart, const int64_t *csize, int maxrank,
        int oversample, double tol, int nthreads, int *rank, double *U,
        const int64_t *uoff, double *V, const int64_t *voff, double *S,
        int smax)
{
    return oracle_compress_engine(0, kd, xr, cr, xc, cc, nfar, brow, bcol,
            rstart, rsize, cstart, csize, maxrank, oversample, tol, nthreads,
            rank, U, uoff, V, voff, S, smax);
}

/* ---- openmp/blrm/dfe.c:16-148.  far tiles carry (brow, bcol, rank, U, V);
 * near tiles are regenerated.  Off-diagonal tiles count twice when symm is
 * 'S' (sqrt(2) factors).  Returns the global relative error; per-tile
 * |A| and |A - UV^T| go to far_norm / far_diff when not NULL. */
double oracle_dfe(const oracle_kernel *kd, const double *xr, int64_t cr,
        const double *xc, int64_t cc, char symm, int64_t nfar,
        const int64_t *fbrow, const int64_t *fbcol, const int *rank,
        const double *U, const int64_t *uoff, const double *V,
        const int64_t *voff, int64_t nnear, const int64_t *nbrow,
        const int64_t *nbcol, const int64_t *rstart, const int64_t *rsize,
        const int64_t *cstart, const int64_t *csize, int nthreads,
        double *far_norm, double *far_diff)
{
    if(nthreads > 0)
        omp_set_num_threads(nthreads);
    double *bn = calloc(nfar+nnear+1, sizeof(double));
    double *bd = calloc(nfar+1, sizeof(double));
    #pragma omp parallel for schedule(dynamic,1)
    for(int64_t bi = 0; bi < nfar+nnear; bi++)
    {
        int far = bi < nfar;
        int64_t i = far ? fbrow[bi] : nbrow[bi-nfar];
        int64_t j = far ? fbcol[bi] : nbcol[bi-nfar];
        int m = (int)rsize[i], n = (int)csize[j];
        int64_t *ir = malloc(sizeof(int64_t)*m);
        int64_t *ic = malloc(sizeof(int64_t)*n);
        double *D = malloc(sizeof(double)*(size_t)m*n);
        for(int a = 0; a < m; a++)
            ir[a] = rstart[i]+a;
        for(int b = 0; b < n; b++)
            ic[b] = cstart[j]+b;
        oracle_kernel_block(kd, xr, cr, xc, cc, m, n, ir, ic, D, m);
        double s = 0;
        for(size_t e = 0; e < (size_t)m*n; e++)
            s += D[e]*D[e];
        double nrm = sqrt(s), dif = 0;
        if(far)
        {
            const double *u = U+uoff[bi], *v = V+voff[bi];
            int r = rank[bi];
            s = 0;
            for(int b = 0; b < n; b++)
                for(int a = 0; a < m; a++)
                {
                    double acc = D[(size_t)b*m+a];
                    for(int q = 0; q < r; q++)
                        acc -= u[(size_t)q*m+a]*v[(size_t)q*n+b];
                    s += acc*acc;
                }
            dif = sqrt(s);
        }
        if(i != j && symm == 'S')
        {
            nrm *= sqrt(2.);
            dif *= sqrt(2.);
        }
        bn[bi] = nrm;
        if(far)
        {
            bd[bi] = dif;
            if(far_norm)
                far_norm[bi] = nrm;
            if(far_diff)
                far_diff[bi] = dif;
        }
        free(ir);
        free(ic);
        free(D);
    }
    double diff = 0, norm = 0;
    for(int64_t bi = 0; bi < nfar; bi++)
        diff += bd[bi]*bd[bi];
    for(int64_t bi = 0; bi < nfar+nnear; bi++)
        norm += bn[bi]*bn[bi];
    free(bn);
    free(bd);
    return sqrt(diff)/sqrt(norm);
}

/* ---- particles.c:177-201 (RAND) and 341-411 (QUASIUNIFORM1); both consume
 * rand() after srand(seed) in the reference's order.  out: count*ndim SoA,
 * not yet Morton-sorted. */
void oracle_gen_rand(int64_t count, int ndim, unsigned seed, double *out)
{
    srand(seed);
    for(int64_t i = 0; i < count*ndim; i++)
        out[i] = (double)rand()/RAND_MAX;
}

int oracle_gen_quasiuniform1(int64_t count, int ndim, unsigned seed,
        double *out)
{
    srand(seed);
    int64_t side = (int64_t)ceil(pow((double)count, 1.0/ndim));
    double mult = 0.8/side/RAND_MAX, add = 0.1/side;
    int64_t idx[16] = {0};
    if(ndim > 16)
        return -1;
    for(int64_t i = 0; i < count; i++)
    {
        for(int d = 0; d < ndim; d++)
            out[i+d*count] = (double)idx[d]/side+add+mult*rand();
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
    return 0;
}

/* ---- particles.c:1090-1221, recursive as in the reference. */
static void radix_rec(const uint32_t *data, int64_t count, int ndim,
        int64_t *order, int64_t *tmp, int sdim, int sbit, int64_t lo,
        int64_t hi)
{
    int64_t lo_last = lo, hi_last = hi;
    const uint32_t *sd = data+sdim*count;
    uint32_t check = 1u << sbit;
    for(int64_t i = lo; i <= hi; i++)
    {
        if((sd[order[i]] & check) == 0)
            tmp[lo_last++] = order[i];
        else
            tmp[hi_last--] = order[i];
    }
    for(int64_t i = lo; i <= hi; i++)
        order[i] = tmp[i];
    if(sdim > 0)
    {
        if(lo_last-lo > 1)
            radix_rec(data, count, ndim, order, tmp, sdim-1, sbit, lo,
                    lo_last-1);
        if(hi-hi_last > 1)
            radix_rec(data, count, ndim, order, tmp, sdim-1, sbit,
                    hi_last+1, hi);
    }
    else if(sbit > 0)
    {
        if(lo_last-lo > 1)
            radix_rec(data, count, ndim, order, tmp, ndim-1, sbit-1, lo,
                    lo_last-1);
        if(hi-hi_last > 1)
            radix_rec(data, count, ndim, order, tmp, ndim-1, sbit-1,
                    hi_last+1, hi);
    }
}

/* Sorts `point` (SoA) in place; also returns the permutation in `order`
 * when not NULL. */
void oracle_zsort(int64_t count, int ndim, double *point, int64_t *order_out)
{
    uint32_t *u = malloc(sizeof(uint32_t)*count*ndim);
    for(int d = 0; d < ndim; d++)
    {
        double *p = point+d*count, mn = p[0], mx = p[0];
        for(int64_t j = 1; j < count; j++)
        {
            if(mn > p[j])
                mn = p[j];
            else if(mx < p[j])
                mx = p[j];
        }
        double range = mx-mn;
        for(int64_t j = 0; j < count; j++)
            u[d*count+j] = (p[j]-mn)/range*UINT32_MAX;
    }
    int64_t *order = malloc(sizeof(int64_t)*count);
    int64_t *tmp = malloc(sizeof(int64_t)*count);
    for(int64_t j = 0; j < count; j++)
        order[j] = j;
    radix_rec(u, count, ndim, order, tmp, ndim-1, 31, 0, count-1);
    double *np = malloc(sizeof(double)*count*ndim);
    for(int64_t j = 0; j < count; j++)
        for(int d = 0; d < ndim; d++)
            np[d*count+j] = point[d*count+order[j]];
    memcpy(point, np, sizeof(double)*count*ndim);
    if(order_out)
        memcpy(order_out, order, sizeof(int64_t)*count);
    free(u);
    free(order);
    free(tmp);
    free(np);
}
