/* Spatial statistics problem (reference include/starsh-spatial.h:35-260,
 * src/applications/spatial.c, src/applications/spatial/kernel_*.c).
 *
 * STARSH_ssdata keeps the reference layout.  Kernels compute
 * A[j*ld+i] = sigma*f(r) for r = |x_irow[i] - x_icol[j]| with
 * f = exp(-r/beta) (EXP), exp(-r^2/(2 beta^2)) (SQREXP),
 * Matern 2^(1-nu)/Gamma(nu) x^nu K_nu(x) with x = sqrt(2 nu) r/beta (MATERN)
 * or x = r/beta (MATERN2), and sigma+noise where r == 0.  The _simd twins
 * compute the same values.  Matern kernels accept nu in {0.5, 1.5, 2.5},
 * where K_nu has a closed form; starsh_ssdata_get_kernel rejects other nu
 * with STARSH_WRONG_PARAMETER (the reference needs GSL for them). */
#ifndef STARSH_SPATIAL_H
#define STARSH_SPATIAL_H

#include "starsh.h"
#include "starsh-particles.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct starsh_ssdata
{
    STARSH_particles particles;
    double beta;
    double nu;
    double noise;
    double sigma;
    double nu2;
    double sigma2;
    double corr;
} STARSH_ssdata;

enum STARSH_SPATIAL_KERNEL
{
    STARSH_SPATIAL_EXP = 1,
    STARSH_SPATIAL_SQREXP = 2,
    STARSH_SPATIAL_MATERN = 3,
    STARSH_SPATIAL_MATERN2 = 4,
    STARSH_SPATIAL_EXP_SIMD = 11,
    STARSH_SPATIAL_SQREXP_SIMD = 12,
    STARSH_SPATIAL_MATERN_SIMD = 13,
    STARSH_SPATIAL_MATERN2_SIMD = 14
};

enum STARSH_SPATIAL_PARAM
{
    STARSH_SPATIAL_NDIM = 1,
    STARSH_SPATIAL_BETA = 2,
    STARSH_SPATIAL_NU = 3,
    STARSH_SPATIAL_NOISE = 4,
    STARSH_SPATIAL_PLACE = 5,
    STARSH_SPATIAL_SIGMA = 6,
    STARSH_SPATIAL_SIGMA2 = 7,
    STARSH_SPATIAL_NU2 = 8,
    STARSH_SPATIAL_CORR = 9
};

/* spatial.c:20-125 */
int starsh_ssdata_new(STARSH_ssdata **data, STARSH_int count, int ndim);
int starsh_ssdata_init(STARSH_ssdata **data, STARSH_int count, int ndim,
        double *point, double beta, double nu, double noise, double sigma);
/* spatial.c:201-259 */
int starsh_ssdata_generate(STARSH_ssdata **data, STARSH_int count, int ndim,
        double beta, double nu, double noise,
        enum STARSH_PARTICLES_PLACEMENT place, double sigma);
/* spatial.c:332-480: 0-terminated (key, value) pairs; SIGMA2/NU2/CORR
 * (parsimonious bivariate kernels) return STARSH_WRONG_PARAMETER. */
int starsh_ssdata_generate_va(STARSH_ssdata **data, STARSH_int count,
        va_list args);
int starsh_ssdata_generate_el(STARSH_ssdata **data, STARSH_int count, ...);
/* spatial.c:491-893 */
int starsh_ssdata_get_kernel(STARSH_kernel **kernel, STARSH_ssdata *data,
        enum STARSH_SPATIAL_KERNEL type);
/* spatial.c:482-489 */
void starsh_ssdata_free(STARSH_ssdata *data);

/* kernel_exp.c:31-193, kernel_sqrexp.c:31-193, kernel_matern.c:33-200,
 * kernel_matern2.c:33-200 */
void starsh_ssdata_block_exp_kernel_1d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_exp_kernel_1d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_exp_kernel_2d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_exp_kernel_2d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_exp_kernel_3d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_exp_kernel_3d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_exp_kernel_4d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_exp_kernel_4d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_exp_kernel_nd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_exp_kernel_nd_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_1d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_1d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_2d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_2d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_3d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_3d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_4d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_4d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_nd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_sqrexp_kernel_nd_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_1d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_1d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_2d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_2d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_3d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_3d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_4d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_4d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_nd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern_kernel_nd_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_1d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_1d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_2d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_2d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_3d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_3d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_4d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_4d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_nd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_ssdata_block_matern2_kernel_nd_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);

#ifdef __cplusplus
}
#endif

#endif
