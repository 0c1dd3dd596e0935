/* Electrodynamics problem (reference include/starsh-electrodynamics.h:34-180,
 * src/applications/electrodynamics.c, electrodynamics/kernel_sin.c:31-105,
 * electrodynamics/kernel_cos.c:31-105).
 * Kernels: sin(k r)/r and cos(k r)/r, and `diag` where r == 0. */
#ifndef STARSH_ELECTRODYNAMICS_H
#define STARSH_ELECTRODYNAMICS_H

#include "starsh.h"
#include "starsh-particles.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct starsh_eddata
{
    STARSH_particles particles;
    double k;
    double diag;
} STARSH_eddata;

enum STARSH_ELECTRODYNAMICS_KERNEL
{
    STARSH_ELECTRODYNAMICS_SIN = 1,
    STARSH_ELECTRODYNAMICS_COS = 2,
    STARSH_ELECTRODYNAMICS_SIN_SIMD = 11,
    STARSH_ELECTRODYNAMICS_COS_SIMD = 12
};

enum STARSH_ELECTRODYNAMICS_PARAM
{
    STARSH_ELECTRODYNAMICS_NDIM = 1,
    STARSH_ELECTRODYNAMICS_K = 2,
    STARSH_ELECTRODYNAMICS_DIAG = 3,
    STARSH_ELECTRODYNAMICS_PLACE = 4
};

/* electrodynamics.c:17-94.  Unlike the reference, *data is set. */
int starsh_eddata_new(STARSH_eddata **data, STARSH_int count, int ndim);
int starsh_eddata_init(STARSH_eddata **data, STARSH_int count, int ndim,
        double *point, double k, double diag);
/* electrodynamics.c:96-245 (defaults k = 1, diag = 1) */
int starsh_eddata_generate(STARSH_eddata **data, STARSH_int count, int ndim,
        double k, double diag, enum STARSH_PARTICLES_PLACEMENT place);
int starsh_eddata_generate_va(STARSH_eddata **data, STARSH_int count,
        va_list args);
int starsh_eddata_generate_el(STARSH_eddata **data, STARSH_int count, ...);
/* electrodynamics.c:247-417 */
int starsh_eddata_get_kernel(STARSH_kernel **kernel, STARSH_eddata *data,
        enum STARSH_ELECTRODYNAMICS_KERNEL type);
void starsh_eddata_free(STARSH_eddata *data);

void starsh_eddata_block_sin_kernel_1d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_sin_kernel_1d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_sin_kernel_2d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_sin_kernel_2d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_sin_kernel_3d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_sin_kernel_3d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_sin_kernel_4d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_sin_kernel_4d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_sin_kernel_nd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_sin_kernel_nd_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_1d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_1d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_2d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_2d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_3d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_3d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_4d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_4d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_nd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_eddata_block_cos_kernel_nd_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);

#ifdef __cplusplus
}
#endif

#endif
