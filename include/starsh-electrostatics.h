/* Electrostatics problem (reference include/starsh-electrostatics.h:35-121,
 * src/applications/electrostatics.c,
 * src/applications/electrostatics/kernel_coulomb_potential.c:31-99).
 * Kernel: A[j*ld+i] = 1/r, and 0 where r == 0. */
#ifndef STARSH_ELECTROSTATICS_H
#define STARSH_ELECTROSTATICS_H

#include "starsh.h"
#include "starsh-particles.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef STARSH_particles STARSH_esdata;

enum STARSH_ELECTROSTATICS_KERNEL
{
    STARSH_ELECTROSTATICS_COULOMB_POTENTIAL = 1,
    STARSH_ELECTROSTATICS_COULOMB_POTENTIAL_SIMD = 2
};

enum STARSH_ELECTROSTATICS_PARAM
{
    STARSH_ELECTROSTATICS_NDIM = 1,
    STARSH_ELECTROSTATICS_PLACE = 2
};

/* electrostatics.c:17-90.  Unlike the reference, *data is set. */
int starsh_esdata_new(STARSH_esdata **data, STARSH_int count, int ndim);
int starsh_esdata_init(STARSH_esdata **data, STARSH_int count, int ndim,
        double *point);
/* electrostatics.c:92-219 */
int starsh_esdata_generate(STARSH_esdata **data, STARSH_int count, int ndim,
        enum STARSH_PARTICLES_PLACEMENT place);
int starsh_esdata_generate_va(STARSH_esdata **data, STARSH_int count,
        va_list args);
int starsh_esdata_generate_el(STARSH_esdata **data, STARSH_int count, ...);
/* electrostatics.c:230-359 */
int starsh_esdata_get_kernel(STARSH_kernel **kernel, STARSH_esdata *data,
        enum STARSH_ELECTROSTATICS_KERNEL type);
void starsh_esdata_free(STARSH_esdata *data);

void starsh_esdata_block_coulomb_potential_kernel_1d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_esdata_block_coulomb_potential_kernel_1d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_esdata_block_coulomb_potential_kernel_2d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_esdata_block_coulomb_potential_kernel_2d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_esdata_block_coulomb_potential_kernel_3d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_esdata_block_coulomb_potential_kernel_3d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_esdata_block_coulomb_potential_kernel_4d(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_esdata_block_coulomb_potential_kernel_4d_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_esdata_block_coulomb_potential_kernel_nd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);
void starsh_esdata_block_coulomb_potential_kernel_nd_simd(int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, void *row_data, void *col_data,
        void *result, int ld);

#ifdef __cplusplus
}
#endif

#endif
