/* STARS-H C API, MI355X build.
 *
 * The declarations below are the subset of the reference API that the tile
 * low-rank compression path uses.  Names, argument meaning, return codes and
 * structure layouts follow ecrc/stars-h 0.3.1 include/starsh.h so that code
 * written against the reference compiles and links against this library
 * unchanged.  Each group cites the reference lines it replaces.
 *
 * Backends.  The reference's CPU and StarPU/KBLAS backends are not part of
 * this library: STARSH_BACKEND=HIP selects starsh_blrm__drsdd_hip (plain FP64
 * form) and STARSH_BACKEND=HIP_MIXED selects starsh_blrm__drsdd_hip_mixed
 * (batched mixed-precision form).  Both run on one MI355X and are declared
 * in starsh-hip.h. */
#ifndef STARSH_H
#define STARSH_H

#include <sys/types.h>
#include <stdarg.h>
#include "starsh-constants.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Environment (reference include/starsh.h:37-52, src/control/init.c) */

struct starsh_params
{
    enum STARSH_BACKEND backend;
    enum STARSH_LRENGINE lrengine;
    int oversample;
};

extern struct starsh_params starsh_params;

/* Replaces init.c:26-67: reads STARSH_BACKEND, STARSH_LRENGINE and
 * STARSH_OVERSAMPLE; falls back to HIP / RSVD / 10. */
int starsh_init(void);
/* Replaces init.c:69-109. */
int starsh_set_backend(const char *string);
/* Replaces init.c:111-147. */
int starsh_set_lrengine(const char *string);
/* Replaces init.c:149-175. */
int starsh_set_oversample(const char *string);

/* ---- Types (reference include/starsh.h:63-93) */

typedef ssize_t STARSH_int;

typedef void STARSH_kernel(int nrows, int ncols, STARSH_int *irow,
        STARSH_int *icol, void *row_data, void *col_data, void *result,
        int ld);

typedef struct array Array;
typedef struct starsh_problem STARSH_problem;
typedef struct starsh_cluster STARSH_cluster;
typedef struct starsh_blrf STARSH_blrf;
typedef struct starsh_blrm STARSH_blrm;

/* ---- Applications (reference include/starsh.h:103-104,
 * src/applications/common.c:22-116).  SPATIAL, ELECTROSTATICS and
 * ELECTRODYNAMICS are supported; other problem types return
 * STARSH_WRONG_PARAMETER. */
int starsh_application(void **data, STARSH_kernel **kernel, STARSH_int count,
        char dtype, int problem_type, int kernel_type, ...);

/* ---- Array (reference include/starsh.h:120-170, src/control/array.c) */

struct array
{
    int ndim;
    int *shape;
    ssize_t *stride;
    char order;
    size_t size;
    char dtype;
    size_t dtype_size;
    size_t nbytes;
    size_t data_nbytes;
    void *data;
};

int array_from_buffer(Array **A, int ndim, int *shape, char dtype, char order,
        void *buffer);
int array_new(Array **A, int ndim, int *shape, char dtype, char order);
void array_free(Array *A);
void array_info(Array *A);

/* ---- Problem (reference include/starsh.h:186-247, src/control/problem.c) */

struct starsh_problem
{
    int ndim;
    STARSH_int *shape;
    char symm;
    char dtype;
    size_t dtype_size;
    size_t entry_size;
    void *row_data;
    void *col_data;
    STARSH_kernel *kernel;
    char *name;
};

int starsh_problem_new(STARSH_problem **problem, int ndim, STARSH_int *shape,
        char symm, char dtype, void *row_data, void *col_data,
        STARSH_kernel *kernel, char *name);
void starsh_problem_free(STARSH_problem *problem);
void starsh_problem_info(STARSH_problem *problem);
int starsh_problem_get_block(STARSH_problem *problem, int nrows, int ncols,
        STARSH_int *irow, STARSH_int *icol, Array **A);
int starsh_problem_to_array(STARSH_problem *problem, Array **A);

/* ---- Cluster (reference include/starsh.h:263-329, src/control/cluster.c) */

struct starsh_cluster
{
    void *data;
    STARSH_int ndata;
    STARSH_int *pivot;
    STARSH_int nblocks;
    STARSH_int nlevels;
    STARSH_int *level;
    STARSH_int *start;
    STARSH_int *size;
    STARSH_int *parent;
    STARSH_int *child_start;
    STARSH_int *child;
    enum STARSH_CLUSTER_TYPE type;
};

int starsh_cluster_new(STARSH_cluster **cluster, void *data, STARSH_int ndata,
        STARSH_int *pivot, STARSH_int nblocks, STARSH_int nlevels,
        STARSH_int *level, STARSH_int *start, STARSH_int *size,
        STARSH_int *parent, STARSH_int *child_start, STARSH_int *child,
        enum STARSH_CLUSTER_TYPE type);
void starsh_cluster_free(STARSH_cluster *cluster);
void starsh_cluster_info(STARSH_cluster *cluster);
int starsh_cluster_new_plain(STARSH_cluster **cluster, void *data,
        STARSH_int ndata, STARSH_int block_size);

/* ---- Block low-rank format (reference include/starsh.h:345-481,
 * src/control/blrf.c) */

struct starsh_blrf
{
    STARSH_problem *problem;
    char symm;
    STARSH_cluster *row_cluster;
    STARSH_cluster *col_cluster;
    STARSH_int nbrows;
    STARSH_int nbcols;
    STARSH_int nblocks_far;
    STARSH_int *block_far;
    STARSH_int *brow_far_start;
    STARSH_int *brow_far;
    STARSH_int *bcol_far_start;
    STARSH_int *bcol_far;
    STARSH_int nblocks_far_local;
    STARSH_int *block_far_local;
    STARSH_int nblocks_near;
    STARSH_int *block_near;
    STARSH_int *brow_near_start;
    STARSH_int *brow_near;
    STARSH_int *bcol_near_start;
    STARSH_int *bcol_near;
    STARSH_int nblocks_near_local;
    STARSH_int *block_near_local;
    enum STARSH_BLRF_TYPE type;
};

int starsh_blrf_new_from_coo(STARSH_blrf **format, STARSH_problem *problem,
        char symm, STARSH_cluster *row_cluster, STARSH_cluster *col_cluster,
        STARSH_int nblocks_far, STARSH_int *block_far,
        STARSH_int nblocks_near, STARSH_int *block_near,
        enum STARSH_BLRF_TYPE type);
int starsh_blrf_new_tlr(STARSH_blrf **format, STARSH_problem *problem,
        char symm, STARSH_cluster *row_cluster, STARSH_cluster *col_cluster);
void starsh_blrf_free(STARSH_blrf *format);
void starsh_blrf_info(STARSH_blrf *format);
void starsh_blrf_print(STARSH_blrf *format);
int starsh_blrf_get_block(STARSH_blrf *format, STARSH_int i, STARSH_int j,
        int *shape, void **D);

/* ---- Block low-rank matrix (reference include/starsh.h:497-545,
 * src/control/blrm.c) */

struct starsh_blrm
{
    STARSH_blrf *format;
    int *far_rank;
    Array **far_U;
    Array **far_V;
    int onfly;
    Array **near_D;
    void *alloc_U;
    void *alloc_V;
    void *alloc_D;
    char alloc_type;
    size_t nbytes;
    size_t data_nbytes;
};

int starsh_blrm_new(STARSH_blrm **matrix, STARSH_blrf *format, int *far_rank,
        Array **far_U, Array **far_V, int onfly, Array **near_D, void *alloc_U,
        void *alloc_V, void *alloc_D, char alloc_type);
void starsh_blrm_free(STARSH_blrm *matrix);
void starsh_blrm_info(STARSH_blrm *matrix);
int starsh_blrm_get_block(STARSH_blrm *matrix, STARSH_int i, STARSH_int j,
        int *shape, int *rank, void **U, void **V, void **D);

/* ---- Approximation plugin (reference include/starsh.h:563-567).
 * starsh_init() points it at the HIP driver selected by STARSH_BACKEND. */
typedef int STARSH_blrm_approximate(STARSH_blrm **matrix, STARSH_blrf *format,
        int maxrank, double tol, int onfly);
extern STARSH_blrm_approximate *starsh_blrm_approximate;

/* ---- Error measure, product and solver under the reference's own names
 * (include/starsh.h:601-604, 622-623, 671-672), so that reference callers
 * such as testing/spatial.c:101 link unchanged.  In this library they run
 * on the GPU: starsh_blrm__dfe and starsh_blrm__dfe_omp are
 * starsh_blrm__dfe_hip (returns -1 on failure, as openmp/blrm/dfe.c:91-93);
 * starsh_blrm__dmml and starsh_blrm__dmml_omp are starsh_blrm__dmml_hip;
 * starsh_itersolvers__dcg_omp is starsh_itersolvers__dcg_hip. */
double starsh_blrm__dfe(STARSH_blrm *matrix);
double starsh_blrm__dfe_omp(STARSH_blrm *matrix);
int starsh_blrm__dmml(STARSH_blrm *matrix, int nrhs, double alpha,
        double *A, int lda, double beta, double *B, int ldb);
int starsh_blrm__dmml_omp(STARSH_blrm *matrix, int nrhs, double alpha,
        double *A, int lda, double beta, double *B, int ldb);
int starsh_itersolvers__dcg_omp(STARSH_blrm *matrix, int nrhs, double *B,
        int ldb, double *X, int ldx, double tol, double *work);

/* ---- Dense helper (reference include/starsh.h:642,
 * src/backends/sequential/dense/dsvfr.c:16-47): rank from singular values
 * by relative Frobenius tail. */
int starsh_dense_dsvfr(int size, double *S, double tol);

/* The reference's dense engines for one tile (include/starsh.h:644-652,
 * src/backends/sequential/dense/{dsdd,drsdd,dqp3}.c), same signatures and
 * results, computed on the GPU: D (nrows x ncols, ld ldD) is copied to the
 * device, U (ld ldU) and V (ld ldV) receive rank columns when the tile is
 * low-rank, rank = -1 otherwise.  work, lwork and iwork are not used. */
void starsh_dense_dlrsdd(int nrows, int ncols, double *D, int ldD, double *U,
        int ldU, double *V, int ldV, int *rank, int maxrank, double tol,
        double *work, int lwork, int *iwork);
void starsh_dense_dlrrsdd(int nrows, int ncols, double *D, int ldD, double *U,
        int ldU, double *V, int ldV, int *rank, int maxrank, int oversample,
        double tol, double *work, int lwork, int *iwork);
void starsh_dense_dlrqp3(int nrows, int ncols, double *D, int ldD, double *U,
        int ldU, double *V, int ldV, int *rank, int maxrank, int oversample,
        double tol, double *work, int lwork, int *iwork);

#ifdef __cplusplus
}
#endif

#include "starsh-hip.h"

#endif
