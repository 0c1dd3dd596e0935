/* Internal: pieces shared by the plain and mixed-precision HIP drivers. */
#ifndef STARSH_HIP_DRIVER_H
#define STARSH_HIP_DRIVER_H

#include "starsh.h"
#include "starsh-hip.h"

typedef struct hip_problem
{
    starsh_hip_kernel_desc kd;
    double *d_xr, *d_xc;   /* points in cluster order on the device */
    STARSH_int nr, nc;
    int oversample;
} hip_problem;

typedef struct hip_factors
{
    Array **far_U, **far_V;
    int *far_rank;
    double *alloc_U, *alloc_V;
    size_t *uoff, *voff;   /* offsets of each tile in alloc_U / alloc_V */
} hip_factors;

int hip_driver_prepare(STARSH_blrf *F, int maxrank, double tol, int onfly,
        hip_problem *hp);
void hip_driver_release(hip_problem *hp);
int hip_driver_alloc_factors(STARSH_blrf *F, int maxrank, hip_factors *hf);
void hip_driver_free_factors(STARSH_blrf *F, hip_factors *hf);
int hip_driver_finish(STARSH_blrm **matrix, STARSH_blrf *format,
        const hip_problem *hp, hip_factors *hf, int onfly,
        starsh_hip_stats *st);
void hip_driver_set_stats(const starsh_hip_stats *stats);
size_t hip_driver_workspace_budget(void);
double hip_driver_now(void);

#endif
