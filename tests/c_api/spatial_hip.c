/* A C caller of the library written against the reference API only: the
 * call sequence of the reference's spatial-statistics test
 * (testing/spatial.c:55-109 of ecrc/stars-h 0.3.1) -- starsh_init,
 * starsh_application, starsh_problem_new, starsh_cluster_new_plain,
 * starsh_blrf_new_tlr, starsh_blrm_approximate, starsh_blrm__dfe_omp -- and
 * its acceptance rule rel_err / tol <= 10.  The backend comes from
 * STARSH_BACKEND (HIP or HIP_MIXED) as in the reference.  It includes only
 * the reference's public headers and links with -lstarsh_hip.
 *
 * usage: spatial_hip ndim placement kernel beta nu N block_size maxrank tol
 * prints "RANKS min avg max", "TIME TO APPROXIMATE", "RELATIVE ERROR". */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <starsh.h>
#include <starsh-spatial.h>

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec+1e-9*ts.tv_nsec;
}

int main(int argc, char **argv)
{
    if(argc != 10)
    {
        printf("%d arguments provided, but 9 are needed\n", argc-1);
        printf("spatial_hip ndim placement kernel beta nu N block_size "
                "maxrank tol\n");
        return 1;
    }
    int problem_ndim = atoi(argv[1]);
    int place = atoi(argv[2]);
    int kernel_type = atoi(argv[3]);
    double beta = atof(argv[4]);
    double nu = atof(argv[5]);
    int N = atoi(argv[6]);
    int block_size = atoi(argv[7]);
    int maxrank = atoi(argv[8]);
    double tol = atof(argv[9]);
    char symm = 'N', dtype = 'd';
    STARSH_int shape[2] = {N, N};
    srand(0);
    int info = starsh_init();
    if(info != 0)
        return info;
    STARSH_ssdata *data;
    STARSH_kernel *kernel;
    info = starsh_application((void **)&data, &kernel, N, dtype,
            STARSH_SPATIAL, kernel_type, STARSH_SPATIAL_NDIM, problem_ndim,
            STARSH_SPATIAL_BETA, beta, STARSH_SPATIAL_NU, nu,
            STARSH_SPATIAL_NOISE, 0.0, STARSH_SPATIAL_PLACE, place, 0);
    if(info != 0)
    {
        printf("Problem was NOT generated (wrong parameters)\n");
        return info;
    }
    STARSH_problem *P;
    info = starsh_problem_new(&P, 2, shape, symm, dtype, data, data, kernel,
            "Spatial Statistics example");
    if(info != 0)
        return info;
    starsh_problem_info(P);
    STARSH_cluster *C;
    info = starsh_cluster_new_plain(&C, data, N, block_size);
    if(info != 0)
        return info;
    starsh_cluster_info(C);
    STARSH_blrf *F;
    STARSH_blrm *M;
    info = starsh_blrf_new_tlr(&F, P, symm, C, C);
    if(info != 0)
        return info;
    starsh_blrf_info(F);
    double t = now();
    info = starsh_blrm_approximate(&M, F, maxrank, tol, 0);
    if(info != 0)
        return info;
    t = now()-t;
    starsh_blrf_info(F);
    starsh_blrm_info(M);
    int rmin = maxrank, rmax = 0;
    double rsum = 0;
    for(STARSH_int bi = 0; bi < F->nblocks_far; bi++)
    {
        int r = M->far_rank[bi];
        rmin = r < rmin ? r : rmin;
        rmax = r > rmax ? r : rmax;
        rsum += r;
    }
    printf("RANKS %d %.3f %d\n", F->nblocks_far ? rmin : 0,
            F->nblocks_far ? rsum/F->nblocks_far : 0.0, rmax);
    printf("TIME TO APPROXIMATE: %e secs\n", t);
    t = now();
    double rel_err = starsh_blrm__dfe_omp(M);
    t = now()-t;
    printf("TIME TO MEASURE ERROR: %e secs\nRELATIVE ERROR: %e\n", t,
            rel_err);
    if(!(rel_err >= 0) || rel_err/tol > 10.)
    {
        printf("Resulting relative error is too big\n");
        return 1;
    }
    /* the product through the reference name as well: y = M x */
    double *x = malloc(sizeof(double)*N), *y = calloc(N, sizeof(double));
    for(int i = 0; i < N; i++)
        x[i] = 1.0;
    info = starsh_blrm__dmml(M, 1, 1.0, x, N, 0.0, y, N);
    printf("DMML INFO %d Y0 %.17g\n", info, y[0]);
    free(x);
    free(y);
    starsh_blrm_free(M);
    starsh_blrf_free(F);
    starsh_cluster_free(C);
    starsh_problem_free(P);
    starsh_ssdata_free(data);
    return info;
}
