/* Backend and low-rank engine selection.
 *
 * Follows src/control/init.c:26-175 and include/control/init.h:14-190 of
 * ecrc/stars-h 0.3.1: STARSH_BACKEND, STARSH_LRENGINE and STARSH_OVERSAMPLE
 * select the function behind the global pointer starsh_blrm_approximate
 * from a backend x engine table.  This library implements the HIP rows
 * only; the reference's CPU/MPI/StarPU names are recognised and reported as
 * not supported, as the reference does for backends it was built without.
 * Unlike the reference (init.c:106-107 indexes the table while lrengine is
 * still NOTSELECTED), both indices are valid before every lookup. */
#include "common.h"
#include "starsh-hip.h"

STARSH_blrm_approximate *starsh_blrm_approximate = NULL;

struct starsh_params starsh_params =
{
    STARSH_BACKEND_NOTSELECTED, STARSH_LRENGINE_NOTSELECTED, -1
};

#define BACKEND_NUM 11
#define LRENGINE_NUM 5

static const struct { const char *string; enum STARSH_BACKEND backend; }
backend_table[BACKEND_NUM] =
{
    {"SEQUENTIAL", STARSH_BACKEND_NOTSUPPORTED},
    {"OPENMP", STARSH_BACKEND_NOTSUPPORTED},
    {"MPI", STARSH_BACKEND_NOTSUPPORTED},
    {"MPI_OPENMP", STARSH_BACKEND_NOTSUPPORTED},
    {"STARPU", STARSH_BACKEND_NOTSUPPORTED},
    {"MPI_STARPU", STARSH_BACKEND_NOTSUPPORTED},
    {"STARPU_KBLAS", STARSH_BACKEND_NOTSUPPORTED},
    {"STARPU_CUDA", STARSH_BACKEND_NOTSUPPORTED},
    {"MPI_STARPU_KBLAS", STARSH_BACKEND_NOTSUPPORTED},
    {"HIP", STARSH_BACKEND_HIP},
    {"HIP_MIXED", STARSH_BACKEND_HIP_MIXED},
};

static const struct { const char *string; enum STARSH_LRENGINE lrengine; }
lrengine_table[LRENGINE_NUM] =
{
    {"SVD", STARSH_LRENGINE_SVD},
    {"DCSVD", STARSH_LRENGINE_DCSVD},
    {"RRQR", STARSH_LRENGINE_RRQR},
    {"RSVD", STARSH_LRENGINE_RSVD},
    {"CROSS", STARSH_LRENGINE_CROSS},
};

/* Row order SVD, DCSVD, RRQR, RSVD, CROSS as in init.h:111-190 (SVD and
 * DCSVD both the divide-and-conquer SVD engine, CROSS the randomized SVD,
 * as in the reference's tables).  The exact engines compute in FP64 in
 * both HIP backends. */
static STARSH_blrm_approximate *dlr_hip[LRENGINE_NUM] =
{
    starsh_blrm__dsdd_hip, starsh_blrm__dsdd_hip, starsh_blrm__dqp3_hip,
    starsh_blrm__drsdd_hip, starsh_blrm__drsdd_hip
};

static STARSH_blrm_approximate *dlr_hip_mixed[LRENGINE_NUM] =
{
    starsh_blrm__dsdd_hip, starsh_blrm__dsdd_hip, starsh_blrm__dqp3_hip,
    starsh_blrm__drsdd_hip_mixed, starsh_blrm__drsdd_hip_mixed
};

static void select_pointer(void)
{
    enum STARSH_LRENGINE e = starsh_params.lrengine;
    if(e < 0)
        e = STARSH_LRENGINE_RSVD;
    if(starsh_params.backend == STARSH_BACKEND_HIP_MIXED)
        starsh_blrm_approximate = dlr_hip_mixed[e];
    else
        starsh_blrm_approximate = dlr_hip[e];
}

int starsh_init(void)
{
    if(starsh_set_backend(getenv("STARSH_BACKEND")) != STARSH_SUCCESS)
        starsh_set_backend(NULL);
    if(starsh_set_lrengine(getenv("STARSH_LRENGINE")) != STARSH_SUCCESS)
        starsh_set_lrengine(NULL);
    if(starsh_set_oversample(getenv("STARSH_OVERSAMPLE")) != STARSH_SUCCESS)
        starsh_set_oversample(NULL);
    return STARSH_SUCCESS;
}

int starsh_set_backend(const char *string)
{
    int i, selected = -1;
    if(string == NULL)
        selected = 9; /* HIP */
    else
        for(i = 0; i < BACKEND_NUM; i++)
            if(!strcmp(string, backend_table[i].string))
            {
                selected = i;
                break;
            }
    if(selected == -1)
    {
        fprintf(stderr, "Environment variable STARSH_BACKEND=%s is invalid\n",
                string);
        return STARSH_WRONG_PARAMETER;
    }
    if(backend_table[selected].backend == STARSH_BACKEND_NOTSUPPORTED)
    {
        fprintf(stderr, "Specified backend %s is not supported\n", string);
        return STARSH_WRONG_PARAMETER;
    }
    starsh_params.backend = backend_table[selected].backend;
    fprintf(stderr, "Selected backend is %s\n", backend_table[selected].string);
    select_pointer();
    return STARSH_SUCCESS;
}

int starsh_set_lrengine(const char *string)
{
    int i, selected = -1;
    if(string == NULL)
        selected = STARSH_LRENGINE_RSVD;
    else
        for(i = 0; i < LRENGINE_NUM; i++)
            if(!strcmp(string, lrengine_table[i].string))
            {
                selected = i;
                break;
            }
    if(selected == -1)
    {
        fprintf(stderr, "Environment variable STARSH_LRENGINE=%s is "
                "invalid\n", string);
        return STARSH_WRONG_PARAMETER;
    }
    starsh_params.lrengine = lrengine_table[selected].lrengine;
    fprintf(stderr, "Selected low-rank engine is %s\n",
            lrengine_table[selected].string);
    if(starsh_params.backend < 0)
        starsh_params.backend = STARSH_BACKEND_HIP;
    select_pointer();
    return STARSH_SUCCESS;
}

int starsh_set_oversample(const char *string)
{
    int value = string == NULL ? 10 : atoi(string);
    if(value <= 0)
    {
        fprintf(stderr, "Environment variable STARSH_OVERSAMPLE=%s is "
                "invalid\n", string);
        return STARSH_WRONG_PARAMETER;
    }
    fprintf(stderr, "Selected oversample size %d\n", value);
    starsh_params.oversample = value;
    return STARSH_SUCCESS;
}
