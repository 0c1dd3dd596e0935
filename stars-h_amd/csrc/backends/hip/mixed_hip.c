/* Batched mixed-precision HIP driver (placeholder until the bucketed
 * pipeline lands): reports that the form is unavailable. */
#include "common.h"

int starsh_blrm__drsdd_hip_mixed(STARSH_blrm **matrix, STARSH_blrf *format,
        int maxrank, double tol, int onfly)
{
    (void)matrix; (void)format; (void)maxrank; (void)tol; (void)onfly;
    STARSH_ERROR("the mixed-precision form is not available in this build");
    return STARSH_WRONG_PARAMETER;
}
