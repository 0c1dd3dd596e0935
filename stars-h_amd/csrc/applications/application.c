/* starsh_application: varargs dispatcher over problem types
 * (src/applications/common.c:22-116 of ecrc/stars-h 0.3.1).  The spatial,
 * electrostatics and electrodynamics problems are supported; they ignore
 * dtype and work in double precision, as in the reference. */
#include "common.h"
#include "starsh-spatial.h"
#include "starsh-electrostatics.h"
#include "starsh-electrodynamics.h"

int starsh_application(void **data, STARSH_kernel **kernel, STARSH_int count,
        char dtype, int problem_type, int kernel_type, ...)
{
    (void)dtype;
    if(data == NULL || kernel == NULL)
        return STARSH_WRONG_PARAMETER;
    va_list args;
    va_start(args, kernel_type);
    int info;
    switch(problem_type)
    {
        case STARSH_SPATIAL:
            info = starsh_ssdata_generate_va((STARSH_ssdata **)data, count,
                    args);
            if(info == STARSH_SUCCESS)
                info = starsh_ssdata_get_kernel(kernel, *data, kernel_type);
            break;
        case STARSH_ELECTROSTATICS:
            info = starsh_esdata_generate_va((STARSH_esdata **)data, count,
                    args);
            if(info == STARSH_SUCCESS)
                info = starsh_esdata_get_kernel(kernel, *data, kernel_type);
            break;
        case STARSH_ELECTRODYNAMICS:
            info = starsh_eddata_generate_va((STARSH_eddata **)data, count,
                    args);
            if(info == STARSH_SUCCESS)
                info = starsh_eddata_get_kernel(kernel, *data, kernel_type);
            break;
        default:
            STARSH_ERROR("Wrong value of problem_type");
            info = STARSH_WRONG_PARAMETER;
    }
    va_end(args);
    return info;
}
