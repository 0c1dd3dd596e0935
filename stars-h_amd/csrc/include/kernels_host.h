/* Internal: identification of this library's STARSH_kernel functions and
 * the host evaluation shared by all of them. */
#ifndef STARSH_KERNELS_HOST_H
#define STARSH_KERNELS_HOST_H

#include "starsh.h"
#include "starsh-hip.h"

/* Maps a kernel pointer to its family (enum STARSH_HIP_KERNEL_FAMILY) and the
 * dimensionality fixed in its name (0 for the _nd kernels, which read it from
 * the data).  Returns 0 on success, -1 for a pointer this library does not
 * own. */
int starsh_kernel_identify(STARSH_kernel *kernel, int *family, int *ndim);

/* Fills a device kernel descriptor from a recognised kernel and its row
 * data; also returns the point array and count of the data. */
int starsh_kernel_describe(STARSH_kernel *kernel, void *data,
        starsh_hip_kernel_desc *kd, const double **points, STARSH_int *count);

#endif
