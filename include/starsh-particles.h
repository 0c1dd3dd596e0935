/* Point sets for N-body style problems (reference
 * include/starsh-particles.h:36-121, src/applications/particles.c).
 *
 * Layout is structure-of-arrays: point[d*count + i] is coordinate d of
 * particle i.  Generators draw from the C library rand(), as the reference
 * does, so seeding with srand(seed) reproduces the reference's point sets on
 * the same libc.  starsh_particles_generate_spheres is an addition for
 * surface meshes of many small spheres. */
#ifndef STARSH_PARTICLES_H
#define STARSH_PARTICLES_H

#include <stdio.h>
#include <stdint.h>
#include "starsh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct starsh_particles
{
    STARSH_int count;
    int ndim;
    double *point;
} STARSH_particles;

enum STARSH_PARTICLES_PLACEMENT
{
    STARSH_PARTICLES_RAND = 1,
    STARSH_PARTICLES_UNIFORM = 3,
    STARSH_PARTICLES_RANDGRID = 4,
    STARSH_PARTICLES_QUASIUNIFORM1 = 6,
    STARSH_PARTICLES_QUASIUNIFORM2 = 7,
    STARSH_PARTICLES_OBSOLETE1 = -1,
    STARSH_PARTICLES_OBSOLETE2 = -2,
    STARSH_PARTICLES_OBSOLETE3 = -3,
    STARSH_PARTICLES_OBSOLETE4 = -4
};

/* particles.c:19-110 */
int starsh_particles_new(STARSH_particles **data, STARSH_int count, int ndim);
int starsh_particles_init(STARSH_particles **data, STARSH_int count, int ndim,
        double *point);
void starsh_particles_free(STARSH_particles *data);
/* particles.c:112-175; OBSOLETE placements return STARSH_WRONG_PARAMETER. */
int starsh_particles_generate(STARSH_particles **data, STARSH_int count,
        int ndim, enum STARSH_PARTICLES_PLACEMENT ptype);
/* particles.c:177-201 */
int starsh_particles_generate_rand(STARSH_particles **data, STARSH_int count,
        int ndim);
/* particles.c:203-267 */
int starsh_particles_generate_randgrid(STARSH_particles **data,
        STARSH_int count, int ndim);
/* particles.c:269-339 */
int starsh_particles_generate_uniform(STARSH_particles **data,
        STARSH_int count, int ndim);
/* particles.c:341-411 */
int starsh_particles_generate_quasiuniform1(STARSH_particles **data,
        STARSH_int count, int ndim);
/* particles.c:413-480 */
int starsh_particles_generate_quasiuniform2(STARSH_particles **data,
        STARSH_int count, int ndim);
/* No reference counterpart: nspheres spheres of the given radius with
 * centres uniform in [0,1]^3 and npoints points drawn uniformly on each
 * surface (normalised Gaussian directions), then Morton-sorted. */
int starsh_particles_generate_spheres(STARSH_particles **data,
        STARSH_int nspheres, STARSH_int npoints, double radius);
/* particles.c:789-1088 */
int starsh_particles_read_from_file(STARSH_particles **data, const char *fname,
        const enum STARSH_FILE_TYPE ftype);
int starsh_particles_read_from_file_pointer(STARSH_particles **data,
        FILE *fp, const enum STARSH_FILE_TYPE ftype);
int starsh_particles_read_from_file_pointer_ascii(STARSH_particles **data,
        FILE *fp);
int starsh_particles_read_from_file_pointer_binary(STARSH_particles **data,
        FILE *fp);
int starsh_particles_write_to_file(const STARSH_particles *data,
        const char *fname, const enum STARSH_FILE_TYPE ftype);
int starsh_particles_write_to_file_pointer(const STARSH_particles *data,
        FILE *fp, const enum STARSH_FILE_TYPE ftype);
int starsh_particles_write_to_file_pointer_ascii(const STARSH_particles *data,
        FILE *fp);
int starsh_particles_write_to_file_pointer_binary(const STARSH_particles *data,
        FILE *fp);
/* particles.c:1090-1221: Morton (Z-order) sort in place. */
int starsh_particles_zsort_inplace(STARSH_particles *data);
/* The same sort on the GPU (no counterpart in the reference): identical
 * permutation, computed by starsh_hip_zsort. */
int starsh_particles_zsort_inplace_hip(STARSH_particles *data);

#ifdef __cplusplus
}
#endif

#endif
