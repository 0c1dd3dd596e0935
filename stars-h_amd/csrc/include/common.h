/* Internal helpers shared by the host C sources: error printing and
 * allocation with STARSH_ERRNO returns, matching the behaviour of the
 * reference macros (include/common.h:68-114 in ecrc/stars-h 0.3.1). */
#ifndef STARSH_INTERNAL_COMMON_H
#define STARSH_INTERNAL_COMMON_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "starsh.h"

#define STARSH_ERROR(...) do { \
    fprintf(stderr, "STARSH ERROR: %s(): ", __func__); \
    fprintf(stderr, __VA_ARGS__); \
    fprintf(stderr, "\n"); } while(0)

#ifdef SHOW_WARNINGS
#define STARSH_WARNING(...) do { \
    fprintf(stderr, "STARSH WARNING: %s(): ", __func__); \
    fprintf(stderr, __VA_ARGS__); \
    fprintf(stderr, "\n"); } while(0)
#else
#define STARSH_WARNING(...) do { } while(0)
#endif

/* Allocate `n` items of *var or return STARSH_MALLOC_ERROR. */
#define STARSH_MALLOC(var, n) do { \
    (var) = malloc(sizeof(*(var))*(size_t)(n) + ((n) == 0)); \
    if(!(var)) { STARSH_ERROR("line %d: malloc() failed", __LINE__); \
        return STARSH_MALLOC_ERROR; } } while(0)

#define STARSH_REALLOC(var, n) do { \
    void *starsh_tmp_ = realloc((var), sizeof(*(var))*(size_t)(n) + ((n) == 0)); \
    if(!starsh_tmp_) { STARSH_ERROR("realloc() failed"); \
        return STARSH_MALLOC_ERROR; } \
    (var) = starsh_tmp_; } while(0)

#endif
