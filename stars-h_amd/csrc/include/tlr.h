/* Device plan of a tile low-rank product (shared by the C host code that
 * builds it and the HIP kernels of csrc/hip/tlr.hip). */
#ifndef STARSH_HIP_TLR_H
#define STARSH_HIP_TLR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One far-field tile: factors U (m x r, ld m) and V (n x r, ld n) at double
 * offsets of the factor buffer, its rows r0.. and columns c0.. of the
 * vectors, and its coefficients: T = V^T X_j (r x nrhs, ld r) at
 * tmp_off*nrhs of the coefficient buffer and T' = U^T X_i at
 * (tmp_off + r)*nrhs.  twin: the transposed product P = V T' is added as
 * well (symmetric format, i != j); P (n x nrhs, ld n) is at p_off*nrhs.
 * Offsets count 2r units per tile plus n per twin tile. */
typedef struct
{
    int64_t u_off, v_off, tmp_off, p_off, r0, c0;
    int m, n, rank, twin;
} tlr_tile;

/* One contribution to a block row: kind 0 = F T with F (rows x cols, ld) at
 * f_off of the factor buffer and T (cols x nrhs, ld cols) at src_off*nrhs of
 * the coefficients; kind 1 = D X[src_off..] with D (rows x cols, ld);
 * kind 2 = D^T X[src_off..] with D (cols x rows, ld); kind 3 = the rows
 * (rows x nrhs, ld) at src_off*nrhs of the coefficient buffer. */
typedef struct
{
    int64_t f_off, src_off;
    int ld, cols, kind, pad;
} tlr_entry;

/* At most 64 rows of one block row: rows rofs.. of the block, which are
 * rows row0.. of the vectors, and the block row's entries first..last-1. */
typedef struct
{
    int64_t row0;
    int rofs, nrows, first, last;
} tlr_chunk;

/* Y = alpha * M * X + beta * Y with the plan in device memory (d_tmp holds
 * the coefficients: (sum over tiles of 2 r, plus n per twin tile) x nrhs
 * doubles). */
/* Static device scratch of the allocation-free product: its size in
 * doubles, the devices it is set up for, the per-device set-up (called by
 * starsh_hip_init with the device current) and the lookup of its device
 * address on a set-up device. */
#define TLR_SCRATCH_DOUBLES ((int64_t)1 << 22)
#define TLR_MAX_DEVICES 64
int tlr_scratch_setup(int device);
int tlr_scratch(int device, double **ptr);

int starsh_hip_tlr_apply(const tlr_tile *d_tiles, int ntiles,
        const tlr_chunk *d_chunks, int nchunks, const tlr_entry *d_entries,
        const double *d_fac, double *d_tmp, int nrhs, double alpha,
        const double *d_x, int ldx, double beta, double *d_y, int ldy);

/* Seconds per starsh_hip_tlr_apply with alpha 1, beta 0 (mean over reps
 * launches after one warm-up; events on the library stream). */
int starsh_hip_tlr_time(const tlr_tile *d_tiles, int ntiles,
        const tlr_chunk *d_chunks, int nchunks, const tlr_entry *d_entries,
        const double *d_fac, double *d_tmp, int nrhs, const double *d_x,
        int ldx, double *d_y, int ldy, int reps, double *seconds);

/* Vector kernels of conjugate gradients on the library stream:
 * *d_out = a . b (d_part: 256 doubles of scratch), y += a x, y = x,
 * p = r + b p. */
int starsh_hip_vdot(int64_t n, const double *d_a, const double *d_b,
        double *d_part, double *d_out);
int starsh_hip_vaxpy(int64_t n, double a, const double *d_x, double *d_y);
int starsh_hip_vcopy(int64_t n, const double *d_x, double *d_y);
int starsh_hip_vxpby(int64_t n, const double *d_r, double b, double *d_p);

#ifdef __cplusplus
}
#endif

#endif
