// Internal interface between the pipeline stages of the HIP backend.
#pragma once

#include "hip_common.h"

enum GemmMode
{
    GM_SKETCH_GEN = 0,
    GM_SKETCH_MEM,
    GM_PROJ_GEN,
    GM_PROJ_MEM,
    GM_COEF,
    GM_UPDATE,
    GM_R2,
    GM_FORM_U,
    GM_FORM_V
};

struct GemmArgs
{
    double *ws;           // batch workspace base
    double *out;          // output base for U / V
    const double *abase;  // explicit tiles base (dense mode)
    KernelParams kp;
    PointSets ps;
    int j0, jb;           // panel position (COEF / UPDATE)
    int which;            // 0: Q (m rows), 1: Q2 (n rows)
};

int pipeline_gemm(int mode, const TileDesc *d_tiles, int ntiles, int mmax,
        int nmax, const GemmArgs &g, hipStream_t s);

// Column norms of the m x k buffers (which = 0) or of Bt (which = 2) into
// d_norm (ntiles x kmax).
int pipeline_colnorm(const TileDesc *d_tiles, int ntiles, int which,
        double *ws, double *d_norm, int kmax, hipStream_t s);

// One block classical Gram-Schmidt panel pass with intra-panel CGS2 and
// deflation.  which: 0 = Q (width k), 1 = Q2 (width p).  Columns whose
// norm after projection is at most tau * ref (ref = d_ref[t*kmax+j], or 1
// when d_ref is null) are set to zero and flagged in d_keep.
int pipeline_panel(const TileDesc *d_tiles, int ntiles, int which,
        double *ws, int j0, int jb, const double *d_ref, int kmax,
        double tau, int *d_keep, int rows_max, hipStream_t s);

// Moves kept columns of Q to the front; sets kc.
int pipeline_compact(TileDesc *d_tiles, int ntiles, double *ws,
        const int *d_keep, int kmax, hipStream_t s);

// From column norms of Bt: |B|_F^2, truncation p and tail; copies
// Bt[:, :p] into the Q2 buffer.
int pipeline_truncate(TileDesc *d_tiles, int ntiles, double *ws,
        const double *d_bnorm, int kmax, double eta_tol, int nmax,
        hipStream_t s);

// Integer test matrix (n x k bytes) for the int8 sketch.
int pipeline_omega_i8(signed char *d_out, int ncols, int k, hipStream_t s);

// Row scales and Y = A Omega with int8 digit planes (mixed form).
int pipeline_sketch_i8(const TileDesc *d_tiles, int ntiles, int mmax,
        int kmax, const GemmArgs &g, const signed char *d_om8,
        hipStream_t s);

// One-sided Jacobi SVD of R2, singular values, rank by the reference rule,
// W_r and J_r.
int pipeline_core(TileDesc *d_tiles, int ntiles, double *ws, int maxrank,
        double tol, int pmax, hipStream_t s);
