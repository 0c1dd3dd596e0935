// Internal interface between the pipeline stages of the HIP backend.
#pragma once

#include "hip_common.h"

enum GemmMode
{
    GM_SKETCH_MEM = 0,
    GM_PROJ_MEM,
    GM_R2,
    GM_FORM_U,
    GM_FORM_V
};

struct GemmArgs
{
    double *ws;           // batch workspace base
    double *out;          // output base for U / V
    const double *abase;  // explicit tiles base (dense mode)
    int ndim, family;     // kernel of every tile of the batch
    int noisy;            // some problem of the batch has a nonzero noise
};

int pipeline_gemm(int mode, const TileDesc *d_tiles, int ntiles, int mmax,
        int nmax, const GemmArgs &g, hipStream_t s);

// R2 = Q2^T Bt[:, :p] with K (the n rows) split over nz workgroups per
// output block; d_part holds ntiles * nz * pmax^2 doubles of partial
// products, summed in order.  nz = 1 (or no d_part) runs unsplit.
int pipeline_r2_split(int ntiles, int pmax, int kmax_rows);
int pipeline_r2(const TileDesc *d_tiles, int ntiles, int pmax, int nz,
        double *d_part, const GemmArgs &g, hipStream_t s);

// Sketch (proj = 0) or projection (proj = 1) with the kernel tile generated
// in registers (gen_gemm.hip).
int pipeline_gen_gemm(int proj, const TileDesc *d_tiles, int ntiles,
        int mmax, int nmax, const GemmArgs &g, hipStream_t s);

// Column norms of the m x k buffers (which = 0) or of Bt (which = 2) into
// d_norm (ntiles x kmax).
int pipeline_colnorm(const TileDesc *d_tiles, int ntiles, int which,
        double *ws, double *d_norm, int kmax, hipStream_t s);

// Block Gram-Schmidt applied twice for a panel of jb columns at j0 (each
// pass: projection against the previous columns, then CGS2 inside the
// panel with deflation).  which: 0 = Q (width k), 1 = Q2 (width p; the
// panel is read from Bt and written to Q2).  Columns whose norm after
// projection is at most tau_first times their original norm in the first
// pass, or tau_second in the second, are set to zero and flagged in
// d_keep.  A tile whose whole panel was dropped in the first pass is marked
// done; *d_active counts the tiles that still have panels to process.
int pipeline_panel(const TileDesc *d_tiles, int ntiles, int which,
        double *ws, int j0, int jb, int kmax, double tau_first,
        double tau_second, int *d_keep, int rows_max, int *d_active,
        hipStream_t s);

// *d_active = number of tiles with a nonzero width (k or p).
int pipeline_count_active(const TileDesc *d_tiles, int ntiles, int which,
        int *d_active, hipStream_t s);

// Moves kept columns of Q to the front; sets kc.
int pipeline_compact(TileDesc *d_tiles, int ntiles, double *ws,
        const int *d_keep, int kmax, hipStream_t s);

// From column norms of Bt: |B|_F^2, truncation p and tail.  The tail
// bound is (eta*tol)^2 |B|_F^2 with each tile's own tol.
int pipeline_truncate(TileDesc *d_tiles, int ntiles, double *ws,
        const double *d_bnorm, int kmax, double eta, int nmax,
        hipStream_t s);

// Integer test matrix for the int8 sketch: k columns of n bytes, leading
// dimension n rounded up to 64, zero padded.
int pipeline_omega_i8(signed char *d_out, int ncols, int k, hipStream_t s);

// Row scales and Y = A Omega with int8 digit planes (mixed form).
int pipeline_sketch_i8(const TileDesc *d_tiles, int ntiles, int mmax,
        int kmax, const GemmArgs &g, const signed char *d_om8,
        hipStream_t s);

// Bytes of one tile's int8 digit planes of Q and their column scales.
int64_t pipeline_qdigits_bytes(int m, int k);

// Digit planes of the kc kept columns of Q, per tile at qd_off of d_qd.
int pipeline_qdigits(const TileDesc *d_tiles, int ntiles, int kmax,
        const double *ws, signed char *d_qd, hipStream_t s);

// Bt = A^T Q with the tile generated in registers and both operands in
// int8 digit planes (mixed form).
int pipeline_proj_i8(const TileDesc *d_tiles, int ntiles, int nmax, int kmax,
        const GemmArgs &g, const signed char *d_qd, hipStream_t s);

// One-sided Jacobi SVD of R2, singular values, rank by the reference rule
// (each tile's own maxrank and tol) and X_r.
int pipeline_core(TileDesc *d_tiles, int ntiles, double *ws, int pmax,
        hipStream_t s);

// Exact engines (exact.hip): sketch Y = A[:, sel] (RSDD_RULE_SELECT tiles,
// k columns from d_sel at each tile's sel_off), and the first piv_k column
// pivots of B = Q^T A for tiles with kc > piv_k (Bt is overwritten).
int pipeline_select(const TileDesc *d_tiles, int ntiles, int mmax, int kmax,
        int ndim, double *ws, const int *d_sel, const double *abase,
        hipStream_t s);
int pipeline_qrcp(TileDesc *d_tiles, int ntiles, int nmax, int kmax,
        double *ws, int *d_piv, hipStream_t s);
