// Product of a compressed (tile low-rank) matrix with a dense block of
// vectors on the GPU, Y = alpha * M * X + beta * Y, and the vector kernels
// of conjugate gradients.  Follows src/backends/sequential/blrm/dmml.c:16-145
// of ecrc/stars-h 0.3.1: every far-field tile (i, j) adds U (V^T X_j) to Y_i
// and, in a symmetric format with i != j, V (U^T X_i) to Y_j; near-field
// tiles add D X_j (and D^T X_i).
//
// Two launches, both deterministic (no atomics):
//  1. tlr_coef_kernel, one workgroup per far-field tile: the r x nrhs
//     coefficients T = V^T X_j; for a tile with a twin (symmetric format,
//     i != j) it first forms T' = U^T X_i (one read of U), then takes T and
//     the twin's whole contribution P = V T' (n x nrhs) from one read of V;
//  2. tlr_rows_kernel: per chunk of 64 rows of one block row, the sum of all
//     contributions to those rows (U T, P, D X_j, D^T X_i) in a fixed
//     order: lane = row, wave w takes entries w, w+4, ... of the block row,
//     and the four partial sums are added in LDS in wave order.
// A twin tile's U is read twice (for T' and for U T) and its V once, 1.5
// reads per factor instead of 2 for coefficient and row passes that each
// read both factors: a factor of a 1600-row tile does not stay on a CU
// between two uses.
#include <stdio.h>
#include "hip_common.h"
#include "tlr.h"

namespace {

constexpr int RHS_GROUP = 4;
constexpr int COL_BLOCK = 16;

// One pass over a factor F (len x r, ld len) for right-hand sides q0 ..
// q0+nq-1 (nq <= NQ, NQ = 1 or RHS_GROUP): column sums
// S[c][q] = sum_l F[l][c] X[l][q] into S + c + q r, and with Tp (r x nrhs,
// ld r) also the rows P[l][q] = sum_c F[l][c] Tp[c][q] into P + l + q len.
// Columns go in blocks of 16 (8) with 64-row chunks of the factor striped over
// the four waves (lane = row); a block's column sums are reduced over the
// lanes by shuffles and over the waves in LDS in wave order, and each row
// of P is owned by one lane, which adds the blocks in column order.
template <int NQ>
__device__ __forceinline__ void factor_pass(const double *F, int len, int r,
        const double *X, int ldx, int q0, int nq, double *S, const double *Tp,
        double *P, double (*red)[COL_BLOCK][RHS_GROUP], double *tp)
{
    // columns per block: 16, or 8 with four right-hand sides (registers)
    constexpr int CB = NQ == 1 ? COL_BLOCK : COL_BLOCK/2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for(int cb = 0; cb < r; cb += CB)
    {
        const int nc = r-cb < CB ? r-cb : CB;
        if(Tp && threadIdx.x < COL_BLOCK*RHS_GROUP)
        {
            const int u = threadIdx.x % COL_BLOCK, q = threadIdx.x/COL_BLOCK;
            tp[threadIdx.x] = u < nc && q < nq ?
                Tp[cb+u+(int64_t)(q0+q)*r] : 0.0;
        }
        __syncthreads();
        double s[CB][NQ] = {};
        // one chunk at a time (two in flight measured no faster)
#pragma unroll 1
        for(int l = wave*64+lane; l-lane < len; l += 256)
        {
            const bool ok = l < len;
            // loads of the columns that exist only (unconditional loads
            // from clamped addresses measured 35% slower on 1600-row tiles
            // of rank 7-31: the extra loads cost more than the branches)
            double x[NQ], v[CB];
#pragma unroll
            for(int q = 0; q < NQ; q++)
                x[q] = ok && q < nq ? X[l+(int64_t)(q0+q)*ldx] : 0.0;
#pragma unroll
            for(int u = 0; u < CB; u++)
                v[u] = ok && u < nc ? F[(int64_t)(cb+u)*len+l] : 0.0;
#pragma unroll
            for(int u = 0; u < CB; u++)
#pragma unroll
                for(int q = 0; q < NQ; q++)
                    s[u][q] = fma(v[u], x[q], s[u][q]);
            if(Tp && ok)
#pragma unroll
                for(int q = 0; q < NQ; q++)
                    if(q < nq)
                    {
                        double p = 0.0;
#pragma unroll
                        for(int u = 0; u < CB; u++)
                            p = fma(v[u], tp[u+q*COL_BLOCK], p);
                        double *pp = P+l+(int64_t)q*len;
                        *pp = cb == 0 ? p : *pp+p;
                    }
        }
#pragma unroll
        for(int u = 0; u < CB; u++)
#pragma unroll
            for(int q = 0; q < NQ; q++)
                if(u < nc && q < nq)
                {
                    double t = s[u][q];
#pragma unroll
                    for(int o = 32; o > 0; o >>= 1)
                        t += __shfl_xor(t, o, 64);
                    s[u][q] = t;
                }
        if(lane == 0)
#pragma unroll
            for(int u = 0; u < CB; u++)
#pragma unroll
                for(int q = 0; q < NQ; q++)
                    red[wave][u][q] = s[u][q];
        __syncthreads();
        if(threadIdx.x < COL_BLOCK*RHS_GROUP)
        {
            const int u = threadIdx.x % COL_BLOCK, q = threadIdx.x/COL_BLOCK;
            if(u < nc && q < nq)
                S[cb+u+(int64_t)(q0+q)*r] = ((red[0][u][q]+red[1][u][q])+
                        red[2][u][q])+red[3][u][q];
        }
        __syncthreads();
    }
}

// one right-hand side: at most 128 VGPRs, four workgroups per CU, so that
// the 820 tiles of a 40 x 40 symmetric format run in one round
template <int NQ>
__global__ __launch_bounds__(256, NQ == 1 ? 4 : 1) void tlr_coef_kernel(
        const tlr_tile *tiles, const double *fac, const double *x, int ldx,
        int nrhs, double *tmp)
{
    __shared__ double red[4][COL_BLOCK][RHS_GROUP];
    __shared__ double tp[COL_BLOCK*RHS_GROUP];
    const tlr_tile t = tiles[blockIdx.x];
    if(t.rank <= 0)
        return;
    double *T = tmp+t.tmp_off*nrhs;            // V^T X_j, r x nrhs
    double *Tt = tmp+(t.tmp_off+t.rank)*nrhs;  // U^T X_i
    double *P = tmp+t.p_off*nrhs;              // V T', n x nrhs
    for(int q0 = 0; q0 < nrhs; q0 += NQ)
    {
        const int nq = nrhs-q0 < NQ ? nrhs-q0 : NQ;
        if(t.twin)
        {
            factor_pass<NQ>(fac+t.u_off, t.m, t.rank, x+t.r0, ldx, q0, nq,
                    Tt, nullptr, nullptr, red, tp);
            // Tt, written by other threads of this workgroup, is visible
            // after the barrier that ended the pass
            factor_pass<NQ>(fac+t.v_off, t.n, t.rank, x+t.c0, ldx, q0, nq,
                    T, Tt, P+(int64_t)q0*t.n, red, tp);
        }
        else
            factor_pass<NQ>(fac+t.v_off, t.n, t.rank, x+t.c0, ldx, q0, nq,
                    T, nullptr, nullptr, red, tp);
    }
}

template <int NQ>
__global__ __launch_bounds__(256) void tlr_rows_kernel(const tlr_chunk *chunks,
        const tlr_entry *entries, const double *fac, const double *tmp,
        const double *x, int ldx, int nrhs, double alpha, double beta,
        double *y, int ldy)
{
    __shared__ double part[4][RHS_GROUP][64];
    const tlr_chunk ch = chunks[blockIdx.x];
    // the wave index as a wave-uniform value: entries and coefficients
    // are then read through the scalar unit
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool row_ok = lane < ch.nrows;
    for(int q0 = 0; q0 < nrhs; q0 += NQ)
    {
        const int nq = nrhs-q0 < NQ ? nrhs-q0 : NQ;
        double acc[NQ] = {};
        for(int e = ch.first+wave; e < ch.last; e += 4)
        {
            const tlr_entry en = entries[e];
            if(!row_ok)
                continue;
            if(en.kind == 0)
            {
                // four independent partial sums over the factor columns
                const double *F = fac+en.f_off+ch.rofs+lane;
                const double *T = tmp+en.src_off*nrhs;
                double a4[4][NQ] = {};
                int c = 0;
                // sixteen, then eight factor columns in flight at a time
                // (column c always goes to partial sum c mod 4)
                for(; c+16 <= en.cols; c += 16)
                {
                    double f16[16];
#pragma unroll
                    for(int u = 0; u < 16; u++)
                        f16[u] = F[(int64_t)(c+u)*en.ld];
#pragma unroll
                    for(int u = 0; u < 16; u++)
#pragma unroll
                        for(int q = 0; q < NQ; q++)
                            if(NQ == 1 || q < nq)
                                a4[u & 3][q] = fma(f16[u], T[c+u+
                                        (int64_t)(q0+q)*en.cols], a4[u & 3][q]);
                }
                for(; c+8 <= en.cols; c += 8)
                {
                    double f8[8];
#pragma unroll
                    for(int u = 0; u < 8; u++)
                        f8[u] = F[(int64_t)(c+u)*en.ld];
#pragma unroll
                    for(int u = 0; u < 8; u++)
#pragma unroll
                        for(int q = 0; q < NQ; q++)
                            if(NQ == 1 || q < nq)
                                a4[u & 3][q] = fma(f8[u], T[c+u+
                                        (int64_t)(q0+q)*en.cols], a4[u & 3][q]);
                }
                for(; c+4 <= en.cols; c += 4)
#pragma unroll
                    for(int u = 0; u < 4; u++)
                    {
                        const double f = F[(int64_t)(c+u)*en.ld];
#pragma unroll
                        for(int q = 0; q < NQ; q++)
                            if(NQ == 1 || q < nq)
                                a4[u][q] = fma(f, T[c+u+(int64_t)(q0+q)*
                                        en.cols], a4[u][q]);
                    }
                for(; c < en.cols; c++)
                {
                    const double f = F[(int64_t)c*en.ld];
#pragma unroll
                    for(int q = 0; q < NQ; q++)
                        if(NQ == 1 || q < nq)
                            a4[0][q] = fma(f, T[c+(int64_t)(q0+q)*en.cols],
                                    a4[0][q]);
                }
#pragma unroll
                for(int q = 0; q < NQ; q++)
                    acc[q] += (a4[0][q]+a4[1][q])+(a4[2][q]+a4[3][q]);
            }
            else if(en.kind == 3)
            {
                // a twin's contribution V T', formed by the coefficient pass
                const double *P = tmp+en.src_off*nrhs+ch.rofs+lane;
#pragma unroll
                for(int q = 0; q < NQ; q++)
                    if(NQ == 1 || q < nq)
                        acc[q] += P[(int64_t)(q0+q)*en.ld];
            }
            else
            {
                const double *D = fac+en.f_off;
                for(int c = 0; c < en.cols; c++)
                {
                    const int r = ch.rofs+lane;
                    const double f = en.kind == 1 ?
                        D[r+(int64_t)c*en.ld] : D[c+(int64_t)r*en.ld];
#pragma unroll
                    for(int q = 0; q < NQ; q++)
                        if(NQ == 1 || q < nq)
                            acc[q] = fma(f, x[en.src_off+c+
                                    (int64_t)(q0+q)*ldx], acc[q]);
                }
            }
        }
#pragma unroll
        for(int q = 0; q < NQ; q++)
            part[wave][q][lane] = acc[q];
        __syncthreads();
        if(wave == 0 && row_ok)
        {
            for(int q = 0; q < nq; q++)
            {
                const double s = ((part[0][q][lane]+part[1][q][lane])+
                        part[2][q][lane])+part[3][q][lane];
                double *yy = y+(int64_t)(q0+q)*ldy+ch.row0+lane;
                *yy = beta == 0.0 ? alpha*s : fma(alpha, s, beta*(*yy));
            }
        }
        __syncthreads();
    }
}

// ---- vector kernels of conjugate gradients (one column of length n)

constexpr int DOT_BLOCKS = 256;

// part[b] = sum over the block's grid-stride share of a[i]*b[i]
__global__ __launch_bounds__(256) void dot_kernel(const double *a,
        const double *b, int64_t n, double *part)
{
    __shared__ double red[4];
    double s = 0.0;
    for(int64_t i = (int64_t)blockIdx.x*256+threadIdx.x; i < n;
            i += (int64_t)gridDim.x*256)
        s = fma(a[i], b[i], s);
#pragma unroll
    for(int o = 32; o > 0; o >>= 1)
        s += __shfl_xor(s, o, 64);
    if((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = s;
    __syncthreads();
    if(threadIdx.x == 0)
        part[blockIdx.x] = ((red[0]+red[1])+red[2])+red[3];
}

// out[0] = sum of part[0..np) in index order
__global__ void sum_kernel(const double *part, int np, double *out)
{
    if(threadIdx.x == 0 && blockIdx.x == 0)
    {
        double s = 0.0;
        for(int i = 0; i < np; i++)
            s += part[i];
        out[0] = s;
    }
}

// y = a*x + y
__global__ __launch_bounds__(256) void axpy_kernel(int64_t n, double a,
        const double *x, double *y)
{
    for(int64_t i = (int64_t)blockIdx.x*256+threadIdx.x; i < n;
            i += (int64_t)gridDim.x*256)
        y[i] = fma(a, x[i], y[i]);
}

// y = x
__global__ __launch_bounds__(256) void copy_kernel(int64_t n, const double *x,
        double *y)
{
    for(int64_t i = (int64_t)blockIdx.x*256+threadIdx.x; i < n;
            i += (int64_t)gridDim.x*256)
        y[i] = x[i];
}

// p = r + b*p
__global__ __launch_bounds__(256) void xpby_kernel(int64_t n,
        const double *r, double b, double *p)
{
    for(int64_t i = (int64_t)blockIdx.x*256+threadIdx.x; i < n;
            i += (int64_t)gridDim.x*256)
        p[i] = fma(b, p[i], r[i]);
}

inline unsigned vec_grid(int64_t n)
{
    const int64_t g = (n+255)/256;
    return (unsigned)(g < 1024 ? (g > 0 ? g : 1) : 1024);
}

} // namespace


// Scratch of the allocation-free product (starsh_hip_tlr_dmml_noalloc): the
// coefficient blocks V^T X_j and U^T X_i of one call, in a static device
// array that each GPU holds from the load of the library to the end of the
// process.  Host code never uses the array's name as an address: its
// device address is looked up once per device (tlr_scratch_setup, called
// by starsh_hip_init with that device current) and kept in scratch_ptr.
static __device__ __align__(256) double scratch[TLR_SCRATCH_DOUBLES];
static double *scratch_ptr[TLR_MAX_DEVICES];

extern "C" int tlr_scratch_setup(int device)
{
    if(device < 0 || device >= TLR_MAX_DEVICES)
        return STARSH_WRONG_PARAMETER;
    if(scratch_ptr[device] != nullptr)
        return STARSH_SUCCESS;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if(e != hipSuccess || p == nullptr)
    {
        fprintf(stderr, "STARSH HIP ERROR: hipGetSymbolAddress of the "
                "product scratch failed on device %d: %s\n", device,
                hipGetErrorString(e));
        return STARSH_UNKNOWN_ERROR;
    }
    scratch_ptr[device] = (double *)p;
    return STARSH_SUCCESS;
}

extern "C" int tlr_scratch(int device, double **ptr)
{
    if(device < 0 || device >= TLR_MAX_DEVICES ||
            scratch_ptr[device] == nullptr)
        return STARSH_UNKNOWN_ERROR;
    *ptr = scratch_ptr[device];
    return STARSH_SUCCESS;
}

int starsh_hip_tlr_apply(const tlr_tile *d_tiles, int ntiles,
        const tlr_chunk *d_chunks, int nchunks, const tlr_entry *d_entries,
        const double *d_fac, double *d_tmp, int nrhs, double alpha,
        const double *d_x, int ldx, double beta, double *d_y, int ldy)
{
    if(nrhs <= 0)
        return STARSH_SUCCESS;
    hipStream_t s = starsh_hip_stream();
    if(ntiles > 0)
    {
        if(nrhs == 1)
            hipLaunchKernelGGL(tlr_coef_kernel<1>, dim3(ntiles), dim3(256),
                    0, s, d_tiles, d_fac, d_x, ldx, nrhs, d_tmp);
        else
            hipLaunchKernelGGL(tlr_coef_kernel<RHS_GROUP>, dim3(ntiles),
                    dim3(256), 0, s, d_tiles, d_fac, d_x, ldx, nrhs, d_tmp);
        HIP_TRY(hipGetLastError());
    }
    if(nchunks > 0)
    {
        if(nrhs == 1)
            hipLaunchKernelGGL(tlr_rows_kernel<1>, dim3(nchunks), dim3(256),
                    0, s, d_chunks, d_entries, d_fac, d_tmp, d_x, ldx, nrhs,
                    alpha, beta, d_y, ldy);
        else
            hipLaunchKernelGGL(tlr_rows_kernel<RHS_GROUP>, dim3(nchunks),
                    dim3(256), 0, s, d_chunks, d_entries, d_fac, d_tmp, d_x,
                    ldx, nrhs, alpha, beta, d_y, ldy);
        HIP_TRY(hipGetLastError());
    }
    return STARSH_SUCCESS;
}

// *d_out = sum a[i]*b[i] (device scalar; d_part holds DOT_BLOCKS doubles)
int starsh_hip_vdot(int64_t n, const double *d_a, const double *d_b,
        double *d_part, double *d_out)
{
    hipStream_t s = starsh_hip_stream();
    hipLaunchKernelGGL(dot_kernel, dim3(DOT_BLOCKS), dim3(256), 0, s, d_a,
            d_b, n, d_part);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, s, d_part,
            DOT_BLOCKS, d_out);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int starsh_hip_vaxpy(int64_t n, double a, const double *d_x, double *d_y)
{
    hipLaunchKernelGGL(axpy_kernel, dim3(vec_grid(n)), dim3(256), 0,
            starsh_hip_stream(), n, a, d_x, d_y);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int starsh_hip_vcopy(int64_t n, const double *d_x, double *d_y)
{
    hipLaunchKernelGGL(copy_kernel, dim3(vec_grid(n)), dim3(256), 0,
            starsh_hip_stream(), n, d_x, d_y);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int starsh_hip_vxpby(int64_t n, const double *d_r, double b, double *d_p)
{
    hipLaunchKernelGGL(xpby_kernel, dim3(vec_grid(n)), dim3(256), 0,
            starsh_hip_stream(), n, d_r, b, d_p);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

// Seconds per starsh_hip_tlr_apply (mean over reps launches after one
// warm-up), measured with events on the library stream.
int starsh_hip_tlr_time(const tlr_tile *d_tiles, int ntiles,
        const tlr_chunk *d_chunks, int nchunks, const tlr_entry *d_entries,
        const double *d_fac, double *d_tmp, int nrhs, const double *d_x,
        int ldx, double *d_y, int ldy, int reps, double *seconds)
{
    hipStream_t s = starsh_hip_stream();
    int info = starsh_hip_tlr_apply(d_tiles, ntiles, d_chunks, nchunks,
            d_entries, d_fac, d_tmp, nrhs, 1.0, d_x, ldx, 0.0, d_y, ldy);
    if(info != STARSH_SUCCESS)
        return info;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s));
    for(int r = 0; r < reps && info == STARSH_SUCCESS; r++)
        info = starsh_hip_tlr_apply(d_tiles, ntiles, d_chunks, nchunks,
                d_entries, d_fac, d_tmp, nrhs, 1.0, d_x, ldx, 0.0, d_y, ldy);
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *seconds = reps > 0 ? ms*1e-3/reps : 0.0;
    return info;
}

