// Stages of the exact engines (SVD and rank-revealing QR, the reference's
// starsh_dense_dlrsdd and starsh_dense_dlrqp3, src/backends/sequential/
// dense/dsdd.c:16-66 and dqp3.c:16-93) that the randomized pipeline does
// not have:
//
//   select_kernel  the sketch Y = A[:, sel]: the tile's columns chosen by a
//                  pivoted QR, generated from the points (or copied from an
//                  explicit tile), in place of Y = A Omega;
//   qrcp_kernel    the column pivots of B = Q^T A when Q spans the range of
//                  A: B's columns are A's in the orthonormal coordinates of
//                  Q, so the residual norms that steer dgeqp3's pivoting,
//                  and hence its pivot order, are those of B's columns.
//
// With these the engines are the pipeline run with particular inputs: SVD
// is the randomized SVD whose sketch spans the whole range of the tile
// (the range is exhausted, or the sketch is min(m, n) wide); RRQR is the
// randomized SVD of the leading maxrank+oversample pivot columns, whose
// orthonormal basis Q1 gives B1 = Q1^T A = R[:mn2, :] P^T of dgeqp3 up to
// the orthogonal factor, so the same singular values and factors.
#include <float.h>
#include "pipeline.h"

namespace {

template <int NDIM>
__global__ __launch_bounds__(256) void select_kernel(const TileDesc *tiles,
        double *ws, const int *sel, const double *abase)
{
    const TileDesc &t = tiles[blockIdx.y];
    const int m = t.m, k = t.k;
    const int pm = (m+63)/64;
    const int bi = blockIdx.x % pm, bj = blockIdx.x / pm;
    if(bj*64 >= k)
        return;
    const int i = bi*64+(threadIdx.x & 63);
    if(i >= m)
        return;
    const int *cols = sel+t.sel_off;
    double *Y = ws+t.q_off;
    for(int jj = threadIdx.x >> 6; jj < 64; jj += 4)
    {
        const int j = bj*64+jj;
        if(j >= k)
            break;
        const int c = cols[j];
        Y[(int64_t)j*m+i] = t.gen ? kernel_value<NDIM>(t.kp, t.xr, t.nr,
                t.r0+i, t.xc, t.nc, t.c0+c) :
            abase[t.a_off+(int64_t)c*m+i];
    }
}

constexpr int QRCP_THREADS = 1024;

// Block-wide argmax of v[0 .. n-1] (first index of the largest value);
// -1 when every value is negative.  red: 2 * (QRCP_THREADS / 64) slots.
__device__ int block_argmax(const double *v, int n, double *redv, int *redi)
{
    double best = -1.0;
    int bi = -1;
    for(int i = threadIdx.x; i < n; i += blockDim.x)
        if(v[i] > best)
        {
            best = v[i];
            bi = i;
        }
    // within the wave: larger value, then smaller index
    for(int o = 32; o > 0; o >>= 1)
    {
        const double ob = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if(ob > best || (ob == best && oi >= 0 && (bi < 0 || oi < bi)))
        {
            best = ob;
            bi = oi;
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if(lane == 0)
    {
        redv[wave] = best;
        redi[wave] = bi;
    }
    __syncthreads();
    if(threadIdx.x == 0)
    {
        double b = -1.0;
        int x = -1;
        for(int w = 0; w < (int)blockDim.x/64; w++)
            if(redv[w] > b || (redv[w] == b && redi[w] >= 0 &&
                        (x < 0 || redi[w] < x)))
            {
                b = redv[w];
                x = redi[w];
            }
        redi[QRCP_THREADS/64] = x;
    }
    __syncthreads();
    const int r = redi[QRCP_THREADS/64];
    __syncthreads();
    return r;
}

// Column-pivoted QR of B (kc x n), held as Bt (n x kc, ld n), for the
// tiles with kc > piv_k: piv_k steps of modified Gram-Schmidt on the rows
// of Bt (B's columns), each step taking the row of largest residual norm
// (the first one on ties), as dgeqp3 does (LAPACK dlaqp2: pivot by the
// largest partial column norm; norms downdated by
//   vn1 <- vn1 sqrt(max(0, 1 - (|r| / vn1)^2)),
// recomputed from the residual when the downdate has lost more than half
// the digits, (1 - (|r|/vn1)^2) (vn1/vn2)^2 <= sqrt(eps)).  Bt is
// overwritten with the residuals.  LDS: vn1[n], vn2[n], q[kc].
__global__ __launch_bounds__(QRCP_THREADS) void qrcp_kernel(TileDesc *tiles,
        double *ws, int *piv)
{
    extern __shared__ double dyn[];
    __shared__ double redv[QRCP_THREADS/64];
    __shared__ int redi[QRCP_THREADS/64+1];
    const TileDesc &td = tiles[blockIdx.x];
    const int n = td.n, kc = td.kc, steps = td.piv_k;
    if(td.piv_off < 0 || steps <= 0 || kc <= steps)
        return;
    double *vn1 = dyn, *vn2 = dyn+n, *q = dyn+2*n;
    double *Bt = ws+td.bt_off;
    int *out = piv+td.piv_off;
    const double tol3z = sqrt(DBL_EPSILON);
    for(int i = threadIdx.x; i < n; i += blockDim.x)
    {
        double s = 0.0;
        for(int j = 0; j < kc; j++)
        {
            const double x = Bt[(int64_t)j*n+i];
            s += x*x;
        }
        vn1[i] = sqrt(s);
        vn2[i] = vn1[i];
    }
    __syncthreads();
    for(int step = 0; step < steps; step++)
    {
        int p = block_argmax(vn1, n, redv, redi);
        if(p < 0)
        {
            // every column chosen already (steps <= n)
            if(threadIdx.x == 0)
                out[step] = -1;
            continue;
        }
        if(threadIdx.x == 0)
            out[step] = p;
        // direction: the residual of column p, normalised (a zero residual
        // leaves q = 0 and the other columns unchanged)
        double s = 0.0;
        for(int j = threadIdx.x; j < kc; j += blockDim.x)
        {
            const double x = Bt[(int64_t)j*n+p];
            q[j] = x;
            s += x*x;
        }
        s = group_sum_lanes(s, 64);
        if((threadIdx.x & 63) == 0)
            redv[threadIdx.x >> 6] = s;
        __syncthreads();
        double nn = 0.0;
        for(int w = 0; w < (int)blockDim.x/64; w++)
            nn += redv[w];
        const double inv = nn > 0.0 ? 1.0/sqrt(nn) : 0.0;
        for(int j = threadIdx.x; j < kc; j += blockDim.x)
            q[j] *= inv;
        __syncthreads();
        if(threadIdx.x == 0)
        {
            vn1[p] = -1.0;     // taken
            vn2[p] = -1.0;
        }
        // remove the direction from every other column; downdate norms
        for(int i = threadIdx.x; i < n; i += blockDim.x)
        {
            if(i == p || vn1[i] < 0.0)
                continue;
            double c = 0.0;
            for(int j = 0; j < kc; j++)
                c += Bt[(int64_t)j*n+i]*q[j];
            double ss = 0.0;
            for(int j = 0; j < kc; j++)
            {
                const double x = Bt[(int64_t)j*n+i]-c*q[j];
                Bt[(int64_t)j*n+i] = x;
                ss += x*x;
            }
            if(vn1[i] > 0.0)
            {
                const double r = fabs(c)/vn1[i];
                double temp = 1.0-r*r;
                temp = temp > 0.0 ? temp : 0.0;
                const double ratio = vn1[i]/vn2[i];
                if(temp*ratio*ratio <= tol3z)
                {
                    vn1[i] = sqrt(ss);
                    vn2[i] = vn1[i];
                }
                else
                    vn1[i] *= sqrt(temp);
            }
        }
        __syncthreads();
    }
}

} // namespace

int pipeline_select(const TileDesc *d_tiles, int ntiles, int mmax, int kmax,
        int ndim, double *ws, const int *d_sel, const double *abase,
        hipStream_t s)
{
    if(ntiles <= 0)
        return STARSH_SUCCESS;
    const int pm = (mmax+63)/64, pk = (kmax+63)/64;
    const dim3 grid(pm*pk, ntiles);
    switch(ndim)
    {
        case 1: hipLaunchKernelGGL(select_kernel<1>, grid, dim3(256), 0, s,
                        d_tiles, ws, d_sel, abase); break;
        case 2: hipLaunchKernelGGL(select_kernel<2>, grid, dim3(256), 0, s,
                        d_tiles, ws, d_sel, abase); break;
        case 3: hipLaunchKernelGGL(select_kernel<3>, grid, dim3(256), 0, s,
                        d_tiles, ws, d_sel, abase); break;
        default: hipLaunchKernelGGL(select_kernel<4>, grid, dim3(256), 0, s,
                        d_tiles, ws, d_sel, abase); break;
    }
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int pipeline_qrcp(TileDesc *d_tiles, int ntiles, int nmax, int kmax,
        double *ws, int *d_piv, hipStream_t s)
{
    if(ntiles <= 0)
        return STARSH_SUCCESS;
    const size_t bytes = sizeof(double)*(2*(size_t)nmax+(size_t)kmax);
    if(bytes > 160*1024)
        return STARSH_WRONG_PARAMETER;
    HIP_TRY(hipFuncSetAttribute((const void *)qrcp_kernel,
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    hipLaunchKernelGGL(qrcp_kernel, dim3(ntiles), dim3(QRCP_THREADS), bytes,
            s, d_tiles, ws, d_piv);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}
