// Approximation error of compressed tiles on the GPU: for every tile the
// entries are regenerated in FP64 and |A|_F^2 and |A - U V^T|_F^2 are
// accumulated (src/backends/openmp/blrm/dfe.c:49-90 computes the same two
// norms per tile with dnrm2).  Partial sums are reduced in a fixed order,
// so results do not depend on scheduling.
#include "hip_common.h"

namespace {

constexpr int RB = 32;

struct ErrTile
{
    int64_t r0, c0, uoff, voff;
    int m, n, rank, pad;
};

template <int NDIM>
__global__ __launch_bounds__(256) void dfe_kernel(KernelParams kp,
        PointSets ps, const ErrTile *__restrict__ tiles, const double *U,
        const double *V, double *part, int nblk)
{
    __shared__ double Us[RB][64+1];
    __shared__ double Vs[RB][64+1];
    __shared__ double red[2][4];
    const ErrTile t = tiles[blockIdx.y];
    double s_a = 0.0, s_d = 0.0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int pm = (t.m+63)/64;
    bool active = t.m > 0 && t.n > 0;
    int bi = 0, bj = 0;
    if(active)
    {
        bi = blockIdx.x % pm;
        bj = blockIdx.x / pm;
        active = bj*64 < t.n;
    }
    if(active)
    {
        const int i = bi*64+lane;
        double acc[16];
#pragma unroll
        for(int q = 0; q < 16; q++)
            acc[q] = 0.0;
        for(int q0 = 0; q0 < t.rank; q0 += RB)
        {
            for(int e = threadIdx.x; e < RB*64; e += 256)
            {
                int qq = e / 64, rr = e % 64;
                int gi = bi*64+rr, gj = bj*64+rr, gq = q0+qq;
                Us[qq][rr] = (gq < t.rank && gi < t.m) ?
                    U[t.uoff+(int64_t)gq*t.m+gi] : 0.0;
                Vs[qq][rr] = (gq < t.rank && gj < t.n) ?
                    V[t.voff+(int64_t)gq*t.n+gj] : 0.0;
            }
            __syncthreads();
            for(int qq = 0; qq < RB; qq++)
            {
                double u = Us[qq][lane];
#pragma unroll
                for(int c = 0; c < 16; c++)
                    acc[c] += u*Vs[qq][wave+4*c];
            }
            __syncthreads();
        }
        if(i < t.m)
#pragma unroll
            for(int c = 0; c < 16; c++)
            {
                const int j = bj*64+wave+4*c;
                if(j < t.n)
                {
                    double a = kernel_value<NDIM>(kp, ps.xr, ps.nr, t.r0+i,
                            ps.xc, ps.nc, t.c0+j);
                    double d = a-acc[c];
                    s_a += a*a;
                    s_d += d*d;
                }
            }
    }
#pragma unroll
    for(int o = 32; o > 0; o >>= 1)
    {
        s_a += __shfl_xor(s_a, o, 64);
        s_d += __shfl_xor(s_d, o, 64);
    }
    if(lane == 0)
    {
        red[0][wave] = s_a;
        red[1][wave] = s_d;
    }
    __syncthreads();
    if(threadIdx.x == 0)
    {
        double *p = part+2*((int64_t)blockIdx.y*nblk+blockIdx.x);
        p[0] = red[0][0]+red[0][1]+red[0][2]+red[0][3];
        p[1] = red[1][0]+red[1][1]+red[1][2]+red[1][3];
    }
}

__global__ void dfe_reduce(const double *part, int nblk, double *out)
{
    if(threadIdx.x != 0)
        return;
    const double *p = part+2*(int64_t)blockIdx.x*nblk;
    double a = 0.0, d = 0.0;
    for(int b = 0; b < nblk; b++)
    {
        a += p[2*b];
        d += p[2*b+1];
    }
    out[2*blockIdx.x] = a;
    out[2*blockIdx.x+1] = d;
}

} // namespace

// Per tile: norm2[t] = |A_t|_F^2 and diff2[t] = |A_t - U_t V_t^T|_F^2 with
// U_t at d_U + uoff[t] (m x rank, ld m) and V_t at d_V + voff[t].
extern "C" int starsh_hip_tile_errors(const starsh_hip_kernel_desc *kd,
        const double *d_xr, int64_t nr, const double *d_xc, int64_t nc,
        int ntiles, const int64_t *r0, const int64_t *c0, const int *m,
        const int *n, const int *rank, const double *d_U,
        const int64_t *uoff, const double *d_V, const int64_t *voff,
        double *norm2, double *diff2)
{
    if(ntiles <= 0)
        return STARSH_SUCCESS;
    if(kd == nullptr || kd->ndim < 1 || kd->ndim > 4)
        return STARSH_WRONG_PARAMETER;
    hipStream_t s = starsh_hip_stream();
    ErrTile *h = (ErrTile *)malloc(sizeof(ErrTile)*ntiles);
    double *hout = (double *)malloc(sizeof(double)*2*ntiles);
    if(h == nullptr || hout == nullptr)
    {
        free(h);
        free(hout);
        return STARSH_MALLOC_ERROR;
    }
    int mmax = 0, nmax = 0;
    for(int t = 0; t < ntiles; t++)
    {
        if(r0[t] < 0 || c0[t] < 0 || r0[t]+m[t] > nr || c0[t]+n[t] > nc ||
                rank[t] < 0)
        {
            free(h);
            free(hout);
            return STARSH_WRONG_PARAMETER;
        }
        h[t] = ErrTile{r0[t], c0[t], uoff ? uoff[t] : 0, voff ? voff[t] : 0,
            m[t], n[t], rank[t], 0};
        mmax = m[t] > mmax ? m[t] : mmax;
        nmax = n[t] > nmax ? n[t] : nmax;
    }
    const int nblk = ((mmax+63)/64)*((nmax+63)/64);
    ErrTile *d_t = nullptr;
    double *d_part = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc((void **)&d_t, sizeof(ErrTile)*ntiles);
    if(e == hipSuccess)
        e = hipMalloc((void **)&d_part, sizeof(double)*2*(size_t)ntiles*
                (nblk > 0 ? nblk : 1));
    if(e == hipSuccess)
        e = hipMalloc((void **)&d_out, sizeof(double)*2*ntiles);
    if(e == hipSuccess)
        e = hipMemcpyAsync(d_t, h, sizeof(ErrTile)*ntiles,
                hipMemcpyHostToDevice, s);
    if(e == hipSuccess && nblk > 0)
    {
        KernelParams kp = make_params(kd);
        PointSets ps{d_xr, nr, d_xc, nc};
        dim3 grid(nblk, ntiles);
        switch(kd->ndim)
        {
            case 1: hipLaunchKernelGGL(dfe_kernel<1>, grid, dim3(256), 0, s,
                        kp, ps, d_t, d_U, d_V, d_part, nblk); break;
            case 2: hipLaunchKernelGGL(dfe_kernel<2>, grid, dim3(256), 0, s,
                        kp, ps, d_t, d_U, d_V, d_part, nblk); break;
            case 3: hipLaunchKernelGGL(dfe_kernel<3>, grid, dim3(256), 0, s,
                        kp, ps, d_t, d_U, d_V, d_part, nblk); break;
            default: hipLaunchKernelGGL(dfe_kernel<4>, grid, dim3(256), 0, s,
                        kp, ps, d_t, d_U, d_V, d_part, nblk); break;
        }
        e = hipGetLastError();
        if(e == hipSuccess)
        {
            hipLaunchKernelGGL(dfe_reduce, dim3(ntiles), dim3(64), 0, s,
                    d_part, nblk, d_out);
            e = hipGetLastError();
        }
        if(e == hipSuccess)
            e = hipMemcpyAsync(hout, d_out, sizeof(double)*2*ntiles,
                    hipMemcpyDeviceToHost, s);
    }
    hipError_t e2 = hipStreamSynchronize(s);
    (void)hipFree(d_t);
    (void)hipFree(d_part);
    (void)hipFree(d_out);
    if(e == hipSuccess && e2 == hipSuccess)
        for(int t = 0; t < ntiles; t++)
        {
            if(norm2)
                norm2[t] = nblk > 0 ? hout[2*t] : 0.0;
            if(diff2)
                diff2[t] = nblk > 0 ? hout[2*t+1] : 0.0;
        }
    free(h);
    free(hout);
    HIP_TRY(e);
    HIP_TRY(e2);
    return STARSH_SUCCESS;
}
