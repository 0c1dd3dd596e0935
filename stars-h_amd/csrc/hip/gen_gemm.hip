// FP64 MFMA GEMMs whose left operand is a kernel tile generated on the fly:
//   SKETCH  Y[m x k]  = A[m x n] Omega[n x k]
//   PROJ    Bt[n x kc] = A^T[n x m] Q[m x kc]
// A never exists in memory.  Each wave owns 16 rows of the block and all BN
// columns, so every generated entry is produced exactly once per block and
// feeds BN/16 MFMAs; lane l generates A[row = l%16][k = l/16] directly in
// the A-operand layout of v_mfma_f64_16x16x4f64.  The right operand is
// staged through LDS, double buffered, one barrier per K step of 16.
#include "pipeline.h"

typedef double double4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int BM = 64;
constexpr int BK = 16;

// Column block width and 16-row groups per wave (1 or 2), 0 for the
// choice by size below; set by starsh_hip_set_option.
int gen_gemm_bn = 0;
int gen_gemm_rw = 0;
// The same, chosen by the autotuner for the current call, per operation
// ([0] sketch, [1] projection; 0 = by size); the options above win.
int gen_gemm_tuned_bn[2] = {0, 0};
int gen_gemm_tuned_rw[2] = {0, 0};

// PROJ=false: rows are row points, K runs over column points.
// PROJ=true : rows are column points, K runs over row points (A^T).
// gnbm x gnbn is the block grid of the largest tile of the launch.
template <int NDIM, int FAM, int BN, int RW, bool PROJ>
__global__ __launch_bounds__(256) void gen_gemm_kernel(const TileDesc *tiles,
        GemmArgs g, int ntiles, int gnbm, int gnbn)
{
    constexpr int NF = BN/16;
    constexpr int BMR = BM*RW;      // rows per block: RW 16-row groups/wave
    int tile, bm, bn;
    if(!map_block(blockIdx.x, ntiles, gnbm, gnbn, !PROJ, tile, bm, bn))
        return;
    const TileDesc t = tiles[tile];
    int M, N, K, ldb, ldc;
    const double *B;
    double *C;
    int64_t rbase, kbase;
    const double *xrow, *xk;
    int64_t nrow, nk;
    if(!PROJ)
    {
        M = t.m; N = t.k; K = t.n;
        B = g.ws+t.om_off; ldb = t.n;
        C = g.ws+t.q_off; ldc = t.m;
        rbase = t.r0; kbase = t.c0;
        xrow = t.xr; nrow = t.nr;
        xk = t.xc; nk = t.nc;
    }
    else
    {
        M = t.n; N = t.kc; K = t.m;
        B = g.ws+t.q_off; ldb = t.m;
        C = g.ws+t.bt_off; ldc = t.n;
        rbase = t.c0; kbase = t.r0;
        xrow = t.xc; nrow = t.nc;
        xk = t.xr; nk = t.nr;
    }
    if(M <= 0 || N <= 0)
        return;
    if(bm*BMR >= M || bn*BN >= N)
        return;
    const int row0 = bm*BMR, col0 = bn*BN;
    // 16-column fragments that hold columns of this tile (wave-uniform)
    const int nf_live = (N-col0+15)/16 < NF ? (N-col0+15)/16 : NF;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    __shared__ double Bs[2][BK][BN+2];

    // this lane's generated rows (fixed) and its K offset inside a 4-step
    const int grow0 = row0+wave*16*RW+(lane & 15);
    unsigned okmask = 0;      // bit r: row grow0 + 16 r is inside the tile
    double xi[RW][NDIM];
#pragma unroll
    for(int r = 0; r < RW; r++)
    {
        const bool ok = grow0+16*r < M;
        okmask |= ok ? 1u << r : 0u;
#pragma unroll
        for(int d = 0; d < NDIM; d++)
            xi[r][d] = ok ? xrow[d*nrow+rbase+grow0+16*r] : 0.0;
    }
    const int kl = lane >> 4;

    double4_t acc[RW][NF];
#pragma unroll
    for(int r = 0; r < RW; r++)
#pragma unroll
        for(int f = 0; f < NF; f++)
            acc[r][f] = double4_t{0.0, 0.0, 0.0, 0.0};

    // B staging: thread loads (kk = tid%16, j = tid/16 + 16 s)
    const int bk = tid & 15, bj = tid >> 4;
    double breg[NF];
    double areg[RW][4];

    auto load_b = [&](int k0) {
#pragma unroll
        for(int s = 0; s < NF; s++)
        {
            const int j = col0+bj+16*s, k = k0+bk;
            breg[s] = (j < N && k < K) ? B[(int64_t)j*ldb+k] : 0.0;
        }
    };
    auto store_b = [&](int buf) {
#pragma unroll
        for(int s = 0; s < NF; s++)
            Bs[buf][bk][bj+16*s] = breg[s];
    };
    auto gen_a = [&](int k0) {
        if constexpr(BN*RW <= 128)
        {
            // few MFMAs per entry: issue the point loads together
            double yv[4][NDIM];
#pragma unroll
            for(int q = 0; q < 4; q++)
            {
                const int k = k0+4*q+kl;
                const int64_t j = kbase+(k < K ? k : K-1);
#pragma unroll
                for(int d = 0; d < NDIM; d++)
                    yv[q][d] = xk[d*nk+j];
            }
#pragma unroll
            for(int r = 0; r < RW; r++)
#pragma unroll
                for(int q = 0; q < 4; q++)
                {
                    const int k = k0+4*q+kl;
                    const double v = kval_pt<NDIM, FAM>(t.kp, xi[r], yv[q]);
                    areg[r][q] = ((okmask >> r & 1u) && k < K) ? v : 0.0;
                }
        }
        else
        {
            // MFMA-bound: keep registers for the accumulators
#pragma unroll
            for(int q = 0; q < 4; q++)
            {
                const int k = k0+4*q+kl;
#pragma unroll
                for(int r = 0; r < RW; r++)
                {
                    const double v = kval<NDIM, FAM>(t.kp, xi[r], xk, nk,
                            kbase+(k < K ? k : K-1));
                    areg[r][q] = ((okmask >> r & 1u) && k < K) ? v : 0.0;
                }
            }
        }
    };

    load_b(0);
    store_b(0);
    gen_a(0);
    __syncthreads();
    int buf = 0;
    for(int k0 = 0; k0 < K; k0 += BK)
    {
        double acur[RW][4];
#pragma unroll
        for(int r = 0; r < RW; r++)
#pragma unroll
            for(int q = 0; q < 4; q++)
                acur[r][q] = areg[r][q];
        const bool more = k0+BK < K;
        if(more)
        {
            load_b(k0+BK);
            gen_a(k0+BK);
        }
        if constexpr(BN*RW <= 128)
        {
            // fragments past the tile's own width are skipped (uniform)
#pragma unroll
            for(int f = 0; f < NF; f++)
                if(f < nf_live)
                {
#pragma unroll
                    for(int q = 0; q < 4; q++)
                    {
                        const double b = Bs[buf][4*q+kl][f*16+(lane & 15)];
#pragma unroll
                        for(int r = 0; r < RW; r++)
                            acc[r][f] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                    acur[r][q], b, acc[r][f], 0, 0, 0);
                    }
                }
        }
        else
        {
#pragma unroll
            for(int q = 0; q < 4; q++)
            {
#pragma unroll
                for(int f = 0; f < NF; f++)
                {
                    // fragments past the tile's own width are skipped
                    // (uniform); each B value feeds RW MFMAs
                    if(!PROJ || f < nf_live)
                    {
                        const double b = Bs[buf][4*q+kl][f*16+(lane & 15)];
#pragma unroll
                        for(int r = 0; r < RW; r++)
                            acc[r][f] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                    acur[r][q], b, acc[r][f], 0, 0, 0);
                    }
                }
            }
        }
        if(more)
            store_b(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // epilogue: lane holds rows lane/16 + 4 v, column lane%16 of each frag
#pragma unroll
    for(int r = 0; r < RW; r++)
#pragma unroll
        for(int f = 0; f < NF; f++)
        {
            const int gj = col0+f*16+(lane & 15);
            if(gj >= N)
                continue;
#pragma unroll
            for(int v = 0; v < 4; v++)
            {
                const int gi = row0+wave*16*RW+r*16+(lane >> 4)+4*v;
                if(gi < M)
                    C[(int64_t)gj*ldc+gi] = acc[r][f][v];
            }
        }
}

template <int NDIM, int FAM, int BN, int RW, bool PROJ>
int launch_one(const TileDesc *d_tiles, int ntiles, int mmax, int nmax,
        const GemmArgs &g, hipStream_t s)
{
    const int nbm = (mmax+BM*RW-1)/(BM*RW);
    const int nbn = (nmax+BN-1)/BN;
    const int64_t grid = mapped_grid(ntiles, nbm, nbn, !PROJ);
    if(grid > 0x7fffffff)
        return STARSH_WRONG_PARAMETER;
    hipLaunchKernelGGL((gen_gemm_kernel<NDIM, FAM, BN, RW, PROJ>),
            dim3((unsigned)grid), dim3(256), 0, s, d_tiles, g, ntiles, nbm,
            nbn);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

template <int NDIM, int FAM, bool PROJ>
int launch_nd(const TileDesc *d_tiles, int ntiles, int mmax, int nmax,
        const GemmArgs &g, hipStream_t s)
{
    // By size (measured on MI355X at tile 1600): the sketch's wide
    // column blocks reuse each generated entry most; the projection (N =
    // kept columns, often just under a multiple of 256) wastes less on
    // 128-column blocks with two 16-row groups per wave.
    int bn, rw;
    if(PROJ)
    {
        bn = nmax > 64 ? 128 : 64;
        rw = nmax > 64 ? 2 : 1;
    }
    else
    {
        bn = nmax > 128 ? 256 : (nmax > 64 ? 128 : 64);
        rw = 1;
    }
    if(gen_gemm_tuned_bn[PROJ] > 0)
        bn = gen_gemm_tuned_bn[PROJ];
    if(gen_gemm_tuned_rw[PROJ] > 0)
        rw = gen_gemm_tuned_rw[PROJ];
    if(gen_gemm_bn > 0)
        bn = gen_gemm_bn;
    if(gen_gemm_rw > 0)
        rw = gen_gemm_rw;
    if(bn == 256)
        return rw == 2 ? launch_one<NDIM, FAM, 256, 2, PROJ>(d_tiles, ntiles,
                mmax, nmax, g, s) : launch_one<NDIM, FAM, 256, 1, PROJ>(
                d_tiles, ntiles, mmax, nmax, g, s);
    if(bn == 128)
        return rw == 2 ? launch_one<NDIM, FAM, 128, 2, PROJ>(d_tiles, ntiles,
                mmax, nmax, g, s) : launch_one<NDIM, FAM, 128, 1, PROJ>(
                d_tiles, ntiles, mmax, nmax, g, s);
    return rw == 2 ? launch_one<NDIM, FAM, 64, 2, PROJ>(d_tiles, ntiles,
            mmax, nmax, g, s) : launch_one<NDIM, FAM, 64, 1, PROJ>(d_tiles,
            ntiles, mmax, nmax, g, s);
}

template <bool PROJ>
struct GenLaunch
{
    const TileDesc *tiles;
    int ntiles, mmax, nmax;
    const GemmArgs &g;
    hipStream_t s;
    template <int NDIM, int FAM> int operator()() const
    {
        return launch_nd<NDIM, FAM, PROJ>(tiles, ntiles, mmax, nmax, g, s);
    }
};

} // namespace

extern "C" int starsh_hip_gen_gemm_rw(int rw)
{
    if(rw != 0 && rw != 1 && rw != 2)
        return STARSH_WRONG_PARAMETER;
    gen_gemm_rw = rw;
    return STARSH_SUCCESS;
}

extern "C" int hip_set_tuned_gemm(int proj, int bn, int rw)
{
    if((proj != 0 && proj != 1) || (bn != 0 && bn != 64 && bn != 128 &&
                bn != 256) || (rw != 0 && rw != 1 && rw != 2))
        return STARSH_WRONG_PARAMETER;
    gen_gemm_tuned_bn[proj] = bn;
    gen_gemm_tuned_rw[proj] = rw;
    return STARSH_SUCCESS;
}

extern "C" int starsh_hip_gen_gemm_bn(int bn)
{
    if(bn != 0 && bn != 64 && bn != 128 && bn != 256)
        return STARSH_WRONG_PARAMETER;
    gen_gemm_bn = bn;
    return STARSH_SUCCESS;
}

int pipeline_gen_gemm(int proj, const TileDesc *d_tiles, int ntiles,
        int mmax, int nmax, const GemmArgs &g, hipStream_t s)
{
    if(ntiles <= 0 || mmax <= 0 || nmax <= 0)
        return STARSH_SUCCESS;
    if(proj)
        return dispatch_kernel(g.ndim, g.family, GenLaunch<true>{d_tiles,
                ntiles, mmax, nmax, g, s});
    return dispatch_kernel(g.ndim, g.family, GenLaunch<false>{d_tiles,
            ntiles, mmax, nmax, g, s});
}
