// Batched FP64 GEMMs of the randomized-SVD pipeline on MFMA
// (v_mfma_f64_16x16x4f64).  One launch covers every tile of a batch:
// grid.y is the tile, grid.x the 64 x 64 output blocks of the largest tile;
// blocks outside a tile's own output return at once.
//
// Modes (C = alpha*op(A)*op(B) + beta*C):
//   SKETCH_MEM  Y[m x k]  = A[m x n] Omega[n x k], A explicit in memory
//   PROJ_MEM    Bt[n x kc] = A^T[n x m] Q[m x kc], A explicit
// (tiles generated from points use gen_gemm.hip instead)
//   R2          R2[p x p] = Q2^T Bt[:, :p]
//   FORM_U      U[m x r] = Q[:, :p] X_r
//   FORM_V      V[n x r] = Bt[:, :p] X_r
#include "hip_common.h"
#include "pipeline.h"

typedef double double4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int BM = 64, BN = 64, BK = 16, LDA_S = BM+2, LDB_S = BN+2;

struct Op
{
    const double *A;
    const double *B;
    double *C;
    int M, N, K;
    int lda, ldb, ldc;
    double alpha, beta;
    bool valid;
};

template <int MODE>
__device__ inline Op setup(const TileDesc &t, const GemmArgs &g)
{
    Op o;
    o.valid = true;
    o.alpha = 1.0;
    o.beta = 0.0;
    double *ws = g.ws;
    switch(MODE)
    {
        case GM_SKETCH_MEM:
            o.M = t.m; o.N = t.k; o.K = t.n;
            o.A = g.abase+t.a_off; o.lda = t.m;
            o.B = ws+t.om_off; o.ldb = t.n;
            o.C = ws+t.q_off; o.ldc = t.m;
            break;
        case GM_PROJ_MEM:
            o.M = t.n; o.N = t.kc; o.K = t.m;
            o.A = g.abase+t.a_off; o.lda = t.m;
            o.B = ws+t.q_off; o.ldb = t.m;
            o.C = ws+t.bt_off; o.ldc = t.n;
            break;
        case GM_R2:
            o.M = t.p; o.N = t.p; o.K = t.n;
            o.A = ws+t.q2_off; o.lda = t.n;
            o.B = ws+t.bt_off; o.ldb = t.n;
            o.C = ws+t.r_off; o.ldc = t.k;
            o.valid = t.p > 0;
            break;
        case GM_FORM_U:
            o.M = t.m; o.N = t.rank; o.K = t.p;
            o.A = ws+t.q_off; o.lda = t.m;
            o.B = ws+t.w_off; o.ldb = t.p;
            o.C = g.out+t.u_off; o.ldc = t.m;
            o.valid = t.rank > 0;
            break;
        default: // GM_FORM_V
            o.M = t.n; o.N = t.rank; o.K = t.p;
            o.A = ws+t.bt_off; o.lda = t.n;
            o.B = ws+t.w_off; o.ldb = t.p;
            o.C = g.out+t.v_off; o.ldc = t.n;
            o.valid = t.rank > 0;
            break;
    }
    return o;
}

// A operand layouts: 0 = column-major (element (i,k) at A[k*lda+i]),
// 1 = transposed storage (element (i,k) at A[i*lda+k]).
template <int MODE> struct ModeTraits;
template <> struct ModeTraits<GM_SKETCH_MEM> { static constexpr int a = 0; };
template <> struct ModeTraits<GM_PROJ_MEM> { static constexpr int a = 1; };
template <> struct ModeTraits<GM_R2> { static constexpr int a = 1; };
template <> struct ModeTraits<GM_FORM_U> { static constexpr int a = 0; };
template <> struct ModeTraits<GM_FORM_V> { static constexpr int a = 0; };

// grid.z > 1 splits K: block z covers K steps [z*ks, (z+1)*ks) and writes
// its partial product to part + (tile*grid.z + z)*pld*pld (ld pld); the
// partials are summed by ksum_kernel.
template <int MODE>
__global__ __launch_bounds__(256) void gemm_kernel(const TileDesc *tiles,
        GemmArgs g, double *part, int pld)
{
    constexpr int AL = ModeTraits<MODE>::a;
    const TileDesc t = tiles[blockIdx.y];
    Op o = setup<MODE>(t, g);
    if(!o.valid || o.M <= 0 || o.N <= 0)
        return;
    int kbeg = 0;
    if(gridDim.z > 1)
    {
        const int ks = ((o.K+gridDim.z-1)/gridDim.z+BK-1)/BK*BK;
        kbeg = blockIdx.z*ks;
        const int kend = kbeg+ks < o.K ? kbeg+ks : o.K;
        o.C = part+((int64_t)blockIdx.y*gridDim.z+blockIdx.z)*pld*pld;
        o.ldc = pld;
        o.beta = 0.0;
        o.K = kend;
    }
    const int nbm = (o.M+BM-1)/BM;
    const int bm = blockIdx.x % nbm, bn = blockIdx.x / nbm;
    if(bn*BN >= o.N)
        return;
    const int row0 = bm*BM, col0 = bn*BN;

    __shared__ double As[BK][LDA_S];
    __shared__ double Bs[BK][LDB_S];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    double4_t acc[2][2];
#pragma unroll
    for(int a = 0; a < 2; a++)
#pragma unroll
        for(int b = 0; b < 2; b++)
            acc[a][b] = double4_t{0.0, 0.0, 0.0, 0.0};

    // operands of K step k0 loaded into registers (ra, rb) one step ahead
    // of the MFMAs that use them, so the loads of the next step overlap the
    // products of this one
    double ra[4], rb[4];
    auto load_ab = [&](int k0) {
        if(AL == 0)
        {
            const int i = tid & 63, gi = row0+i;
#pragma unroll
            for(int s = 0; s < 4; s++)
            {
                const int kk = (tid >> 6)+4*s, gk = k0+kk;
                ra[s] = gi < o.M && gk < o.K ?
                    o.A[(int64_t)gk*o.lda+gi] : 0.0;
            }
        }
        else
        {
            const int kk = tid & 15, gk = k0+kk;
#pragma unroll
            for(int s = 0; s < 4; s++)
            {
                const int i = (tid >> 4)+16*s, gi = row0+i;
                ra[s] = gi < o.M && gk < o.K ?
                    o.A[(int64_t)gi*o.lda+gk] : 0.0;
            }
        }
        // B is column-major K x N
        const int kk = tid & 15, gk = k0+kk;
#pragma unroll
        for(int s = 0; s < 4; s++)
        {
            const int j = (tid >> 4)+16*s, gj = col0+j;
            rb[s] = gj < o.N && gk < o.K ? o.B[(int64_t)gj*o.ldb+gk] : 0.0;
        }
    };
    // A (BM x BK) as As[k][i], B (BK x BN) as Bs[k][j]
    auto store_ab = [&]() {
#pragma unroll
        for(int s = 0; s < 4; s++)
        {
            if(AL == 0)
                As[(tid >> 6)+4*s][tid & 63] = ra[s];
            else
                As[tid & 15][(tid >> 4)+16*s] = ra[s];
            Bs[tid & 15][(tid >> 4)+16*s] = rb[s];
        }
    };
    if(kbeg < o.K)
        load_ab(kbeg);
    for(int k0 = kbeg; k0 < o.K; k0 += BK)
    {
        store_ab();
        __syncthreads();
        if(k0+BK < o.K)
            load_ab(k0+BK);
#pragma unroll
        for(int k4 = 0; k4 < BK; k4 += 4)
        {
            const int kr = k4+(lane >> 4), c = lane & 15;
            double a0 = As[kr][wm*32+c], a1 = As[kr][wm*32+16+c];
            double b0 = Bs[kr][wn*32+c], b1 = Bs[kr][wn*32+16+c];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0,
                    acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1,
                    acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0,
                    acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1,
                    acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
    // ---- epilogue: lane holds rows lane/16 + 4*v, column lane%16 of each
    // 16 x 16 block (v_mfma_f64_16x16x4f64 output layout).
#pragma unroll
    for(int a = 0; a < 2; a++)
#pragma unroll
        for(int b = 0; b < 2; b++)
        {
            const int gj = col0+wn*32+b*16+(lane & 15);
            if(gj >= o.N)
                continue;
#pragma unroll
            for(int v = 0; v < 4; v++)
            {
                const int gi = row0+wm*32+a*16+(lane >> 4)+4*v;
                if(gi >= o.M)
                    continue;
                double *c = o.C+(int64_t)gj*o.ldc+gi;
                double r = o.alpha*acc[a][b][v];
                if(o.beta != 0.0)
                    r += o.beta*(*c);
                *c = r;
            }
        }
}

// R2[:p, :p] of every tile = sum of its nz partials (in order).
__global__ __launch_bounds__(256) void ksum_kernel(const TileDesc *tiles,
        const double *part, int pld, int nz, double *ws)
{
    const TileDesc t = tiles[blockIdx.y];
    const int p = t.p;
    const int64_t e = (int64_t)blockIdx.x*256+threadIdx.x;
    if(p <= 0 || e >= (int64_t)p*p)
        return;
    const int i = (int)(e % p), j = (int)(e / p);
    const double *src = part+(int64_t)blockIdx.y*nz*pld*pld+(int64_t)j*pld+i;
    double sum = 0.0;
    for(int z = 0; z < nz; z++)
        sum += src[(int64_t)z*pld*pld];
    ws[t.r_off+(int64_t)j*t.k+i] = sum;
}

template <int MODE>
int launch(const TileDesc *d_tiles, int ntiles, int mmax, int nmax,
        const GemmArgs &g, hipStream_t s, int nz = 1, double *part = nullptr,
        int pld = 0)
{
    int nbm = (mmax+BM-1)/BM, nbn = (nmax+BN-1)/BN;
    if(nbm <= 0 || nbn <= 0 || ntiles <= 0)
        return STARSH_SUCCESS;
    hipLaunchKernelGGL((gemm_kernel<MODE>), dim3(nbm*nbn, ntiles, nz),
            dim3(256), 0, s, d_tiles, g, part, pld);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

} // namespace

int pipeline_r2_split(int ntiles, int pmax, int kmax_rows)
{
    // at least 128 rows of K per block, at most 8 blocks, and the partials
    // within 1 GiB
    int nz = kmax_rows/128;
    nz = nz < 1 ? 1 : (nz > 8 ? 8 : nz);
    const double per = 8.0*ntiles*(double)pmax*pmax;
    while(nz > 1 && per*nz > (double)(1 << 30))
        nz--;
    return nz;
}

int pipeline_r2(const TileDesc *d_tiles, int ntiles, int pmax, int nz,
        double *d_part, const GemmArgs &g, hipStream_t s)
{
    if(ntiles <= 0 || pmax <= 0)
        return STARSH_SUCCESS;
    if(nz <= 1 || d_part == nullptr)
        return launch<GM_R2>(d_tiles, ntiles, pmax, pmax, g, s);
    const int info = launch<GM_R2>(d_tiles, ntiles, pmax, pmax, g, s, nz,
            d_part, pmax);
    if(info != STARSH_SUCCESS)
        return info;
    const int64_t cnt = (int64_t)pmax*pmax;
    hipLaunchKernelGGL(ksum_kernel, dim3((unsigned)((cnt+255)/256), ntiles),
            dim3(256), 0, s, d_tiles, d_part, pmax, nz, g.ws);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int pipeline_gemm(int mode, const TileDesc *d_tiles, int ntiles, int mmax,
        int nmax, const GemmArgs &g, hipStream_t s)
{
    switch(mode)
    {
        case GM_SKETCH_MEM: return launch<GM_SKETCH_MEM>(d_tiles, ntiles,
                                    mmax, nmax, g, s);
        case GM_PROJ_MEM: return launch<GM_PROJ_MEM>(d_tiles, ntiles, mmax,
                                  nmax, g, s);
        case GM_R2: return launch<GM_R2>(d_tiles, ntiles, mmax, nmax, g, s);
        case GM_FORM_U: return launch<GM_FORM_U>(d_tiles, ntiles, mmax, nmax,
                                g, s);
        case GM_FORM_V: return launch<GM_FORM_V>(d_tiles, ntiles, mmax, nmax,
                                g, s);
        default: return STARSH_WRONG_PARAMETER;
    }
}
