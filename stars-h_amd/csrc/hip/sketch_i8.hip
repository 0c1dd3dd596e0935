// Mixed-precision sketch Y = A Omega on int8 MFMA (v_mfma_i32_16x16x64_i8)
// with exact integer accumulation inside each K chunk of 64 columns.
//
// For every row i and chunk c of 64 columns the entries are scaled by a
// power of two 2^e >= max_{j in c} |A_ij| and rounded to an integer
// X = rint(A_ij 2^(27-e)), |X| <= 2^27, which is split into four balanced
// base-128 digits (Ozaki splitting):
//   X = d1 2^21 + d2 2^14 + d3 2^7 + d4,   |d| <= 64,
// so A_ij = 2^(e-27) X up to 2^-28 of the chunk's largest entry.  The test
// matrix is integer: Omega_ij = round(32 x_ij) clamped to [-127, 127], x_ij
// the dlarnv normal stream of the reference.  Each digit plane's product
// with Omega over one chunk is exact in int32 (|sum| <= 64 * 64 * 127 <
// 2^19); the planes are recombined exactly as 2^14 (2^7 s1 + s2) +
// (2^7 s3 + s4) and added to the FP64 row of Y with the chunk's scale.
// Entries are generated in FP64 from the points, as in the plain form; only
// the products run in int8.
#include "pipeline.h"

typedef int int4v __attribute__((ext_vector_type(4)));

namespace {

constexpr int RB = 64;           // rows per block
constexpr int CB = 64;           // sketch columns per block
constexpr int KB = 64;           // K chunk (one MFMA depth)
constexpr int APAD = 80;         // bytes per LDS row (64 + 16 padding)

__constant__ uint64_t c_apow_i8[64];

// Integer test matrix from the dlarnv stream (see runtime.hip), stored
// column by column with the leading dimension ncols rounded up to 64 and
// zeros in the padding, so that every 16-byte K slice of a column is an
// aligned vector load.
__global__ void omega_i8_kernel(signed char *out, int ncols, int ldo,
        int64_t count)
{
    const uint64_t mask = (1ULL << 48)-1;
    const uint64_t a = 33952834046453ULL;
    const double twopi = 6.2831853071795864769252867663;
    int64_t p = (int64_t)blockIdx.x*blockDim.x+threadIdx.x;
    if(p >= count)
        return;
    const int64_t col = p/ldo, row = p-col*ldo;
    if(row >= ncols)
    {
        out[p] = 0;
        return;
    }
    const int64_t e = col*ncols+row;   // index in the reference's n x k
    uint64_t q = 2*(uint64_t)e+1, s = 1;
    for(int b = 0; q != 0; b++, q >>= 1)
        if(q & 1)
            s = (s*c_apow_i8[b]) & mask;
    double u1 = (double)s*0x1p-48;
    s = (s*a) & mask;
    double u2 = (double)s*0x1p-48;
    double x = sqrt(-2.0*log(u1))*cos(twopi*u2);
    double v = rint(32.0*x);
    v = v > 127.0 ? 127.0 : (v < -127.0 ? -127.0 : v);
    out[p] = (signed char)v;
}

__device__ inline int pack4(int a, int b, int c, int d)
{
    return (a & 255) | ((b & 255) << 8) | ((c & 255) << 16) |
        ((unsigned)(d & 255) << 24);
}

// One block: RB rows x CB sketch columns of one tile.  Generation: thread
// t owns row t/4 and 16 consecutive columns (t%4)*16.. of each chunk, so a
// row's chunk maximum is a reduction over four neighbouring lanes.  MFMA:
// wave w owns rows 16w.. and all CB columns.
template <int NDIM, int FAM, int CB>
__global__ __launch_bounds__(256) void sketch_i8_kernel(const TileDesc *tiles,
        double *ws, const signed char *om8)
{
    constexpr int NF = CB/16;
    __shared__ __align__(16) signed char As[4][RB*APAD];
    __shared__ __align__(16) signed char Bs[CB*APAD];
    __shared__ double csc[RB];
    const TileDesc t = tiles[blockIdx.y];
    const int nbm = (t.m+RB-1)/RB;
    const int bm = blockIdx.x % nbm, bn = blockIdx.x / nbm;
    if(t.m <= 0 || bn*CB >= t.k)
        return;
    const int row0 = bm*RB, col0 = bn*CB;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int gi = tid >> 2, gk = (tid & 3)*16;
    const bool row_ok = row0+gi < t.m;
    double xi[NDIM];
#pragma unroll
    for(int d = 0; d < NDIM; d++)
        xi[d] = row_ok ? t.xr[d*t.nr+t.r0+row0+gi] : 0.0;
    const signed char *om = om8+t.om8_off;
    const int ldo = (t.n+63)/64*64;
    double yacc[NF][4];
#pragma unroll
    for(int c = 0; c < NF; c++)
#pragma unroll
        for(int v = 0; v < 4; v++)
            yacc[c][v] = 0.0;
    for(int k0 = 0; k0 < t.n; k0 += KB)
    {
        // ---- A: FP64 entries of the chunk, scale, four int8 digit planes
        // column points of this thread's 16 entries, loaded together
        double yv[16][NDIM];
#pragma unroll
        for(int e = 0; e < 16; e++)
        {
            const int kk = k0+gk+e;
            const int64_t jj = t.c0+(kk < t.n ? kk : t.n-1);
#pragma unroll
            for(int d = 0; d < NDIM; d++)
                yv[e][d] = t.xc[d*t.nc+jj];
        }
        double a[16];
        double mx = 0.0;
#pragma unroll
        for(int e = 0; e < 16; e++)
        {
            const int kk = k0+gk+e;
            const double v = kval_pt<NDIM, FAM>(t.kp, xi, yv[e]);
            a[e] = (row_ok && kk < t.n) ? v : 0.0;
            const double f = fabs(a[e]);
            mx = f > mx ? f : mx;
        }
        double o = __shfl_xor(mx, 1, 64);
        mx = o > mx ? o : mx;
        o = __shfl_xor(mx, 2, 64);
        mx = o > mx ? o : mx;
        int ex = 0;
        if(mx > 0.0)
            frexp(mx, &ex);     // mx = f 2^ex, f in [0.5, 1)
        const double mul = mx > 0.0 ? ldexp(1.0, 27-ex) : 0.0;
        if((tid & 3) == 0)
            csc[gi] = mx > 0.0 ? ldexp(1.0, ex-27) : 0.0;
        int dg[4][16];
#pragma unroll
        for(int e = 0; e < 16; e++)
        {
            int x = (int)rint(a[e]*mul);
            const int d4 = ((x+64) & 127)-64;
            x = (x-d4) >> 7;
            const int d3 = ((x+64) & 127)-64;
            x = (x-d3) >> 7;
            const int d2 = ((x+64) & 127)-64;
            dg[3][e] = d4;
            dg[2][e] = d3;
            dg[1][e] = d2;
            dg[0][e] = (x-d2) >> 7;
        }
#pragma unroll
        for(int s = 0; s < 4; s++)
        {
            int4v v;
            v.x = pack4(dg[s][0], dg[s][1], dg[s][2], dg[s][3]);
            v.y = pack4(dg[s][4], dg[s][5], dg[s][6], dg[s][7]);
            v.z = pack4(dg[s][8], dg[s][9], dg[s][10], dg[s][11]);
            v.w = pack4(dg[s][12], dg[s][13], dg[s][14], dg[s][15]);
            *(int4v *)&As[s][gi*APAD+gk] = v;
        }
        // ---- B: Omega columns col0.., K rows k0.. (column-major, leading
        // dimension ldo = n rounded up to 64, zero padded)
#pragma unroll
        for(int h = 0; h < CB/64; h++)
        {
            const int j = (tid >> 2)+64*h, kq = (tid & 3)*16;
            int4v v = int4v{0, 0, 0, 0};
            if(col0+j < t.k)
                v = *(const int4v *)(om+(int64_t)(col0+j)*ldo+k0+kq);
            *(int4v *)&Bs[j*APAD+kq] = v;
        }
        __syncthreads();
        // ---- MFMA over the chunk, then exact recombination into FP64.
        // The i32 output of v_mfma_i32_16x16x64_i8 holds rows
        // 4*(lane/16)+v, column lane%16.
        const int ar = wave*16+(lane & 15), kq = (lane >> 4)*16;
        int4v as[4];
#pragma unroll
        for(int s = 0; s < 4; s++)
            as[s] = *(const int4v *)&As[s][ar*APAD+kq];
        double sc[4];
#pragma unroll
        for(int v = 0; v < 4; v++)
            sc[v] = csc[wave*16+4*(lane >> 4)+v];
#pragma unroll
        for(int c = 0; c < NF; c++)
        {
            const int4v b = *(const int4v *)&Bs[(c*16+(lane & 15))*APAD+kq];
            int4v acc[4];
#pragma unroll
            for(int s = 0; s < 4; s++)
                acc[s] = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[s], b,
                        int4v{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
            for(int v = 0; v < 4; v++)
            {
                const int hi = acc[0][v]*128+acc[1][v];
                const int lo = acc[2][v]*128+acc[3][v];
                yacc[c][v] = fma((double)hi*16384.0+(double)lo, sc[v],
                        yacc[c][v]);
            }
        }
        __syncthreads();
    }
    // ---- epilogue: Y = sum over chunks / 32 (Omega's integer scale)
    double *Y = ws+t.q_off;
#pragma unroll
    for(int c = 0; c < NF; c++)
    {
        const int gj = col0+c*16+(lane & 15);
        if(gj >= t.k)
            continue;
#pragma unroll
        for(int v = 0; v < 4; v++)
        {
            const int gr = row0+wave*16+4*(lane >> 4)+v;
            if(gr < t.m)
                Y[(int64_t)gj*t.m+gr] = yacc[c][v]*(1.0/32.0);
        }
    }
}

} // namespace

int pipeline_omega_i8(signed char *d_out, int ncols, int k, hipStream_t s)
{
    static bool ready = false;
    if(!ready)
    {
        uint64_t tb[64], mask = (1ULL << 48)-1, x = 33952834046453ULL;
        for(int b = 0; b < 64; b++)
        {
            tb[b] = x;
            x = (x*x) & mask;
        }
        HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_apow_i8), tb, sizeof(tb)));
        ready = true;
    }
    const int ldo = (ncols+63)/64*64;
    int64_t count = (int64_t)ldo*k;
    if(count <= 0)
        return STARSH_SUCCESS;
    hipLaunchKernelGGL(omega_i8_kernel, dim3((unsigned)((count+255)/256)),
            dim3(256), 0, s, d_out, ncols, ldo, count);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

namespace {

struct SketchLaunch
{
    const TileDesc *tiles;
    int ntiles, nbm, kmax;
    const GemmArgs &g;
    const signed char *om8;
    hipStream_t s;
    template <int NDIM, int FAM> int operator()() const
    {
        if(kmax <= 64)
            hipLaunchKernelGGL((sketch_i8_kernel<NDIM, FAM, 64>),
                    dim3(nbm, ntiles), dim3(256), 0, s, tiles, g.ws, om8);
        else
            hipLaunchKernelGGL((sketch_i8_kernel<NDIM, FAM, 128>),
                    dim3(nbm*((kmax+127)/128), ntiles), dim3(256), 0, s,
                    tiles, g.ws, om8);
        HIP_TRY(hipGetLastError());
        return STARSH_SUCCESS;
    }
};

} // namespace

int pipeline_sketch_i8(const TileDesc *d_tiles, int ntiles, int mmax,
        int kmax, const GemmArgs &g, const signed char *d_om8, hipStream_t s)
{
    if(ntiles <= 0 || mmax <= 0 || kmax <= 0)
        return STARSH_SUCCESS;
    const int nbm = (mmax+RB-1)/RB;
    return dispatch_kernel(g.ndim, g.family,
            SketchLaunch{d_tiles, ntiles, nbm, kmax, g, d_om8, s});
}
