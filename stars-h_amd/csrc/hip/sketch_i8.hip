// Mixed-precision sketch Y = A Omega on int8 MFMA (v_mfma_i32_16x16x64_i8)
// with exact integer accumulation inside each K chunk of 64 columns.
//
// For every row i and chunk c of 64 columns the entries are scaled by a
// power of two 2^e > max_{j in c} |A_ij| and rounded to an integer
// X = rint(A_ij 2^(30-e)), |X| <= 2^30, so A_ij = 2^(e-30) X up to 2^-31 of
// the chunk's largest entry.  X is split into four signed byte digits,
//   X = d3 2^24 + d2 2^16 + d1 2^8 + d0,  d0..d2 in [-128, 127], |d3| <= 64,
// read off the bytes of Z = (X + 0x808080) ^ 0x808080: adding 128 to each
// low byte moves its range to [0, 255] with the carry into the next byte,
// and the exclusive or turns c in [0, 255] into the int8 pattern of c - 128.
// The test matrix is integer: Omega_ij = round(32 x_ij) clamped to
// [-127, 127], x_ij the dlarnv normal stream of the reference.  Each digit
// plane's product with Omega over one chunk is exact in int32
// (|sum| <= 64 * 128 * 127 < 2^21); the planes are recombined exactly as
// 2^16 (2^8 s3 + s2) + (2^8 s1 + s0) and added to the FP64 row of Y with
// the chunk's scale.  Entries are generated in FP64 from the points (with
// the table-driven exp for the exponential families); only the products
// run in int8.
#include "pipeline.h"

typedef int int4v __attribute__((ext_vector_type(4)));

namespace {

constexpr int RB = 64;           // rows per block (16 per wave)
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

// One block: RB rows x CB sketch columns of one tile, as four independent
// waves of 16 rows each (no block-wide barrier after the set-up).  In a
// wave, lane l generates row l/4 and the 16 consecutive columns
// (l%4)*16.. of each chunk, so a row's chunk maximum is a reduction over
// four neighbouring lanes; the column points of the chunk are staged in
// the wave's LDS area one chunk ahead.  The wave then multiplies its own
// 16 x 64 digit planes with the CB x 64 slice of Omega, read straight from
// global memory (it is shared by all tiles and stays in the caches).
template <int NDIM, int FAM, int CB>
constexpr int sketch_waves()
{
    // three waves per SIMD fit without spills up to three dimensions
    return CB == 64 && NDIM < 4 ? 3 : 2;
}

template <int NDIM, int FAM, int CB, bool NOISE>
__global__ __launch_bounds__(256, (sketch_waves<NDIM, FAM, CB>()))
void sketch_i8_kernel(const TileDesc *tiles,
        double *ws, const signed char *om8)
{
    constexpr int NF = CB/16;
    __shared__ __align__(16) signed char As[4][4][16*APAD];
    __shared__ __align__(16) double xs[4][2][NDIM][KB];
    __shared__ double csc[4][16];
    __shared__ double tab[exp_family<FAM>() ? EXP_TAB : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const TileDesc t = tiles[blockIdx.y];
    if constexpr(exp_family<FAM>())
        exp_table(tab, t.kp.sigma, tid, 256);
    __syncthreads();
    const double cs = coord_scale<FAM>(t.kp);
    const int nbm = (t.m+RB-1)/RB;
    const int bm = blockIdx.x % nbm, bn = blockIdx.x / nbm;
    const int row0 = bm*RB+wave*16, col0 = bn*CB;
    if(t.m <= 0 || t.n <= 0 || bn*CB >= t.k || row0 >= t.m)
        return;
    const int gi = lane >> 2, gk = (lane & 3)*16;
    const bool row_ok = row0+gi < t.m;
    double xi[NDIM];
#pragma unroll
    for(int d = 0; d < NDIM; d++)
        xi[d] = row_ok ? cs*t.xr[d*t.nr+t.r0+row0+gi] : 0.0;
    const signed char *om = om8+t.om8_off;
    const int ldo = (t.n+63)/64*64;
    signed char *Aw[4] = {As[wave][0], As[wave][1], As[wave][2],
        As[wave][3]};
    // column points of chunk 0: lane l holds column l of the chunk
    const int64_t cbase = t.c0;
#pragma unroll
    for(int d = 0; d < NDIM; d++)
        xs[wave][0][d][lane] = cs*t.xc[d*t.nc+cbase+(lane < t.n ? lane :
                t.n-1)];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double yacc[NF][4];
#pragma unroll
    for(int c = 0; c < NF; c++)
#pragma unroll
        for(int v = 0; v < 4; v++)
            yacc[c][v] = 0.0;
    int buf = 0;
    for(int k0 = 0; k0 < t.n; k0 += KB, buf ^= 1)
    {
        // ---- loads for later in the iteration: this chunk's Omega
        // fragments (column col0+16c+lane%16, K bytes kq..kq+15) and the
        // next chunk's column points
        const int kq = (lane >> 4)*16;
        int4v ob[NF];
#pragma unroll
        for(int c = 0; c < NF; c++)
        {
            const int j = col0+c*16+(lane & 15);
            ob[c] = j < t.k ? *(const int4v *)(om+(int64_t)j*ldo+k0+kq) :
                int4v{0, 0, 0, 0};
        }
        double xn[NDIM];
        const int kn = k0+KB+lane;
#pragma unroll
        for(int d = 0; d < NDIM; d++)
            xn[d] = cs*t.xc[d*t.nc+cbase+(kn < t.n ? kn : t.n-1)];
        // ---- A: FP64 entries of the chunk, scale, four byte planes
        double a[16];
        double mx = 0.0;
#pragma unroll
        for(int e = 0; e < 16; e++)
        {
            double dist = 0.0;
#pragma unroll
            for(int d = 0; d < NDIM; d++)
            {
                const double q = xi[d]-xs[wave][buf][d][gk+e];
                dist = fma(q, q, dist);
            }
            // no masking: a row past the tile only feeds its own output
            // row, which is not stored, and a column past the tile repeats
            // the last column's point (so it leaves the chunk maximum
            // unchanged) and meets a zero row of Omega
            a[e] = kval_scaled<FAM, NOISE>(t.kp, dist, tab);
            mx = fmax(mx, fabs(a[e]));
        }
        mx = fmax(mx, __shfl_xor(mx, 1, 64));
        mx = fmax(mx, __shfl_xor(mx, 2, 64));
        int ex = 0;
        if(mx > 0.0)
            frexp(mx, &ex);     // mx = f 2^ex, f in [0.5, 1)
        if((lane & 3) == 0)
            csc[wave][gi] = digit_unit(mx, ex);
        const double dsc = digit_scale(mx, ex);
#pragma unroll
        for(int g = 0; g < 4; g++)
        {
            unsigned z[4];
#pragma unroll
            for(int e = 0; e < 4; e++)
                z[e] = scaled_byte_digits(a[4*g+e], dsc);
            unsigned pl[4];
            byte_planes(z[0], z[1], z[2], z[3], pl);
#pragma unroll
            for(int s = 0; s < 4; s++)
                *(unsigned *)&Aw[s][gi*APAD+gk+4*g] = pl[s];
        }
#pragma unroll
        for(int d = 0; d < NDIM; d++)
            xs[wave][buf^1][d][lane] = xn[d];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // ---- MFMA over the chunk, then exact recombination into FP64.
        // The i32 output of v_mfma_i32_16x16x64_i8 holds rows
        // 4*(lane/16)+v, column lane%16.
        int4v as[4];
#pragma unroll
        for(int s = 0; s < 4; s++)
            as[s] = *(const int4v *)&Aw[s][(lane & 15)*APAD+kq];
        double sc[4];
#pragma unroll
        for(int v = 0; v < 4; v++)
            sc[v] = csc[wave][4*(lane >> 4)+v];
#pragma unroll
        for(int c = 0; c < NF; c++)
        {
            int4v acc[4];
#pragma unroll
            for(int s = 0; s < 4; s++)
                acc[s] = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[s], ob[c],
                        int4v{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
            for(int v = 0; v < 4; v++)
            {
                const int hi = acc[3][v]*256+acc[2][v];
                const int lo = acc[1][v]*256+acc[0][v];
                yacc[c][v] = fma(fma((double)hi, 65536.0, (double)lo), sc[v],
                        yacc[c][v]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
            const int gr = row0+4*(lane >> 4)+v;
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
    template <int NDIM, int FAM, bool NOISE> void launch() const
    {
        if(kmax <= 64)
            hipLaunchKernelGGL((sketch_i8_kernel<NDIM, FAM, 64, NOISE>),
                    dim3(nbm, ntiles), dim3(256), 0, s, tiles, g.ws, om8);
        else
            hipLaunchKernelGGL((sketch_i8_kernel<NDIM, FAM, 128, NOISE>),
                    dim3(nbm*((kmax+127)/128), ntiles), dim3(256), 0, s,
                    tiles, g.ws, om8);
    }
    template <int NDIM, int FAM> int operator()() const
    {
        // the select of sigma + noise at distance 0 only where a problem
        // of the batch has a nonzero noise (other families always select)
        if(exp_family<FAM>() && !g.noisy)
            launch<NDIM, FAM, false>();
        else
            launch<NDIM, FAM, true>();
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
