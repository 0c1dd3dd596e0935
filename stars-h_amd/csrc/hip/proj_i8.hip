// Mixed-precision projection Bt = A^T Q on int8 MFMA (v_mfma_i32_16x16x64_i8),
// both operands split into signed byte digits (byte_digits):
//  - Q (m x kc, orthonormal columns) once per tile by qdigits_kernel: each
//    column c is scaled by a power of two 2^f > max_i |Q_ic| and rounded to
//    integers |X| <= 2^30 (2^-31 of the column's largest entry);
//  - A^T in the projection kernel, generated in FP64 from the points: each
//    entry A_ij is scaled per (column j, chunk of 64 rows i) like the sketch.
// A product of two four-digit numbers has digit products of weight
// 2^(8(s+t)); the products with s + t >= 2 (13 of 16) are accumulated in
// int32, grouped by weight; the three left out amount to at most 2^-37 of
// the product of the two scales per term, against 2^-30 for the rounding
// of the operands.  Each group sums at most four products of
// 64 terms bounded by 128 * 128, so |sum| < 2^22 is exact.  Per chunk the
// groups are recombined in FP64 and added with the entry's scale; the
// column scale of Q is applied once at the end.
#include "pipeline.h"

typedef int int4v __attribute__((ext_vector_type(4)));

namespace {

constexpr int KB = 64;           // K chunk (one MFMA depth)

// Per tile, at byte offset qd_off of the digit buffer: plane s of column c
// at (s*k16 + c)*m64 (k16 = k rounded up to 16, m64 = m rounded up to 64,
// rows past m zero), then the column scales 2^(f-30) as doubles at
// 4*k16*m64.  One block per (tile, column).
__global__ __launch_bounds__(256) void qdigits_kernel(const TileDesc *tiles,
        const double *ws, signed char *qd)
{
    __shared__ double red[4];
    __shared__ double sh_scale;
    const TileDesc t = tiles[blockIdx.x];
    const int c = blockIdx.y;
    if(c >= t.kc || t.m <= 0)
        return;
    const int tid = threadIdx.x;
    const int m64 = (t.m+63)/64*64, k16 = (t.k+15)/16*16;
    const double *q = ws+t.q_off+(int64_t)c*t.m;
    double mx = 0.0;
    for(int i = tid; i < t.m; i += 256)
        mx = fmax(mx, fabs(q[i]));
    for(int o = 32; o > 0; o >>= 1)
        mx = fmax(mx, __shfl_xor(mx, o, 64));
    if((tid & 63) == 0)
        red[tid >> 6] = mx;
    __syncthreads();
    if(tid == 0)
    {
        const double v = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        int ex = 0;
        if(v > 0.0)
            frexp(v, &ex);      // v = f 2^ex, f in [0.5, 1)
        double *scale = (double *)(qd+t.qd_off+4*(int64_t)k16*m64);
        scale[c] = digit_unit(v, ex);
        sh_scale = digit_scale(v, ex);
    }
    __syncthreads();
    const double dsc = sh_scale;
    signed char *base = qd+t.qd_off;
    // four consecutive rows per thread and step
    for(int i0 = 4*tid; i0 < m64; i0 += 1024)
    {
        unsigned z[4];
#pragma unroll
        for(int e = 0; e < 4; e++)
        {
            const int i = i0+e;
            z[e] = scaled_byte_digits(i < t.m ? q[i] : 0.0, dsc);
        }
        unsigned pl[4];
        byte_planes(z[0], z[1], z[2], z[3], pl);
#pragma unroll
        for(int s = 0; s < 4; s++)
            *(unsigned *)(base+((int64_t)s*k16+c)*m64+i0) = pl[s];
    }
}

template <int NDIM, int FAM, int CB>
constexpr int proj_waves()
{
    // three waves per SIMD fit without spills except for sin/cos
    return CB == 64 && NDIM < 3 && exp_family<FAM>() ? 4 :
        CB == 64 && NDIM < 4 && FAM != STARSH_HIP_SIN &&
        FAM != STARSH_HIP_COS ? 3 : 2;
}

// One block: 64 rows j (columns of the tile) x CB columns c of Bt, as four
// independent waves of 16 rows.  In a wave, lane l generates A_ij for
// j = j0 + l%16 and the 16 rows i = i0 + 16(l/16).. of each chunk, which is
// the A-operand layout of the MFMA, so the digits go from registers
// straight into it; the maximum over the chunk's rows for one j is a
// reduction over lanes l, l^16, l^32, l^48.  The row points of the chunk
// are staged in the wave's LDS area one chunk ahead.  The digit planes of
// Q are read from global memory (all blocks of a tile run on one XCD and
// share them through its L2).
template <int NDIM, int FAM, int CB, bool NOISE>
__global__ __launch_bounds__(256, (proj_waves<NDIM, FAM, CB>()))
void proj_i8_kernel(const TileDesc *tiles, double *ws,
        const signed char *qd, int ntiles, int gnbm, int gnbn)
{
    constexpr int NF = CB/16;
    __shared__ __align__(16) double xs[4][2][NDIM][KB];
    __shared__ double tab[exp_family<FAM>() ? EXP_TAB : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int tile, bm, bn;
    if(!map_block(blockIdx.x, ntiles, gnbm, gnbn, false, tile, bm, bn))
        return;     // the whole block
    const TileDesc t = tiles[tile];
    if constexpr(exp_family<FAM>())
        exp_table(tab, t.kp.sigma, tid, 256);
    __syncthreads();
    const double cs = coord_scale<FAM>(t.kp);
    const int N = t.kc;
    const int j0 = bm*64+wave*16, c0 = bn*CB;
    if(t.m <= 0 || t.n <= 0 || N <= 0 || c0 >= N || j0 >= t.n)
        return;
    const int nf_live = (N-c0+15)/16 < NF ? (N-c0+15)/16 : NF;
    const int jl = lane & 15, ig = lane >> 4;
    // column point of this lane's j (the last one past the tile)
    const int jj = j0+jl < t.n ? j0+jl : t.n-1;
    double yj[NDIM];
#pragma unroll
    for(int d = 0; d < NDIM; d++)
        yj[d] = cs*t.xc[d*t.nc+t.c0+jj];
    // columns up to k16 - 1 exist in the digit planes (the live fragments
    // end there); those from kc on hold no digits and give unstored sums
    const int m64 = (t.m+63)/64*64, k16 = (t.k+15)/16*16;
    const signed char *qb = qd+t.qd_off;
    const double *qscale = (const double *)(qb+4*(int64_t)k16*m64);
    // row points of chunk 0: lane l holds row l of the chunk
#pragma unroll
    for(int d = 0; d < NDIM; d++)
        xs[wave][0][d][lane] = cs*t.xr[d*t.nr+t.r0+(lane < t.m ? lane :
                t.m-1)];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double yacc[NF][4];
#pragma unroll
    for(int f = 0; f < NF; f++)
#pragma unroll
        for(int v = 0; v < 4; v++)
            yacc[f][v] = 0.0;
    // digit planes of Q: column c0 + 16 f + l%16, rows i0 + 16(l/16)..
    auto load_q = [&](int f, int i0, int4v (&b)[4]) {
        const int64_t col = c0+16*f+jl;
#pragma unroll
        for(int s = 0; s < 4; s++)
            b[s] = *(const int4v *)(qb+((int64_t)s*k16+col)*m64+i0+16*ig);
    };
    int buf = 0;
    for(int i0 = 0; i0 < t.m; i0 += KB, buf ^= 1)
    {
        int4v bq[4];
        load_q(0, i0, bq);
        double xn[NDIM];
        const int in = i0+KB+lane;
#pragma unroll
        for(int d = 0; d < NDIM; d++)
            xn[d] = cs*t.xr[d*t.nr+t.r0+(in < t.m ? in : t.m-1)];
        // ---- A^T entries (j, i0 + 16 ig + e); rows past the tile repeat
        // the last row's point (no effect on the maximum) and meet zero
        // digits of Q; a j past the tile only feeds unstored outputs
        double a[16];
        double mx = 0.0;
#pragma unroll
        for(int e = 0; e < 16; e++)
        {
            double dist = 0.0;
#pragma unroll
            for(int d = 0; d < NDIM; d++)
            {
                const double q = xs[wave][buf][d][16*ig+e]-yj[d];
                dist = fma(q, q, dist);
            }
            a[e] = kval_scaled<FAM, NOISE>(t.kp, dist, tab);
            mx = fmax(mx, fabs(a[e]));
        }
        mx = fmax(mx, __shfl_xor(mx, 16, 64));
        mx = fmax(mx, __shfl_xor(mx, 32, 64));
        int ex = 0;
        if(mx > 0.0)
            frexp(mx, &ex);
        const double dsc = digit_scale(mx, ex);
        int4v as[4];
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
                as[s][g] = (int)pl[s];
        }
#pragma unroll
        for(int d = 0; d < NDIM; d++)
            xs[wave][buf^1][d][lane] = xn[d];
        // scale of output row 4(l/16) + v (held by the lanes with
        // l%16 = that row); 2^16 for the weight of the lowest group kept
        const double sa = digit_unit(mx, ex)*65536.0;
        double sc[4];
#pragma unroll
        for(int v = 0; v < 4; v++)
            sc[v] = __shfl(sa, 4*ig+v, 64);
        // ---- MFMA by weight groups, then FP64 recombination
#pragma unroll
        for(int f = 0; f < NF; f++)
        {
            if(f >= nf_live)
                continue;
            // fragment 0's planes were loaded before the generation
            int4v b[4];
            if(f == 0)
            {
#pragma unroll
                for(int s = 0; s < 4; s++)
                    b[s] = bq[s];
            }
            else
                load_q(f, i0, b);
            const int4v z4 = int4v{0, 0, 0, 0};
            int4v g2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[0], b[2],
                    z4, 0, 0, 0);
            g2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[1], b[1], g2, 0, 0,
                    0);
            g2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[2], b[0], g2, 0, 0,
                    0);
            int4v g3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[0], b[3],
                    z4, 0, 0, 0);
            g3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[1], b[2], g3, 0, 0,
                    0);
            g3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[2], b[1], g3, 0, 0,
                    0);
            g3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[3], b[0], g3, 0, 0,
                    0);
            int4v g4 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[1], b[3],
                    z4, 0, 0, 0);
            g4 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[2], b[2], g4, 0, 0,
                    0);
            g4 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[3], b[1], g4, 0, 0,
                    0);
            int4v g5 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[2], b[3],
                    z4, 0, 0, 0);
            g5 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[3], b[2], g5, 0, 0,
                    0);
            const int4v g6 = __builtin_amdgcn_mfma_i32_16x16x64_i8(as[3],
                    b[3], z4, 0, 0, 0);
#pragma unroll
            for(int v = 0; v < 4; v++)
            {
                // sum = 2^16 (g2 + 2^8 (g3 + 2^8 g4) + 2^24 (g5 + 2^8 g6))
                const int p1 = g6[v]*256+g5[v];
                const int p2 = g4[v]*256+g3[v];
                const double w = fma((double)p1, 16777216.0,
                        fma((double)p2, 256.0, (double)g2[v]));
                yacc[f][v] = fma(w, sc[v], yacc[f][v]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // ---- epilogue: Bt[j, c] = sum * 2^(f_c - 30); lane holds rows
    // 4(l/16) + v, column l%16 of each fragment
    double *Bt = ws+t.bt_off;
#pragma unroll
    for(int f = 0; f < NF; f++)
    {
        const int c = c0+16*f+jl;
        if(c >= N)
            continue;
        const double qs = qscale[c];
#pragma unroll
        for(int v = 0; v < 4; v++)
        {
            const int j = j0+4*ig+v;
            if(j < t.n)
                Bt[(int64_t)c*t.n+j] = yacc[f][v]*qs;
        }
    }
}

struct ProjLaunch
{
    const TileDesc *tiles;
    int ntiles, nmax, kmax;
    const GemmArgs &g;
    const signed char *qd;
    hipStream_t s;
    template <int NDIM, int FAM, bool NOISE> int launch() const
    {
        const int nbm = (nmax+63)/64;
        if(kmax <= 64)
        {
            const int64_t grid = mapped_grid(ntiles, nbm, 1, false);
            if(grid > 0x7fffffff)
                return STARSH_WRONG_PARAMETER;
            hipLaunchKernelGGL((proj_i8_kernel<NDIM, FAM, 64, NOISE>),
                    dim3((unsigned)grid), dim3(256), 0, s, tiles, g.ws, qd,
                    ntiles, nbm, 1);
        }
        else
        {
            const int nbn = (kmax+127)/128;
            const int64_t grid = mapped_grid(ntiles, nbm, nbn, false);
            if(grid > 0x7fffffff)
                return STARSH_WRONG_PARAMETER;
            hipLaunchKernelGGL((proj_i8_kernel<NDIM, FAM, 128, NOISE>),
                    dim3((unsigned)grid), dim3(256), 0, s, tiles, g.ws, qd,
                    ntiles, nbm, nbn);
        }
        HIP_TRY(hipGetLastError());
        return STARSH_SUCCESS;
    }
    template <int NDIM, int FAM> int operator()() const
    {
        // the select of sigma + noise at distance 0 only where a problem
        // of the batch has a nonzero noise (other families always select)
        if(exp_family<FAM>() && !g.noisy)
            return launch<NDIM, FAM, false>();
        return launch<NDIM, FAM, true>();
    }
};

} // namespace

int64_t pipeline_qdigits_bytes(int m, int k)
{
    const int64_t m64 = (m+63)/64*64, k16 = (k+15)/16*16;
    return 4*k16*m64+8*k16;
}

int pipeline_qdigits(const TileDesc *d_tiles, int ntiles, int kmax,
        const double *ws, signed char *d_qd, hipStream_t s)
{
    if(ntiles <= 0 || kmax <= 0)
        return STARSH_SUCCESS;
    hipLaunchKernelGGL(qdigits_kernel, dim3(ntiles, kmax), dim3(256), 0, s,
            d_tiles, ws, d_qd);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int pipeline_proj_i8(const TileDesc *d_tiles, int ntiles, int nmax, int kmax,
        const GemmArgs &g, const signed char *d_qd, hipStream_t s)
{
    if(ntiles <= 0 || nmax <= 0 || kmax <= 0)
        return STARSH_SUCCESS;
    return dispatch_kernel(g.ndim, g.family,
            ProjLaunch{d_tiles, ntiles, nmax, kmax, g, d_qd, s});
}
