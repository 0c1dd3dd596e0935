// Orthogonalisation stages of the randomized SVD: block Gram-Schmidt
// applied twice (BCGS2) with column deflation, compaction of the kept
// columns, and truncation of the projected matrix.
//
// The reference factors the sketch with Householder QR (dgeqrf + dorgqr,
// src/backends/sequential/dense/drsdd.c:64-67).  The sketch of a smooth
// kernel tile is numerically rank deficient far below the sketch width, and
// plain block Gram-Schmidt loses orthogonality on such columns.  Columns
// whose residual after projection falls to tau times their original norm
// are therefore dropped (set to zero): they carry no information above
// rounding, and leaving them out changes the singular values of B = Q^T A
// only at the level of tau*|A|.  Once a whole panel is dropped, the range
// of the sketch is exhausted and the remaining panels of that tile are
// skipped; a device counter of tiles still in progress lets the host stop
// launching panels when it reaches zero.
#include "pipeline.h"

namespace {


__device__ inline double wave_sum(double v)
{
    return group_sum_lanes(v, 64);
}

// Column norms squared: which = 0 -> Q buffer (m x k), 2 -> Bt (n x kc).
__global__ __launch_bounds__(256) void colnorm_kernel(const TileDesc *tiles,
        int which, const double *ws, double *out, int kmax)
{
    const TileDesc t = tiles[blockIdx.y];
    const int rows = which == 0 ? t.m : t.n;
    const int cols = which == 0 ? t.k : t.kc;
    const double *base = ws+(which == 0 ? t.q_off : t.bt_off);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for(int c = blockIdx.x*4+wave; c < cols; c += gridDim.x*4)
    {
        const double *col = base+(int64_t)c*rows;
        double s = 0.0;
        for(int i = lane; i < rows; i += 64)
            s += col[i]*col[i];
        s = wave_sum(s);
        if(lane == 0)
            out[(int64_t)blockIdx.y*kmax+c] = s;
    }
}

// Sum over the 64 lanes of a wave of each of V values (V a power of two,
// at most 64) by recursive halving: at every level O (32, 16, ..., 1) a
// lane keeps half of its values, the upper half if lane & O, and adds its
// partner's copy of them (partners as in lane_partner / swap_sum), so
// V-1+log2(64/V) exchanges reduce all V.  On return the lane holds the
// total of value index halving_index<V>(lane).  The steps are unrolled by
// template recursion.
template <int H, int O>
struct Halve
{
    template <int V>
    static __device__ inline void run(double (&v)[V], int lane)
    {
        if constexpr(O >= 16)
        {
            // the swap hands the lower (even-row) lane's v[i] and the
            // upper lane's v[i+H] to the partner that keeps them
#pragma unroll
            for(int i = 0; i < H; i++)
                v[i] = swap_sum<O>(v[i], v[i+H]);
        }
        else
        {
            const bool up = (lane & O) != 0;
#pragma unroll
            for(int i = 0; i < H; i++)
            {
                const double send = up ? v[i] : v[i+H];
                const double keep = up ? v[i+H] : v[i];
                v[i] = keep+lane_partner<O>(send);
            }
        }
        Halve<H/2, O/2>::run(v, lane);
    }
};

template <int O>
struct Halve<0, O>
{
    template <int V>
    static __device__ inline void run(double (&v)[V], int lane)
    {
        (void)lane;
        if constexpr(O >= 1)
        {
            v[0] = pair_sum<O>(v[0]);
            Halve<0, O/2>::run(v, lane);
        }
    }
};

template <int V>
__device__ inline double halving_sum(double (&v)[V], int lane)
{
    Halve<V/2, 32>::run(v, lane);
    return v[0];
}

template <int V>
__device__ inline int halving_index(int lane)
{
    int idx = 0;
#pragma unroll
    for(int st = 0; (V >> (st+1)) >= 1; st++)
        if(lane & (32 >> st))
            idx += V >> (st+1);
    return idx;
}

// Block Gram-Schmidt applied twice (BCGS2) to the panel W = columns
// j0..j0+PB-1 of Q (which = 0, m rows, width k) or Q2 (which = 1, n rows,
// width p), fused in one workgroup per tile.  Each of the two passes:
//   1. projection against the previous columns, CH at a time (block
//      modified Gram-Schmidt): C = Q_chᵀ W, W -= Q_ch C; each chunk of Q is
//      read once, and its CH x PB dot products are reduced by
//      halving_sum, then across the waves in LDS;
//   2. CGS2 inside the panel with deflation: a column whose norm after
//      projection is at most tau times its reference norm (in the first
//      pass the norm of the original column, computed when the panel is
//      loaded; 1 in the second pass, the column having been normalised) is
//      set to zero and flagged in keep.
// For Q2 the panel is read from Bt (Q2 starts as Bt[:, :p]) and written to
// Q2.
// The panel lives in registers: thread t holds rows t, t+NT, ... (RPT of
// them).  A tile whose whole panel is dropped in the first pass is marked
// done (its range is exhausted); *active counts the tiles that still have
// panels to process.
template <int NT, int RPT, int PB>
__global__ __launch_bounds__(NT) void bcgs_kernel(TileDesc *tiles, int which,
        double *ws, int j0, int kmax, double tau2_first, double tau2_second,
        int *keep, int *active)
{
    constexpr int NW = NT/64, CH = 4, V = CH*PB;
    __shared__ double red[NW][V];
    __shared__ double fin[V];
    __shared__ double pred[2][NW][PB];
    TileDesc &td = tiles[blockIdx.x];
    const bool q2 = which != 0;
    const int done_bit = q2 ? 2 : 1;
    const int rows = q2 ? td.n : td.m;
    const int width = q2 ? td.p : td.k;
    if(j0 >= width)
        return;
    const int jb = width-j0 < PB ? width-j0 : PB;
    int *kp = keep+(int64_t)blockIdx.x*kmax+j0;
    if(td.flags & done_bit)
    {
        // range already exhausted: every remaining column is dropped
        if(threadIdx.x < jb)
            kp[threadIdx.x] = 0;
        return;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *Qb = ws+(q2 ? td.q2_off : td.q_off);
    double *W = Qb+(int64_t)j0*rows;
    // Q2 starts as Bt[:, :p]: its panel is read from Bt and written to Q2
    const double *Wsrc = q2 ? ws+td.bt_off+(int64_t)j0*rows : W;
    double w[RPT][PB];
#pragma unroll
    for(int r = 0; r < RPT; r++)
    {
        const int i = threadIdx.x+r*NT;
#pragma unroll
        for(int c = 0; c < PB; c++)
            w[r][c] = (i < rows && c < jb) ? Wsrc[(int64_t)c*rows+i] : 0.0;
    }
    // squared norms of the panel's original columns (the deflation
    // reference of the first pass)
    double ref2[PB];
    {
        double v[PB];
#pragma unroll
        for(int c = 0; c < PB; c++)
        {
            double d = 0.0;
#pragma unroll
            for(int r = 0; r < RPT; r++)
                d += w[r][c]*w[r][c];
            v[c] = d;
        }
        const double tot = halving_sum<PB>(v, lane);
        if((lane & ((64/PB)-1)) == 0)
            red[wave][halving_index<PB>(lane)] = tot;
        __syncthreads();
#pragma unroll
        for(int c = 0; c < PB; c++)
        {
            double sum = 0.0;
#pragma unroll
            for(int q2w = 0; q2w < NW; q2w++)
                sum += red[q2w][c];
            ref2[c] = sum;
        }
        __syncthreads();
    }
    const int my_idx = halving_index<V>(lane);
    const bool writer = (lane & ((64/V)-1)) == 0;
    int rb = 0;   // partial buffer of the next in-panel reduction
    for(int pass = 0; pass < 2; pass++)
    {
        // ---- 1. projection against columns 0..j0-1 (j0 is a multiple of
        // CH)
        for(int l0 = 0; l0 < j0; l0 += CH)
        {
            double q[RPT][CH];
#pragma unroll
            for(int r = 0; r < RPT; r++)
            {
                const int i = threadIdx.x+r*NT;
#pragma unroll
                for(int c = 0; c < CH; c++)
                    q[r][c] = i < rows ? Qb[(int64_t)(l0+c)*rows+i] : 0.0;
            }
            double v[V];
#pragma unroll
            for(int c = 0; c < CH; c++)
#pragma unroll
                for(int b = 0; b < PB; b++)
                {
                    double d = 0.0;
#pragma unroll
                    for(int r = 0; r < RPT; r++)
                        d += q[r][c]*w[r][b];
                    v[c*PB+b] = d;
                }
            const double tot = halving_sum<V>(v, lane);
            if(writer)
                red[wave][my_idx] = tot;
            __syncthreads();
            if(threadIdx.x < V)
            {
                double sum = 0.0;
#pragma unroll
                for(int q2w = 0; q2w < NW; q2w++)
                    sum += red[q2w][threadIdx.x];
                fin[threadIdx.x] = sum;
            }
            __syncthreads();
#pragma unroll
            for(int c = 0; c < CH; c++)
#pragma unroll
                for(int b = 0; b < PB; b++)
                {
                    const double coef = fin[c*PB+b];
#pragma unroll
                    for(int r = 0; r < RPT; r++)
                        w[r][b] -= q[r][c]*coef;
                }
        }
        // ---- 2. CGS2 inside the panel with deflation
        int ndropped = 0;
#pragma unroll
        for(int c = 0; c < PB; c++)
        {
            if(c < jb)
            {
#pragma unroll
                for(int sub = 0; sub < 2; sub++)
                {
                    if(c > 0)
                    {
#pragma unroll
                        for(int l = 0; l < PB; l++)
                            if(l < c)
                            {
                                double d = 0.0;
#pragma unroll
                                for(int r = 0; r < RPT; r++)
                                    d += w[r][l]*w[r][c];
                                d = wave_sum(d);
                                if(lane == 0)
                                    pred[rb][wave][l] = d;
                            }
                        __syncthreads();
#pragma unroll
                        for(int l = 0; l < PB; l++)
                            if(l < c)
                            {
                                double coef = 0.0;
#pragma unroll
                                for(int q2w = 0; q2w < NW; q2w++)
                                    coef += pred[rb][q2w][l];
#pragma unroll
                                for(int r = 0; r < RPT; r++)
                                    w[r][c] -= coef*w[r][l];
                            }
                        rb ^= 1;
                    }
                }
                double d = 0.0;
#pragma unroll
                for(int r = 0; r < RPT; r++)
                    d += w[r][c]*w[r][c];
                d = wave_sum(d);
                if(lane == 0)
                    pred[rb][wave][0] = d;
                __syncthreads();
                double n2 = 0.0;
#pragma unroll
                for(int q2w = 0; q2w < NW; q2w++)
                    n2 += pred[rb][q2w][0];
                rb ^= 1;
                const double r2 = pass == 0 ? ref2[c] : 1.0;
                const double tau2 = pass == 0 ? tau2_first : tau2_second;
                const bool drop = !(n2 > tau2*r2) || n2 == 0.0;
                const double inv = drop ? 0.0 : 1.0/sqrt(n2);
#pragma unroll
                for(int r = 0; r < RPT; r++)
                    w[r][c] *= inv;
                if(drop)
                    ndropped++;
                if(threadIdx.x == 0)
                {
                    if(pass == 0)
                        kp[c] = drop ? 0 : 1;
                    else if(drop)
                        kp[c] = 0;
                }
            }
        }
        // (a selection sketch A[:, sel] is not a random combination of the
        // columns: a dropped panel does not mean its range is exhausted)
        if(pass == 0 && ndropped == jb && !(td.rule & RSDD_RULE_SELECT))
        {
            // the range is exhausted: the panel is zero, later panels are
            // skipped.  Q2 is used in place (no compaction), so the
            // columns of those panels are zeroed here and Q2^T Bt ignores
            // them.
#pragma unroll
            for(int r = 0; r < RPT; r++)
            {
                const int i = threadIdx.x+r*NT;
                if(i < rows)
#pragma unroll
                    for(int c = 0; c < PB; c++)
                        if(c < jb)
                            W[(int64_t)c*rows+i] = 0.0;
            }
            if(q2)
            {
                double *rest = W+(int64_t)jb*rows;
                const int64_t cnt = (int64_t)(width-j0-jb)*rows;
                for(int64_t e = threadIdx.x; e < cnt; e += NT)
                    rest[e] = 0.0;
            }
            if(threadIdx.x == 0)
            {
                td.flags |= done_bit;
                atomicSub(active, 1);
            }
            return;
        }
    }
#pragma unroll
    for(int r = 0; r < RPT; r++)
    {
        const int i = threadIdx.x+r*NT;
        if(i < rows)
#pragma unroll
            for(int c = 0; c < PB; c++)
                if(c < jb)
                    W[(int64_t)c*rows+i] = w[r][c];
    }
    if(threadIdx.x == 0 && j0+jb >= width)
        atomicSub(active, 1);   // finished all panels
}

// Moves kept columns of Q to the front, in order; sets kc.
__global__ __launch_bounds__(256) void compact_kernel(TileDesc *tiles,
        double *ws, const int *keep, int kmax)
{
    TileDesc &td = tiles[blockIdx.x];
    const int *kp = keep+(int64_t)blockIdx.x*kmax;
    double *Q = ws+td.q_off;
    const int m = td.m;
    int c = 0;
    for(int j = 0; j < td.k; j++)
    {
        if(!kp[j])
            continue;
        if(c != j)
        {
            for(int i = threadIdx.x; i < m; i += blockDim.x)
                Q[(int64_t)c*m+i] = Q[(int64_t)j*m+i];
            __syncthreads();
        }
        c++;
    }
    if(threadIdx.x == 0)
        td.kc = c;
}

// |B|_F^2, truncation point p (smallest p >= 1 whose row tail of B is at
// most (eta*tol)^2 |B|_F^2) and the tail.
__global__ __launch_bounds__(256) void truncate_kernel(TileDesc *tiles,
        int ntiles, const double *bnorm2, int kmax, double eta)
{
    const int t = blockIdx.x*blockDim.x+threadIdx.x;
    if(t >= ntiles)
        return;
    TileDesc &td = tiles[t];
    const double *nb = bnorm2+(int64_t)t*kmax;
    double total = 0.0;
    for(int c = 0; c < td.kc; c++)
        total += nb[c];
    const double eta_tol = eta*td.tol;
    double limit = eta_tol*eta_tol*total;
    int p = td.kc;
    double tail = 0.0;
    while(p > 1 && tail+nb[p-1] <= limit)
    {
        tail += nb[p-1];
        p--;
    }
    if(total == 0.0)
    {
        p = 0;
        tail = 0.0;
    }
    td.p = p;
    td.tail2 = tail;
    td.bnorm2 = total;
}

// Number of tiles that take part in a BCGS2 run (width > 0).
__global__ void count_active_kernel(const TileDesc *tiles, int ntiles,
        int which, int *active)
{
    int c = 0;
    for(int t = threadIdx.x; t < ntiles; t += blockDim.x)
        c += (which ? tiles[t].p : tiles[t].k) > 0;
    for(int o = 32; o > 0; o >>= 1)
        c += __shfl_xor(c, o, 64);
    __shared__ int part[16];
    if((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = c;
    __syncthreads();
    if(threadIdx.x == 0)
    {
        int s = 0;
        for(int q = 0; q < (int)blockDim.x/64; q++)
            s += part[q];
        *active = s;
    }
}

} // namespace

int pipeline_colnorm(const TileDesc *d_tiles, int ntiles, int which,
        double *ws, double *d_norm, int kmax, hipStream_t s)
{
    if(ntiles <= 0)
        return STARSH_SUCCESS;
    int gx = (kmax+3)/4;
    if(gx > 32)
        gx = 32;
    hipLaunchKernelGGL(colnorm_kernel, dim3(gx, ntiles), dim3(256), 0, s,
            d_tiles, which, ws, d_norm, kmax);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int pipeline_count_active(const TileDesc *d_tiles, int ntiles, int which,
        int *d_active, hipStream_t s)
{
    hipLaunchKernelGGL(count_active_kernel, dim3(1), dim3(1024), 0, s,
            d_tiles, ntiles, which, d_active);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int pipeline_panel(const TileDesc *d_tiles, int ntiles, int which,
        double *ws, int j0, int jb, int kmax, double tau_first,
        double tau_second, int *d_keep, int rows_max, int *d_active,
        hipStream_t s)
{
    if(ntiles <= 0)
        return STARSH_SUCCESS;
    TileDesc *tiles = const_cast<TileDesc *>(d_tiles);
    const double tau2a = tau_first*tau_first, tau2b = tau_second*tau_second;
#define BCGS_LAUNCH(NT, RPT, PB) \
        hipLaunchKernelGGL((bcgs_kernel<NT, RPT, PB>), dim3(ntiles), \
                dim3(NT), 0, s, tiles, which, ws, j0, kmax, tau2a, tau2b, \
                d_keep, d_active)
    if(jb == 4 && rows_max <= 1024)
        BCGS_LAUNCH(256, 4, 4);
    else if(jb == 4 && rows_max <= 2048)
        BCGS_LAUNCH(256, 8, 4);
    else if(jb == 4 && rows_max <= 4096)
        BCGS_LAUNCH(512, 8, 4);
    else if(jb == 8 && rows_max <= 1024)
        BCGS_LAUNCH(256, 4, 8);
    else if(jb == 8 && rows_max <= 2048)
        BCGS_LAUNCH(512, 4, 8);
#undef BCGS_LAUNCH
    else
        return STARSH_WRONG_PARAMETER;
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int pipeline_compact(TileDesc *d_tiles, int ntiles, double *ws,
        const int *d_keep, int kmax, hipStream_t s)
{
    if(ntiles <= 0)
        return STARSH_SUCCESS;
    hipLaunchKernelGGL(compact_kernel, dim3(ntiles), dim3(256), 0, s,
            d_tiles, ws, d_keep, kmax);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

int pipeline_truncate(TileDesc *d_tiles, int ntiles, double *ws,
        const double *d_bnorm, int kmax, double eta, int nmax,
        hipStream_t s)
{
    (void)nmax;
    if(ntiles <= 0)
        return STARSH_SUCCESS;
    (void)ws;
    hipLaunchKernelGGL(truncate_kernel, dim3((ntiles+255)/256), dim3(256), 0,
            s, d_tiles, ntiles, d_bnorm, kmax, eta);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}
