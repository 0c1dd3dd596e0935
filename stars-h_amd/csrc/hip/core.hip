// Core of the randomized SVD: singular values and vectors of the small
// matrix R2 (p x p, with B[:p, :]^T = Q2 R2), the rank by the reference rule
// (src/backends/sequential/dense/dsvfr.c:16-47) and the matrix X from which
// U and V are formed.
//
// One-sided (Hestenes) Jacobi on the columns of R2^T gives R2^T J = W with
// orthogonal columns; the column norms of W are the singular values of B
// and X = W diag(sigma)^-1 holds B's left singular vectors.  Hence, with
// B = Us S Vs^T,
//   U = Q[:, :p] X[:, :r],   V = Vs S = B^T Us = Bt[:, :p] X[:, :r]
// (V = S Vs^T rows as in drsdd.c:83-87), and no rotation is accumulated.
// R2 comes from a QR of B's leading rows, which are graded like the
// singular values; Jacobi on R2^T converges in fewer sweeps than on R2.
#include <float.h>
#include "pipeline.h"
#include "jacobi_rotation.h"

namespace {

constexpr int CORE_THREADS = 256;     // tiles whose W fits in LDS
constexpr int CORE_THREADS_BIG = 1024; // W in global memory
constexpr int LDS_PMAX = 128;
constexpr int JACOBI_SWEEPS_DEFAULT = 60;

__device__ inline double group_sum(double v, int g)
{
    return group_sum_lanes(v, g);
}

// One-sided Jacobi step on the column pair (wa, wb) of length p by a group
// of g threads (gl = index in the group): returns true if it rotated.
__device__ inline bool rotate_pair(double *wa, double *wb, int p, int g,
        int gl, double ctol)
{
    double al = 0.0, be = 0.0, ga = 0.0;
    for(int i = gl; i < p; i += g)
    {
        double u = wa[i], v = wb[i];
        al += u*u;
        be += v*v;
        ga += u*v;
    }
    al = group_sum(al, g);
    be = group_sum(be, g);
    ga = group_sum(ga, g);
    if(!jr_rotates(al, be, ga, ctol))
        return false;
    double c, s;
    jacobi_rotation(al, be, ga, c, s);
    for(int i = gl; i < p; i += g)
    {
        double u = wa[i], v = wb[i];
        wa[i] = c*u-s*v;
        wb[i] = s*u+c*v;
    }
    return true;
}

// rotate_pair with both columns read into registers once (E rows per
// thread, p <= E g): the loads of a rotation are issued together and the
// update reads no memory.  The same sums in the same order (rows past p
// hold zeros, which leave every sum unchanged), so the same rotation.
template <int E>
__device__ inline bool rotate_pair_reg(double *wa, double *wb, int p, int g,
        int gl, double ctol)
{
    double u[E], v[E];
    double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
    for(int t = 0; t < E; t++)
    {
        const int i = gl+t*g;
        u[t] = i < p ? wa[i] : 0.0;
        v[t] = i < p ? wb[i] : 0.0;
    }
#pragma unroll
    for(int t = 0; t < E; t++)
    {
        al += u[t]*u[t];
        be += v[t]*v[t];
        ga += u[t]*v[t];
    }
    al = group_sum(al, g);
    be = group_sum(be, g);
    ga = group_sum(ga, g);
    if(!jr_rotates(al, be, ga, ctol))
        return false;
    double c, s;
    jacobi_rotation(al, be, ga, c, s);
#pragma unroll
    for(int q = 0; q < E; q++)
    {
        const int i = gl+q*g;
        if(i < p)
        {
            wa[i] = c*u[q]-s*v[q];
            wb[i] = s*u[q]+c*v[q];
        }
    }
    return true;
}

// Columns of up to CREG*64 rows held in registers by one wave (lane l holds
// rows l, l+64, ...).
constexpr int CREG = 16;

// rotate_pair with column a in the wave's registers (ca) and column b in
// memory: the same sums in the same order, so the same rotation; b is read
// once and written once.
__device__ inline bool rotate_reg(double (&ca)[CREG], double *wb, int p,
        int lane, double ctol)
{
    double cb[CREG];
    double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
    for(int t = 0; t < CREG; t++)
    {
        const int i = lane+64*t;
        cb[t] = i < p ? wb[i] : 0.0;
        if(i < p)
        {
            double u = ca[t], v = cb[t];
            al += u*u;
            be += v*v;
            ga += u*v;
        }
    }
    al = group_sum(al, 64);
    be = group_sum(be, 64);
    ga = group_sum(ga, 64);
    if(!jr_rotates(al, be, ga, ctol))
        return false;
    double c, s;
    jacobi_rotation(al, be, ga, c, s);
#pragma unroll
    for(int q = 0; q < CREG; q++)
    {
        const int i = lane+64*q;
        if(i < p)
        {
            double u = ca[q], v = cb[q];
            ca[q] = c*u-s*v;
            wb[i] = s*u+c*v;
        }
    }
    return true;
}

// W: p x p working matrix (ld = ldw), in LDS or global memory.  A group
// of g threads handles one column pair, g consecutive rows at a time.  In
// LDS the groups are as small as needed to give every pair of a round its
// own group; in global memory a whole wave takes a pair, so that each
// access of a column is one coalesced 512-byte segment.
template <int NT>
__device__ void jacobi(double *W, int ldw, int p, bool in_lds,
        int max_sweeps, int *sweeps_out)
{
    __shared__ int rotations;
    const int pe = p+(p & 1);
    const int npairs = pe/2;
    int g = 64;
    if(in_lds)
        while(g > 1 && g*npairs > NT)
            g >>= 1;
    const int ngroups = NT/g;
    const int grp = threadIdx.x/g, gl = threadIdx.x % g;
    const double ctol = (double)p*DBL_EPSILON;
    int sweep;
    for(sweep = 0; sweep < max_sweeps; sweep++)
    {
        if(threadIdx.x == 0)
            rotations = 0;
        __syncthreads();
        for(int round = 0; round < pe-1; round++)
        {
            for(int pi = grp; pi < npairs; pi += ngroups)
            {
                // Circle method: position 0 fixed, others rotate.
                int x = pi, y = pe-1-pi;
                int a = x == 0 ? 0 : 1+(x-1+round) % (pe-1);
                int b = 1+(y-1+round) % (pe-1);
                if(a > b)
                {
                    int tmp = a;
                    a = b;
                    b = tmp;
                }
                if(b >= p)
                    continue;
                double *wa = W+(int64_t)a*ldw, *wb = W+(int64_t)b*ldw;
                const bool rot = p <= 4*g ?
                    rotate_pair_reg<4>(wa, wb, p, g, gl, ctol) :
                    p <= 8*g ? rotate_pair_reg<8>(wa, wb, p, g, gl, ctol) :
                    rotate_pair(wa, wb, p, g, gl, ctol);
                if(rot && gl == 0)
                    atomicAdd(&rotations, 1);
            }
            __syncthreads();
        }
        if(rotations == 0)
            break;
        __syncthreads();
    }
    *sweeps_out = sweep;
}

// One-sided Jacobi for W (p x p, ld p) in global memory, one wave per
// column pair.  The columns are split into blocks of BW; a sweep visits
// every column pair once, in block-cyclic order: the pairs inside each
// block, then for every block pair (I, J), I < J, the BW x BW cross pairs
// in BW rounds of BW disjoint pairs.  The blocks being worked on sit in LDS
// (L: 2 BW columns of p doubles), so each block pair costs one load and
// one store of its columns instead of one pass over W per round.
__device__ void jacobi_blocked(double *W, int p, double *L, int bw,
        int max_sweeps, int *sweeps_out)
{
    __shared__ int rotations;
    const int nb = (p+bw-1)/bw;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nwaves = blockDim.x >> 6;
    const double ctol = (double)p*DBL_EPSILON;
    auto width = [&](int blk) { return min(bw, p-blk*bw); };
    auto load = [&](int blk, double *dst) {
        const int cnt = width(blk)*p;
        const double *src = W+(int64_t)blk*bw*p;
        for(int e = threadIdx.x; e < cnt; e += blockDim.x)
            dst[e] = src[e];
    };
    auto store = [&](int blk, const double *src) {
        const int cnt = width(blk)*p;
        double *dst = W+(int64_t)blk*bw*p;
        for(int e = threadIdx.x; e < cnt; e += blockDim.x)
            dst[e] = src[e];
    };
    int sweep;
    for(sweep = 0; sweep < max_sweeps; sweep++)
    {
        if(threadIdx.x == 0)
            rotations = 0;
        __syncthreads();
        for(int I = 0; I < nb; I++)
        {
            const int wi = width(I);
            load(I, L);
            __syncthreads();
            // pairs inside block I (circle method)
            const int we = wi+(wi & 1);
            for(int round = 0; round < we-1; round++)
            {
                for(int pi = wave; pi < we/2; pi += nwaves)
                {
                    int x = pi, y = we-1-pi;
                    int a = x == 0 ? 0 : 1+(x-1+round) % (we-1);
                    int b = 1+(y-1+round) % (we-1);
                    if(a > b)
                    {
                        int tmp = a;
                        a = b;
                        b = tmp;
                    }
                    if(b < wi && rotate_pair(L+(int64_t)a*p, L+(int64_t)b*p,
                                p, 64, lane, ctol) && lane == 0)
                        atomicAdd(&rotations, 1);
                }
                __syncthreads();
            }
            // cross pairs with every later block: wave a keeps column a of
            // block I in its registers for all of them (bw <= waves)
            double *LJ = L+(int64_t)bw*p;
            const int a = wave;
            const bool own = a < wi;
            double ca[CREG];
#pragma unroll
            for(int t = 0; t < CREG; t++)
            {
                const int i = lane+64*t;
                ca[t] = own && i < p ? L[(int64_t)a*p+i] : 0.0;
            }
            for(int J = I+1; J < nb; J++)
            {
                const int wj = width(J);
                load(J, LJ);
                __syncthreads();
                for(int round = 0; round < bw; round++)
                {
                    const int b = (a+round) % bw;
                    if(own && b < wj && rotate_reg(ca, LJ+(int64_t)b*p, p,
                                lane, ctol) && lane == 0)
                        atomicAdd(&rotations, 1);
                    __syncthreads();
                }
                store(J, LJ);
                __syncthreads();
            }
            if(own)
#pragma unroll
                for(int t = 0; t < CREG; t++)
                {
                    const int i = lane+64*t;
                    if(i < p)
                        L[(int64_t)a*p+i] = ca[t];
                }
            __syncthreads();
            store(I, L);
            __syncthreads();
        }
        if(rotations == 0)
            break;
        __syncthreads();
    }
    *sweeps_out = sweep;
}

// Column pairs handled by half-waves (32 lanes) for p <= CR32*32.
constexpr int CR32 = 17;

// rotate_reg for a group of 32 lanes (gl = lane in the group; rows gl,
// gl+32, ...).
__device__ inline bool rotate_reg32(double (&ca)[CR32], double *wb, int p,
        int gl, double ctol)
{
    double cb[CR32];
    double al = 0.0, be = 0.0, ga = 0.0;
#pragma unroll
    for(int t = 0; t < CR32; t++)
    {
        const int i = gl+32*t;
        cb[t] = i < p ? wb[i] : 0.0;
        if(i < p)
        {
            double u = ca[t], v = cb[t];
            al += u*u;
            be += v*v;
            ga += u*v;
        }
    }
    al = group_sum(al, 32);
    be = group_sum(be, 32);
    ga = group_sum(ga, 32);
    if(!jr_rotates(al, be, ga, ctol))
        return false;
    double c, s;
    jacobi_rotation(al, be, ga, c, s);
#pragma unroll
    for(int q = 0; q < CR32; q++)
    {
        const int i = gl+32*q;
        if(i < p)
        {
            double u = ca[q], v = cb[q];
            ca[q] = c*u-s*v;
            wb[i] = s*u+c*v;
        }
    }
    return true;
}

// jacobi_blocked with blocks of 32 columns, one column pair per half-wave
// (32 pairs at a time in 1024 threads), for p <= 544.  Only one block sits
// in LDS (L: 32 p doubles): block I is rotated there within itself, then
// half-wave a takes its column a into registers, and the later blocks J
// pass through L one after the other; column a goes back to W at the end.
// The same pairs in the same order as jacobi_blocked with bw = 32.
__device__ void jacobi_blocked32(double *W, int p, double *L, int max_sweeps,
        int *sweeps_out)
{
    constexpr int BW = 32;
    __shared__ int rotations;
    const int nb = (p+BW-1)/BW;
    const int hg = threadIdx.x >> 5, gl = threadIdx.x & 31;
    const int ngroups = blockDim.x >> 5;
    const double ctol = (double)p*DBL_EPSILON;
    auto width = [&](int blk) { return min(BW, p-blk*BW); };
    auto load = [&](int blk) {
        const int cnt = width(blk)*p;
        const double *src = W+(int64_t)blk*BW*p;
        for(int e = threadIdx.x; e < cnt; e += blockDim.x)
            L[e] = src[e];
    };
    auto store = [&](int blk) {
        const int cnt = width(blk)*p;
        double *dst = W+(int64_t)blk*BW*p;
        for(int e = threadIdx.x; e < cnt; e += blockDim.x)
            dst[e] = L[e];
    };
    int sweep;
    for(sweep = 0; sweep < max_sweeps; sweep++)
    {
        if(threadIdx.x == 0)
            rotations = 0;
        __syncthreads();
        for(int I = 0; I < nb; I++)
        {
            const int wi = width(I);
            load(I);
            __syncthreads();
            // pairs inside block I (circle method)
            const int we = wi+(wi & 1);
            for(int round = 0; round < we-1; round++)
            {
                for(int pi = hg; pi < we/2; pi += ngroups)
                {
                    int x = pi, y = we-1-pi;
                    int a = x == 0 ? 0 : 1+(x-1+round) % (we-1);
                    int b = 1+(y-1+round) % (we-1);
                    if(a > b)
                    {
                        int tmp = a;
                        a = b;
                        b = tmp;
                    }
                    if(b < wi && rotate_pair_reg<CR32>(L+(int64_t)a*p,
                                L+(int64_t)b*p, p, 32, gl, ctol) && gl == 0)
                        atomicAdd(&rotations, 1);
                }
                __syncthreads();
            }
            // cross pairs: half-wave a keeps column a of block I
            const int a = hg;
            const bool own = a < wi;
            double ca[CR32];
#pragma unroll
            for(int t = 0; t < CR32; t++)
            {
                const int i = gl+32*t;
                ca[t] = own && i < p ? L[(int64_t)a*p+i] : 0.0;
            }
            __syncthreads();
            for(int J = I+1; J < nb; J++)
            {
                const int wj = width(J);
                load(J);
                __syncthreads();
                for(int round = 0; round < BW; round++)
                {
                    const int b = (a+round) % BW;
                    if(own && b < wj && rotate_reg32(ca, L+(int64_t)b*p, p,
                                gl, ctol) && gl == 0)
                        atomicAdd(&rotations, 1);
                    __syncthreads();
                }
                store(J);
                __syncthreads();
            }
            if(own)
#pragma unroll
                for(int t = 0; t < CR32; t++)
                {
                    const int i = gl+32*t;
                    if(i < p)
                        W[((int64_t)I*BW+a)*p+i] = ca[t];
                }
            __syncthreads();
        }
        if(rotations == 0)
            break;
        __syncthreads();
    }
    *sweeps_out = sweep;
}

// Dynamic LDS: sig[pmax] doubles, then the Jacobi buffer (W, lds_p^2
// doubles, for the small tiles; two column blocks, 2 bw pmax doubles, for
// the big ones, or one block of 32 columns with bw = 32), then perm[pmax]
// ints.  With big == 0 the kernel takes the
// tiles with p <= lds_p (W in LDS); with big != 0 those with p > LDS_PMAX
// (W in their workspace, blocked Jacobi through LDS).
template <int NT>
__global__ __launch_bounds__(NT) void core_kernel(TileDesc *tiles,
        double *ws, int pmax, int lds_p, int big, int bw, int max_sweeps)
{
    extern __shared__ double dyn[];
    double *sig = dyn;
    double *lds = dyn+pmax;
    // bw = 32: one block in LDS (jacobi_blocked32); else two
    int *perm = (int *)(lds+(big ? (int64_t)(bw == 32 ? 1 : 2)*bw*pmax :
                (int64_t)lds_p*lds_p));
    __shared__ int s_rank;
    __shared__ unsigned long long s_wmax;
    TileDesc &td = tiles[blockIdx.x];
    const int p = td.p, k = td.k, maxrank = td.maxrank;
    const double tol = td.tol;
    if(big ? p <= LDS_PMAX : p > lds_p)
        return;
    if(p <= 0)
    {
        // zero sketch: all singular values vanish and the rank is 0
        for(int j = threadIdx.x; j < k; j += blockDim.x)
            ws[td.s_off+j] = 0.0;
        if(threadIdx.x == 0)
            td.rank = 0;
        return;
    }
    const double *R2 = ws+td.r_off;
    double *W;
    int ldw;
    if(p <= lds_p)
    {
        W = lds;
        ldw = p;
    }
    else
    {
        W = ws+td.r_off+(int64_t)k*k;
        ldw = p;
    }
    // W = R2^T, scaled by the power of two 2^-e that brings its largest
    // entry into [1/2, 1).  The scaling is exact, so the rotations are
    // those of R2^T itself, but squared norms and their squares inside the
    // rotation stay far from the bottom of the FP64 range for tiles whose
    // entries are tiny (far tiles of fast-decaying kernels).
    if(threadIdx.x == 0)
        s_wmax = 0;
    __syncthreads();
    double wmax = 0.0;
    for(int e = threadIdx.x; e < p*p; e += blockDim.x)
    {
        int i = e % p, j = e / p;
        const double v = R2[(int64_t)i*k+j];
        W[(int64_t)j*ldw+i] = v;
        wmax = fmax(wmax, fabs(v));
    }
    // non-negative doubles order like their bit patterns
    atomicMax(&s_wmax, (unsigned long long)__double_as_longlong(wmax));
    __syncthreads();
    int wexp = 0;
    const double wm = __longlong_as_double((long long)s_wmax);
    if(wm > 0.0 && wm < INFINITY)
        frexp(wm, &wexp);
    const double wscale = ldexp(1.0, -wexp), wunscale = ldexp(1.0, wexp);
    if(wexp != 0)
        for(int e = threadIdx.x; e < p*p; e += blockDim.x)
        {
            int i = e % p, j = e / p;
            W[(int64_t)j*ldw+i] *= wscale;
        }
    __syncthreads();
    int sweeps = 0;
    if(big)
    {
        // the blocked sweeps give each wave (bw <= 16) or half-wave
        // (bw = 32) one column of a block, so they run only in the 16-wave
        // kernel
        if constexpr(NT == CORE_THREADS_BIG)
        {
            if(bw == 32)
                jacobi_blocked32(W, p, lds, max_sweeps, &sweeps);
            else
                jacobi_blocked(W, p, lds, bw, max_sweeps, &sweeps);
        }
    }
    else
        jacobi<NT>(W, ldw, p, true, max_sweeps, &sweeps);
    __syncthreads();
    // Singular values: column norms of W.
    for(int j = threadIdx.x; j < p; j += blockDim.x)
    {
        double s = 0.0;
        for(int i = 0; i < p; i++)
            s += W[(int64_t)j*ldw+i]*W[(int64_t)j*ldw+i];
        sig[j] = sqrt(s);
    }
    __syncthreads();
    // Descending order, ties by index.
    for(int j = threadIdx.x; j < p; j += blockDim.x)
    {
        int pos = 0;
        double v = sig[j];
        for(int l = 0; l < p; l++)
            pos += sig[l] > v || (sig[l] == v && l < j);
        perm[pos] = j;
    }
    __syncthreads();
    double *S = ws+td.s_off;
    for(int j = threadIdx.x; j < k; j += blockDim.x)
    {
        double v = 0.0;
        if(j < p)
            v = sig[perm[j]]*wunscale;
        else if(j == p)
            v = sqrt(td.tail2);
        S[j] = v;
    }
    __syncthreads();
    if(threadIdx.x == 0)
    {
        // dsvfr over the k values, in the reference's order.
        int i;
        double err_tol = 0;
        for(i = 0; i < k; i++)
            err_tol += S[i]*S[i];
        int r;
        if(err_tol == 0)
            r = 0;
        else
        {
            i = k;
            err_tol *= tol*tol;
            double tmp = S[k-1]*S[k-1];
            while(i > 1 && err_tol >= tmp)
            {
                i--;
                tmp += S[i-1]*S[i-1];
            }
            r = i;
        }
        if(r > p)
            r = p;
        // a rank above maxrank, or a core SVD that did not converge within
        // the sweep limit (the counterpart of info != 0 from dgesdd), marks
        // a false far-field tile (drsdd.c:78,89-92)
        const bool unconverged = sweeps >= max_sweeps;
        // the exact engines also require rank < min(m, n)/2 (dsdd.c:53,
        // dqp3.c:84)
        const int mn = td.m < td.n ? td.m : td.n;
        if(r > maxrank || unconverged ||
                ((td.rule & RSDD_RULE_HALF) && r >= mn/2))
            r = -1;
        s_rank = r;
        td.rank = r;
        // bit 2: not converged; bits 8-15: sweeps, 16-26: p (diagnostics)
        td.flags |= (unconverged ? 4 : 0) | (min(sweeps, 255) << 8) |
            (p << 16);
    }
    __syncthreads();
    const int r = s_rank;
    if(r <= 0)
        return;
    // X (p x r): the r leading columns of W scaled to unit length
    double *X = ws+td.w_off;
    for(int e = threadIdx.x; e < p*r; e += blockDim.x)
    {
        int i = e % p, c = e / p;
        X[e] = W[(int64_t)perm[c]*ldw+i]/sig[perm[c]];
    }
}

} // namespace

// Sweep limit of the Jacobi core: a tile whose core has not converged
// after this many sweeps is reported as a false far-field tile.
static int jacobi_max_sweeps = JACOBI_SWEEPS_DEFAULT;

extern "C" int starsh_hip_jacobi_sweeps(int sweeps)
{
    if(sweeps < 0 || sweeps > JACOBI_SWEEPS_DEFAULT)
        return STARSH_WRONG_PARAMETER;
    jacobi_max_sweeps = sweeps == 0 ? JACOBI_SWEEPS_DEFAULT : sweeps;
    return STARSH_SUCCESS;
}

int pipeline_core(TileDesc *d_tiles, int ntiles, double *ws, int pmax,
        hipStream_t s)
{
    if(ntiles <= 0)
        return STARSH_SUCCESS;
    if(pmax > 1024)
        return STARSH_WRONG_PARAMETER;
    const int pm = pmax > 0 ? pmax : 1;
    const int lds_p = pm <= LDS_PMAX ? pm : LDS_PMAX;
    const size_t bytes = sizeof(double)*((size_t)pm+(size_t)lds_p*lds_p)+
        sizeof(int)*(size_t)pm;
    HIP_TRY(hipFuncSetAttribute((const void *)core_kernel<CORE_THREADS>,
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    hipLaunchKernelGGL(core_kernel<CORE_THREADS>, dim3(ntiles),
            dim3(CORE_THREADS), bytes, s, d_tiles, ws, pm, lds_p, 0, 0,
            jacobi_max_sweeps);
    HIP_TRY(hipGetLastError());
    if(pm > LDS_PMAX)
    {
        // one block of 32 columns in LDS up to 32*CR32 rows (half-waves
        // hold their column in registers), else two blocks of 16 columns
        // up to 576 rows, 8 beyond (at most 160 KiB with sig and perm)
        const int bw = pm <= 32*CR32 ? 32 : (pm <= 576 ? 16 : 8);
        const size_t bb = sizeof(double)*((size_t)pm+
                (size_t)(bw == 32 ? 1 : 2)*bw*pm)+sizeof(int)*(size_t)pm;
        HIP_TRY(hipFuncSetAttribute(
                    (const void *)core_kernel<CORE_THREADS_BIG>,
                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)bb));
        hipLaunchKernelGGL(core_kernel<CORE_THREADS_BIG>, dim3(ntiles),
                dim3(CORE_THREADS_BIG), bb, s, d_tiles, ws, pm, 0, 1, bw,
                jacobi_max_sweeps);
        HIP_TRY(hipGetLastError());
    }
    return STARSH_SUCCESS;
}
