// Morton (Z-order) sort of particles on the GPU, with the exact order of
// the reference's starsh_particles_zsort_inplace
// (src/applications/particles.c:1090-1221 of ecrc/stars-h 0.3.1).
//
// The reference scales every coordinate to u = (x - min)/(max - min) *
// UINT32_MAX (truncated to uint32) and sorts by a recursive binary
// partition that visits, from bit 31 down to bit 0, dimensions ndim-1 down
// to 0, moving 0-bit points to the front in order and 1-bit points to the
// back in reverse order.  The result is the order of the interleaved key
//   key = sum_{b, d} bit b of u_d at position b*ndim + d,
// with points of equal key in their original order when the key has an
// even number of 1 bits and in reverse order when odd (each 1 bit reverses
// them once).  Here that order is produced by a stable LSD radix sort on
// the composite key (Morton key, then the original index or its reverse by
// that parity), 8 bits per pass:
//   zs_minmax_*   per-dimension min and max (exact, any reduction order);
//   zs_key        the key words of every point;
//   zs_hist       per-tile digit counts;  zs_scan  offsets (digit-major);
//   zs_scatter    stable scatter: each tile ranks its points by digit in
//                 their order, with wave ballots for ranks inside a wave;
//   zs_gather     the permuted coordinates.
#include <stdio.h>
#include <algorithm>
#include <utility>
#include "hip_common.h"

namespace {

constexpr int ZT = 256;          // threads per block
constexpr int ZI = 4;            // points per thread per tile
constexpr int TILE = ZT*ZI;      // points per tile
constexpr int NDIG = 256;        // 8-bit digits

__global__ __launch_bounds__(ZT) void zs_minmax_partial(const double *x,
        int64_t n, int ndim, double *part)
{
    // part[(b * ndim + d) * 2 + {0, 1}] = min, max of block b
    __shared__ double smin[ZT], smax[ZT];
    for(int d = 0; d < ndim; d++)
    {
        double lo = INFINITY, hi = -INFINITY;
        for(int64_t i = (int64_t)blockIdx.x*ZT+threadIdx.x; i < n;
                i += (int64_t)gridDim.x*ZT)
        {
            const double v = x[(int64_t)d*n+i];
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
        smin[threadIdx.x] = lo;
        smax[threadIdx.x] = hi;
        __syncthreads();
        for(int o = ZT/2; o > 0; o >>= 1)
        {
            if(threadIdx.x < o)
            {
                smin[threadIdx.x] = fmin(smin[threadIdx.x],
                        smin[threadIdx.x+o]);
                smax[threadIdx.x] = fmax(smax[threadIdx.x],
                        smax[threadIdx.x+o]);
            }
            __syncthreads();
        }
        if(threadIdx.x == 0)
        {
            part[((int64_t)blockIdx.x*ndim+d)*2] = smin[0];
            part[((int64_t)blockIdx.x*ndim+d)*2+1] = smax[0];
        }
        __syncthreads();
    }
}

__global__ void zs_minmax_final(const double *part, int nb, int ndim,
        double *mm)
{
    // one thread per dimension: mm[d] = min, mm[ndim + d] = max
    const int d = threadIdx.x;
    if(d >= ndim)
        return;
    double lo = INFINITY, hi = -INFINITY;
    for(int b = 0; b < nb; b++)
    {
        lo = fmin(lo, part[((int64_t)b*ndim+d)*2]);
        hi = fmax(hi, part[((int64_t)b*ndim+d)*2+1]);
    }
    mm[d] = lo;
    mm[ndim+d] = hi;
}

// Key words of point i (SoA, nw words of n): word 0 = tie-break (original
// index, or its reverse when the key has an odd number of 1 bits), words
// 1..ndim = the Morton key, least significant word first.
__global__ __launch_bounds__(ZT) void zs_key(const double *x, int64_t n,
        int ndim, const double *mm, uint32_t *key, uint32_t *idx)
{
    const int64_t i = (int64_t)blockIdx.x*ZT+threadIdx.x;
    if(i >= n)
        return;
    uint32_t u[4];
    for(int d = 0; d < ndim; d++)
    {
        const double lo = mm[d], range = mm[ndim+d]-lo;
        // the reference's expression: (x - min)/range*UINT32_MAX, then
        // conversion to uint32 (truncation)
        const double v = (x[(int64_t)d*n+i]-lo)/range*4294967295.0;
        u[d] = (uint32_t)v;
    }
    // bit b of u_d goes to key bit b*ndim + d; key word w holds bits
    // 32w .. 32w+31
    uint32_t w[4] = {0, 0, 0, 0};
    for(int b = 0; b < 32; b++)
        for(int d = 0; d < ndim; d++)
        {
            const int pos = b*ndim+d;
            w[pos >> 5] |= ((u[d] >> b) & 1u) << (pos & 31);
        }
    int ones = 0;
    for(int q = 0; q < ndim; q++)
        ones += __popc(w[q]);
    key[i] = (ones & 1) ? (uint32_t)(n-1-i) : (uint32_t)i;
    for(int q = 0; q < ndim; q++)
        key[(int64_t)(q+1)*n+i] = w[q];
    idx[i] = (uint32_t)i;
}

__device__ inline uint32_t digit_of(const uint32_t *key, int64_t n, int word,
        int shift, int64_t i)
{
    return (key[(int64_t)word*n+i] >> shift) & (NDIG-1);
}

__global__ __launch_bounds__(ZT) void zs_hist(const uint32_t *key,
        int64_t n, int word, int shift, uint32_t *hist, int ntile)
{
    __shared__ uint32_t cnt[NDIG];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t t0 = (int64_t)blockIdx.x*TILE;
    for(int r = 0; r < ZI; r++)
    {
        const int64_t i = t0+r*ZT+threadIdx.x;
        if(i < n)
            atomicAdd(&cnt[digit_of(key, n, word, shift, i)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x*ntile+blockIdx.x] = cnt[threadIdx.x];
}

// Exclusive scan of m values in one block of 1024 threads (sequential
// chunks, running total carried between them).
__global__ __launch_bounds__(1024) void zs_scan(uint32_t *v, int64_t m)
{
    __shared__ uint32_t s[1024];
    __shared__ uint32_t carry;
    if(threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    for(int64_t c0 = 0; c0 < m; c0 += 1024)
    {
        const int64_t i = c0+threadIdx.x;
        const uint32_t x = i < m ? v[i] : 0;
        s[threadIdx.x] = x;
        __syncthreads();
        for(int o = 1; o < 1024; o <<= 1)
        {
            const uint32_t y = threadIdx.x >= (unsigned)o ?
                s[threadIdx.x-o] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        if(i < m)
            v[i] = carry+s[threadIdx.x]-x;
        __syncthreads();
        if(threadIdx.x == 1023)
            carry += s[1023];
        __syncthreads();
    }
}

__global__ __launch_bounds__(ZT) void zs_scatter(const uint32_t *key,
        const uint32_t *idx, uint32_t *key2, uint32_t *idx2, int64_t n,
        int nw, int word, int shift, const uint32_t *off, int ntile)
{
    __shared__ uint32_t base[NDIG];      // tile offset + points placed so far
    __shared__ uint32_t wcnt[ZT/64][NDIG];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    base[threadIdx.x] = off[(int64_t)threadIdx.x*ntile+blockIdx.x];
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64-lane));
    const int64_t t0 = (int64_t)blockIdx.x*TILE;
    for(int r = 0; r < ZI; r++)
    {
        for(int w = 0; w < ZT/64; w++)
            wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        // points of this round, in order: thread t takes t0 + r*ZT + t
        const int64_t i = t0+r*ZT+threadIdx.x;
        const bool valid = i < n;
        const uint32_t dg = valid ? digit_of(key, n, word, shift, i) : 0;
        // lanes of this wave with the same digit
        uint64_t peers = __ballot(valid);
        for(int b = 0; b < 8; b++)
        {
            const uint64_t m = __ballot(((dg >> b) & 1u) != 0);
            peers &= ((dg >> b) & 1u) ? m : ~m;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt);
        if(valid && before == 0)
            wcnt[wave][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        if(valid)
        {
            uint32_t pos = base[dg]+before;
            for(int w = 0; w < wave; w++)
                pos += wcnt[w][dg];
            // pos < n by construction (offsets of a full histogram); the
            // test keeps a wrong offset from writing out of bounds
            if(pos >= (uint64_t)n)
                pos = (uint32_t)(n-1);
            for(int q = 0; q < nw; q++)
                key2[(int64_t)q*n+pos] = key[(int64_t)q*n+i];
            idx2[pos] = idx[i];
        }
        __syncthreads();
        uint32_t add = 0;
        for(int w = 0; w < ZT/64; w++)
            add += wcnt[w][threadIdx.x];
        base[threadIdx.x] += add;
        __syncthreads();
    }
}

__global__ __launch_bounds__(ZT) void zs_gather(const double *x, int64_t n,
        int ndim, const uint32_t *idx, double *y, int64_t *order)
{
    const int64_t i = (int64_t)blockIdx.x*ZT+threadIdx.x;
    if(i >= n)
        return;
    uint32_t j = idx[i];
    j = j < (uint64_t)n ? j : 0;
    for(int d = 0; d < ndim; d++)
        y[(int64_t)d*n+i] = x[(int64_t)d*n+j];
    if(order)
        order[i] = j;
}

} // namespace

// Sorts n points (SoA, ndim x n, at d_x) into Morton order: d_y gets the
// permuted coordinates, d_order (optional) the permutation (d_y[., i] =
// d_x[., d_order[i]]).  d_x and d_y must not overlap.
extern "C" int starsh_hip_zsort(const double *d_x, int64_t n, int ndim,
        double *d_y, int64_t *d_order)
{
    if(n < 0 || n > 0x7fffffff || ndim < 1 || ndim > 4 ||
            (n > 0 && (d_x == nullptr || d_y == nullptr)))
        return STARSH_WRONG_PARAMETER;
    if(n == 0)
        return STARSH_SUCCESS;
    int info = starsh_hip_init(0);
    if(info != STARSH_SUCCESS)
        return info;
    hipStream_t s = starsh_hip_stream();
    const int nw = ndim+1;
    const int ntile = (int)((n+TILE-1)/TILE);
    const int nbmm = (int)std::min<int64_t>((n+ZT-1)/ZT, 1024);
    // scratch: min/max partials, key words and indices twice, histogram
    size_t bytes = sizeof(double)*((size_t)nbmm*ndim*2+2*ndim)+
        sizeof(uint32_t)*(2*(size_t)n*nw+2*(size_t)n+(size_t)NDIG*ntile);
    char *buf = nullptr;
    HIP_TRY(hipMallocAsync((void **)&buf, bytes, s));
    double *part = (double *)buf;
    double *mm = part+(size_t)nbmm*ndim*2;
    uint32_t *key = (uint32_t *)(mm+2*ndim);
    uint32_t *key2 = key+(size_t)n*nw;
    uint32_t *idx = key2+(size_t)n*nw;
    uint32_t *idx2 = idx+n;
    uint32_t *hist = idx2+n;
    const unsigned nb = (unsigned)((n+ZT-1)/ZT);
    hipLaunchKernelGGL(zs_minmax_partial, dim3(nbmm), dim3(ZT), 0, s, d_x, n,
            ndim, part);
    hipLaunchKernelGGL(zs_minmax_final, dim3(1), dim3(64), 0, s, part, nbmm,
            ndim, mm);
    hipLaunchKernelGGL(zs_key, dim3(nb), dim3(ZT), 0, s, d_x, n, ndim, mm,
            key, idx);
    // word 0 (tie-break) needs only the bits of indices below n
    int tbits = 0;
    while(tbits < 32 && ((int64_t)1 << tbits) < n)
        tbits++;
    for(int word = 0; word < nw; word++)
        for(int shift = 0; shift < 32; shift += 8)
        {
            if(word == 0 && shift >= tbits)
                break;
            hipLaunchKernelGGL(zs_hist, dim3(ntile), dim3(ZT), 0, s, key, n,
                    word, shift, hist, ntile);
            hipLaunchKernelGGL(zs_scan, dim3(1), dim3(1024), 0, s, hist,
                    (int64_t)NDIG*ntile);
            hipLaunchKernelGGL(zs_scatter, dim3(ntile), dim3(ZT), 0, s, key,
                    idx, key2, idx2, n, nw, word, shift, hist, ntile);
            std::swap(key, key2);
            std::swap(idx, idx2);
        }
    hipLaunchKernelGGL(zs_gather, dim3(nb), dim3(ZT), 0, s, d_x, n, ndim,
            idx, d_y, d_order);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipFreeAsync(buf, s));
    HIP_TRY(hipStreamSynchronize(s));
    return STARSH_SUCCESS;
}
