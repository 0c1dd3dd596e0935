// Device selection, the library stream, memory helpers, the dlarnv test
// matrix and dense tile generation.
#include <stdio.h>
#include "hip_common.h"
#include "tlr.h"

static hipStream_t g_stream = nullptr;
static int g_device = -1;

extern "C" hipStream_t starsh_hip_stream(void)
{
    return g_stream;
}

extern "C" int starsh_hip_init(int device)
{
    if(g_stream != nullptr && device == g_device)
        return STARSH_SUCCESS;
    int count = 0;
    if(hipGetDeviceCount(&count) != hipSuccess || count == 0)
    {
        fprintf(stderr, "STARSH HIP ERROR: no HIP device is available\n");
        return STARSH_UNKNOWN_ERROR;
    }
    if(device < 0 || device >= count)
        return STARSH_WRONG_PARAMETER;
    HIP_TRY(hipSetDevice(device));
    if(g_stream == nullptr)
        HIP_TRY(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    // device address of the allocation-free product's scratch, once per
    // device, with the device current
    const int info = tlr_scratch_setup(device);
    if(info != STARSH_SUCCESS)
        return info;
    g_device = device;
    return STARSH_SUCCESS;
}

extern "C" int starsh_hip_current_device(void)
{
    return g_device;
}

extern "C" int starsh_hip_malloc(void **d_ptr, size_t bytes)
{
    if(d_ptr == nullptr)
        return STARSH_WRONG_PARAMETER;
    int info = starsh_hip_init(g_device < 0 ? 0 : g_device);
    if(info != STARSH_SUCCESS)
        return info;
    *d_ptr = nullptr;
    HIP_TRY(hipMalloc(d_ptr, bytes ? bytes : 1));
    return STARSH_SUCCESS;
}

extern "C" int starsh_hip_free(void *d_ptr)
{
    if(d_ptr != nullptr)
        HIP_TRY(hipFree(d_ptr));
    return STARSH_SUCCESS;
}

extern "C" int starsh_hip_memcpy_h2d(void *d_dst, const void *src,
        size_t bytes)
{
    HIP_TRY(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice,
                g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return STARSH_SUCCESS;
}

extern "C" int starsh_hip_memcpy_d2h(void *dst, const void *d_src,
        size_t bytes)
{
    HIP_TRY(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost,
                g_stream));
    HIP_TRY(hipStreamSynchronize(g_stream));
    return STARSH_SUCCESS;
}

extern "C" int starsh_hip_synchronize(void)
{
    if(g_stream != nullptr)
        HIP_TRY(hipStreamSynchronize(g_stream));
    return STARSH_SUCCESS;
}

// ---- Test matrix of the randomized SVD: dlarnv(3, {0,0,0,1}, n*k).
// Uniform number q (1-based) of the stream is (a^q mod 2^48) * 2^-48 with
// a = 33952834046453; normal number e (0-based) pairs uniforms 2e+1, 2e+2
// into sqrt(-2 log u1) cos(2 pi u2).  Each thread jumps ahead with a table
// of a^(2^b) mod 2^48.
__constant__ uint64_t c_apow[64];

__global__ void omega_kernel(double *__restrict__ out, int64_t count)
{
    const uint64_t mask = (1ULL << 48)-1;
    const uint64_t a = 33952834046453ULL;
    const double twopi = 6.2831853071795864769252867663;
    int64_t e = (int64_t)blockIdx.x*blockDim.x+threadIdx.x;
    if(e >= count)
        return;
    uint64_t q = 2*(uint64_t)e+1, s = 1;
    for(int b = 0; q != 0; b++, q >>= 1)
        if(q & 1)
            s = (s*c_apow[b]) & mask;
    double u1 = (double)s*0x1p-48;
    s = (s*a) & mask;
    double u2 = (double)s*0x1p-48;
    out[e] = sqrt(-2.0*log(u1))*cos(twopi*u2);
}

extern "C" int starsh_hip_omega(double *d_omega, int ncols, int k)
{
    int info = starsh_hip_init(g_device < 0 ? 0 : g_device);
    if(info != STARSH_SUCCESS)
        return info;
    if(ncols < 0 || k < 0 || (ncols > 0 && k > 0 && d_omega == nullptr))
        return STARSH_WRONG_PARAMETER;
    static bool table_ready = false;
    if(!table_ready)
    {
        uint64_t t[64], mask = (1ULL << 48)-1, x = 33952834046453ULL;
        for(int b = 0; b < 64; b++)
        {
            t[b] = x;
            x = (x*x) & mask;
        }
        HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_apow), t, sizeof(t)));
        table_ready = true;
    }
    int64_t count = (int64_t)ncols*k;
    if(count == 0)
        return STARSH_SUCCESS;
    int threads = 256;
    int64_t blocks = (count+threads-1)/threads;
    hipLaunchKernelGGL(omega_kernel, dim3((unsigned)blocks), dim3(threads), 0,
            g_stream, d_omega, count);
    HIP_TRY(hipGetLastError());
    return STARSH_SUCCESS;
}

// ---- Dense tile generation.  One block of 256 threads covers a 64 x 64
// patch of one tile; grid.y indexes tiles.
struct GenTile
{
    int64_t r0, c0, off;
    int m, n;
};

template <int NDIM>
__global__ __launch_bounds__(256) void generate_kernel(KernelParams kp,
        PointSets ps, const GenTile *__restrict__ tiles, int ntiles,
        double *__restrict__ out)
{
    const GenTile t = tiles[blockIdx.y];
    if(t.m <= 0 || t.n <= 0)
        return;
    int pm = (t.m+63)/64;
    int bi = blockIdx.x % pm, bj = blockIdx.x / pm;
    if(bj*64 >= t.n)
        return;
    int i = bi*64+(threadIdx.x & 63);
    if(i >= t.m)
        return;
    for(int jj = threadIdx.x >> 6; jj < 64; jj += 4)
    {
        int j = bj*64+jj;
        if(j >= t.n)
            break;
        out[t.off+(int64_t)j*t.m+i] = kernel_value<NDIM>(kp, ps.xr, ps.nr,
                t.r0+i, ps.xc, ps.nc, t.c0+j);
    }
}

extern "C" int starsh_hip_generate_tiles(const starsh_hip_kernel_desc *kd,
        const double *d_rowpts, int64_t nrowpts, const double *d_colpts,
        int64_t ncolpts, int ntiles, const int64_t *r0, const int64_t *c0,
        const int *m, const int *n, const int64_t *off, double *d_out)
{
    if(kd == nullptr || ntiles < 0 || kd->ndim < 1 || kd->ndim > 4)
        return STARSH_WRONG_PARAMETER;
    if(ntiles == 0)
        return STARSH_SUCCESS;
    int info = starsh_hip_init(g_device < 0 ? 0 : g_device);
    if(info != STARSH_SUCCESS)
        return info;
    GenTile *h = (GenTile *)malloc(sizeof(GenTile)*ntiles);
    if(h == nullptr)
        return STARSH_MALLOC_ERROR;
    int mmax = 0, nmax = 0;
    for(int t = 0; t < ntiles; t++)
    {
        if(m[t] < 0 || n[t] < 0 || r0[t] < 0 || c0[t] < 0 ||
                r0[t]+m[t] > nrowpts || c0[t]+n[t] > ncolpts)
        {
            free(h);
            return STARSH_WRONG_PARAMETER;
        }
        h[t] = GenTile{r0[t], c0[t], off[t], m[t], n[t]};
        mmax = m[t] > mmax ? m[t] : mmax;
        nmax = n[t] > nmax ? n[t] : nmax;
    }
    GenTile *d = nullptr;
    hipError_t e = hipMalloc((void **)&d, sizeof(GenTile)*ntiles);
    if(e == hipSuccess)
        e = hipMemcpyAsync(d, h, sizeof(GenTile)*ntiles,
                hipMemcpyHostToDevice, g_stream);
    if(e == hipSuccess && mmax > 0 && nmax > 0)
    {
        KernelParams kp = make_params(kd);
        PointSets ps{d_rowpts, nrowpts, d_colpts, ncolpts};
        // Blocks are laid out per tile as pm x pn patches; tiles smaller
        // than the largest one leave trailing blocks idle.
        int pm = (mmax+63)/64, pn = (nmax+63)/64;
        dim3 grid(pm*pn, ntiles);
        switch(kd->ndim)
        {
            case 1: hipLaunchKernelGGL(generate_kernel<1>, grid, dim3(256), 0,
                        g_stream, kp, ps, d, ntiles, d_out); break;
            case 2: hipLaunchKernelGGL(generate_kernel<2>, grid, dim3(256), 0,
                        g_stream, kp, ps, d, ntiles, d_out); break;
            case 3: hipLaunchKernelGGL(generate_kernel<3>, grid, dim3(256), 0,
                        g_stream, kp, ps, d, ntiles, d_out); break;
            default: hipLaunchKernelGGL(generate_kernel<4>, grid, dim3(256), 0,
                        g_stream, kp, ps, d, ntiles, d_out); break;
        }
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(g_stream);
    if(d != nullptr)
        (void)hipFree(d);
    free(h);
    HIP_TRY(e);
    HIP_TRY(e2);
    return STARSH_SUCCESS;
}
