// Attainable FP64 rates of the GPU, measured, for reading the roofline
// fractions of the FP64 kernels against what the chip sustains rather than
// its published peak: a grid of 4096 workgroups of independent
// v_mfma_f64_16x16x4f64 chains (eight accumulators per wave), then the same
// grid of FP64 fused multiply-adds on the vector ALUs.
#include "hip_common.h"

typedef double double4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int BLOCKS = 4096, CHAINS = 8;

__global__ __launch_bounds__(256) void mfma_probe(double *out, int iters)
{
    const double a = threadIdx.x*1e-3, b = 1.0+blockIdx.x*1e-6;
    double4_t acc[CHAINS];
#pragma unroll
    for(int i = 0; i < CHAINS; i++)
        acc[i] = double4_t{0.0, 0.0, 0.0, 0.0};
    for(int it = 0; it < iters; it++)
#pragma unroll
        for(int i = 0; i < CHAINS; i++)
            acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0,
                    0);
    double s = 0.0;
#pragma unroll
    for(int i = 0; i < CHAINS; i++)
        s += acc[i][0]+acc[i][1]+acc[i][2]+acc[i][3];
    out[(int64_t)blockIdx.x*blockDim.x+threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void valu_probe(double *out, int iters)
{
    const double a = threadIdx.x*1e-3, b = 1.0-blockIdx.x*1e-9;
    double x[CHAINS];
#pragma unroll
    for(int i = 0; i < CHAINS; i++)
        x[i] = i;
    for(int it = 0; it < iters; it++)
#pragma unroll
        for(int i = 0; i < CHAINS; i++)
            x[i] = fma(x[i], b, a);
    double s = 0.0;
#pragma unroll
    for(int i = 0; i < CHAINS; i++)
        s += x[i];
    out[(int64_t)blockIdx.x*blockDim.x+threadIdx.x] = s;
}

template <typename K>
int time_kernel(K kernel, double *d, int iters, hipStream_t s, float *ms)
{
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    hipLaunchKernelGGL(kernel, dim3(BLOCKS), dim3(256), 0, s, d, 16);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e0, s));
    hipLaunchKernelGGL(kernel, dim3(BLOCKS), dim3(256), 0, s, d, iters);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return STARSH_SUCCESS;
}

} // namespace

extern "C" int starsh_hip_fp64_probe(double *mfma_tflops, double *valu_tflops)
{
    if(mfma_tflops == nullptr || valu_tflops == nullptr)
        return STARSH_WRONG_PARAMETER;
    hipStream_t s = starsh_hip_stream();
    double *d = nullptr;
    HIP_TRY(hipMalloc((void **)&d, sizeof(double)*BLOCKS*256));
    float ms_m = 0.f, ms_v = 0.f;
    const int it_m = 2000, it_v = 16000;
    int info = time_kernel(mfma_probe, d, it_m, s, &ms_m);
    if(info == STARSH_SUCCESS)
        info = time_kernel(valu_probe, d, it_v, s, &ms_v);
    (void)hipFree(d);
    if(info != STARSH_SUCCESS)
        return info;
    // 2*16*16*4 flop per MFMA and wave, four waves per workgroup; two flop
    // per fused multiply-add and lane
    const double f_m = 2.0*1024*CHAINS*(double)it_m*BLOCKS*4;
    const double f_v = 2.0*CHAINS*(double)it_v*BLOCKS*256;
    *mfma_tflops = f_m/(ms_m*1e-3)/1e12;
    *valu_tflops = f_v/(ms_v*1e-3)/1e12;
    return STARSH_SUCCESS;
}
