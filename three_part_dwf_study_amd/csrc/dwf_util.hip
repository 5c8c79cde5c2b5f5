// dwf_util.hip -- device memory, timing and copy-kernel helpers of the C ABI.

#include "../../include/dwf_hip.h"
#include "dwf_context.hpp"

struct dwf_timer {
    hipEvent_t a = nullptr, b = nullptr;
};

namespace {

__global__ void k_f64_to_f32(const double* a, float* b, const long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

// Plain copy of n 16-byte words with streaming (non-temporal) loads and stores: the
// achievable-bandwidth reference the stencil is compared with.
__global__ __launch_bounds__(256) void k_copy16(const float4* __restrict__ src, float4* __restrict__ dst,
                                                long long n)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        __builtin_nontemporal_store(__builtin_nontemporal_load(reinterpret_cast<const f4*>(src) + i),
                                    reinterpret_cast<f4*>(dst) + i);
}

int fail(hipError_t e, const char* what)
{
    dwf_set_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
    return DWF_ERR_HIP;
}

}  // namespace

extern "C" {

void* dwf_device_alloc(unsigned long long bytes)
{
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        fail(e, "dwf_device_alloc");
        return nullptr;
    }
    return p;
}

int dwf_device_free(void* ptr)
{
    const hipError_t e = hipFree(ptr);
    return e == hipSuccess ? DWF_OK : fail(e, "dwf_device_free");
}

int dwf_memcpy_htod(void* dst, const void* src, unsigned long long bytes)
{
    const hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? DWF_OK : fail(e, "dwf_memcpy_htod");
}

int dwf_memcpy_dtoh(void* dst, const void* src, unsigned long long bytes)
{
    const hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? DWF_OK : fail(e, "dwf_memcpy_dtoh");
}

int dwf_ctx_synchronize(dwf_ctx* ctx)
{
    if (!ctx) return DWF_ERR_ARG;
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? DWF_OK : fail(e, "dwf_ctx_synchronize");
}

dwf_timer* dwf_timer_create(void)
{
    dwf_timer* t = new dwf_timer();
    if (hipEventCreate(&t->a) != hipSuccess || hipEventCreate(&t->b) != hipSuccess) {
        dwf_set_error("hipEventCreate failed");
        delete t;
        return nullptr;
    }
    return t;
}

void dwf_timer_destroy(dwf_timer* t)
{
    if (!t) return;
    if (t->a) (void)hipEventDestroy(t->a);
    if (t->b) (void)hipEventDestroy(t->b);
    delete t;
}

int dwf_timer_start(dwf_timer* t, dwf_ctx* ctx)
{
    if (!t || !ctx) return DWF_ERR_ARG;
    const hipError_t e = hipEventRecord(t->a, ctx->stream);
    return e == hipSuccess ? DWF_OK : fail(e, "dwf_timer_start");
}

int dwf_timer_stop(dwf_timer* t, dwf_ctx* ctx)
{
    if (!t || !ctx) return DWF_ERR_ARG;
    const hipError_t e = hipEventRecord(t->b, ctx->stream);
    return e == hipSuccess ? DWF_OK : fail(e, "dwf_timer_stop");
}

double dwf_timer_elapsed_ms(dwf_timer* t)
{
    if (!t) return -1.0;
    if (hipEventSynchronize(t->b) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t->a, t->b) != hipSuccess) return -1.0;
    return ms;
}

int dwf_convert_f64_f32_device(dwf_ctx* ctx, const double* src, float* dst, unsigned long long n)
{
    if (!ctx || !dst || !src) {
        dwf_set_error("dwf_convert_f64_f32_device: bad argument");
        return DWF_ERR_ARG;
    }
    const long long blocks = std::min<long long>(((long long)n + 255) / 256, 256LL * 64);
    if (n) k_f64_to_f32<<<(unsigned)blocks, 256, 0, ctx->stream>>>(src, dst, (long long)n);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? DWF_OK : fail(e, "dwf_convert_f64_f32_device");
}

int dwf_copy_device(dwf_ctx* ctx, void* dst, const void* src, unsigned long long bytes)
{
    if (!ctx || !dst || !src || bytes % 16) {
        dwf_set_error("dwf_copy_device: bad argument");
        return DWF_ERR_ARG;
    }
    const long long n = (long long)(bytes / 16);
    const long long blocks = std::min<long long>((n + 255) / 256, 65536LL);
    k_copy16<<<(unsigned)blocks, 256, 0, ctx->stream>>>((const float4*)src, (float4*)dst, n);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? DWF_OK : fail(e, "dwf_copy_device");
}

}  // extern "C"
