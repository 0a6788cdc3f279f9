// Streaming calibration of the device's practical HBM ceiling: three full-grid kernels (one float4 per thread) over
// buffers far larger than the caches, timed with HIP events.  The benchmark reports the gather kernels' rates against
// these next to the datasheet peak.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
// sum of the buffer (one float4 per thread, a block sum written per block so the loads are not dead)
__global__ void __launch_bounds__(256) k_stream_read(const float4 *x, int64_t n4, float *part)
{
    __shared__ float sh[4];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float s = 0.0f;
    if (i < n4)
    {
        const float4 v = x[i];
        s = (v.x + v.y) + (v.z + v.w);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}
__global__ void __launch_bounds__(256) k_stream_copy(const float4 *x, int64_t n4, float4 *y)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n4) y[i] = x[i];
}
__global__ void __launch_bounds__(256) k_stream_triad(const float4 *b, const float4 *c, int64_t n4, float s, float4 *a)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n4)
    {
        const float4 u = b[i], v = c[i];
        a[i] = make_float4(u.x + s * v.x, u.y + s * v.y, u.z + s * v.z, u.w + s * v.w);
    }
}
void chk(hipError_t e, const char *w)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}
} // namespace

// mhdg_stream_calibration (include/mhdg.h): kind 0 read, 1 copy, 2 triad over arrays of `bytes` each; median of
// `repeats` launches in microseconds and the bytes each launch moves
extern "C" int mhdg_stream_calibration(int kind, int64_t bytes, int repeats, double *median_us, double *bytes_moved)
{
    float *x = nullptr, *y = nullptr, *z = nullptr, *part = nullptr;
    hipStream_t st = nullptr;
    std::vector<hipEvent_t> ev;
    int rc = 0;
    try
    {
        if (kind < 0 || kind > 2 || bytes < 1024 || repeats < 1 || repeats > 100) throw std::runtime_error("stream calibration: bad arguments");
        const int64_t n4 = bytes / 16;
        const int64_t blocks = (n4 + 255) / 256;
        if (blocks > INT32_MAX) throw std::runtime_error("stream calibration: buffer too large");
        chk(hipStreamCreate(&st), "stream");
        chk(hipMalloc(&x, size_t(n4) * 16), "hipMalloc");
        chk(hipMalloc(&y, size_t(n4) * 16), "hipMalloc");
        if (kind == 2) chk(hipMalloc(&z, size_t(n4) * 16), "hipMalloc");
        chk(hipMalloc(&part, size_t(blocks) * 4), "hipMalloc");
        chk(hipMemsetAsync(x, 0, size_t(n4) * 16, st), "memset");
        chk(hipMemsetAsync(y, 0, size_t(n4) * 16, st), "memset");
        if (z) chk(hipMemsetAsync(z, 0, size_t(n4) * 16, st), "memset");
        ev.resize(size_t(repeats) + 1);
        for (auto &e : ev) chk(hipEventCreate(&e), "event");
        auto launch = [&]() {
            const float4 *x4 = reinterpret_cast<const float4 *>(x);
            if (kind == 0) hipLaunchKernelGGL(k_stream_read, dim3(uint32_t(blocks)), dim3(256), 0, st, x4, n4, part);
            else if (kind == 1) hipLaunchKernelGGL(k_stream_copy, dim3(uint32_t(blocks)), dim3(256), 0, st, x4, n4, reinterpret_cast<float4 *>(y));
            else hipLaunchKernelGGL(k_stream_triad, dim3(uint32_t(blocks)), dim3(256), 0, st, x4, reinterpret_cast<const float4 *>(y), n4, 0.5f,
                                    reinterpret_cast<float4 *>(z));
        };
        launch(); // warm-up
        chk(hipEventRecord(ev[0], st), "event");
        for (int r = 0; r < repeats; ++r)
        {
            launch();
            chk(hipEventRecord(ev[size_t(r) + 1], st), "event");
        }
        chk(hipStreamSynchronize(st), "sync");
        std::vector<double> us(static_cast<size_t>(repeats));
        for (int r = 0; r < repeats; ++r)
        {
            float ms = 0;
            chk(hipEventElapsedTime(&ms, ev[size_t(r)], ev[size_t(r) + 1]), "elapsed");
            us[size_t(r)] = 1e3 * ms;
        }
        std::sort(us.begin(), us.end());
        *median_us = us[us.size() / 2];
        *bytes_moved = double(n4) * 16.0 * (kind == 0 ? 1 : kind == 1 ? 2 : 3);
    }
    catch (const std::exception &) { rc = 1; }
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    if (z) (void)hipFree(z);
    if (part) (void)hipFree(part);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}
