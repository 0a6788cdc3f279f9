// The allocation-free form of the div B diagnostic (mhdg_div_b_noalloc, include/mhdg.h), for callers that cannot
// allocate device memory: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
// Inputs and outputs are the caller's device arrays; the per-block partial results of the reduction live in
// `scratch`, a fixed device array of this file (kScratchBytes on every GPU the library runs on, held until the process
// ends).  Its device address is looked up once per device at context creation (noalloc_setup) and only that address
// is passed to kernels.  Every call on one GPU shares the array: calls that overlap on one GPU (two contexts' streams,
// or two host threads) race on it.
#include <array>
#include <mutex>
#include <stdexcept>
#include <string>

#include "ctx.h"
#include "kernels.h"

namespace mhdg
{
namespace
{
constexpr int32_t kNoAllocMaxBlocks = 65536; // blocks of 256 cells: at most 16,777,216 cells
constexpr size_t kScratchBytes = size_t(kNoAllocMaxBlocks) * (8 + 8 + 4);
constexpr int kMaxDevices = 64;
} // namespace

// per-block max |d2 B|, max relative value and first cell attaining the max (k_div_b's partials)
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace
{
std::mutex g_mu;
std::array<void *, kMaxDevices> g_ptr{};
std::array<bool, kMaxDevices> g_done{};
std::array<std::string, kMaxDevices> g_err;

void *scratch_of(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("mhdg_div_b_noalloc: device id out of range");
    if (!g_done[size_t(device)]) throw std::runtime_error("mhdg_div_b_noalloc: scratch not set up on this device");
    if (!g_ptr[size_t(device)]) throw std::runtime_error("mhdg_div_b_noalloc: scratch unavailable: " + g_err[size_t(device)]);
    return g_ptr[size_t(device)];
}
} // namespace

// Called by context creation with `device` current: the device address of `scratch` on that device, once.
void noalloc_setup(int device)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= kMaxDevices || g_done[size_t(device)]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) g_ptr[size_t(device)] = p;
    else g_err[size_t(device)] = e == hipSuccess ? "null address" : hipGetErrorString(e);
    g_done[size_t(device)] = true;
}

// One block: the partials of k_div_b in a fixed order, the largest |d2 B| with the smallest cell id among equal
// values (the host reduction of mhdg_div_b, so both entry points return the same numbers); out = {max |d2 B|, max
// relative value, worst cell or -1}.
__global__ void __launch_bounds__(256) k_div_b_final(const double *pa, const int32_t *pc, const double *pr, int32_t nb, double *out)
{
    __shared__ double sa[256], sr[256];
    __shared__ int32_t sc[256];
    const int t = threadIdx.x;
    double a = 0.0, r = 0.0;
    int32_t c = INT32_MAX;
    for (int32_t i = t; i < nb; i += 256)
    {
        if (pa[i] > a || (pa[i] == a && pa[i] > 0.0 && pc[i] < c)) { a = pa[i]; c = pc[i]; }
        r = mx(r, pr[i]);
    }
    sa[t] = a; sr[t] = r; sc[t] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (t < s)
        {
            if (sa[t + s] > sa[t] || (sa[t + s] == sa[t] && sa[t + s] > 0.0 && sc[t + s] < sc[t])) { sa[t] = sa[t + s]; sc[t] = sc[t + s]; }
            sr[t] = mx(sr[t], sr[t + s]);
        }
        __syncthreads();
    }
    if (t == 0)
    {
        out[0] = sa[0];
        out[1] = sr[0];
        out[2] = sc[0] == INT32_MAX ? -1.0 : double(sc[0]);
    }
}
} // namespace mhdg

using namespace mhdg;

extern "C" int mhdg_div_b_noalloc(mhdg_ctx *c, const double *B, const double *extra, double *out)
{
    try
    {
        if (!B || !out) throw std::runtime_error("mhdg_div_b_noalloc: null device pointer");
        const int64_t nb = (int64_t(c->nc) + 255) / 256;
        if (nb > kNoAllocMaxBlocks)
            throw std::runtime_error("mhdg_div_b_noalloc: " + std::to_string(c->nc) + " cells exceed the maximum of " +
                                     std::to_string(int64_t(kNoAllocMaxBlocks) * 256) + " (fixed scratch)");
        unsigned char *base = static_cast<unsigned char *>(scratch_of(c->device));
        double *pa = reinterpret_cast<double *>(base);
        double *pr = pa + kNoAllocMaxBlocks;
        int32_t *pc = reinterpret_cast<int32_t *>(pr + kNoAllocMaxBlocks);
        if (nb > 0) hipLaunchKernelGGL(k_div_b, dim3(uint32_t(nb)), dim3(256), 0, c->stream, c->d, B, extra, pa, pc, pr);
        hipLaunchKernelGGL(k_div_b_final, dim3(1), dim3(256), 0, c->stream, static_cast<const double *>(pa), static_cast<const int32_t *>(pc),
                           static_cast<const double *>(pr), int32_t(nb), out);
        c->check_launch("div_b_noalloc");
        return 0;
    }
    catch (const std::exception &e)
    {
        c->err = e.what();
        return 1;
    }
}
