// FP64 GMRES with the mixed-precision solve's orthogonalisation (precision 3): the reference system and iteration
// (ImplicitInduction::solve, Common/induction/ImplicitInduction.cpp:299-390: Jacobi, warm start, restarts, true
// residual per restart) with classical Gram-Schmidt in place of modified Gram-Schmidt: one multi-dot pass and one
// multi-axpy pass per Arnoldi step (DGKS re-projection when the first removed more than 1 - kDgks64 of the norm),
// all in FP64.  It is the FP64 counterpart of the mixed inner solve, so that the mixed path's gain over the reference
// iteration can be split into its precision part and its orthogonalisation part.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "ctx.h"

namespace mhdg
{
namespace
{
constexpr int kChunk = 16;       // basis vectors per multi-dot pass
constexpr int kBlocks = 1024;    // grid of the orthogonalisation kernels
constexpr double kDgks64 = 0.1;  // as the mixed solve

__device__ __forceinline__ double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
inline int nbk(int64_t n) { return static_cast<int>((n + 255) / 256); }
void chk(hipError_t e, const char *w)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}
} // namespace

// per-block partials of <V_{i0+k}, w> (k < nv) and <w, w> (slot kChunk), two elements per thread
__global__ void __launch_bounds__(256) k_multidot_f64(const double *const *V, int32_t i0, int32_t nv, const double *w, int32_t n, double *part)
{
    __shared__ double sh[4][kChunk + 1];
    double acc[kChunk + 1];
#pragma unroll
    for (int k = 0; k <= kChunk; ++k) acc[k] = 0.0;
    const double *Vk[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) Vk[k] = k < nv ? V[i0 + k] : V[i0];
    const int32_t n2 = n >> 1;
    const double2 *w2 = reinterpret_cast<const double2 *>(w);
    for (int32_t q = blockIdx.x * 256 + threadIdx.x; q < n2; q += gridDim.x * 256)
    {
        const double2 a = w2[q];
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (k < nv)
            {
                const double2 b = reinterpret_cast<const double2 *>(Vk[k])[q];
                acc[k] += a.x * b.x + a.y * b.y;
            }
        acc[kChunk] += a.x * a.x + a.y * a.y;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1))
    {
        const double a = w[n - 1];
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (k < nv) acc[k] += a * Vk[k][n - 1];
        acc[kChunk] += a * a;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= kChunk; ++k)
    {
        const double s = wsum(acc[k]);
        if (lane == 0) sh[wv][k] = s;
    }
    __syncthreads();
    if (threadIdx.x <= kChunk)
    {
        const int k = threadIdx.x;
        part[size_t(blockIdx.x) * (kChunk + 1) + k] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
    }
}
// out[i0 + k] = sum of the partials (k < nv), out[63] = <w, w> when with_ww (one block, fixed order)
__global__ void __launch_bounds__(256) k_multidot_f64_final(const double *part, int32_t nblocks, int32_t i0, int32_t nv, int32_t with_ww,
                                                            double *out)
{
    __shared__ double sh[4][kChunk + 1];
    double t[kChunk + 1];
#pragma unroll
    for (int k = 0; k <= kChunk; ++k) t[k] = 0.0;
    for (int32_t bl = threadIdx.x; bl < nblocks; bl += 256)
#pragma unroll
        for (int k = 0; k <= kChunk; ++k) t[k] += part[size_t(bl) * (kChunk + 1) + k];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= kChunk; ++k)
    {
        const double s = wsum(t[k]);
        if (lane == 0) sh[wv][k] = s;
    }
    __syncthreads();
    if (threadIdx.x <= kChunk)
    {
        const int k = threadIdx.x;
        const double s = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
        if (k < nv) out[i0 + k] = s;
        if (k == kChunk && with_ww) out[63] = s;
    }
}
// w -= sum_i h_i V_i (i < nv); then V_next = w / |w| with |w|^2 = <w,w> - sum h_i^2 (Pythagoras) when that keeps at
// least kDgks64 of the norm (h[nv] = |w|, h[62] = 0), else w is kept for a second projection and the per-block
// partials of |w|^2 go to part (h[62] = 1)
__global__ void __launch_bounds__(256) k_multiaxpy_f64(const double *const *V, int32_t nv, double *h, double *w, int32_t n, int32_t allow_direct,
                                                       double *vnext, double *part)
{
    __shared__ double sh[4];
    __shared__ double hs[64];
    __shared__ const double *vp[64];
    if (threadIdx.x < nv) { hs[threadIdx.x] = h[threadIdx.x]; vp[threadIdx.x] = V[threadIdx.x]; }
    __syncthreads();
    double ss = 0.0;
    for (int32_t i = 0; i < nv; ++i) ss += hs[i] * hs[i];
    const double ww = h[63];
    const double hn2 = ww - ss;
    const double hp = hn2 > 0 ? sqrt(hn2) : 0.0;
    const bool direct = allow_direct && hp >= kDgks64 * sqrt(ww > 0 ? ww : 0.0) && hp > 1e-300;
    const double inv = direct ? 1.0 / hp : 0.0;
    double acc = 0.0;
    for (int32_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256)
    {
        double x = w[e];
        for (int32_t i = 0; i < nv; ++i) x -= hs[i] * vp[i][e];
        if (direct) vnext[e] = x * inv;
        else
        {
            w[e] = x;
            acc += x * x;
        }
    }
    if (direct)
    {
        if (blockIdx.x == 0 && threadIdx.x == 0) { h[nv] = hp; h[62] = 0.0; }
        return;
    }
    const double s = wsum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
        if (blockIdx.x == 0) h[62] = 1.0;
    }
}
// |w| from the partials of k_multiaxpy_f64 into h[slot] (one block, fixed order)
__global__ void __launch_bounds__(256) k_norm_final(const double *part, int32_t nblocks, double *h, int32_t slot)
{
    __shared__ double sh[4];
    double t = 0.0;
    for (int32_t bl = threadIdx.x; bl < nblocks; bl += 256) t += part[bl];
    const double s = wsum(t);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) h[slot] = sqrt(((sh[0] + sh[1]) + sh[2]) + sh[3]);
}
} // namespace mhdg

using namespace mhdg;

#define LAUNCH(NAME, ...)                                                                                                    \
    do                                                                                                                       \
    {                                                                                                                        \
        prof_begin(NAME);                                                                                                    \
        hipLaunchKernelGGL(__VA_ARGS__);                                                                                     \
        prof_end();                                                                                                          \
    } while (0)

int mhdg_ctx::solve64_cgs(const double *Phi_n, const double *Eex, const int8_t *dir, double dt, double *Eout, double *rel_res)
{
    const int m = restart;
    const int max_it = cfg.ind_max_it;
    const double rtol = rtol_override > 0 ? rtol_override : cfg.ind_rtol;
    if (!cgs_part)
    {
        cgs_part = static_cast<double *>(alloc(size_t(kBlocks) * (kChunk + 1) * sizeof(double)));
        cgs_h = static_cast<double *>(alloc(64 * sizeof(double)));
    }
    LAUNCH("solve_init", k_solve_init, dim3(nbk(ne)), dim3(256), 0, stream, Eex, have_corr ? corr : nullptr, dir, d.m_e, d.d2, dt,
           dte_cur, ne, x0, invdiag);
    rhs_residual(Phi_n, Eex, x0, dir, dt); // b = E_expl + F(Phi^n), r = b - A x0
    chk(hipMemsetAsync(delta, 0, size_t(ne) * 8, stream), "memset");
    const double bb = dot(bvec, bvec, 0, 0);
    const double bnorm = std::sqrt(std::max(bb, 1e-300));
    double beta = dot(rvec, rvec, 1, 1);
    int total = 0;
    std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> cs(m), sn(m), g(m + 1), yv(m), hc(64), c2(64);
    const double *const *Vp = Vptrs;
    while (beta > rtol * bnorm && total < max_it)
    {
        LAUNCH("scale", k_div_scalar, dim3(nbk(ne)), dim3(256), 0, stream, rvec, scal + 1, ne, V[0]);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m && total < max_it; ++j, ++total)
        {
            apply_op(nullptr, V[j], invdiag, dir, dt, 1, 1, nullptr, wvec);
            for (int i0 = 0; i0 <= j; i0 += kChunk)
            {
                const int nv = std::min(kChunk, j + 1 - i0);
                LAUNCH("multidot64", k_multidot_f64, dim3(kBlocks), dim3(256), 0, stream, Vp, i0, nv, static_cast<const double *>(wvec), ne,
                       cgs_part);
                LAUNCH("multidot64_final", k_multidot_f64_final, dim3(1), dim3(256), 0, stream, static_cast<const double *>(cgs_part), kBlocks,
                       i0, nv, i0 == 0 ? 1 : 0, cgs_h);
            }
            LAUNCH("multiaxpy64", k_multiaxpy_f64, dim3(kBlocks), dim3(256), 0, stream, Vp, j + 1, cgs_h, wvec, ne, 1, V[j + 1], cgs_part);
            d2h(hc.data(), cgs_h, sizeof(double) * 64);
            double hn = hc[j + 1];
            if (hc[62] != 0.0)
            {
                // DGKS: a second projection of the kept w, then its norm
                for (int i0 = 0; i0 <= j; i0 += kChunk)
                {
                    const int nv = std::min(kChunk, j + 1 - i0);
                    LAUNCH("multidot64", k_multidot_f64, dim3(kBlocks), dim3(256), 0, stream, Vp, i0, nv, static_cast<const double *>(wvec), ne,
                           cgs_part);
                    LAUNCH("multidot64_final", k_multidot_f64_final, dim3(1), dim3(256), 0, stream, static_cast<const double *>(cgs_part),
                           kBlocks, i0, nv, 0, cgs_h);
                }
                LAUNCH("multiaxpy64", k_multiaxpy_f64, dim3(kBlocks), dim3(256), 0, stream, Vp, j + 1, cgs_h, wvec, ne, 0, V[j + 1], cgs_part);
                LAUNCH("norm64", k_norm_final, dim3(1), dim3(256), 0, stream, static_cast<const double *>(cgs_part), kBlocks, cgs_h, j + 1);
                d2h(c2.data(), cgs_h, sizeof(double) * 64);
                for (int i = 0; i <= j; ++i) hc[i] += c2[i];
                hn = c2[j + 1];
                LAUNCH("scale", k_scale_div<double>, dim3(nbk(ne)), dim3(256), 0, stream, static_cast<const double *>(wvec),
                       static_cast<const double *>(cgs_h + j + 1), ne, V[j + 1]);
            }
            check_launch("arnoldi (CGS)");
            for (int i = 0; i <= j; ++i) H[i][j] = hc[i];
            H[j + 1][j] = hn;
            for (int i = 0; i < j; ++i)
            {
                const double tt = cs[i] * H[i][j] + sn[i] * H[i + 1][j];
                H[i + 1][j] = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
                H[i][j] = tt;
            }
            const double den = std::hypot(H[j][j], H[j + 1][j]);
            cs[j] = den > 0 ? H[j][j] / den : 1.0; sn[j] = den > 0 ? H[j + 1][j] / den : 0.0;
            H[j][j] = den; H[j + 1][j] = 0.0;
            g[j + 1] = -sn[j] * g[j]; g[j] = cs[j] * g[j];
            if (std::abs(g[j + 1]) <= rtol * bnorm || !(hn > 1e-300)) { ++j; ++total; break; }
        }
        for (int i = j - 1; i >= 0; --i)
        {
            double s = g[i];
            for (int k = i + 1; k < j; ++k) s -= H[i][k] * yv[k];
            yv[i] = s / H[i][i];
        }
        chk(hipMemcpyAsync(ydev, yv.data(), sizeof(double) * std::max(j, 1), hipMemcpyHostToDevice, stream), "H2D");
        LAUNCH("gmres_update", k_gmres_update, dim3(nbk(ne)), dim3(256), 0, stream, static_cast<const double *const *>(Vptrs), ydev, j,
               invdiag, dir, ne, delta);
        LAUNCH("axpy", k_add, dim3(nbk(ne)), dim3(256), 0, stream, x0, delta, ne, tmp);
        apply_op(nullptr, tmp, nullptr, dir, dt, 1, 0, bvec, rvec);
        beta = dot(rvec, rvec, 1, 1);
    }
    LAUNCH("solve_final", k_solve_final, dim3(nbk(ne)), dim3(256), 0, stream, x0, delta, Eex, dir, ne, Eout, corr);
    check_launch("solve final");
    have_corr = true;
    if (rel_res) *rel_res = beta / bnorm;
    return total;
}
