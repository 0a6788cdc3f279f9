// Mixed-precision implicit induction solve (ImplicitInduction::solve's system, Common/induction/
// ImplicitInduction.cpp:299-390, solved to the same FP64 relative residual):
//   outer FP64 iterative refinement   r = b - A x (FP64 operator), x += d
//   inner GMRES(m) on A d = r         FP32 operator and FP32 basis, right Jacobi preconditioner,
//                                     classical Gram-Schmidt with FP64 accumulation (one multi-dot pass and
//                                     one multi-axpy pass per step, re-orthogonalised when the DGKS test
//                                     asks for it), Hessenberg and Givens in FP64 on the host.
#include <algorithm>
#include <array>
#include <utility>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "ctx.h"

namespace mhdg
{
namespace
{
constexpr int kPairUnroll = 3; // row pairs loaded together by the FP32 edge kernel (rows of up to six entries in one group)
struct F3 { float x, y, z; };
__device__ __forceinline__ F3 f3(const float *p, long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ F3 fcross(const F3 &a, const F3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ __forceinline__ float fdot(const F3 &a, const F3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float fsg(int32_t v) { return v >= 0 ? 1.0f : -1.0f; }
// a float4 load through a pointer known to be global memory (a pointer read from LDS or an array is generic, and its
// loads become flat loads that also count against the LDS wait counter)
typedef float F4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) F4v *GlobalF4;
__device__ __forceinline__ float4 ldg4(const float *p, int32_t q)
{
    const F4v v = ((GlobalF4)p)[q];
    return make_float4(v.x, v.y, v.z, v.w);
}
typedef int I4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) I4v *GlobalI4;
typedef const __attribute__((address_space(1))) int32_t *GlobalI32;
// Row pairs of the edge kernel, kept (mask -1) or made a zero pad (mask 0) with no branch: a conditional load, or a
// select on its result, became a branch around the load and a wait at its end.  Pairs past a slice's count read the
// following slots (the word arrays end in one zeroed group of slots).
// wide pair: one 16-byte load at word w (a multiple of 4)
__device__ __forceinline__ int4 ld_wide(const int32_t *words, int32_t w, int32_t mask)
{
    const I4v v = ((GlobalI4)words)[w >> 2];
    return make_int4(v.x & mask, v.y & mask, v.z & mask, v.w & mask);
}
// compact pair: three words from w (one 12-byte load): x = the two 16-bit offsets, y and z = the weights
__device__ __forceinline__ int3 ld_compact(const int32_t *words, int32_t w, int32_t mask)
{
    const GlobalI32 p = (GlobalI32)words + w;
    return make_int3(p[0] & mask, p[1] & mask, p[2] & mask);
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
} // namespace

// Row range of a CSR table: fixed degree (deg > 0, every row the same length) or the ptr array
__device__ __forceinline__ void rows(int32_t deg, const int32_t *ptr, int32_t r, int32_t &k0, int32_t &k1)
{
    if (deg > 0) { k0 = r * deg; k1 = k0 + deg; }
    else { k0 = ptr[r]; k1 = ptr[r + 1]; }
}

// Phi'_f = sum_k f2e_k pv_e with pv = P v (0 on Dirichlet edges), times dt_e with a per-edge step
__global__ void __launch_bounds__(256) k_d1_f32(DevMesh d, DevMesh32 s, const float *pv, const float *dte, float *phi)
{
    const int32_t f = gid();
    if (f >= d.nf) return;
    if (s.fdeg == 3 && !dte)
    {
        // triangles: the face's three edge entries in one 12-byte load, then the three gathers, summed in entry order
        const int3 q = reinterpret_cast<const int3 *>(d.f2e)[f];
        const float x0 = pv[pid(q.x)], x1 = pv[pid(q.y)], x2 = pv[pid(q.z)];
        float sum = 0.0f;
        sum += fsg(q.x) * x0;
        sum += fsg(q.y) * x1;
        sum += fsg(q.z) * x2;
        phi[f] = sum;
        return;
    }
    int32_t k0, k1;
    rows(s.fdeg, d.f2e_ptr, f, k0, k1);
    float sum = 0.0f;
    for (int32_t k = k0; k < k1; k += kRowUnroll)
    {
        int32_t q[kRowUnroll];
        float x[kRowUnroll], t[kRowUnroll];
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k + u < k1) q[u] = d.f2e[k + u];
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k + u < k1)
            {
                x[u] = pv[pid(q[u])];
                if (dte) t[u] = dte[pid(q[u])];
            }
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k + u < k1) sum += fsg(q[u]) * (dte ? t[u] * x[u] : x[u]);
    }
    phi[f] = sum;
}

// What the edge kernel writes from y = dt F(Phi'), F = m_e sum_f s star2_f Phi'_f + sum_c w_ec g_c . t_e (MODE, a template
// argument, so the Krylov product is the same code as before):
//   kOpShifted: (A P - I) v = (pv - v) + y       (pv = P v, the Jacobi-preconditioned Krylov vector)
//   kOpApply:   A x = x + y                       (x = pv, zero on Dirichlet edges)
//   kOpSmooth:  x + omega D^-1 (b - A x)          (a damped-Jacobi sweep of A x = b, D^-1 = invdiag)
//   kOpResid:   b - A x
// and 0 on Dirichlet edges in every mode.
// The per-edge inputs of the output.  The FP32 edge kernel loads them after its row sums: loading them ahead of the
// rows took four more VGPRs and measured no faster.
struct EdgeIn
{
    float me, pve, ve, be; // m_e, pv_e, v_e (kOpShifted), b_e (kOpSmooth, kOpResid)
};
template <int MODE>
__device__ __forceinline__ EdgeIn edge_in(const DevMesh32 &s, int32_t e, const float *v, const float *pv, const float *b)
{
    return {s.m_e[e], pv[e], MODE == kOpShifted ? v[e] : 0.0f, MODE == kOpSmooth || MODE == kOpResid ? b[e] : 0.0f};
}
template <int MODE>
__device__ __forceinline__ float edge_out(const DevMesh32 &s, int32_t e, float y, const EdgeIn &x, float omega)
{
    if (MODE == kOpShifted) return (x.pve - x.ve) + y;
    const float ax = x.pve + y;
    if (MODE == kOpApply) return ax;
    if (MODE == kOpSmooth) return x.pve + omega * s.invdiag[e] * (x.be - ax);
    return x.be - ax;
}

// One high-valence edge (the polar axes) per block: terms in parallel, tree sums (order-free in FP32)
template <int MODE>
__device__ __forceinline__ void edge_apply_f32_heavy(const DevMesh &d, const DevMesh32 &s, int32_t resistive, int32_t hall, int32_t e,
                                                     const float *phi, const float *g, const float *v, const float *pv, const int8_t *dir,
                                                     float dt, float *out, const float *b, float omega)
{
    __shared__ float sh[2][kEdgeBlock / 64];
    float p0 = 0.0f, p1 = 0.0f;
    if (resistive)
        for (int32_t k = d.e2f_ptr[e] + threadIdx.x; k < d.e2f_ptr[e + 1]; k += kEdgeBlock)
        {
            const int32_t q = d.e2f[k], f = pid(q);
            p0 += fsg(q) * (s.star2[f] * phi[f]);
        }
    if (hall)
    {
        const F3 t = f3(s.evec, e);
        for (int32_t k = d.ec_ptr[e] + threadIdx.x; k < d.ec_ptr[e + 1]; k += kEdgeBlock) p1 += s.ecw[k] * fdot(f3(g, d.ec_cell[k]), t);
    }
    for (int o = 32; o > 0; o >>= 1) { p0 += __shfl_down(p0, o, 64); p1 += __shfl_down(p1, o, 64); }
    if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = p0; sh[1][threadIdx.x >> 6] = p1; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        float P0 = 0.0f, P1 = 0.0f;
        for (int w = 0; w < kEdgeBlock / 64; ++w) { P0 += sh[0][w]; P1 += sh[1][w]; }
        if (dir[e]) { out[e] = 0.0f; return; }
        const EdgeIn x = edge_in<MODE>(s, e, v, pv, b);
        out[e] = edge_out<MODE>(s, e, dt * (x.me * P0 + P1), x, omega);
    }
}

// Sum of one edge's row in its slice (descriptor dsc): KIND 0 the face row, sum_f w_f x_f; KIND 1 the cell row,
// sum_c w_c g_c . t.  A group of the row's pairs is loaded, then gathered, then summed in entry order.
template <int KIND>
__device__ __forceinline__ float row_sum(const int4 &dsc, const int32_t *words, int32_t lane, const float *x, const F3 &t)
{
    float acc = 0.0f;
    const int32_t np = dsc.y;
    if (dsc.w == 0)
    {
        for (int32_t p = 0; p < np; p += kPairUnroll)
        {
            int3 q[kPairUnroll];
            float xs[KIND == 0 ? 2 * kPairUnroll : 1];
            F3 gs[KIND == 1 ? 2 * kPairUnroll : 1];
#pragma unroll
            for (int u = 0; u < kPairUnroll; ++u) q[u] = ld_compact(words, dsc.x + 3 * (64 * (p + u) + lane), p + u < np ? -1 : 0);
#pragma unroll
            for (int u = 0; u < kPairUnroll; ++u)
            {
                const int32_t i0 = dsc.z + (q[u].x & 0xffff), i1 = dsc.z + int32_t(uint32_t(q[u].x) >> 16);
                if constexpr (KIND == 0) { xs[2 * u] = x[i0]; xs[2 * u + 1] = x[i1]; }
                else { gs[2 * u] = f3(x, i0); gs[2 * u + 1] = f3(x, i1); }
            }
#pragma unroll
            for (int u = 0; u < kPairUnroll; ++u)
            {
                if constexpr (KIND == 0)
                {
                    acc += __int_as_float(q[u].y) * xs[2 * u];
                    acc += __int_as_float(q[u].z) * xs[2 * u + 1];
                }
                else
                {
                    acc += __int_as_float(q[u].y) * fdot(gs[2 * u], t);
                    acc += __int_as_float(q[u].z) * fdot(gs[2 * u + 1], t);
                }
            }
        }
    }
    else
    {
        for (int32_t p = 0; p < np; p += kPairUnroll)
        {
            int4 q[kPairUnroll];
            float xs[KIND == 0 ? 2 * kPairUnroll : 1];
            F3 gs[KIND == 1 ? 2 * kPairUnroll : 1];
#pragma unroll
            for (int u = 0; u < kPairUnroll; ++u) q[u] = ld_wide(words, dsc.x + 4 * (64 * (p + u) + lane), p + u < np ? -1 : 0);
#pragma unroll
            for (int u = 0; u < kPairUnroll; ++u)
            {
                if constexpr (KIND == 0) { xs[2 * u] = x[q[u].x]; xs[2 * u + 1] = x[q[u].z]; }
                else { gs[2 * u] = f3(x, q[u].x); gs[2 * u + 1] = f3(x, q[u].z); }
            }
#pragma unroll
            for (int u = 0; u < kPairUnroll; ++u)
            {
                if constexpr (KIND == 0)
                {
                    acc += __int_as_float(q[u].y) * xs[2 * u];
                    acc += __int_as_float(q[u].w) * xs[2 * u + 1];
                }
                else
                {
                    acc += __int_as_float(q[u].y) * fdot(gs[2 * u], t);
                    acc += __int_as_float(q[u].w) * fdot(gs[2 * u + 1], t);
                }
            }
        }
    }
    return acc;
}

// out = dir ? 0 : (A P - I) v = (pv - v) + dt F(Phi') with F = m_e sum_f s star2_f Phi'_f + sum_c w_ec g_c . t_e
// (the shifted operator: same Krylov space as A P, and the O(1) identity part never enters the FP32 vector)
// The rows come in slices of 64 edges (DevMesh32::fdesc/fwords, cdesc/cwords): each pair load of a wavefront reads one
// contiguous range, 768 bytes of 12-byte pairs with 16-bit offsets in a compact slice (the per-edge CSR rows, 40 bytes
// apart, touched a cache line per lane), a row's pair loads are issued before its gathers, padding slots are gathered
// and added as zero terms without a branch, and the sums run in entry order.  The kernel's time follows its L2
// requests, so the compact pairs (a quarter fewer row bytes) took it from 278 to 240 us per launch on the 8M shell.
// Face row, then cell row: both rows' loads together took 74 VGPRs (six waves per SIMD) and ran slower.
// Blocks from n_light on take one high-valence edge each (heavy[block - n_light]) in the same launch.
template <int MODE>
__global__ void __launch_bounds__(kEdgeBlock) k_edge_apply_f32(DevMesh d, DevMesh32 s, int32_t resistive, int32_t hall, const float *phi,
                                                        const float *g, const float *v, const float *pv, const int8_t *dir, float dt,
                                                        float *out, const int32_t *heavy, int32_t n_light, const float *b, float omega)
{
    if (int32_t(blockIdx.x) >= n_light)
    {
        edge_apply_f32_heavy<MODE>(d, s, resistive, hall, heavy[blockIdx.x - n_light], phi, g, v, pv, dir, dt, out, b, omega);
        return;
    }
    const int32_t e = xcd_block(uint32_t(n_light)) * kEdgeBlock + int32_t(threadIdx.x);
    if (e >= d.ne) return;
    const int8_t de = dir[e], he = s.eheavy[e]; // both loads ahead of one test
    if (de | he)
    {
        if (!he) out[e] = 0.0f; // a Dirichlet edge; a high-valence edge is one of the heavy blocks
        return;
    }
    const int32_t sl = __builtin_amdgcn_readfirstlane(e >> 6), lane = e & 63; // a wavefront is one slice
    const F3 t = hall ? f3(s.evec, e) : F3{0.0f, 0.0f, 0.0f};
    const float p0 = resistive ? row_sum<0>(s.fdesc[sl], s.fwords, lane, phi, t) : 0.0f;
    const float p1 = hall ? row_sum<1>(s.cdesc[sl], s.cwords, lane, g, t) : 0.0f;
    const EdgeIn x = edge_in<MODE>(s, e, v, pv, b);
    out[e] = edge_out<MODE>(s, e, dt * (x.me * p0 + p1), x, omega);
}
template __global__ void k_edge_apply_f32<kOpShifted>(DevMesh, DevMesh32, int32_t, int32_t, const float *, const float *, const float *,
                                                      const float *, const int8_t *, float, float *, const int32_t *, int32_t, const float *,
                                                      float);
template __global__ void k_edge_apply_f32<kOpApply>(DevMesh, DevMesh32, int32_t, int32_t, const float *, const float *, const float *,
                                                    const float *, const int8_t *, float, float *, const int32_t *, int32_t, const float *,
                                                    float);
template __global__ void k_edge_apply_f32<kOpSmooth>(DevMesh, DevMesh32, int32_t, int32_t, const float *, const float *, const float *,
                                                     const float *, const int8_t *, float, float *, const int32_t *, int32_t, const float *,
                                                     float);
template __global__ void k_edge_apply_f32<kOpResid>(DevMesh, DevMesh32, int32_t, int32_t, const float *, const float *, const float *,
                                                    const float *, const int8_t *, float, float *, const int32_t *, int32_t, const float *,
                                                    float);

// Cell kernels with G lanes per cell: the lanes read the cell's (cell, face) table entries side by side (coalesced
// across the wavefront) and combine their partial sums with shuffles (G = 4 measured 1.2-1.3x faster than one
// thread per cell on the 8M shell; one thread per edge stays faster for the edge kernel).
// b_c = sum_k recon_k Phi'_{f_k}
template <int G> __device__ __forceinline__ float group_sum(float v)
{
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}
template <int G> __global__ void __launch_bounds__(256) k_bcell_f32_g(DevMesh d, DevMesh32 s, const float *phi, float *b)
{
    const int32_t t = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    const int32_t c = t / G, lane = t % G;
    const bool ok = c < d.nc;
    float bx = 0.0f, by = 0.0f, bz = 0.0f;
    if (ok && s.c2f16 && s.cdeg == G)
    {
        // one entry per lane, the face from its 16-bit offset (G == 4: a wavefront is 16 cells)
        const int32_t k = c * G + lane;
        const float v = phi[s.c2fbase[t >> 6] + s.c2f16[k]];
        bx = s.recon[3 * size_t(k)] * v; by = s.recon[3 * size_t(k) + 1] * v; bz = s.recon[3 * size_t(k) + 2] * v;
    }
    else if (ok)
    {
        int32_t k0, k1;
        rows(s.cdeg, d.c2f_ptr, c, k0, k1);
        for (int32_t k = k0 + lane; k < k1; k += G)
        {
            const float v = phi[pid(d.c2f[k])];
            bx += s.recon[3 * size_t(k)] * v; by += s.recon[3 * size_t(k) + 1] * v; bz += s.recon[3 * size_t(k) + 2] * v;
        }
    }
    bx = group_sum<G>(bx); by = group_sum<G>(by); bz = group_sum<G>(bz);
    if (ok && lane == 0) { b[3 * size_t(c)] = bx; b[3 * size_t(c) + 1] = by; b[3 * size_t(c) + 2] = bz; }
}
// g_c = J_c x (alpha_c B0_c) with J_c = sum_k (s A_f / V_c)_k x b_f, b_f the face average with the neighbour;
// zero where alpha_c B0_c is
template <int G> __global__ void __launch_bounds__(256) k_hall_g_f32_g(DevMesh d, DevMesh32 s, const float *b, float *g)
{
    const int32_t t = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    const int32_t c = t / G, lane = t % G;
    const bool ok = c < d.nc;
    // alpha B0 gates only the result: a test of it ahead of the neighbour gathers made every load of the cell wait
    // for it (174 against 151 us per launch on the 8M shell)
    const F3 ab = ok ? f3(s.aB0, c) : F3{0.0f, 0.0f, 0.0f};
    F3 J{0.0f, 0.0f, 0.0f};
    if (ok && s.cdeg == G)
    {
        // one entry per lane (tetrahedra with four lanes): no row bounds to wait for
        // (a 16-bit neighbour offset from the cell measured 6.31 against 6.06 ms per 8M step: its decode sits between
        // the load and the neighbour gather)
        const int32_t k = c * G + lane, nb = s.nbr[k];
        const F3 av = f3(s.aV, k), bc = f3(b, c);
        F3 bf = bc;
        if (nb >= 0)
        {
            const F3 bo = f3(b, nb);
            bf = {0.5f * (bc.x + bo.x), 0.5f * (bc.y + bo.y), 0.5f * (bc.z + bo.z)};
        }
        J = fcross(av, bf);
    }
    else if (ok)
    {
        int32_t k0, k1;
        rows(s.cdeg, d.c2f_ptr, c, k0, k1);
        const F3 bc = f3(b, c);
        for (int32_t k = k0 + lane; k < k1; k += G)
        {
            const int32_t nb = s.nbr[k];
            const F3 av = f3(s.aV, k);
            F3 bf = bc;
            if (nb >= 0)
            {
                const F3 bo = f3(b, nb);
                bf = {0.5f * (bc.x + bo.x), 0.5f * (bc.y + bo.y), 0.5f * (bc.z + bo.z)};
            }
            const F3 c3 = fcross(av, bf);
            J.x += c3.x; J.y += c3.y; J.z += c3.z;
        }
    }
    J.x = group_sum<G>(J.x); J.y = group_sum<G>(J.y); J.z = group_sum<G>(J.z);
    if (ok && lane == 0)
    {
        const bool hall = ab.x != 0.0f || ab.y != 0.0f || ab.z != 0.0f;
        const F3 out = hall ? fcross(J, ab) : F3{0.0f, 0.0f, 0.0f};
        g[3 * size_t(c)] = out.x; g[3 * size_t(c) + 1] = out.y; g[3 * size_t(c) + 2] = out.z;
    }
}
template __global__ void k_bcell_f32_g<4>(DevMesh, DevMesh32, const float *, float *);
template __global__ void k_hall_g_f32_g<4>(DevMesh, DevMesh32, const float *, float *);

// The same for the high-valence edges, one block per edge (tree sums: order-free in FP32)
// Operator table of the FP32 path: per (cell, face) s A_f / V_c and the neighbour across the face
__global__ void __launch_bounds__(256) k_tables_f32(DevMesh d, float *aV, int32_t *nbr)
{
    const int32_t c = gid();
    if (c < d.nc)
    {
        const double iv = 1.0 / d.cvol[c];
        for (int32_t k = d.c2f_ptr[c]; k < d.c2f_ptr[c + 1]; ++k)
        {
            const int32_t v = d.c2f[k], f = pid(v);
            const double sg = psg(v) * iv;
            aV[3 * size_t(k)] = static_cast<float>(sg * d.farea[3 * size_t(f)]);
            aV[3 * size_t(k) + 1] = static_cast<float>(sg * d.farea[3 * size_t(f) + 1]);
            aV[3 * size_t(k) + 2] = static_cast<float>(sg * d.farea[3 * size_t(f) + 2]);
            const int32_t l = d.fL[f], r = d.fR[f];
            nbr[k] = (l >= 0 && r >= 0) ? (l == c ? r : l) : -1;
        }
    }
}

// h[i0+k] = <w, V_{i0+k}> for k < NV (FP64 accumulation, 4 elements per thread and iteration); with_ww
// also <w, w> into out[63].  NV is a template argument so that the NV loads of an iteration are issued together (a
// run-time count guarded each load by a branch and kept one load in flight per thread).
template <int NV>
__global__ void __launch_bounds__(256) k_multidot_f32(const float *const *V, int32_t i0, const float *w, int32_t n, double *part,
                                                      int32_t with_ww, RedScratch rs, double *out)
{
    __shared__ double sh[4][kDotChunk + 1];
    __shared__ int last;
    double acc[kDotChunk + 1];
#pragma unroll
    for (int k = 0; k <= kDotChunk; ++k) acc[k] = 0.0;
    const float *Vk[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) Vk[k] = V[i0 + k];
    const int32_t n4 = n >> 2;
    for (int32_t q = blockIdx.x * 256 + threadIdx.x; q < n4; q += gridDim.x * 256)
    {
        const float4 a = ldg4(w, q);
        float4 b[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) b[k] = ldg4(Vk[k], q);
#pragma unroll
        for (int k = 0; k < NV; ++k)
            acc[k] += (double(a.x) * b[k].x + double(a.y) * b[k].y) + (double(a.z) * b[k].z + double(a.w) * b[k].w);
        acc[kDotChunk] += (double(a.x) * a.x + double(a.y) * a.y) + (double(a.z) * a.z + double(a.w) * a.w);
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3))
    {
        const int32_t e = 4 * n4 + threadIdx.x;
        const double a = w[e];
#pragma unroll
        for (int k = 0; k < NV; ++k) acc[k] += a * Vk[k][e];
        acc[kDotChunk] += a * a;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= kDotChunk; ++k)
        if (k < NV || k == kDotChunk)
        {
            const double sm = wave_sum(acc[k]);
            if (lane == 0) sh[wv][k] = sm;
        }
    __syncthreads();
    if (threadIdx.x < NV || threadIdx.x == kDotChunk)
    {
        const int k = threadIdx.x;
        part[size_t(blockIdx.x) * (kDotChunk + 1) + k] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
    }
    if (!rs.ticket) return; // partials only: k_multidot_final sums them
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(rs.ticket, 1u) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t[kDotChunk + 1];
#pragma unroll
    for (int k = 0; k <= kDotChunk; ++k) t[k] = 0.0;
    for (uint32_t bl = threadIdx.x; bl < gridDim.x; bl += 256)
#pragma unroll
        for (int k = 0; k <= kDotChunk; ++k)
            if (k < NV || k == kDotChunk)
                t[k] += __hip_atomic_load(&part[size_t(bl) * (kDotChunk + 1) + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int k = 0; k <= kDotChunk; ++k)
        if (k < NV || k == kDotChunk)
        {
            const double s2 = wave_sum(t[k]);
            if (lane == 0) sh[wv][k] = s2;
        }
    __syncthreads();
    if (threadIdx.x < NV || threadIdx.x == kDotChunk)
    {
        const int k = threadIdx.x;
        const double s2 = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
        if (k < NV) out[i0 + k] = s2;
        if (k == kDotChunk && with_ww) out[63] = s2;
    }
    if (threadIdx.x == 0) *rs.ticket = 0u;
}
using MultidotFn = void (*)(const float *const *, int32_t, const float *, int32_t, double *, int32_t, RedScratch, double *);
template <int... I> constexpr std::array<MultidotFn, sizeof...(I) + 1> multidot_table(std::integer_sequence<int, I...>)
{
    return {nullptr, &k_multidot_f32<I + 1>...};
}
// kMultidot[nv]: the multi-dot pass over nv basis vectors, 1 <= nv <= kDotChunk
const std::array<MultidotFn, kDotChunk + 1> kMultidot = multidot_table(std::make_integer_sequence<int, kDotChunk>{});

// x -= h_i V_i for the G vectors from i0 on, in order (their G loads issued together)
template <int G>
__device__ __forceinline__ void axpy_group(const float *const *vp, const double *hs, int32_t i0, int32_t q, double &x, double &y, double &z,
                                           double &u)
{
    float4 b[G];
#pragma unroll
    for (int k = 0; k < G; ++k) b[k] = ldg4(vp[i0 + k], q);
#pragma unroll
    for (int k = 0; k < G; ++k)
    {
        const double hi = hs[i0 + k];
        x -= hi * b[k].x; y -= hi * b[k].y; z -= hi * b[k].z; u -= hi * b[k].w;
    }
}
constexpr int kAxpyGroup = 8;
// x -= sum_{i < nv} h_i V_i in order of i: groups of kAxpyGroup loads, then the remainder's loads together
__device__ __forceinline__ void axpy_all(const float *const *vp, const double *hs, int32_t nv, int32_t q, double &x, double &y, double &z,
                                         double &u)
{
    int32_t i = 0;
    for (; i + kAxpyGroup <= nv; i += kAxpyGroup) axpy_group<kAxpyGroup>(vp, hs, i, q, x, y, z, u);
    switch (nv - i)
    {
    case 7: axpy_group<7>(vp, hs, i, q, x, y, z, u); break;
    case 6: axpy_group<6>(vp, hs, i, q, x, y, z, u); break;
    case 5: axpy_group<5>(vp, hs, i, q, x, y, z, u); break;
    case 4: axpy_group<4>(vp, hs, i, q, x, y, z, u); break;
    case 3: axpy_group<3>(vp, hs, i, q, x, y, z, u); break;
    case 2: axpy_group<2>(vp, hs, i, q, x, y, z, u); break;
    case 1: axpy_group<1>(vp, hs, i, q, x, y, z, u); break;
    default: break;
    }
}

// w -= sum_{i < nv} h_i V_i (FP64 arithmetic, FP32 storage, 4 elements per thread); nrm[0] = |w| after
__global__ void __launch_bounds__(256) k_multiaxpy_f32(const float *const *V, int32_t nv, const double *h, float *w, int32_t n,
                                                       RedScratch rs, double *nrm)
{
    __shared__ double sh[4];
    __shared__ int last;
    __shared__ double hs[64];
    __shared__ const float *vp[64];
    if (threadIdx.x < nv) { hs[threadIdx.x] = h[threadIdx.x]; vp[threadIdx.x] = V[threadIdx.x]; }
    __syncthreads();
    double acc = 0.0;
    const int32_t n4 = n >> 2;
    float4 *w4 = reinterpret_cast<float4 *>(w);
    for (int32_t q = blockIdx.x * 256 + threadIdx.x; q < n4; q += gridDim.x * 256)
    {
        const float4 a = w4[q];
        double x = a.x, y = a.y, z = a.z, u = a.w;
        axpy_all(vp, hs, nv, q, x, y, z, u);
        const float4 r = make_float4(float(x), float(y), float(z), float(u));
        w4[q] = r;
        acc += (double(r.x) * r.x + double(r.y) * r.y) + (double(r.z) * r.z + double(r.w) * r.w);
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3))
    {
        const int32_t e = 4 * n4 + threadIdx.x;
        double we = w[e];
        for (int32_t i = 0; i < nv; ++i) we -= hs[i] * static_cast<double>(vp[i][e]);
        const float wf = static_cast<float>(we);
        w[e] = wf;
        acc += double(wf) * wf;
    }
    const double s = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        rs.partial[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
        __threadfence();
        last = atomicAdd(rs.ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t = 0.0;
    for (uint32_t bl = threadIdx.x; bl < gridDim.x; bl += 256)
        t += __hip_atomic_load(&rs.partial[bl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double s2 = wave_sum(t);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s2;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        nrm[0] = sqrt(((sh[0] + sh[1]) + sh[2]) + sh[3]);
        *rs.ticket = 0u;
    }
}

// out[i0 + k] (k < nv) and out[63] (with_ww) from the per-block partials of k_multidot_f32 (one block)
__global__ void __launch_bounds__(256) k_multidot_final(const double *part, int32_t nblocks, int32_t i0, int32_t nv, int32_t with_ww,
                                                        double *out)
{
    __shared__ double sh[4][kDotChunk + 1];
    double t[kDotChunk + 1];
#pragma unroll
    for (int k = 0; k <= kDotChunk; ++k) t[k] = 0.0;
    for (int32_t bl = threadIdx.x; bl < nblocks; bl += 256)
#pragma unroll
        for (int k = 0; k <= kDotChunk; ++k) t[k] += part[size_t(bl) * (kDotChunk + 1) + k];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= kDotChunk; ++k)
    {
        const double s = wave_sum(t[k]);
        if (lane == 0) sh[wv][k] = s;
    }
    __syncthreads();
    if (threadIdx.x <= kDotChunk)
    {
        const int k = threadIdx.x;
        const double s = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
        if (k < nv) out[i0 + k] = s;
        if (k == kDotChunk && with_ww) out[63] = s;
    }
}

// Classical Gram-Schmidt update fused with the normalisation: with h[0..nv) = <w, V_i> and h[63] = <w, w> from the
// multi-dot pass, hn = sqrt(<w,w> - sum h_i^2) (Pythagoras, FP64) is the norm of w' = w - sum h_i V_i.  When
// hn >= dgks |w| the kernel writes V_nv = w' / hn and pv = dir ? 0 : V_nv invdiag directly and h[nv] = hn; else
// (severe cancellation: a second projection follows) it writes w' into w and h[nv] = |w'| computed from w'.
// h[62] = 1 in the second case, 0 in the first.  Every block takes the same branch (same h).
__global__ void __launch_bounds__(256) k_multiaxpy_scale_f32(const float *const *V, int32_t nv, double *h, float *w, int32_t n, double dgks,
                                                             const float *invdiag, const int8_t *dir, float *vnext, float *pv, RedScratch rs)
{
    __shared__ double sh[4];
    __shared__ int last;
    __shared__ double hs[64];
    __shared__ const float *vp[64];
    if (threadIdx.x < nv) { hs[threadIdx.x] = h[threadIdx.x]; vp[threadIdx.x] = V[threadIdx.x]; }
    __syncthreads();
    double ss = 0.0;
    for (int32_t i = 0; i < nv; ++i) ss += hs[i] * hs[i];
    const double ww = h[63];
    const double hn2 = ww - ss;
    const double hp = hn2 > 0 ? sqrt(hn2) : 0.0;
    const bool direct = hp >= dgks * sqrt(ww > 0 ? ww : 0.0) && hp > 1e-300;
    const double inv = direct ? 1.0 / hp : 0.0;
    double acc = 0.0;
    const int32_t n4 = n >> 2;
    float4 *w4 = reinterpret_cast<float4 *>(w);
    for (int32_t q = blockIdx.x * 256 + threadIdx.x; q < n4; q += gridDim.x * 256)
    {
        const float4 a = w4[q];
        double x = a.x, y = a.y, z = a.z, u = a.w;
        axpy_all(vp, hs, nv, q, x, y, z, u);
        if (direct)
        {
            const float4 v = make_float4(float(x * inv), float(y * inv), float(z * inv), float(u * inv));
            reinterpret_cast<float4 *>(vnext)[q] = v;
            const float4 iv = reinterpret_cast<const float4 *>(invdiag)[q];
            const int32_t e = 4 * q;
            reinterpret_cast<float4 *>(pv)[q] = make_float4(dir[e] ? 0.0f : v.x * iv.x, dir[e + 1] ? 0.0f : v.y * iv.y,
                                                            dir[e + 2] ? 0.0f : v.z * iv.z, dir[e + 3] ? 0.0f : v.w * iv.w);
        }
        else
        {
            const float4 r = make_float4(float(x), float(y), float(z), float(u));
            w4[q] = r;
            acc += (double(r.x) * r.x + double(r.y) * r.y) + (double(r.z) * r.z + double(r.w) * r.w);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < (n & 3))
    {
        const int32_t e = 4 * n4 + threadIdx.x;
        double we = w[e];
        for (int32_t i = 0; i < nv; ++i) we -= hs[i] * static_cast<double>(vp[i][e]);
        if (direct)
        {
            const float v = static_cast<float>(we * inv);
            vnext[e] = v;
            pv[e] = dir[e] ? 0.0f : v * invdiag[e];
        }
        else
        {
            const float wf = static_cast<float>(we);
            w[e] = wf;
            acc += double(wf) * wf;
        }
    }
    if (direct)
    {
        if (blockIdx.x == 0 && threadIdx.x == 0) { h[nv] = hp; h[62] = 0.0; }
        return;
    }
    const double s = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        rs.partial[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
        __threadfence();
        last = atomicAdd(rs.ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t = 0.0;
    for (uint32_t bl = threadIdx.x; bl < gridDim.x; bl += 256)
        t += __hip_atomic_load(&rs.partial[bl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const double s2 = wave_sum(t);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s2;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        h[nv] = sqrt(((sh[0] + sh[1]) + sh[2]) + sh[3]);
        h[62] = 1.0;
        *rs.ticket = 0u;
    }
}

// V_0 = r / beta in FP32 and its preconditioned copy pv = dir ? 0 : V_0 invdiag
__global__ void __launch_bounds__(256) k_to_f32_scaled(const double *x, const double *s, const float *invdiag, const int8_t *dir, int32_t n,
                                                       float *out, float *pv)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float v = static_cast<float>(x[e] / s[0]);
    out[e] = v;
    pv[e] = dir[e] ? 0.0f : v * invdiag[e];
}
// V_{j+1} = w / h and pv = dir ? 0 : V_{j+1} invdiag
__global__ void __launch_bounds__(256) k_scale_f32(const float *w, const double *h, const float *invdiag, const int8_t *dir, int32_t n,
                                                   float *out, float *pv)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const double hv = h[0];
    if (e >= n || !(hv > 1e-300)) return;
    const float v = static_cast<float>(static_cast<double>(w[e]) / hv);
    out[e] = v;
    pv[e] = dir[e] ? 0.0f : v * invdiag[e];
}
// delta += P sum_{i < j} y_i V_i on the non-Dirichlet edges, four edges per thread (the grid covers n / 4 + 1
// threads), the sum in order of i with the loads of a group of vectors issued together
__global__ void __launch_bounds__(256) k_update_mixed(const float *const *V, const double *y, int32_t j, const float *invdiag,
                                                      const int8_t *dir, int32_t n, double *delta)
{
    __shared__ double ny[64];
    __shared__ const float *vp[64];
    if (threadIdx.x < j) { ny[threadIdx.x] = -y[threadIdx.x]; vp[threadIdx.x] = V[threadIdx.x]; }
    __syncthreads();
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x, n4 = n >> 2;
    if (q < n4)
    {
        // s - (-y_i) v = s + y_i v exactly
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        axpy_all(vp, ny, j, q, s[0], s[1], s[2], s[3]);
        const float4 iv = ldg4(invdiag, q);
        const float ivs[4] = {iv.x, iv.y, iv.z, iv.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (!dir[4 * q + k]) delta[4 * q + k] += s[k] * static_cast<double>(ivs[k]);
    }
    else if (q == n4)
        for (int32_t e = 4 * n4; e < n; ++e)
        {
            if (dir[e]) continue;
            double s = 0.0;
            for (int32_t i = 0; i < j; ++i) s += y[i] * static_cast<double>(vp[i][e]);
            delta[e] += s * static_cast<double>(invdiag[e]);
        }
}
// Slices of the edge kernel's rows (DevMesh32::fdesc/fwords, cdesc/cwords).  Per 64-edge slice: the pairs per
// edge its longest light row needs and the smallest and largest id in its light rows ({pairs, min, max, 0}; a slice
// with none has max -1)
__global__ void __launch_bounds__(256) k_slice_scan(const int32_t *ptr, const int32_t *idx, int32_t signed_ids, const int32_t *e2f_ptr,
                                                    int32_t ne, int32_t n_slices, int4 *out)
{
    const int32_t e = gid();
    int32_t p = 0, lo = INT32_MAX, hi = -1;
    if (e < ne && e2f_ptr[e + 1] - e2f_ptr[e] <= kLightMax)
    {
        p = (ptr[e + 1] - ptr[e] + 1) / 2;
        for (int32_t k = ptr[e]; k < ptr[e + 1]; ++k)
        {
            const int32_t i = signed_ids ? pid(idx[k]) : idx[k];
            lo = min(lo, i);
            hi = max(hi, i);
        }
    }
    for (int o = 32; o > 0; o >>= 1)
    {
        p = max(p, __shfl_xor(p, o, 64));
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
    }
    if ((threadIdx.x & 63) == 0 && (e >> 6) < n_slices) out[e >> 6] = make_int4(p, lo, hi, 0);
}
// each light edge's entries at their slots: face rows (star2 != nullptr: entry (face, signed star2 of the face) from the
// signed e2f ids) or cell rows ((cell, weight) from idx and w); slots past a row keep the zero fill
__global__ void __launch_bounds__(256) k_slice_fill(const int32_t *ptr, const int32_t *e2f_ptr, const int32_t *idx, const float *w,
                                                    const float *star2, int32_t ne, const int4 *desc, int32_t *words)
{
    const int32_t e = gid();
    if (e >= ne || e2f_ptr[e + 1] - e2f_ptr[e] > kLightMax) return;
    const int4 dsc = desc[e >> 6];
    const int32_t lane = e & 63;
    for (int32_t k = ptr[e], j = 0; k < ptr[e + 1]; ++k, ++j)
    {
        int32_t id, wb;
        if (star2)
        {
            const int32_t q = idx[k];
            id = pid(q);
            wb = __float_as_int(fsg(q) * star2[id]);
        }
        else
        {
            id = idx[k];
            wb = __float_as_int(w[k]);
        }
        const size_t slot = 64 * size_t(j >> 1) + size_t(lane);
        if (dsc.w)
        {
            int32_t *q = words + size_t(dsc.x) + 4 * slot + 2 * (j & 1);
            q[0] = id;
            q[1] = wb;
        }
        else
        {
            int32_t *q = words + size_t(dsc.x) + 3 * slot;
            // the offset fits 16 bits (the host chose this slice compact)
            q[0] = int32_t(uint32_t(q[0]) | (uint32_t(id - dsc.z) << (16 * (j & 1))));
            q[1 + (j & 1)] = wb;
        }
    }
}
// 16-bit face offsets of the four-lane FP32 cell-vector kernel on tetrahedra (DevMesh32::c2f16, c2fbase): per
// wavefront of 16 cells the smallest face id and whether its faces span less than 2^16
__global__ void __launch_bounds__(256) k_cell16_scan(DevMesh d, int32_t nc, int32_t n_groups, int32_t *base, int8_t *fits)
{
    const int32_t t = gid(), c = t >> 2;
    int32_t lo = INT32_MAX, hi = -1;
    if (c < nc) lo = hi = pid(d.c2f[t]);
    for (int o = 32; o > 0; o >>= 1)
    {
        lo = min(lo, __shfl_xor(lo, o, 64));
        hi = max(hi, __shfl_xor(hi, o, 64));
    }
    if ((threadIdx.x & 63) == 0 && (t >> 6) < n_groups)
    {
        base[t >> 6] = hi < 0 ? 0 : lo;
        fits[t >> 6] = int8_t(hi < 0 || int64_t(hi) - lo <= 0xffff);
    }
}
__global__ void __launch_bounds__(256) k_cell16_fill(DevMesh d, int32_t nc, const int32_t *base, uint16_t *c2f16)
{
    const int32_t t = gid(), c = t >> 2;
    if (c < nc) c2f16[t] = uint16_t(pid(d.c2f[t]) - base[t >> 6]);
}
__global__ void __launch_bounds__(256) k_heavy_flags(const int32_t *e2f_ptr, int32_t ne, int8_t *flag)
{
    const int32_t e = gid();
    if (e < ne) flag[e] = e2f_ptr[e + 1] - e2f_ptr[e] > kLightMax;
}
__global__ void __launch_bounds__(256) k_f64_to_f32(const double *x, int64_t n, float *out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(x[i]);
}
// alpha_c B0_c in FP32 (formed in FP64): the FP32 Hall kernel reads one 3-vector per cell instead of alpha and B0
__global__ void __launch_bounds__(256) k_alpha_f32(const double *alpha, const double *B0, int32_t nc, float *aB0)
{
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    for (int k = 0; k < 3; ++k) aB0[3 * size_t(c) + k] = static_cast<float>(alpha[c] * B0[3 * size_t(c) + k]);
}
} // namespace mhdg

using namespace mhdg;

namespace
{
#ifndef MHDG_INNER_RTOL
#define MHDG_INNER_RTOL 1e-6
#endif
#ifndef MHDG_DGKS
#define MHDG_DGKS 0.1
#endif
// inner FP32 GMRES stops at this fraction of the outer FP64 residual
constexpr double kInnerRtol = MHDG_INNER_RTOL;
// DGKS test: re-orthogonalise when the projection removed more than 1 - kDgks of the vector's norm
constexpr double kDgks = MHDG_DGKS;
// An inner solve also stops once its estimate is below kStopFloor * rtol |b|: the last refinement cycle then ends just
// under the outer tolerance (FP32 rounding at that residual is far below the margin).  Swept 0.25 .. 0.85 on the 2M and
// 8M shells: 0.85 took 44 -> 40 iterations on the 8M first step with no third cycle anywhere.
constexpr double kStopFloor = 0.85;
inline int nb(int64_t n) { return static_cast<int>((n + 255) / 256); }
void chk(hipError_t e, const char *w)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}
} // namespace

#define LAUNCH(NAME, ...)                                                                                                    \
    do                                                                                                                       \
    {                                                                                                                        \
        prof_begin(NAME);                                                                                                    \
        hipLaunchKernelGGL(__VA_ARGS__);                                                                                     \
        prof_end();                                                                                                          \
    } while (0)

void mhdg_ctx::mixed_build()
{
    if (!mixed) mixed = new MixedSolver();
    MixedSolver &M = *mixed;
    if (M.built) return;
    const size_t E = ne, F = nf, C = nc;
    auto f32 = [&](const double *src, size_t n) {
        float *p = static_cast<float *>(alloc(n * sizeof(float)));
        hipLaunchKernelGGL(k_f64_to_f32, dim3(nb(n)), dim3(256), 0, stream, src, int64_t(n), p);
        return p;
    };
    const size_t nc2f = c2f_nnz, nec = ec_nnz;
    M.s.star2 = f32(d.star2, F);
    M.s.m_e = f32(d.m_e, E);
    M.s.recon = f32(d.recon, 3 * nc2f);
    M.s.ecw = f32(d.ecw, nec);
    {
        // slice layout of the edge kernel's rows: per-slice scans on the device, the descriptors here (set-up only).
        // MHDG_WIDE_IDS=1 keeps every slice wide and the cell kernels on 32-bit ids (the same sums; tests compare).
        const int32_t n_slices = int32_t((E + 63) / 64);
        const char *wide_env = std::getenv("MHDG_WIDE_IDS");
        const bool force_wide = wide_env && wide_env[0] == '1';
        auto slices = [&](const int32_t *ptr, const int32_t *idx, int32_t signed_ids, const float *w, const float *star2,
                          const int4 *&desc_out, const int32_t *&words_out) {
            int4 *desc = static_cast<int4 *>(alloc(size_t(std::max(n_slices, 1)) * sizeof(int4)));
            hipLaunchKernelGGL(k_slice_scan, dim3(nb(E)), dim3(256), 0, stream, ptr, idx, signed_ids, d.e2f_ptr, int32_t(E), n_slices, desc);
            std::vector<int4> h(static_cast<size_t>(n_slices));
            chk(hipMemcpyAsync(h.data(), desc, h.size() * sizeof(int4), hipMemcpyDeviceToHost, stream), "D2H");
            chk(hipStreamSynchronize(stream), "sync");
            int64_t acc = 0;
            for (int4 &q : h)
            {
                const int32_t np = q.x, lo = q.y, hi = q.z;
                const bool wide = force_wide || (hi >= 0 && int64_t(hi) - lo > 0xffff);
                if (wide) acc = (acc + 3) & ~int64_t(3); // a wide slice's 16-byte loads are aligned
                q = make_int4(int32_t(acc), np, wide || hi < 0 ? 0 : lo, wide ? 1 : 0);
                acc += int64_t(np) * 64 * (wide ? 4 : 3);
                if (acc > INT32_MAX / 2) throw std::runtime_error("edge rows too large for 32-bit slice offsets");
            }
            chk(hipMemcpyAsync(desc, h.data(), h.size() * sizeof(int4), hipMemcpyHostToDevice, stream), "H2D");
            // + one group of zeroed slots: the kernel loads whole groups of pairs and drops those past a slice's count
            const size_t nw = size_t(acc) + 64 * 4 * size_t(kPairUnroll);
            int32_t *words = static_cast<int32_t *>(alloc(nw * sizeof(int32_t)));
            chk(hipMemsetAsync(words, 0, nw * sizeof(int32_t), stream), "memset");
            hipLaunchKernelGGL(k_slice_fill, dim3(nb(E)), dim3(256), 0, stream, ptr, d.e2f_ptr, idx, w, star2, int32_t(E), desc, words);
            chk(hipStreamSynchronize(stream), "sync"); // h is freed on return
            desc_out = desc;
            words_out = words;
        };
        slices(d.e2f_ptr, d.e2f, 1, nullptr, M.s.star2, M.s.fdesc, M.s.fwords);
        slices(d.ec_ptr, d.ec_cell, 0, M.s.ecw, nullptr, M.s.cdesc, M.s.cwords);
        int8_t *eh = static_cast<int8_t *>(alloc(E));
        hipLaunchKernelGGL(k_heavy_flags, dim3(nb(E)), dim3(256), 0, stream, d.e2f_ptr, int32_t(E), eh);
        M.s.eheavy = eh;
    }
    M.s.evec = f32(d.evec, 3 * E);

    {
        float *aV = static_cast<float *>(alloc(3 * nc2f * sizeof(float)));
        int32_t *nbr = static_cast<int32_t *>(alloc(nc2f * sizeof(int32_t)));
        hipLaunchKernelGGL(k_tables_f32, dim3(nb(nc)), dim3(256), 0, stream, d, aV, nbr);
        M.s.aV = aV; M.s.nbr = nbr;
        // 16-bit face offsets for tetrahedra when every wavefront's faces fit (set-up only; MHDG_WIDE_IDS=1 keeps 32-bit ids)
        const char *wide_env = std::getenv("MHDG_WIDE_IDS");
        M.s.c2f16 = nullptr; M.s.c2fbase = nullptr;
        if (cdeg == kCellLanes && !(wide_env && wide_env[0] == '1'))
        {
            const int32_t n_groups = int32_t((int64_t(nc) * kCellLanes + 63) / 64);
            int32_t *base = static_cast<int32_t *>(alloc(size_t(n_groups) * sizeof(int32_t)));
            int8_t *fits = static_cast<int8_t *>(alloc(size_t(n_groups)));
            hipLaunchKernelGGL(k_cell16_scan, dim3(nb(int64_t(nc) * kCellLanes)), dim3(256), 0, stream, d, int32_t(nc), n_groups, base, fits);
            std::vector<int8_t> hf(static_cast<size_t>(n_groups));
            chk(hipMemcpyAsync(hf.data(), fits, hf.size(), hipMemcpyDeviceToHost, stream), "D2H");
            chk(hipStreamSynchronize(stream), "sync");
            if (std::all_of(hf.begin(), hf.end(), [](int8_t v) { return v != 0; }))
            {
                uint16_t *c16 = static_cast<uint16_t *>(alloc(nc2f * sizeof(uint16_t)));
                hipLaunchKernelGGL(k_cell16_fill, dim3(nb(int64_t(nc) * kCellLanes)), dim3(256), 0, stream, d, int32_t(nc), base, c16);
                M.s.c2f16 = c16; M.s.c2fbase = base;
            }
        }
        M.s.cdeg = cdeg; M.s.fdeg = fdeg;
    }
    M.pv = static_cast<float *>(alloc(E * sizeof(float)));
    M.s.invdiag = static_cast<float *>(alloc(E * sizeof(float)));
    M.s.aB0 = static_cast<float *>(alloc(3 * C * sizeof(float)));
    M.phi = static_cast<float *>(alloc(F * sizeof(float)));
    M.b = static_cast<float *>(alloc(3 * C * sizeof(float)));
    M.g = static_cast<float *>(alloc(3 * C * sizeof(float)));
    M.w = static_cast<float *>(alloc(E * sizeof(float)));
    M.m = restart;
    M.V.resize(M.m + 1);
    for (auto &v : M.V) v = static_cast<float *>(alloc(E * sizeof(float)));
    M.Vptrs = static_cast<float **>(alloc(sizeof(float *) * M.V.size()));
    chk(hipMemcpy(M.Vptrs, M.V.data(), sizeof(float *) * M.V.size(), hipMemcpyHostToDevice), "upload");
    M.h = static_cast<double *>(alloc(64 * sizeof(double)));
    // partial rows of the multi-dot: kOrthoBlocks in the main pass, kRedBlocks in the DGKS re-projection
    M.part = static_cast<double *>(alloc(size_t(std::max(kOrthoBlocks, kRedBlocks)) * (kDotChunk + 1) * sizeof(double)));
    M.part_axpy = static_cast<double *>(alloc(size_t(kOrthoBlocks) * sizeof(double)));
    chk(hipStreamSynchronize(stream), "sync");
    M.built = true;
}

int mhdg_ctx::solve_mixed(const double *Phi_n, const double *Eex, const int8_t *dir, double dt, double *Eout, double *rel_res, bool use_mg,
                          const int8_t *dir_host)
{
    mixed_build();
    MixedSolver &M = *mixed;
    const int m = M.m;
    const int max_it = cfg.ind_max_it;
    const double rtol = cfg.ind_rtol;
    const double inner_rtol = kInnerRtol;
    if (ideal_cfg && u_ideal) throw std::runtime_error("mixed-precision solve: the implicit ideal part has no FP32 operator; use precision 0");
    if (use_mg)
    {
        // the coarse operators are built for a scalar dt
        if (dte_cur) throw std::runtime_error("multigrid preconditioner: per-edge time steps are not supported; use precision 1");
        if (!dir_host) throw std::runtime_error("multigrid preconditioner: no host copy of the Dirichlet set");
        mg_build(dir_host);
    }
    const float *dte32 = nullptr;
    if (dte_cur)
    {
        if (!M.dte32) M.dte32 = static_cast<float *>(alloc(size_t(ne) * sizeof(float)));
        LAUNCH("f64_to_f32", k_f64_to_f32, dim3(nb(ne)), dim3(256), 0, stream, dte_cur, int64_t(ne), M.dte32);
        dte32 = M.dte32;
    }
    // frozen state and Jacobi diagonal in FP32
    LAUNCH("alpha_f32", k_alpha_f32, dim3(nb(nc)), dim3(256), 0, stream, alpha, B0, nc, M.s.aB0);
    LAUNCH("solve_init", k_solve_init, dim3(nb(ne)), dim3(256), 0, stream, Eex, have_corr ? corr : nullptr, dir, d.m_e, d.d2, dt, dte_cur, ne, x0,
           invdiag);
    LAUNCH("f64_to_f32", k_f64_to_f32, dim3(nb(ne)), dim3(256), 0, stream, invdiag, int64_t(ne), const_cast<float *>(M.s.invdiag));
    if (use_mg)
    {
        mg_set_solve(dir, static_cast<float>(dt));
        mg_prepare(dt);
    }
    // b = E_expl + F(Phi^n) and the first cycle's residual b - A x0, in one pass
    rhs_residual(Phi_n, Eex, x0, dir, dt);
    chk(hipMemsetAsync(delta, 0, size_t(ne) * 8, stream), "memset");
    const double bb = dot(bvec, bvec, 0, 0);
    const double bnorm = std::sqrt(std::max(bb, 1e-300));
    const float dtf = dte32 ? 1.0f : static_cast<float>(dt); // with dt_e, d1 carries the step
    int total = 0;
    double beta = 0;
    std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> cs(m), sn(m), g(m + 1), yv(m), hc(64), c2(64);
    const float *const *Vp = M.Vptrs;
    // MHDG_SOLVE_TRACE=1: one stderr line per refinement cycle (FP64 residual, inner iterations, inner estimate)
    static const bool trace = std::getenv("MHDG_SOLVE_TRACE") && std::getenv("MHDG_SOLVE_TRACE")[0] == '1';
    int cycle_start = 0;
    double est = 0;
    for (int outer = 0; outer < 200; ++outer)
    {
        // FP64 residual of x0 + delta (delta = 0 in the first cycle: rhs_residual formed it)
        if (outer > 0)
        {
            LAUNCH("axpy", k_add, dim3(nb(ne)), dim3(256), 0, stream, x0, delta, ne, tmp);
            apply_op(nullptr, tmp, nullptr, dir, dt, 1, 0, bvec, rvec);
        }
        beta = dot(rvec, rvec, 1, 1);
        if (trace)
            std::fprintf(stderr, "[mixed] cycle %d: residual %.3e |b|, previous cycle %d iterations, inner estimate %.3e |b|\n", outer,
                         beta / bnorm, total - cycle_start, est / bnorm);
        cycle_start = total;
        if (!(beta > rtol * bnorm) || total >= max_it) break;
        LAUNCH("scale", k_to_f32_scaled, dim3(nb(ne)), dim3(256), 0, stream, rvec, scal + 1, M.s.invdiag, dir, ne, M.V[0], M.pv);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        const double stop = std::max(inner_rtol * beta, kStopFloor * rtol * bnorm);
        int j = 0;
        for (; j < m && total < max_it; ++j, ++total)
        {
            if (use_mg)
            {
                // flexible: Z_j = M V_j (one V-cycle), w = A Z_j
                mg_cycle(0, M.V[j], M.Z[j]);
                op32(kOpApply, op_hall ? 1 : 0, M.Z[j], nullptr, nullptr, 0.0f, M.w);
            }
            else
            {
                // w = (A P - I) V_j in FP32
                LAUNCH("d1_f32", k_d1_f32, dim3(nb(nf)), dim3(256), 0, stream, d, M.s, M.pv, dte32, M.phi);
                if (op_hall)
                {
                    LAUNCH("bcell_f32", k_bcell_f32_g<kCellLanes>, dim3(nb(kCellLanes * int64_t(nc))), dim3(256), 0, stream, d, M.s, M.phi, M.b);
                    LAUNCH("hall_g_f32", k_hall_g_f32_g<kCellLanes>, dim3(nb(kCellLanes * int64_t(nc))), dim3(256), 0, stream, d, M.s, M.b, M.g);
                }
                const int32_t n_light = int32_t((int64_t(ne) + kEdgeBlock - 1) / kEdgeBlock);
                LAUNCH("edge_apply_f32", k_edge_apply_f32<kOpShifted>, dim3(n_light + n_heavy), dim3(kEdgeBlock), 0, stream, d, M.s,
                       int32_t(op_res), int32_t(op_hall), M.phi, M.g, M.V[j], M.pv, dir, dtf, M.w, static_cast<const int32_t *>(heavy), n_light,
                       static_cast<const float *>(nullptr), 0.0f);
            }
            // classical Gram-Schmidt against V_0..V_j; the update also normalises V_{j+1} unless the DGKS test asks
            // for a second projection
            for (int i0 = 0; i0 <= j; i0 += kDotChunk)
            {
                const int nv = std::min(kDotChunk, j + 1 - i0);
                LAUNCH("multidot", kMultidot[nv], dim3(kOrthoBlocks), dim3(256), 0, stream, Vp, i0, M.w, ne, M.part, i0 == 0 ? 1 : 0,
                       RedScratch{nullptr, nullptr}, M.h);
                LAUNCH("multidot_final", k_multidot_final, dim3(1), dim3(256), 0, stream, static_cast<const double *>(M.part), kOrthoBlocks,
                       i0, nv, i0 == 0 ? 1 : 0, M.h);
            }
            LAUNCH("multiaxpy", k_multiaxpy_scale_f32, dim3(kOrthoBlocks), dim3(256), 0, stream, Vp, j + 1, M.h, M.w, ne, double(kDgks),
                   M.s.invdiag, dir, M.V[j + 1], M.pv, RedScratch{M.part_axpy, rs.ticket});
            d2h(hc.data(), M.h, sizeof(double) * 64);
            double hn = hc[j + 1];
            if (hc[62] != 0.0)
            {
                // DGKS: one more projection
                for (int i0 = 0; i0 <= j; i0 += kDotChunk)
                    LAUNCH("multidot", kMultidot[std::min(kDotChunk, j + 1 - i0)], dim3(kRedBlocks), dim3(256), 0, stream, Vp, i0, M.w,
                           ne, M.part, 0, rs, M.h);
                LAUNCH("multiaxpy", k_multiaxpy_f32, dim3(kRedBlocks), dim3(256), 0, stream, Vp, j + 1, M.h, M.w, ne, rs, M.h + j + 1);
                d2h(c2.data(), M.h, sizeof(double) * 64);
                for (int i = 0; i <= j; ++i) hc[i] += c2[i];
                hn = c2[j + 1];
                chk(hipMemcpyAsync(M.h + j + 1, &hn, 8, hipMemcpyHostToDevice, stream), "H2D");
                LAUNCH("scale", k_scale_f32, dim3(nb(ne)), dim3(256), 0, stream, M.w, M.h + j + 1, M.s.invdiag, dir, ne, M.V[j + 1], M.pv);
            }
            for (int i = 0; i <= j; ++i) H[i][j] = hc[i];
            if (!use_mg) H[j][j] += 1.0; // Arnoldi ran on A P - I
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
            est = std::abs(g[j + 1]);
            if (std::abs(g[j + 1]) <= stop || !(hn > 1e-300)) { ++j; ++total; break; }
        }
        for (int i = j - 1; i >= 0; --i)
        {
            double s = g[i];
            for (int k = i + 1; k < j; ++k) s -= H[i][k] * yv[k];
            yv[i] = s / H[i][i];
        }
        chk(hipMemcpyAsync(ydev, yv.data(), sizeof(double) * std::max(j, 1), hipMemcpyHostToDevice, stream), "H2D");
        if (use_mg)
            LAUNCH("update_flex", k_update_flex, dim3(nb(ne)), dim3(256), 0, stream, static_cast<const float *const *>(M.Zptrs), ydev, j, dir, ne, delta);
        else
            LAUNCH("update_mixed", k_update_mixed, dim3(nb(ne / 4 + 1)), dim3(256), 0, stream, Vp, ydev, j, M.s.invdiag, dir, ne, delta);
    }
    LAUNCH("solve_final", k_solve_final, dim3(nb(ne)), dim3(256), 0, stream, x0, delta, Eex, dir, ne, Eout, corr);
    check_launch("mixed solve");
    have_corr = true;
    if (rel_res) *rel_res = beta / bnorm;
    return total;
}
