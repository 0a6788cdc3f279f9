// Implicit Hall/resistive induction operator (ImplicitInduction::apply_F / apply_A,
// Common/induction/ImplicitInduction.cpp:127-198) as face, cell and edge gathers, plus the GMRES vector
// kernels.  Reductions are deterministic: a fixed grid, a fixed per-block tree, and one block combines the
// block partials in a fixed tree (k_finish, a second launch), so reruns are bit-identical.  (A last-block
// ticket in the same launch costs ~90 us per reduction here: the 1024 blocks finish together and their
// fenced atomics on one counter serialise.)
#include "induction.h"

namespace mhdg
{
// ------------------------------------------------------------------------------------ reductions
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1)
    {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Called by every thread of every block with the block's partial value (valid in thread 0).  The last
// block writes op(sum of partials) to *out.  op: 0 = sum, 1 = sqrt(sum).
__device__ void grid_finish(double part, RedScratch rs, double *out, int op)
{
    __shared__ double sh[kRedThreads];
    __shared__ int last;
    if (!rs.ticket)
    {
        // partials only; k_finish combines them
        if (threadIdx.x == 0) rs.partial[blockIdx.x] = part;
        return;
    }
    if (threadIdx.x == 0)
    {
        rs.partial[blockIdx.x] = part;
        __threadfence();
        const unsigned prev = atomicAdd(rs.ticket, 1u);
        last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double v = 0.0;
    for (int i = threadIdx.x; i < static_cast<int>(gridDim.x); i += kRedThreads)
        v += __hip_atomic_load(&rs.partial[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = block_sum(v, sh);
    if (threadIdx.x == 0)
    {
        *out = op == 1 ? sqrt(v) : v;
        *rs.ticket = 0u;
    }
}

// out = op(sum of the n block partials) in the fixed tree of grid_finish's last block (one block)
__global__ void __launch_bounds__(kRedThreads) k_finish(const double *partial, int32_t n, double *out, int op)
{
    __shared__ double sh[kRedThreads];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += kRedThreads) v += partial[i];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) *out = op == 1 ? sqrt(v) : v;
}

template <class T>
__global__ void __launch_bounds__(kRedThreads) k_dot(const T *a, const T *b, int32_t n, RedScratch rs, double *out, int op)
{
    __shared__ double sh[kRedThreads];
    double s = 0.0;
    for (int32_t i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads)
        s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
    grid_finish(block_sum(s, sh), rs, out, op);
}
template __global__ void k_dot<double>(const double *, const double *, int32_t, RedScratch, double *, int);
template __global__ void k_dot<float>(const float *, const float *, int32_t, RedScratch, double *, int);

// Modified Gram-Schmidt step i (ImplicitInduction.cpp:341-347): w -= h[i-1] V_{i-1} (i > 0), then
// h[i] = <w, V_i>; with `norm` set: w -= h[i-1] V_{i-1}, then h[i] = |w|.
template <class T>
__global__ void __launch_bounds__(kRedThreads) k_mgs(T *w, const T *Vprev, const T *Vi, int32_t n, double *h, int32_t i, int32_t norm,
                                                    RedScratch rs)
{
    __shared__ double sh[kRedThreads];
    const double hp = i > 0 ? h[i - 1] : 0.0;
    double s = 0.0;
    for (int32_t e = blockIdx.x * kRedThreads + threadIdx.x; e < n; e += gridDim.x * kRedThreads)
    {
        double we = static_cast<double>(w[e]);
        if (i > 0)
        {
            we -= hp * static_cast<double>(Vprev[e]);
            w[e] = static_cast<T>(we);
            we = static_cast<double>(w[e]);
        }
        s += norm ? we * we : we * static_cast<double>(Vi[e]);
    }
    grid_finish(block_sum(s, sh), rs, h + i, norm ? 1 : 0);
}
template __global__ void k_mgs<double>(double *, const double *, const double *, int32_t, double *, int32_t, int32_t, RedScratch);
template __global__ void k_mgs<float>(float *, const float *, const float *, int32_t, double *, int32_t, int32_t, RedScratch);

// V_{j+1} = w / h (reference divides, ImplicitInduction.cpp:342)
template <class T>
__global__ void __launch_bounds__(256) k_scale_div(const T *w, const double *h, int32_t n, T *out)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const double hv = h[0];
    if (e >= n || !(hv > 1e-300)) return;
    out[e] = static_cast<T>(static_cast<double>(w[e]) / hv);
}
template __global__ void k_scale_div<double>(const double *, const double *, int32_t, double *);
template __global__ void k_scale_div<float>(const float *, const double *, int32_t, float *);

// r / beta into V_0 (ImplicitInduction.cpp:330)
__global__ void __launch_bounds__(256) k_div_scalar(const double *r, const double *beta, int32_t n, double *out)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = r[e] / beta[0];
}

// ------------------------------------------------------------------------------------ operator
// Phi'_f = sum_k f2e_k x_e, x = dir ? 0 : v * invdiag (right Jacobi preconditioner) or x = v; with a per-edge
// pseudo-time step dte (nullable) the terms are (f2e_k dte_e) x_e (ImplicitInduction.cpp:187-194).
__global__ void __launch_bounds__(256) k_d1(DevMesh d, const double *v, const double *invdiag, const int8_t *dir, const double *dte,
                                            double *phi)
{
    const int32_t f = gid();
    if (f >= d.nf) return;
    double s = 0.0;
    if (d.fdeg == 3 && !dte)
    {
        // triangles: the face's three edge entries in one 12-byte load, then the gathers, summed in entry order
        const int3 q = reinterpret_cast<const int3 *>(d.f2e)[f];
        const int32_t e[3] = {pid(q.x), pid(q.y), pid(q.z)};
        double x[3];
#pragma unroll
        for (int u = 0; u < 3; ++u)
        {
            const double ve = v[e[u]];
            if (invdiag)
            {
                const double iv = invdiag[e[u]];
                x[u] = dir[e[u]] ? 0.0 : ve * iv;
            }
            else x[u] = ve;
        }
        s += psg(q.x) * x[0];
        s += psg(q.y) * x[1];
        s += psg(q.z) * x[2];
        phi[f] = s;
        return;
    }
    const int32_t k0 = d.f2e_ptr[f], k1 = d.f2e_ptr[f + 1];
    for (int32_t k = k0; k < k1; k += kRowUnroll)
    {
        // loads of the group first, sums in order
        int32_t q[kRowUnroll];
        double x[kRowUnroll], t[kRowUnroll];
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k + u < k1) q[u] = d.f2e[k + u];
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k + u < k1)
            {
                // v and invdiag gathered alongside dir, not behind its test
                const int32_t e = pid(q[u]);
                const double ve = v[e];
                if (invdiag)
                {
                    const double iv = invdiag[e];
                    x[u] = dir[e] ? 0.0 : ve * iv;
                }
                else x[u] = ve;
                if (dte) t[u] = dte[e];
            }
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k + u < k1) s += dte ? psg(q[u]) * t[u] * x[u] : psg(q[u]) * x[u];
    }
    phi[f] = s;
}

// b_c = sum_k recon_k Phi_{f_k} (ImplicitInduction.cpp:150-160)
__global__ void __launch_bounds__(256) k_bcell(DevMesh d, const double *phi, double *b)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    double bx = 0.0, by = 0.0, bz = 0.0;
    const int32_t k0 = d.c2f_ptr[c], k1 = d.c2f_ptr[c + 1];
    for (int32_t k = k0; k < k1; k += kRowUnroll)
    {
        double x[kRowUnroll];
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k + u < k1) x[u] = phi[pid(d.c2f[k + u])];
#pragma unroll
        for (int u = 0; u < kRowUnroll; ++u)
            if (k + u < k1)
            {
                const size_t r = 3 * size_t(k + u);
                bx += d.recon[r] * x[u]; by += d.recon[r + 1] * x[u]; bz += d.recon[r + 2] * x[u];
            }
    }
    b[3 * size_t(c)] = bx; b[3 * size_t(c) + 1] = by; b[3 * size_t(c) + 2] = bz;
}

// g_c = alpha_c (J_c[b] x B0_c), zero where alpha_c = 0 (ImplicitInduction.cpp:162-174)
__global__ void __launch_bounds__(256) k_hall_g(DevMesh d, const double *b, const double *alpha, const double *B0, double *g)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    const double a = alpha[c];
    V3 out;
    if (a != 0.0)
    {
        V3 sum;
        const V3 bc = L3(b, c);
        const int32_t k0 = d.c2f_ptr[c], k1 = d.c2f_ptr[c + 1];
        for (int32_t k = k0; k < k1; k += kRowUnroll)
        {
            int32_t q[kRowUnroll], o[kRowUnroll];
            V3 A[kRowUnroll], bo[kRowUnroll];
#pragma unroll
            for (int u = 0; u < kRowUnroll; ++u)
                if (k + u < k1) q[u] = d.c2f[k + u];
#pragma unroll
            for (int u = 0; u < kRowUnroll; ++u)
                if (k + u < k1)
                {
                    const int32_t f = pid(q[u]), l = d.fL[f], r = d.fR[f];
                    o[u] = (l >= 0 && r >= 0) ? (l == c ? r : l) : -1;
                    A[u] = L3(d.farea, f);
                }
#pragma unroll
            for (int u = 0; u < kRowUnroll; ++u)
                if (k + u < k1 && o[u] >= 0) bo[u] = L3(b, o[u]);
#pragma unroll
            for (int u = 0; u < kRowUnroll; ++u)
                if (k + u < k1)
                {
                    V3 bf = bc;
                    if (o[u] >= 0) bf = 0.5 * (bc + bo[u]);
                    sum += cross(psg(q[u]) * A[u], bf);
                }
        }
        const V3 J = (1.0 / d.cvol[c]) * sum;
        out = a * cross(J, L3(B0, c));
    }
    st3(g, c, out);
}

// k_bcell and k_hall_g with G lanes per cell: lane q forms the terms of entries base + q (the lanes' table and
// gather loads side by side), and every lane adds the group's terms in entry order through shuffles, so each sum
// runs in the order of the one-thread kernels and the results are the same bit for bit.
template <int G> __global__ void __launch_bounds__(256) k_bcell_g(DevMesh d, const double *phi, double *b)
{
    const int32_t t = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    const int32_t c = t / G, lane = t % G;
    const bool ok = c < d.nc;
    // a fixed row length (tetrahedra) spares the row-bound loads ahead of the gathers
    const int32_t k0 = !ok ? 0 : d.cdeg > 0 ? c * d.cdeg : d.c2f_ptr[c];
    const int32_t k1 = !ok ? 0 : d.cdeg > 0 ? k0 + d.cdeg : d.c2f_ptr[c + 1];
    double bx = 0.0, by = 0.0, bz = 0.0;
    for (int32_t base = k0; base < k1; base += G)
    {
        const int32_t k = base + lane;
        double px = 0.0, py = 0.0, pz = 0.0;
        if (k < k1)
        {
            const double v = phi[pid(d.c2f[k])];
            px = d.recon[3 * size_t(k)] * v; py = d.recon[3 * size_t(k) + 1] * v; pz = d.recon[3 * size_t(k) + 2] * v;
        }
#pragma unroll
        for (int q = 0; q < G; ++q)
        {
            const double qx = __shfl(px, q, G), qy = __shfl(py, q, G), qz = __shfl(pz, q, G);
            if (base + q < k1) { bx += qx; by += qy; bz += qz; }
        }
    }
    if (ok && lane == 0) { b[3 * size_t(c)] = bx; b[3 * size_t(c) + 1] = by; b[3 * size_t(c) + 2] = bz; }
}
template <int G> __global__ void __launch_bounds__(256) k_hall_g_g(DevMesh d, const double *b, const double *alpha, const double *B0, double *g)
{
    const int32_t t = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    const int32_t c = t / G, lane = t % G;
    const bool ok = c < d.nc;
    const double a = ok ? alpha[c] : 0.0;
    V3 out;
    // the sum's loads do not wait for alpha (a load behind the alpha test waited for it); only the result does
    if (ok)
    {
        V3 sum;
        const V3 bc = L3(b, c);
        const int32_t k0 = d.cdeg > 0 ? c * d.cdeg : d.c2f_ptr[c], k1 = d.cdeg > 0 ? k0 + d.cdeg : d.c2f_ptr[c + 1];
        for (int32_t base = k0; base < k1; base += G)
        {
            const int32_t k = base + lane;
            V3 t3;
            if (k < k1)
            {
                // the neighbour from the per-entry table: its gather waits on one load, not on the face's f2c
                const int32_t v = d.c2f[k], f = pid(v), o = d.c2n[k];
                V3 bf = bc;
                if (o >= 0) bf = 0.5 * (bc + L3(b, o));
                t3 = cross(psg(v) * L3(d.farea, f), bf);
            }
#pragma unroll
            for (int q = 0; q < G; ++q)
            {
                const V3 tq(__shfl(t3.x, q, G), __shfl(t3.y, q, G), __shfl(t3.z, q, G));
                if (base + q < k1) sum += tq;
            }
        }
        if (a != 0.0)
        {
            const V3 J = (1.0 / d.cvol[c]) * sum;
            out = a * cross(J, L3(B0, c));
        }
    }
    if (ok && lane == 0) st3(g, c, out);
}
template __global__ void k_bcell_g<kCellLanes64>(DevMesh, const double *, double *);
template __global__ void k_hall_g_g<kCellLanes64>(DevMesh, const double *, const double *, const double *, double *);

// Edge assembly of apply_F (ImplicitInduction.cpp:139-182): part0 = sum_f f2e star2 Phi (faces ascending),
// part1 = sum_c g_c . (w_ec t_e) (cells ascending), part2 = imask ? ide : 0 (implicit ideal part, ide nullable),
// F = m part0 + part1 + part2.
// mode 0: out = F.  mode 1 (apply_A): out = x + dt F with x = dir ? 0 : v*invdiag (invdiag non-null) or v;
// with `mask` set, Dirichlet rows of the result are zeroed (Arnoldi vector, ImplicitInduction.cpp:340).
__global__ void __launch_bounds__(256) k_edge_apply(DevMesh d, int32_t resistive, int32_t hall, const double *phi, const double *g,
                                                    const double *ide, const int8_t *imask, const double *v, const double *invdiag,
                                                    const int8_t *dir, double dt, int32_t mode, int32_t mask, const double *bvec, double *out)
{
    const int32_t e = gid();
    if (e >= d.ne) return;
    const int32_t f0 = d.e2f_ptr[e], f1 = d.e2f_ptr[e + 1];
    if (f1 - f0 > kLightMax) return; // k_edge_apply_heavy
    double p0 = 0.0, p1 = 0.0;
    // each group's index loads, then its gathers, then the sums in entry order
    if (resistive)
        for (int32_t k = f0; k < f1; k += kEdgeUnroll)
        {
            int32_t q[kEdgeUnroll];
            double st[kEdgeUnroll], x[kEdgeUnroll];
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < f1) q[u] = d.e2f[k + u];
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < f1) { st[u] = d.star2[pid(q[u])]; x[u] = phi[pid(q[u])]; }
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < f1)
                {
                    const double val = st[u] * x[u];
                    p0 += psg(q[u]) * val;
                }
        }
    if (hall)
    {
        const V3 t = L3(d.evec, e);
        const int32_t c0 = d.ec_ptr[e], c1 = d.ec_ptr[e + 1];
        for (int32_t k = c0; k < c1; k += kEdgeUnroll)
        {
            int32_t q[kEdgeUnroll];
            double w[kEdgeUnroll];
            V3 x[kEdgeUnroll];
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < c1) { q[u] = d.ec_cell[k + u]; w[u] = d.ecw[k + u]; }
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < c1) x[u] = L3(g, q[u]);
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < c1)
                {
                    const V3 wt = w[u] * t;
                    p1 += dot(x[u], wt);
                }
        }
    }
    const double p2 = ide ? (imask[e] ? ide[e] : 0.0) : 0.0;
    const double F = d.m_e[e] * p0 + p1 + p2;
    double r;
    if (mode == 0) r = F;
    else
    {
        const double x = invdiag ? (dir[e] ? 0.0 : v[e] * invdiag[e]) : v[e];
        r = x + dt * F;
    }
    if (mask && dir[e]) r = 0.0;
    if (bvec) r = dir[e] ? 0.0 : bvec[e] - r; // residual form r = b - A x
    out[e] = r;
}

// k_edge_apply for the high-valence edges, one block per edge: terms in parallel into LDS, sums sequential
// in the reference's order (faces ascending, cells ascending).
__global__ void __launch_bounds__(256) k_edge_apply_heavy(DevMesh d, int32_t resistive, int32_t hall, const int32_t *heavy, const double *phi,
                                                          const double *g, const double *ide, const int8_t *imask, const double *v,
                                                          const double *invdiag, const int8_t *dir, double dt, int32_t mode, int32_t mask,
                                                          const double *bvec, double *out)
{
    __shared__ double t0[kHeavyMax], t1[2 * kHeavyMax];
    const int32_t e = heavy[blockIdx.x];
    const int32_t f0 = d.e2f_ptr[e], nf = d.e2f_ptr[e + 1] - f0;
    const int32_t c0 = d.ec_ptr[e], ncl = d.ec_ptr[e + 1] - c0;
    if (resistive)
        for (int32_t i = threadIdx.x; i < nf; i += blockDim.x)
        {
            const int32_t q = d.e2f[f0 + i], f = pid(q);
            const double val = d.star2[f] * phi[f];
            t0[i] = psg(q) * val;
        }
    if (hall)
    {
        const V3 t = L3(d.evec, e);
        for (int32_t i = threadIdx.x; i < ncl; i += blockDim.x)
        {
            const V3 wt = d.ecw[c0 + i] * t;
            t1[i] = dot(L3(g, d.ec_cell[c0 + i]), wt);
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double p0 = 0.0, p1 = 0.0;
    if (resistive) for (int32_t i = 0; i < nf; ++i) p0 += t0[i];
    if (hall) for (int32_t i = 0; i < ncl; ++i) p1 += t1[i];
    const double p2 = ide ? (imask[e] ? ide[e] : 0.0) : 0.0;
    const double F = d.m_e[e] * p0 + p1 + p2;
    double r;
    if (mode == 0) r = F;
    else
    {
        const double x = invdiag ? (dir[e] ? 0.0 : v[e] * invdiag[e]) : v[e];
        r = x + dt * F;
    }
    if (mask && dir[e]) r = 0.0;
    if (bvec) r = dir[e] ? 0.0 : bvec[e] - r;
    out[e] = r;
}

// ---------------------------------------------------------------- right-hand side and first residual, fused
// A solve starts with two products of F on different vectors, b = E_expl + F(Phi^n) and r0 = b - A x0 (A x0 formed
// from Phi' = d1 x0). The kernels below evaluate both in one pass over the tables: every row index and table value
// is loaded once for the two vectors, and each vector's terms are the same operations in the same order as in
// k_bcell_g, k_hall_g_g and k_edge_apply, so b and r0 are the same bits as the two separate products give.
template <int G> __global__ void __launch_bounds__(256) k_bcell2_g(DevMesh d, const double *phi1, const double *phi2, double *b1, double *b2)
{
    const int32_t t = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    const int32_t c = t / G, lane = t % G;
    const bool ok = c < d.nc;
    const int32_t k0 = !ok ? 0 : d.cdeg > 0 ? c * d.cdeg : d.c2f_ptr[c];
    const int32_t k1 = !ok ? 0 : d.cdeg > 0 ? k0 + d.cdeg : d.c2f_ptr[c + 1];
    double bx[2] = {0.0, 0.0}, by[2] = {0.0, 0.0}, bz[2] = {0.0, 0.0};
    for (int32_t base = k0; base < k1; base += G)
    {
        const int32_t k = base + lane;
        double px[2] = {0.0, 0.0}, py[2] = {0.0, 0.0}, pz[2] = {0.0, 0.0};
        if (k < k1)
        {
            const int32_t f = pid(d.c2f[k]);
            const double v[2] = {phi1[f], phi2[f]};
            const double r0 = d.recon[3 * size_t(k)], r1 = d.recon[3 * size_t(k) + 1], r2 = d.recon[3 * size_t(k) + 2];
#pragma unroll
            for (int n = 0; n < 2; ++n) { px[n] = r0 * v[n]; py[n] = r1 * v[n]; pz[n] = r2 * v[n]; }
        }
#pragma unroll
        for (int q = 0; q < G; ++q)
#pragma unroll
            for (int n = 0; n < 2; ++n)
            {
                const double qx = __shfl(px[n], q, G), qy = __shfl(py[n], q, G), qz = __shfl(pz[n], q, G);
                if (base + q < k1) { bx[n] += qx; by[n] += qy; bz[n] += qz; }
            }
    }
    if (ok && lane == 0)
    {
        b1[3 * size_t(c)] = bx[0]; b1[3 * size_t(c) + 1] = by[0]; b1[3 * size_t(c) + 2] = bz[0];
        b2[3 * size_t(c)] = bx[1]; b2[3 * size_t(c) + 1] = by[1]; b2[3 * size_t(c) + 2] = bz[1];
    }
}
template <int G>
__global__ void __launch_bounds__(256) k_hall_g2_g(DevMesh d, const double *b1, const double *b2, const double *alpha, const double *B0,
                                                   double *g1, double *g2)
{
    const int32_t t = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    const int32_t c = t / G, lane = t % G;
    const bool ok = c < d.nc;
    const double a = ok ? alpha[c] : 0.0;
    V3 out[2];
    if (ok)
    {
        V3 sum[2];
        const V3 bc[2] = {L3(b1, c), L3(b2, c)};
        const int32_t k0 = d.cdeg > 0 ? c * d.cdeg : d.c2f_ptr[c], k1 = d.cdeg > 0 ? k0 + d.cdeg : d.c2f_ptr[c + 1];
        for (int32_t base = k0; base < k1; base += G)
        {
            const int32_t k = base + lane;
            V3 t3[2];
            if (k < k1)
            {
                const int32_t v = d.c2f[k], f = pid(v), o = d.c2n[k];
                const V3 A = psg(v) * L3(d.farea, f);
                V3 bf[2] = {bc[0], bc[1]};
                if (o >= 0)
                {
                    bf[0] = 0.5 * (bc[0] + L3(b1, o));
                    bf[1] = 0.5 * (bc[1] + L3(b2, o));
                }
                t3[0] = cross(A, bf[0]);
                t3[1] = cross(A, bf[1]);
            }
#pragma unroll
            for (int q = 0; q < G; ++q)
#pragma unroll
                for (int n = 0; n < 2; ++n)
                {
                    const V3 tq(__shfl(t3[n].x, q, G), __shfl(t3[n].y, q, G), __shfl(t3[n].z, q, G));
                    if (base + q < k1) sum[n] += tq;
                }
        }
        if (a != 0.0)
        {
            const V3 B = L3(B0, c);
            const double iv = 1.0 / d.cvol[c];
#pragma unroll
            for (int n = 0; n < 2; ++n)
            {
                const V3 J = iv * sum[n];
                out[n] = a * cross(J, B);
            }
        }
    }
    if (ok && lane == 0) { st3(g1, c, out[0]); st3(g2, c, out[1]); }
}
template __global__ void k_bcell2_g<kCellLanes64>(DevMesh, const double *, const double *, double *, double *);
template __global__ void k_hall_g2_g<kCellLanes64>(DevMesh, const double *, const double *, const double *, const double *, double *,
                                                   double *);

// The edge sums of F for (phi1, g1) and (phi2, g2), then b = E_expl + F1 and r0 = dir ? 0 : b - (x + dt F2)
// (k_edge_apply's mode 0 and mode 1 with bvec, no implicit ideal part)
__device__ __forceinline__ void rhs_residual_out(const DevMesh &d, int32_t e, double P0a, double P1a, double P0b, double P1b,
                                                 const double *E_expl, const double *x, const int8_t *dir, double dt, double *bvec,
                                                 double *r)
{
    const double p2 = 0.0;
    const double F1 = d.m_e[e] * P0a + P1a + p2, F2 = d.m_e[e] * P0b + P1b + p2;
    const double bv = E_expl[e] + F1;
    double rr = x[e] + dt * F2;
    rr = dir[e] ? 0.0 : bv - rr;
    bvec[e] = bv;
    r[e] = rr;
}
__global__ void __launch_bounds__(256) k_edge_rhs_residual(DevMesh d, int32_t resistive, int32_t hall, const double *phi1, const double *g1,
                                                           const double *phi2, const double *g2, const double *E_expl, const double *x,
                                                           const int8_t *dir, double dt, double *bvec, double *r)
{
    const int32_t e = gid();
    if (e >= d.ne) return;
    const int32_t f0 = d.e2f_ptr[e], f1 = d.e2f_ptr[e + 1];
    if (f1 - f0 > kLightMax) return; // k_edge_rhs_residual_heavy
    double p0[2] = {0.0, 0.0}, p1[2] = {0.0, 0.0};
    if (resistive)
        for (int32_t k = f0; k < f1; k += kEdgeUnroll)
        {
            int32_t q[kEdgeUnroll];
            double st[kEdgeUnroll], xa[kEdgeUnroll], xb[kEdgeUnroll];
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < f1) q[u] = d.e2f[k + u];
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < f1) { const int32_t f = pid(q[u]); st[u] = d.star2[f]; xa[u] = phi1[f]; xb[u] = phi2[f]; }
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < f1)
                {
                    const double va = st[u] * xa[u], vb = st[u] * xb[u];
                    p0[0] += psg(q[u]) * va;
                    p0[1] += psg(q[u]) * vb;
                }
        }
    if (hall)
    {
        const V3 t = L3(d.evec, e);
        const int32_t c0 = d.ec_ptr[e], c1 = d.ec_ptr[e + 1];
        for (int32_t k = c0; k < c1; k += kEdgeUnroll)
        {
            int32_t q[kEdgeUnroll];
            double w[kEdgeUnroll];
            V3 xa[kEdgeUnroll], xb[kEdgeUnroll];
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < c1) { q[u] = d.ec_cell[k + u]; w[u] = d.ecw[k + u]; }
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < c1) { xa[u] = L3(g1, q[u]); xb[u] = L3(g2, q[u]); }
#pragma unroll
            for (int u = 0; u < kEdgeUnroll; ++u)
                if (k + u < c1)
                {
                    const V3 wt = w[u] * t;
                    p1[0] += dot(xa[u], wt);
                    p1[1] += dot(xb[u], wt);
                }
        }
    }
    rhs_residual_out(d, e, p0[0], p1[0], p0[1], p1[1], E_expl, x, dir, dt, bvec, r);
}
// the high-valence edges, one block per edge (terms in parallel into LDS, sums sequential, as k_edge_apply_heavy)
__global__ void __launch_bounds__(256) k_edge_rhs_residual_heavy(DevMesh d, int32_t resistive, int32_t hall, const int32_t *heavy,
                                                                 const double *phi1, const double *g1, const double *phi2, const double *g2,
                                                                 const double *E_expl, const double *x, const int8_t *dir, double dt,
                                                                 double *bvec, double *r)
{
    __shared__ double t0[2][kHeavyMax], t1[2][2 * kHeavyMax];
    const int32_t e = heavy[blockIdx.x];
    const int32_t f0 = d.e2f_ptr[e], nf = d.e2f_ptr[e + 1] - f0;
    const int32_t c0 = d.ec_ptr[e], ncl = d.ec_ptr[e + 1] - c0;
    if (resistive)
        for (int32_t i = threadIdx.x; i < nf; i += blockDim.x)
        {
            const int32_t q = d.e2f[f0 + i], f = pid(q);
            const double va = d.star2[f] * phi1[f], vb = d.star2[f] * phi2[f];
            t0[0][i] = psg(q) * va;
            t0[1][i] = psg(q) * vb;
        }
    if (hall)
    {
        const V3 t = L3(d.evec, e);
        for (int32_t i = threadIdx.x; i < ncl; i += blockDim.x)
        {
            const V3 wt = d.ecw[c0 + i] * t;
            const int32_t c = d.ec_cell[c0 + i];
            t1[0][i] = dot(L3(g1, c), wt);
            t1[1][i] = dot(L3(g2, c), wt);
        }
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double p0[2] = {0.0, 0.0}, p1[2] = {0.0, 0.0};
    for (int n = 0; n < 2; ++n)
    {
        if (resistive) for (int32_t i = 0; i < nf; ++i) p0[n] += t0[n][i];
        if (hall) for (int32_t i = 0; i < ncl; ++i) p1[n] += t1[n][i];
    }
    rhs_residual_out(d, e, p0[0], p1[0], p0[1], p1[1], E_expl, x, dir, dt, bvec, r);
}

// delta += dir ? 0 : (sum_i y_i V_i) * invdiag  (ImplicitInduction.cpp:366-376)
__global__ void __launch_bounds__(256) k_gmres_update(const double *const *V, const double *y, int32_t j, const double *invdiag,
                                                      const int8_t *dir, int32_t n, double *delta)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double s = 0.0;
    for (int32_t i = 0; i < j; ++i) s += y[i] * V[i][e];
    delta[e] += dir[e] ? 0.0 : s * invdiag[e];
}

// elementwise helpers of solve()
__global__ void __launch_bounds__(256) k_solve_init(const double *E_expl, const double *corr, const int8_t *dir, const double *m_e,
                                                    const double *d2, double dt, const double *dte, int32_t n, double *x0, double *invdiag)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double x = E_expl[e];
    if (corr && !dir[e]) x += corr[e];
    x0[e] = x;
    invdiag[e] = 1.0 / (1.0 + (dte ? dte[e] : dt) * m_e[e] * d2[e]);
}
__global__ void __launch_bounds__(256) k_add(const double *a, const double *b, int32_t n, double *out)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = a[e] + b[e];
}
__global__ void __launch_bounds__(256) k_solve_final(const double *x0, const double *delta, const double *E_expl, const int8_t *dir,
                                                     int32_t n, double *E, double *corr)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double v = x0[e] + delta[e];
    E[e] = v;
    corr[e] = dir[e] ? 0.0 : v - E_expl[e];
}
} // namespace mhdg
