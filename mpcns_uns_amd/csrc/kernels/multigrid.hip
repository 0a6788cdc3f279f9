// Multigrid preconditioner of the mixed-precision implicit induction solve (precision 2): one FP32 V-cycle per
// flexible-GMRES step, on the agglomeration hierarchy of host/multigrid.h.
//
//   fine level     the operator A = I + dt F d1 applied matrix-free by the FP32 kernels of the inner solve (d1, cell
//                  field, Hall term, edge gather), the full operator or its resistive part (mg_fine_hall), with
//                  damped-Jacobi sweeps x += omega D^-1 (b - A x), D the Jacobi diagonal 1 + dt m_e d2_e
//   coarse levels  A_l = M_l + dt K_l (Galerkin products of the identity and the resistive curl-curl, built once per
//                  mesh): one thread per row of the CSR, damped Jacobi with D_l = M_l + dt diag(K_l)
//   transfers      restriction b_{l+1} = P^T r_l (a signed gather of the finer unknowns of each coarse unknown),
//                  prolongation x_l += P x_{l+1} (one signed entry per finer unknown)
//   cycle          mg_pre sweeps (the first from x = 0 is omega D^-1 b), residual, restriction, the next level,
//                  prolongation, mg_post sweeps; mg_coarse sweeps on the coarsest level
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../host/multigrid.h"
#include "ctx.h"

namespace mhdg
{
struct MgLevelDev
{
    int32_t n = 0, n_fine = 0;
    const int32_t *map = nullptr;                  // per finer unknown: signed index on this level or kMgNone
    const int32_t *pt_ptr = nullptr, *pt = nullptr; // P^T rows (signed finer unknowns)
    const float *mass = nullptr;
    const int32_t *k_ptr = nullptr, *k_col = nullptr, *k_diag = nullptr;
    const float *k_val = nullptr;
    float *dinv = nullptr, *x = nullptr, *y = nullptr, *b = nullptr, *r = nullptr;
};

struct MgDevice
{
    std::vector<MgLevelDev> lv; // lv[0]: the fine level (no arrays of its own)
    std::vector<void *> mem;
    float *r0 = nullptr;        // fine residual
    int pre = 2, post = 2, coarse = 4, fine_hall = 1;
    float omega = 0.7f;
    const int8_t *dir_dev = nullptr; // Dirichlet set and time step of the current solve
    float dt = 0.0f;
    ~MgDevice()
    {
        for (void *p : mem) (void)hipFree(p);
    }
    template <class T> T *buf(size_t n)
    {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc (multigrid)");
        mem.push_back(p);
        return static_cast<T *>(p);
    }
    template <class T> T *up(const std::vector<T> &v)
    {
        T *p = buf<T>(v.size());
        if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy (multigrid)");
        return p;
    }
};

namespace
{
inline int nbk(int64_t n) { return static_cast<int>((n + 255) / 256); }
void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}

// D_l^-1 = 1 / (M_l + dt K_l[i][i]) (0 where the diagonal vanishes)
__global__ void __launch_bounds__(256) k_mg_dinv(int32_t n, const float *mass, const float *kval, const int32_t *kdiag, float dt, float *dinv)
{
    const int32_t i = gid();
    if (i >= n) return;
    const float dg = mass[i] + (kdiag[i] >= 0 ? dt * kval[kdiag[i]] : 0.0f);
    dinv[i] = dg != 0.0f ? 1.0f / dg : 0.0f;
}
// fine-level Jacobi diagonal 1 / (1 + dt m_e d2_e), formed in FP64 as the solve's own
__global__ void __launch_bounds__(256) k_mg_fine_invdiag(int32_t n, const double *m_e, const double *d2, double dt, float *invdiag)
{
    const int32_t e = gid();
    if (e < n) invdiag[e] = static_cast<float>(1.0 / (1.0 + dt * m_e[e] * d2[e]));
}
// the first sweep from x = 0: x = omega D^-1 b (0 on Dirichlet unknowns)
__global__ void __launch_bounds__(256) k_mg_first(int32_t n, const float *dinv, const int8_t *dir, const float *b, float omega, float *x)
{
    const int32_t i = gid();
    if (i >= n) return;
    x[i] = (dir && dir[i]) ? 0.0f : omega * dinv[i] * b[i];
}
// coarse level, one thread per row: A x = M x + dt K x; resid: out = b - A x, else out = x + omega D^-1 (b - A x)
__global__ void __launch_bounds__(256) k_mg_sweep(int32_t n, const int32_t *kp, const int32_t *kc, const float *kv, const float *mass,
                                                  const float *dinv, float dt, float omega, const float *x, const float *b, float *out,
                                                  int32_t resid)
{
    const int32_t i = gid();
    if (i >= n) return;
    const int32_t k0 = kp[i], k1 = kp[i + 1];
    float s = 0.0f;
    for (int32_t k = k0; k < k1; k += kEdgeUnroll)
    {
        int32_t c[kEdgeUnroll];
        float a[kEdgeUnroll], v[kEdgeUnroll];
#pragma unroll
        for (int u = 0; u < kEdgeUnroll; ++u)
            if (k + u < k1) { c[u] = kc[k + u]; a[u] = kv[k + u]; }
#pragma unroll
        for (int u = 0; u < kEdgeUnroll; ++u)
            if (k + u < k1) v[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < kEdgeUnroll; ++u)
            if (k + u < k1) s += a[u] * v[u];
    }
    const float ax = mass[i] * x[i] + dt * s;
    out[i] = resid ? b[i] - ax : x[i] + omega * dinv[i] * (b[i] - ax);
}
// b_c[I] = sum over the finer unknowns of I of +-r_f
__global__ void __launch_bounds__(256) k_mg_restrict(int32_t n, const int32_t *pp, const int32_t *pt, const float *r, float *b)
{
    const int32_t I = gid();
    if (I >= n) return;
    float s = 0.0f;
    for (int32_t q = pp[I]; q < pp[I + 1]; ++q)
    {
        const int32_t v = pt[q];
        s += v >= 0 ? r[v] : -r[~v];
    }
    b[I] = s;
}
// x_f[i] += +-x_c[map[i]]
__global__ void __launch_bounds__(256) k_mg_prolong(int32_t n_fine, const int32_t *map, const float *xc, float *xf)
{
    const int32_t i = gid();
    if (i >= n_fine) return;
    const int32_t v = map[i];
    if (v == kMgNone) return;
    xf[i] += v >= 0 ? xc[v] : -xc[~v];
}
// flexible GMRES update: delta += sum_i y_i Z_i (Z already preconditioned, zero on Dirichlet edges)
__global__ void __launch_bounds__(256) k_update_flex(const float *const *Z, const double *y, int32_t j, const int8_t *dir, int32_t n,
                                                     double *delta)
{
    const int32_t e = gid();
    if (e >= n || dir[e]) return;
    double s = 0.0;
    for (int32_t i = 0; i < j; ++i) s += y[i] * static_cast<double>(Z[i][e]);
    delta[e] += s;
}
__global__ void __launch_bounds__(256) k_f32_to_f64(const float *x, int64_t n, double *out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<double>(x[i]);
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

void mhdg_ctx::mg_build(const int8_t *dir_host)
{
    if (mg && mg_dir_h.size() == size_t(ne) && std::equal(dir_host, dir_host + ne, mg_dir_h.begin())) return;
    if (mgh.n_nodes <= 0) throw std::runtime_error("multigrid preconditioner: the mesh descriptor carried no e2n incidence");
    const int levels = cfg.mg_levels > 0 ? cfg.mg_levels : 4, passes = cfg.mg_agglom_passes > 0 ? cfg.mg_agglom_passes : 3;
    const MgHierarchy H = build_mg_hierarchy(mgh.n_nodes, ne, mgh.tail.data(), mgh.head.data(), mgh.star1.data(), nf, mgh.f2e_ptr.data(),
                                             mgh.f2e.data(), mgh.star2.data(), mgh.m_e.data(), dir_host, levels, passes);
    delete mg;
    mg = nullptr;
    auto *G = new MgDevice();
    try
    {
        G->lv.resize(H.lv.size());
        G->lv[0].n = ne;
        for (size_t l = 1; l < H.lv.size(); ++l)
        {
            const MgLevel &L = H.lv[l];
            MgLevelDev &D = G->lv[l];
            D.n = L.n;
            D.n_fine = H.lv[l - 1].n;
            D.map = G->up(L.map);
            D.pt_ptr = G->up(L.pt_ptr);
            D.pt = G->up(L.pt);
            D.mass = G->up(std::vector<float>(L.mass.begin(), L.mass.end()));
            D.k_ptr = G->up(L.k_ptr);
            D.k_col = G->up(L.k_col);
            D.k_val = G->up(std::vector<float>(L.k_val.begin(), L.k_val.end()));
            std::vector<int32_t> kd(size_t(L.n), -1);
            for (int32_t i = 0; i < L.n; ++i)
                for (int32_t k = L.k_ptr[size_t(i)]; k < L.k_ptr[size_t(i) + 1]; ++k)
                    if (L.k_col[size_t(k)] == i) kd[size_t(i)] = k;
            D.k_diag = G->up(kd);
            D.dinv = G->buf<float>(size_t(L.n));
            D.x = G->buf<float>(size_t(L.n));
            D.y = G->buf<float>(size_t(L.n));
            D.b = G->buf<float>(size_t(L.n));
            D.r = G->buf<float>(size_t(L.n));
        }
        G->r0 = G->buf<float>(size_t(ne));
        G->pre = std::max(1, cfg.mg_pre > 0 ? cfg.mg_pre : 2);
        G->post = std::max(0, cfg.mg_post > 0 ? cfg.mg_post : 2);
        G->coarse = std::max(1, cfg.mg_coarse > 0 ? cfg.mg_coarse : 4);
        G->omega = static_cast<float>(cfg.mg_relax > 0 ? cfg.mg_relax : 0.7);
        G->fine_hall = cfg.mg_fine_hall;
    }
    catch (...)
    {
        delete G;
        throw;
    }
    mg = G;
    mg_dir_h.assign(dir_host, dir_host + ne);
    // the flexible basis Z
    MixedSolver &M = *mixed;
    if (M.Z.empty())
    {
        M.Z.resize(M.V.size());
        for (auto &z : M.Z) z = static_cast<float *>(alloc(size_t(ne) * sizeof(float)));
        M.Zptrs = static_cast<float **>(alloc(sizeof(float *) * M.Z.size()));
        hip_check(hipMemcpy(M.Zptrs, M.Z.data(), sizeof(float *) * M.Z.size(), hipMemcpyHostToDevice), "upload");
    }
}

void mhdg_ctx::mg_set_solve(const int8_t *dir_dev, float dt)
{
    mg->dir_dev = dir_dev;
    mg->dt = dt;
}

void mhdg_ctx::mg_prepare(double dt)
{
    MgDevice &G = *mg;
    G.dt = static_cast<float>(dt);
    for (size_t l = 1; l < G.lv.size(); ++l)
    {
        const MgLevelDev &D = G.lv[l];
        LAUNCH("mg_dinv", k_mg_dinv, dim3(nbk(D.n)), dim3(256), 0, stream, D.n, D.mass, D.k_val, D.k_diag, G.dt, D.dinv);
    }
}

void mhdg_ctx::op32(int mode, int hall_on, const float *pv, const float *v, const float *b, float omega, float *out)
{
    MixedSolver &M = *mixed;
    const int8_t *dir = mg->dir_dev;
    const float dtf = mg->dt;
    LAUNCH("d1_f32", k_d1_f32, dim3(nbk(nf)), dim3(256), 0, stream, d, M.s, pv, static_cast<const float *>(nullptr), M.phi);
    if (hall_on)
    {
        LAUNCH("bcell_f32", k_bcell_f32_g<kCellLanes>, dim3(nbk(kCellLanes * int64_t(nc))), dim3(256), 0, stream, d, M.s, M.phi, M.b);
        LAUNCH("hall_g_f32", k_hall_g_f32_g<kCellLanes>, dim3(nbk(kCellLanes * int64_t(nc))), dim3(256), 0, stream, d, M.s, M.b, M.g);
    }
    const int32_t n_light = int32_t((int64_t(ne) + kEdgeBlock - 1) / kEdgeBlock);
    const dim3 grid(n_light + n_heavy), block(kEdgeBlock);
    const int32_t *hv = static_cast<const int32_t *>(heavy);
    switch (mode)
    {
    case kOpApply:
        LAUNCH("mg_apply_f32", k_edge_apply_f32<kOpApply>, grid, block, 0, stream, d, M.s, int32_t(op_res), int32_t(hall_on), M.phi, M.g, v, pv,
               dir, dtf, out, hv, n_light, b, omega);
        break;
    case kOpSmooth:
        LAUNCH("mg_smooth_f32", k_edge_apply_f32<kOpSmooth>, grid, block, 0, stream, d, M.s, int32_t(op_res), int32_t(hall_on), M.phi, M.g, v,
               pv, dir, dtf, out, hv, n_light, b, omega);
        break;
    case kOpResid:
        LAUNCH("mg_resid_f32", k_edge_apply_f32<kOpResid>, grid, block, 0, stream, d, M.s, int32_t(op_res), int32_t(hall_on), M.phi, M.g, v,
               pv, dir, dtf, out, hv, n_light, b, omega);
        break;
    default: throw std::runtime_error("op32: unknown mode");
    }
}

void mhdg_ctx::mg_cycle(int l, const float *b, float *x)
{
    MgDevice &G = *mg;
    const int nl = static_cast<int>(G.lv.size());
    const float w = G.omega;
    if (l == 0)
    {
        MixedSolver &M = *mixed;
        const int hall_on = (G.fine_hall && op_hall) ? 1 : 0;
        LAUNCH("mg_first", k_mg_first, dim3(nbk(ne)), dim3(256), 0, stream, ne, static_cast<const float *>(M.s.invdiag), G.dir_dev, b, w, x);
        for (int s = 1; s < G.pre; ++s) op32(kOpSmooth, hall_on, x, nullptr, b, w, x);
        if (nl > 1)
        {
            op32(kOpResid, hall_on, x, nullptr, b, 0.0f, G.r0);
            const MgLevelDev &C = G.lv[1];
            LAUNCH("mg_restrict", k_mg_restrict, dim3(nbk(C.n)), dim3(256), 0, stream, C.n, C.pt_ptr, C.pt, static_cast<const float *>(G.r0), C.b);
            mg_cycle(1, C.b, C.x);
            LAUNCH("mg_prolong", k_mg_prolong, dim3(nbk(ne)), dim3(256), 0, stream, ne, C.map, static_cast<const float *>(C.x), x);
        }
        for (int s = 0; s < G.post; ++s) op32(kOpSmooth, hall_on, x, nullptr, b, w, x);
        return;
    }
    MgLevelDev &L = G.lv[size_t(l)];
    const bool coarsest = l == nl - 1;
    LAUNCH("mg_first", k_mg_first, dim3(nbk(L.n)), dim3(256), 0, stream, L.n, static_cast<const float *>(L.dinv),
           static_cast<const int8_t *>(nullptr), b, w, x);
    // Jacobi reads the old iterate everywhere: sweeps alternate between x and L.y and end in x
    auto sweeps = [&](int count) {
        float *cur = x, *oth = L.y;
        for (int s = 0; s < count; ++s)
        {
            LAUNCH("mg_sweep", k_mg_sweep, dim3(nbk(L.n)), dim3(256), 0, stream, L.n, L.k_ptr, L.k_col, L.k_val, L.mass,
                   static_cast<const float *>(L.dinv), G.dt, w, static_cast<const float *>(cur), b, oth, 0);
            std::swap(cur, oth);
        }
        if (cur != x) hip_check(hipMemcpyAsync(x, cur, size_t(L.n) * sizeof(float), hipMemcpyDeviceToDevice, stream), "copy");
    };
    if (coarsest)
    {
        sweeps(G.coarse - 1);
        return;
    }
    sweeps(G.pre - 1);
    LAUNCH("mg_resid", k_mg_sweep, dim3(nbk(L.n)), dim3(256), 0, stream, L.n, L.k_ptr, L.k_col, L.k_val, L.mass,
           static_cast<const float *>(L.dinv), G.dt, 0.0f, static_cast<const float *>(x), b, L.r, 1);
    const MgLevelDev &C = G.lv[size_t(l) + 1];
    LAUNCH("mg_restrict", k_mg_restrict, dim3(nbk(C.n)), dim3(256), 0, stream, C.n, C.pt_ptr, C.pt, static_cast<const float *>(L.r), C.b);
    mg_cycle(l + 1, C.b, C.x);
    LAUNCH("mg_prolong", k_mg_prolong, dim3(nbk(L.n)), dim3(256), 0, stream, L.n, C.map, static_cast<const float *>(C.x), x);
    sweeps(G.post);
}

// One V-cycle z = M v with the frozen state of the last mhdg_ind_set_frozen and the Jacobi diagonals of step dt (the
// preconditioner exactly as the flexible solve applies it), for tests of the cycle against a host restatement.
extern "C" int mhdg_mg_apply(mhdg_ctx *c, const int8_t *dir, double dt, const double *v, double *z)
{
    try
    {
        if (!(dt > 0)) throw std::runtime_error("mhdg_mg_apply: dt must be positive");
        c->mixed_build();
        c->mg_build(dir);
        MixedSolver &M = *c->mixed;
        int8_t *dd = reinterpret_cast<int8_t *>(c->stage(2, (size_t(c->ne) + 7) / 8));
        hip_check(hipMemcpyAsync(dd, dir, size_t(c->ne), hipMemcpyHostToDevice, c->stream), "H2D");
        const double *vd = c->stage_in(0, v, c->ne);
        double *zd = c->stage(3, c->ne);
        hipLaunchKernelGGL(k_alpha_f32, dim3(nbk(c->nc)), dim3(256), 0, c->stream, c->alpha, c->B0, c->nc, M.s.aB0);
        hipLaunchKernelGGL(k_mg_fine_invdiag, dim3(nbk(c->ne)), dim3(256), 0, c->stream, c->ne, c->d.m_e, c->d.d2, dt,
                           const_cast<float *>(M.s.invdiag));
        c->mg_set_solve(dd, static_cast<float>(dt));
        c->mg_prepare(dt);
        hipLaunchKernelGGL(k_f64_to_f32, dim3(nbk(c->ne)), dim3(256), 0, c->stream, vd, int64_t(c->ne), M.V[0]);
        c->mg_cycle(0, M.V[0], M.Z[0]);
        hipLaunchKernelGGL(k_f32_to_f64, dim3(nbk(c->ne)), dim3(256), 0, c->stream, static_cast<const float *>(M.Z[0]), int64_t(c->ne), zd);
        c->check_launch("mhdg_mg_apply");
        c->d2h(z, zd, size_t(c->ne) * 8);
        return 0;
    }
    catch (const std::exception &e)
    {
        c->err = e.what();
        return 1;
    }
}

// Levels of the hierarchy built for `dir` (building it if needed) and the unknowns of each (n_per_level holds up to
// max_levels entries).
extern "C" int mhdg_mg_info(mhdg_ctx *c, const int8_t *dir, int32_t max_levels, int32_t *levels, int32_t *n_per_level)
{
    try
    {
        c->mixed_build();
        c->mg_build(dir);
        *levels = static_cast<int32_t>(c->mg->lv.size());
        for (int32_t l = 0; l < *levels && l < max_levels; ++l) n_per_level[l] = c->mg->lv[size_t(l)].n;
        return 0;
    }
    catch (const std::exception &e)
    {
        c->err = e.what();
        return 1;
    }
}
