// Context and C-ABI of the device path (include/mhdg.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mhdg.h"
#include "../host/multigrid.h"
#include "../host/tables.h"
#include "ctx.h"

using namespace mhdg;

namespace
{
thread_local std::string g_create_err;

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
inline int nblk(int64_t n, int bs = 256) { return static_cast<int>((n + bs - 1) / bs); }
} // namespace

// kernel launch on the context stream, bracketed by timing events when profiling is on
#define LAUNCH(NAME, ...)                                                                                                    \
    do                                                                                                                       \
    {                                                                                                                        \
        prof_begin(NAME);                                                                                                    \
        hipLaunchKernelGGL(__VA_ARGS__);                                                                                     \
        prof_end();                                                                                                          \
    } while (0)

void *mhdg_ctx::alloc(size_t bytes)
{
    void *p = nullptr;
    hip_check(hipMalloc(&p, bytes ? bytes : 8), "hipMalloc");
    owned.push_back(p);
    return p;
}
template <class T> T *mhdg_ctx::upload(const std::vector<T> &v)
{
    T *p = static_cast<T *>(alloc(v.size() * sizeof(T)));
    if (!v.empty()) hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    return p;
}
template <class T> T *mhdg_ctx::upload(const T *v, size_t n)
{
    T *p = static_cast<T *>(alloc(n * sizeof(T)));
    if (n) hip_check(hipMemcpy(p, v, n * sizeof(T), hipMemcpyHostToDevice), "upload");
    return p;
}
double *mhdg_ctx::stage(int slot, size_t n)
{
    auto &s = staging[slot];
    if (s.second < n)
    {
        if (s.first) hip_check(hipFree(s.first), "hipFree");
        hip_check(hipMalloc(&s.first, n * sizeof(double)), "hipMalloc");
        s.second = n;
    }
    return static_cast<double *>(s.first);
}
double *mhdg_ctx::stage_in(int slot, const void *host, size_t n_doubles)
{
    double *p = stage(slot, n_doubles);
    hip_check(hipMemcpyAsync(p, host, n_doubles * sizeof(double), hipMemcpyHostToDevice, stream), "H2D");
    return p;
}
void mhdg_ctx::d2h(void *host, const void *dev, size_t bytes)
{
    if (host_pin && bytes <= kPinStageBytes)
    {
        // small read-backs (reduction results, Hessenberg columns: one per Krylov step) through pinned memory,
        // which saves the runtime's staging of a copy into pageable memory
        char *pin = static_cast<char *>(host_pin) + kPinStageOffset;
        hip_check(hipMemcpyAsync(pin, dev, bytes, hipMemcpyDeviceToHost, stream), "D2H");
        hip_check(hipStreamSynchronize(stream), "sync");
        std::memcpy(host, pin, bytes);
        return;
    }
    hip_check(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, stream), "D2H");
    hip_check(hipStreamSynchronize(stream), "sync");
}
void mhdg_ctx::check_launch(const char *what)
{
    hip_check(hipGetLastError(), what);
}

void mhdg_ctx::prof_begin(const char *name)
{
    prof_armed = false;
    if (!prof_on || (!prof_only.empty() && prof_only != name)) return;
    while (prof_pool.size() < prof_used + 2)
    {
        hipEvent_t ev;
        hip_check(hipEventCreate(&ev), "event");
        prof_pool.push_back(ev);
    }
    ProfRec r{name, prof_pool[prof_used], prof_pool[prof_used + 1]};
    prof_used += 2;
    hip_check(hipEventRecord(r.a, stream), "event");
    prof_pending.push_back(r);
    prof_armed = true;
}
void mhdg_ctx::prof_end()
{
    if (!prof_armed) return;
    prof_armed = false;
    hip_check(hipEventRecord(prof_pending.back().b, stream), "event");
}
void mhdg_ctx::prof_flush()
{
    if (prof_pending.empty()) return;
    hip_check(hipStreamSynchronize(stream), "sync");
    for (auto &r : prof_pending)
    {
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, r.a, r.b), "elapsed");
        auto &acc = prof_acc[r.name];
        acc.first += 1;
        acc.second += ms;
    }
    prof_pending.clear();
    prof_used = 0;
}

mhdg_ctx::~mhdg_ctx()
{
    delete mixed;
    delete mg;
    for (double *p : aw.ck) (void)hipFree(p);
    for (hipEvent_t ev : prof_pool) (void)hipEventDestroy(ev);
    for (void *p : owned) (void)hipFree(p);
    for (auto &kv : staging) if (kv.second.first) (void)hipFree(kv.second.first);
    if (host_pin) (void)hipHostFree(host_pin);
    if (stream) (void)hipStreamDestroy(stream);
}

// ------------------------------------------------------------------------------------ creation
static void make_params(const mhdg_config &c, Params &P)
{
    std::memset(&P, 0, sizeof(P));
    P.gamma = c.gamma; P.inver_MA2 = c.inver_MA2; P.M_ref = c.M_ref; P.M_H = c.M_H; P.M_Na = c.M_Na; P.ne_floor = c.ne_floor;
    P.hall_coef = c.hall_coef; P.cfl = c.cfl; P.rel_limit = c.recon_rel_limit; P.rho_floor = c.rho_floor; P.p_floor = c.p_floor;
    P.rho_floor_dt = c.rho_floor_dt; P.cuct = c.cuct;
    P.second_order = c.second_order; P.lorentz = c.lorentz_force; P.hall = c.hall && c.hall_coef > 0;
    P.resistive = c.resistive && c.inver_Rem > 0;
    if (c.n_species > kMaxSpecies || c.n_species < 0 || c.n_species == 1) throw std::runtime_error("n_species must be 0 or 2..10");
    P.ns = c.n_species > 2 ? c.n_species : 2;
    if (P.ns == 2)
    {
        P.M_s[0] = c.M_H;
        P.M_s[1] = c.M_Na;
        for (int s = 0; s < 2; ++s)
            for (int m = 0; m < 5; ++m) P.fs[s][m] = c.fs[s][m];
    }
    else
        for (int s = 0; s < P.ns; ++s)
        {
            if (!(c.ion_mass[s] > 0)) throw std::runtime_error("ion_mass must be positive");
            P.M_s[s] = c.ion_mass[s];
            for (int m = 0; m < 5; ++m) P.fs[s][m] = c.ion_fs[s][m];
        }
    bool src = false;
    for (int s = 0; s < P.ns; ++s) { P.src_rate[s] = c.src_rate[s]; src = src || c.src_rate[s] != 0.0; }
    if (src && !(c.src_scale_height > 0)) throw std::runtime_error("src_scale_height must be positive with ion production");
    P.src_H = c.src_scale_height; P.src_r0 = c.src_r0; P.src_T = c.src_T;
    if (c.induce_coeff < 0) throw std::runtime_error("induce_coeff must be >= 0");
    if (c.induce_coeff > 0 && P.ns != 2) throw std::runtime_error("the gyration split needs the two-ion fluid");
    P.induce_coeff = c.induce_coeff;
    // E_inf = -1 * (u_inf x B_imf) (MercuryUns/driver/Solver.cpp:162)
    const V3 u(c.u_inf[0], c.u_inf[1], c.u_inf[2]), B(c.B_imf[0], c.B_imf[1], c.B_imf[2]);
    const V3 E = -1.0 * cross(u, B);
    P.E_inf[0] = E.x; P.E_inf[1] = E.y; P.E_inf[2] = E.z;
}

void mhdg_ctx::build(const mhdg_mesh_desc &m, const mhdg_config &c)
{
    cfg = c;
    make_params(c, P);
    op_hall = P.hall != 0;
    op_res = P.resistive != 0;
    ideal_cfg = c.implicit_ideal != 0;
    hip_check(hipGetDevice(&device), "hipGetDevice");
    noalloc_setup(device);
    hip_check(hipStreamCreate(&stream), "hipStreamCreate");
    Tables t = build_tables(m, c, kLightMax);
    if (t.max_valence > kHeavyMax) throw std::runtime_error("edge valence exceeds the supported maximum");
    nc = t.nc; nf = t.nf; ne = t.ne;
    n_heavy = static_cast<int32_t>(t.heavy.size());
    c2f_nnz = t.c2f.size();
    ec_nnz = t.ec_cell.size();
    e2f_nnz = t.e2f.size();
    auto uniform = [](const std::vector<int32_t> &ptr) {
        const int32_t deg = ptr.size() > 1 ? ptr[1] - ptr[0] : 0;
        for (size_t r = 1; r < ptr.size(); ++r)
            if (ptr[r] - ptr[r - 1] != deg) return 0;
        return deg;
    };
    cdeg = uniform(t.c2f_ptr);
    fdeg = uniform(t.f2e_ptr);
    if (c.ind_restart > 60) throw std::runtime_error("ind_restart above 60 is not supported");
    const size_t E = ne, F = nf, C = nc;
    d.nc = nc; d.nf = nf; d.ne = ne;
    d.cdeg = cdeg; d.fdeg = fdeg;
    d.c2f_ptr = upload(t.c2f_ptr); d.c2f = upload(t.c2f); d.cfs = upload(t.cfs);
    d.f2e_ptr = upload(t.f2e_ptr); d.f2e = upload(t.f2e);
    d.e2f_ptr = upload(t.e2f_ptr); d.e2f = upload(t.e2f);
    d.fL = upload(t.fL); d.fR = upload(t.fR);
    {
        std::vector<int32_t> c2n(t.c2f.size());
        for (int32_t c = 0; c < nc; ++c)
            for (int64_t k = t.c2f_ptr[c]; k < t.c2f_ptr[c + 1]; ++k)
            {
                const int32_t v = t.c2f[k], f = v >= 0 ? v : ~v, l = t.fL[f], r = t.fR[f];
                c2n[k] = (l >= 0 && r >= 0) ? (l == c ? r : l) : -1;
            }
        d.c2n = upload(c2n);
    }
    d.ec_ptr = upload(t.ec_ptr); d.ec_cell = upload(t.ec_cell);
    d.mat = upload(t.mat8);
    d.kind = upload(m.face_kind, F); d.emode = upload(m.edge_mode, E);
    d.farea = upload(m.face_area, 3 * F); d.fcen = upload(m.face_center, 3 * F); d.famag = upload(m.face_area_mag, F);
    d.ccen = upload(m.cell_center, 3 * C); d.cvol = upload(m.cell_volume, C);
    d.evec = upload(m.edge_vec, 3 * E); d.emid = upload(m.edge_mid, 3 * E); d.elen = upload(m.edge_len, E);
    d.ecw = upload(m.ec_weight, size_t(m.ec_ptr[E])); d.efw = upload(m.ef_weight, t.e2f.size());
    d.star2 = upload(m.star2, F);
    d.snb = upload(t.snb); d.sw = upload(t.sw); d.sd = upload(t.sd); d.sinv = upload(t.sinv); d.svalid = upload(t.svalid);
    d.recon = upload(t.recon); d.m_e = upload(t.m_e); d.d2 = upload(t.d2); d.taper = upload(t.taper);
    heavy = upload(t.heavy);
    taper_h = t.taper;
    d.atm = nullptr;
    bool any_src = false;
    for (int s = 0; s < P.ns; ++s) any_src = any_src || P.src_rate[s] != 0.0;
    if (any_src)
    {
        // exp(-(|x_c| - r0) / H), evaluated once on the host
        std::vector<double> atm(C);
        for (size_t i = 0; i < C; ++i)
        {
            const double *x = m.cell_center + 3 * i;
            atm[i] = std::exp(-(std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) - P.src_r0) / P.src_H);
        }
        d.atm = upload(atm);
    }
    std::vector<int8_t> dir(E);
    for (size_t e = 0; e < E; ++e) dir[e] = c.wall_dirichlet ? m.edge_mode[e] != 0 : m.edge_mode[e] == 2;
    dir_mercury = upload(dir);
    dir_mercury_h = dir;
    if (m.e2n.ptr && m.e2n.idx && m.e2n.sgn && m.n_nodes > 0)
    {
        // what the multigrid hierarchy is built from, at its first use
        edge_nodes(m, mgh.tail, mgh.head);
        mgh.n_nodes = m.n_nodes;
        mgh.star1.assign(m.star1, m.star1 + E);
        mgh.star2.assign(m.star2, m.star2 + F);
        mgh.f2e_ptr = t.f2e_ptr;
        mgh.f2e = t.f2e;
        mgh.m_e = t.m_e;
    }

    auto D = [&](size_t n) { return static_cast<double *>(alloc(n * sizeof(double))); };
    Phi = D(F); PhiAdd = D(F); Badd = D(3 * C);
    hip_check(hipMemset(PhiAdd, 0, F * 8), "memset"); hip_check(hipMemset(Badd, 0, 3 * C * 8), "memset");
    hip_check(hipMemset(Phi, 0, F * 8), "memset");
    for (int s = 0; s < P.ns; ++s)
    {
        U[s] = D(5 * C); W[s] = D(5 * C); grad[s] = D(15 * C);
        hip_check(hipMemset(U[s], 0, 5 * C * 8), "memset");
    }
    Bind = D(3 * C); Btot = D(3 * C); uplus = D(3 * C); J = D(3 * C); alpha = D(C); B0 = D(3 * C);
    hip_check(hipMemset(alpha, 0, C * 8), "memset"); hip_check(hipMemset(B0, 0, 3 * C * 8), "memset");
    FA = D(5 * size_t(P.ns) * F); Ef = D(3 * F); E_expl = D(E); E_last = D(E);
    dt_part = D(size_t(nblk(nc)) + 1); dt_dev = D(1);
    // induction work
    phi_w = D(F); bcell = D(3 * C); gcell = D(3 * C); corr = D(E); have_corr = false;
    x0 = D(E); bvec = D(E); rvec = D(E); wvec = D(E); invdiag = D(E); delta = D(E); tmp = D(E);
    scal = D(64);
    rs.partial = D(kRedBlocks);
    rs.ticket = static_cast<unsigned *>(alloc(64));
    hip_check(hipMemset(rs.ticket, 0, 64), "memset");
    restart = std::max(2, c.ind_restart);
    V.resize(restart + 1);
    for (auto &v : V) v = D(E);
    Vptrs = static_cast<double **>(alloc(sizeof(double *) * V.size()));
    hip_check(hipMemcpy(Vptrs, V.data(), sizeof(double *) * V.size(), hipMemcpyHostToDevice), "upload");
    hcol = D(restart + 2);
    ydev = D(restart + 1);
    hip_check(hipHostMalloc(&host_pin, kPinBytes), "hipHostMalloc");
    hip_check(hipDeviceSynchronize(), "sync");
}

// ------------------------------------------------------------------------------------ pieces
void mhdg_ctx::launch_edge_emf(const double *Efp, const double *u, const double *phi_ind, double cuct, const double *Ein, double *Eout,
                               int bc)
{
    LAUNCH("edge_emf", k_edge_emf<double>, dim3(nblk(ne)), dim3(256), 0, stream, d, P, Efp, u, phi_ind, cuct, kLightMax, Ein, Eout, bc);
    if (n_heavy)
        LAUNCH("edge_emf_heavy", k_edge_emf_heavy<double>, dim3(n_heavy), dim3(256), 0, stream, d, P, Efp, u, phi_ind, cuct,
                           static_cast<const int32_t *>(heavy), Ein, Eout, bc);
    check_launch("edge emf");
}

double mhdg_ctx::dot(const double *a, const double *b, int op, int slot)
{
    LAUNCH("dot", k_dot<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, stream, a, b, ne, RedScratch{rs.partial, nullptr}, scal + slot, op);
    LAUNCH("finish", k_finish, dim3(1), dim3(kRedThreads), 0, stream, static_cast<const double *>(rs.partial), kRedBlocks, scal + slot, op);
    check_launch("dot");
    double v;
    d2h(&v, scal + slot, 8);
    return v;
}

// F(phi) for a face vector phi into out (mode 0), or A x (mode 1) with x = precond(v) / v, optional mask /
// residual form.  ImplicitInduction::apply_F / apply_A (ImplicitInduction.cpp:127-198).
void mhdg_ctx::apply_op(const double *phi_in, const double *v, const double *inv, const int8_t *dir, double dt, int mode, int mask,
                        const double *b, double *out)
{
    const double *phi = phi_in;
    if (!phi)
    {
        LAUNCH("d1", k_d1, dim3(nblk(nf)), dim3(256), 0, stream, d, v, inv, dir, dte_cur, phi_w);
        phi = phi_w;
        if (dte_cur) dt = 1.0; // d1 already carries dt_e: y = x + F (ImplicitInduction.cpp:196)
    }
    const bool ideal_part = ideal_cfg && u_ideal;
    const int64_t lanes = kCellLanes64 * int64_t(nc);
    if (op_hall || ideal_part) LAUNCH("bcell", k_bcell_g<kCellLanes64>, dim3(nblk(lanes)), dim3(256), 0, stream, d, phi, bcell);
    if (op_hall && lin_b) LAUNCH("hall_g_tangent", k_hall_g_tangent, dim3(nblk(nc)), dim3(256), 0, stream, d, lin_b, bcell, alpha, Badd, gcell);
    else if (op_hall) LAUNCH("hall_g", k_hall_g_g<kCellLanes64>, dim3(nblk(lanes)), dim3(256), 0, stream, d, bcell, alpha, B0, gcell);
    if (ideal_part)
    {
        // add_ideal_emf of (u, b, phi) with b as total and induced field (ImplicitInduction.cpp:150-157)
        SpC none{};
        LAUNCH("face", k_face<double>, dim3(nblk(nf)), dim3(256), 0, stream, d, P, 0, 0, none, none, none, static_cast<double *>(nullptr), u_ideal,
               bcell, bcell, phi, static_cast<const double *>(nullptr), Ef_i);
        launch_edge_emf(Ef_i, u_ideal, phi, P.cuct, nullptr, ide, 0);
    }
    const double *idp = ideal_part ? ide : nullptr;
    LAUNCH("edge_apply", k_edge_apply, dim3(nblk(ne)), dim3(256), 0, stream, d, int32_t(op_res), int32_t(op_hall), phi, gcell, idp,
           imask, v, inv, dir, dt, mode, mask, b, out);
    if (n_heavy)
        LAUNCH("edge_apply_heavy", k_edge_apply_heavy, dim3(n_heavy), dim3(256), 0, stream, d, int32_t(op_res), int32_t(op_hall),
               static_cast<const int32_t *>(heavy), phi, gcell, idp, imask, v, inv, dir, dt, mode, mask, b, out);
    check_launch("apply_op");
}

void mhdg_ctx::rhs_residual(const double *Phi_n, const double *Eex, const double *x, const int8_t *dir, double dt)
{
    if ((ideal_cfg && u_ideal) || lin_b)
    {
        apply_op(Phi_n, nullptr, nullptr, nullptr, 0.0, 0, 0, nullptr, tmp); // F(Phi^n)
        LAUNCH("axpy", k_add, dim3(nblk(ne)), dim3(256), 0, stream, Eex, tmp, ne, bvec);
        apply_op(nullptr, x, nullptr, dir, dt, 1, 0, bvec, rvec); // r = b - A x
        return;
    }
    const size_t C = nc;
    if (!bcell2) { bcell2 = static_cast<double *>(alloc(3 * C * 8)); gcell2 = static_cast<double *>(alloc(3 * C * 8)); }
    LAUNCH("d1", k_d1, dim3(nblk(nf)), dim3(256), 0, stream, d, x, static_cast<const double *>(nullptr), dir, dte_cur, phi_w);
    if (dte_cur) dt = 1.0; // d1 already carries dt_e (ImplicitInduction.cpp:196)
    const int64_t lanes = kCellLanes64 * int64_t(nc);
    if (op_hall)
    {
        LAUNCH("bcell", k_bcell2_g<kCellLanes64>, dim3(nblk(lanes)), dim3(256), 0, stream, d, Phi_n, static_cast<const double *>(phi_w), bcell,
               bcell2);
        LAUNCH("hall_g", k_hall_g2_g<kCellLanes64>, dim3(nblk(lanes)), dim3(256), 0, stream, d, static_cast<const double *>(bcell),
               static_cast<const double *>(bcell2), alpha, B0, gcell, gcell2);
    }
    LAUNCH("edge_apply", k_edge_rhs_residual, dim3(nblk(ne)), dim3(256), 0, stream, d, int32_t(op_res), int32_t(op_hall), Phi_n,
           static_cast<const double *>(gcell), static_cast<const double *>(phi_w), static_cast<const double *>(gcell2), Eex, x, dir, dt, bvec,
           rvec);
    if (n_heavy)
        LAUNCH("edge_apply_heavy", k_edge_rhs_residual_heavy, dim3(n_heavy), dim3(256), 0, stream, d, int32_t(op_res), int32_t(op_hall),
               static_cast<const int32_t *>(heavy), Phi_n, static_cast<const double *>(gcell), static_cast<const double *>(phi_w),
               static_cast<const double *>(gcell2), Eex, x, dir, dt, bvec, rvec);
    check_launch("rhs_residual");
}

// ImplicitInduction::solve (Common/induction/ImplicitInduction.cpp:299-390), FP64, Jacobi, MGS Arnoldi:
// the reference's control flow with the Hessenberg / Givens arithmetic on the host.
int mhdg_ctx::solve64(const double *Phi_n, const double *Eex, const int8_t *dir, double dt, double *Eout, double *rel_res)
{
    const int m = restart;
    const int max_it = cfg.ind_max_it;
    const double rtol = rtol_override > 0 ? rtol_override : cfg.ind_rtol;
    LAUNCH("solve_init", k_solve_init, dim3(nblk(ne)), dim3(256), 0, stream, Eex, have_corr ? corr : nullptr, dir, d.m_e, d.d2, dt,
           dte_cur, ne, x0, invdiag);
    rhs_residual(Phi_n, Eex, x0, dir, dt); // b = E_expl + F(Phi^n), r = b - A x0
    hip_check(hipMemsetAsync(delta, 0, size_t(ne) * 8, stream), "memset");
    const double bb = dot(bvec, bvec, 0, 0);
    const double bnorm = std::sqrt(std::max(bb, 1e-300));
    double beta = dot(rvec, rvec, 1, 1);
    int total = 0;
    std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> cs(m), sn(m), g(m + 1), yv(m), hc(m + 2);
    while (beta > rtol * bnorm && total < max_it)
    {
        LAUNCH("scale", k_div_scalar, dim3(nblk(ne)), dim3(256), 0, stream, rvec, scal + 1, ne, V[0]);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m && total < max_it; ++j, ++total)
        {
            apply_op(nullptr, V[j], invdiag, dir, dt, 1, 1, nullptr, wvec);
            const RedScratch parts{rs.partial, nullptr};
            for (int i = 0; i <= j; ++i)
            {
                LAUNCH("mgs", k_mgs<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, stream, wvec, i > 0 ? V[i - 1] : nullptr, V[i], ne,
                       hcol, i, 0, parts);
                LAUNCH("finish", k_finish, dim3(1), dim3(kRedThreads), 0, stream, static_cast<const double *>(rs.partial), kRedBlocks, hcol + i, 0);
            }
            LAUNCH("mgs", k_mgs<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, stream, wvec, V[j], nullptr, ne, hcol, j + 1, 1, parts);
            LAUNCH("finish", k_finish, dim3(1), dim3(kRedThreads), 0, stream, static_cast<const double *>(rs.partial), kRedBlocks, hcol + j + 1, 1);
            LAUNCH("scale", k_scale_div<double>, dim3(nblk(ne)), dim3(256), 0, stream, wvec, hcol + j + 1, ne, V[j + 1]);
            check_launch("arnoldi");
            d2h(hc.data(), hcol, sizeof(double) * (j + 2));
            for (int i = 0; i <= j + 1; ++i) H[i][j] = hc[i];
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
            if (std::abs(g[j + 1]) <= rtol * bnorm) { ++j; ++total; break; }
        }
        for (int i = j - 1; i >= 0; --i)
        {
            double s = g[i];
            for (int k = i + 1; k < j; ++k) s -= H[i][k] * yv[k];
            yv[i] = s / H[i][i];
        }
        hip_check(hipMemcpyAsync(ydev, yv.data(), sizeof(double) * std::max(j, 1), hipMemcpyHostToDevice, stream), "H2D");
        LAUNCH("gmres_update", k_gmres_update, dim3(nblk(ne)), dim3(256), 0, stream, static_cast<const double *const *>(Vptrs), ydev, j,
                           invdiag, dir, ne, delta);
        LAUNCH("axpy", k_add, dim3(nblk(ne)), dim3(256), 0, stream, x0, delta, ne, tmp);
        apply_op(nullptr, tmp, nullptr, dir, dt, 1, 0, bvec, rvec);
        beta = dot(rvec, rvec, 1, 1);
    }
    LAUNCH("solve_final", k_solve_final, dim3(nblk(ne)), dim3(256), 0, stream, x0, delta, Eex, dir, ne, Eout, corr);
    check_launch("solve final");
    have_corr = true;
    if (rel_res) *rel_res = beta / bnorm;
    return total;
}

// mercury::Solver::step_first_order (MercuryUns/driver/Solver.cpp:228-277) on the resident state
double mhdg_ctx::step(double dt_max, int precision, int *iters, double *res)
{
    // refusals before any launch, so that a throw leaves the resident state as it was
    if (precision < 0 || precision > 3)
        throw std::runtime_error("step: precision must be 0 (FP64), 1 (mixed), 2 (mixed with multigrid) or 3 (FP64 with classical Gram-Schmidt)");
    if (precision == 2 && mgh.n_nodes <= 0) throw std::runtime_error("multigrid preconditioner: the mesh descriptor carried no e2n incidence");
    if ((precision == 1 || precision == 2) && ideal_cfg && u_ideal)
        throw std::runtime_error("mixed-precision solve: the implicit ideal part has no FP32 operator; use precision 0");
    const int nb = nblk(nc);
    SpC Uc{}, Wc{};
    SpW Ww{};
    for (int s = 0; s < P.ns; ++s) { Uc.p[s] = U[s]; Wc.p[s] = W[s]; Ww.p[s] = W[s]; }
    LAUNCH("prepare", k_prepare<double>, dim3(nb), dim3(256), 0, stream, d, P, Uc, static_cast<const double *>(Phi),
           static_cast<const double *>(Badd), Ww, Bind, Btot, uplus, alpha,
                       dt_part);
    LAUNCH("dt_final", k_dt_final<double>, dim3(1), dim3(256), 0, stream, dt_part, nb, dt_max, dt_dev);
    double *dth = static_cast<double *>(host_pin);
    hip_check(hipMemcpyAsync(dth, dt_dev, 8, hipMemcpyDeviceToHost, stream), "D2H");
    hipEvent_t ev;
    hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
    hip_check(hipEventRecord(ev, stream), "event");
    // Strang split of the ions' gyration (Solver.cpp:237-241): half a step, then prepare again.  The rotation changes
    // U, so the time step is checked first, as the reference throws before it touches the state (Solver.cpp:236).
    const bool split = P.induce_coeff > 0 && P.lorentz;
    if (split)
    {
        hip_check(hipEventSynchronize(ev), "event sync");
        if (!std::isfinite(dth[0]) || dth[0] <= 0)
        {
            (void)hipEventDestroy(ev);
            throw std::runtime_error("invalid time step");
        }
        LAUNCH("gyrate", k_gyrate, dim3(nb), dim3(256), 0, stream, d, P, U[0], U[1], static_cast<const double *>(Btot),
               static_cast<const double *>(dt_dev), 0.5);
        LAUNCH("prepare", k_prepare<double>, dim3(nb), dim3(256), 0, stream, d, P, Uc, static_cast<const double *>(Phi),
               static_cast<const double *>(Badd), Ww, Bind, Btot, uplus, alpha, dt_part);
    }
    SpC Ws{}, none{}, gradsC{};
    SpW grads{}, Us{}, noneW{};
    for (int s = 0; s < P.ns; ++s)
    {
        Ws.p[s] = W[s];
        grads.p[s] = grad[s];
        gradsC.p[s] = grad[s];
        Us.p[s] = U[s];
    }
    if (P.second_order)
    {
        // many species: the stencil rows loaded once for all of them
        auto kgrad = P.ns > 2 ? k_grad_current<double, true> : k_grad_current<double, false>;
        LAUNCH("grad_current", kgrad, dim3(nb), dim3(256), 0, stream, d, P, P.ns, 0, Ws, none, grads, 1, Bind, J);
    }
    else
    {
        for (int s = 0; s < P.ns; ++s) hip_check(hipMemsetAsync(grad[s], 0, 15 * size_t(nc) * 8, stream), "memset");
        LAUNCH("grad_current", k_grad_current<double>, dim3(nb), dim3(256), 0, stream, d, P, 0, 0, Ws, none, grads, 1, Bind, J);
    }
    if (P.ns > 2)
    {
        // many species (for two, one thread per face with LDS-staged rows measured faster: 2.03 against 2.28 ms at 8M): one thread per (face, species) for the fluxes, then the face EMF candidates alone
        const int per = 256 / P.ns;
        LAUNCH("face", k_face_species<double>, dim3(nblk(nf, per)), dim3(256), 0, stream, d, P, P.ns, Ws, gradsC, FA);
        LAUNCH("face_emf", k_face<double>, dim3(nblk(nf)), dim3(256), 0, stream, d, P, 0, 0, Ws, gradsC, none, static_cast<double *>(nullptr),
               uplus, Btot, Bind, Phi, PhiAdd, Ef);
    }
    else
        LAUNCH("face", k_face<double>, dim3(nblk(nf)), dim3(256), 0, stream, d, P, P.ns, 0, Ws, gradsC, none, FA, uplus, Btot, Bind, Phi, PhiAdd, Ef);
    check_launch("explicit stage");
    hip_check(hipEventSynchronize(ev), "event sync");
    (void)hipEventDestroy(ev);
    const double gdt = dth[0];
    // nothing above wrote U or Phi (only derived fields), so the state is unchanged when this throws
    if (!std::isfinite(gdt) || gdt <= 0) throw std::runtime_error("invalid time step");
    if (P.ns > 2)
        LAUNCH("cell_update", k_cell_update_many<double>, dim3(nb), dim3(256), 0, stream, d, P, P.ns, static_cast<const double *>(FA), Wc,
               static_cast<const double *>(Btot), static_cast<const double *>(J), static_cast<const double *>(dt_dev), Us);
    else
        LAUNCH("cell_update", k_cell_update<double>, dim3(nb), dim3(256), 0, stream, d, P, P.ns, 0, static_cast<const double *>(FA), none, Wc,
               static_cast<const double *>(Btot), static_cast<const double *>(J), static_cast<const double *>(dt_dev), 0.0, 1, 1, Us, noneW);
    launch_edge_emf(Ef, uplus, Phi, P.cuct, nullptr, E_expl, 1);
    *iters = 0;
    *res = 0.0;
    const double *Euse = E_expl;
    if ((op_hall || op_res || ideal_cfg) && !skip_implicit) // ImplicitInduction::active()
    {
        hip_check(hipMemcpyAsync(B0, Btot, 3 * size_t(nc) * 8, hipMemcpyDeviceToDevice, stream), "copy");
        if (precision == 0) *iters = solve64(Phi, E_expl, dir_mercury, gdt, E_last, res);
        else if (precision == 3) *iters = solve64_cgs(Phi, E_expl, dir_mercury, gdt, E_last, res);
        else *iters = solve_mixed(Phi, E_expl, dir_mercury, gdt, E_last, res, precision == 2, dir_mercury_h.data());
        Euse = E_last;
    }
    else
        hip_check(hipMemcpyAsync(E_last, E_expl, size_t(ne) * 8, hipMemcpyDeviceToDevice, stream), "copy");
    LAUNCH("faraday", k_faraday<double>, dim3(nblk(nf)), dim3(256), 0, stream, d, Euse, nullptr, dt_dev, 0.0, Phi);
    if (split) // the second half of the gyration, with the field of the step's prepare (Solver.cpp:272)
        LAUNCH("gyrate", k_gyrate, dim3(nb), dim3(256), 0, stream, d, P, U[0], U[1], static_cast<const double *>(Btot),
               static_cast<const double *>(dt_dev), 0.5);
    check_launch("faraday");
    if (prof_on) prof_flush();
    return gdt;
}

// ------------------------------------------------------------------------------------ C-ABI
#define MHDG_TRY(ctx, ...)                                                                                                   \
    try                                                                                                                      \
    {                                                                                                                        \
        __VA_ARGS__;                                                                                                         \
        return 0;                                                                                                            \
    }                                                                                                                        \
    catch (const std::exception &e)                                                                                          \
    {                                                                                                                        \
        (ctx)->err = e.what();                                                                                               \
        return 1;                                                                                                            \
    }

extern "C" {
void mhdg_config_defaults(mhdg_config *c)
{
    std::memset(c, 0, sizeof(*c));
    c->gamma = 5.0 / 3.0; c->inver_MA2 = 1.0; c->M_ref = 1.0; c->M_H = 1.0; c->M_Na = 1.0; c->ne_floor = 0.01;
    c->hall_coef = 0.0; c->inver_Rem = 0.0;
    c->cfl = 1.0; c->recon_rel_limit = 0.0; c->rho_floor = 1e-10; c->p_floor = 1e-12; c->rho_floor_dt = 1e-2;
    c->second_order = 1; c->lorentz_force = 1; c->hall = 0; c->resistive = 0;
    c->eta_r_inner = 0.84; c->eta_r_outer = 1.04; c->eta_width = 0.02; c->hall_r_min = 1.0; c->hall_r_max = 1.012;
    c->ind_rtol = 1e-4; c->ind_max_it = 100; c->ind_restart = 20; c->cuct = 0.5;
    const double fs[5] = {1, -1, 0, 0, 1};
    for (int m = 0; m < 5; ++m) { c->fs[0][m] = fs[m]; c->fs[1][m] = fs[m]; }
    c->fs[1][0] = 1e-6; c->fs[1][4] = 1e-6;
    c->u_inf[0] = -1; c->u_inf[1] = 0; c->u_inf[2] = 0;
    c->mg_levels = 4; c->mg_agglom_passes = 3; c->mg_pre = 2; c->mg_post = 2; c->mg_coarse = 4; c->mg_relax = 0.7; c->mg_fine_hall = 1;
}

int mhdg_create(const mhdg_mesh_desc *mesh, const mhdg_config *cfg, mhdg_ctx **out)
{
    auto *c = new mhdg_ctx();
    try
    {
        if (!mesh || !cfg) throw std::runtime_error("mhdg_create: null descriptor");
        check_desc(*mesh); // before any device call: a bad descriptor is reported as such, with or without a GPU
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) throw std::runtime_error("no HIP device available");
        c->build(*mesh, *cfg);
        *out = c;
        return 0;
    }
    catch (const std::exception &e)
    {
        g_create_err = e.what();
        delete c;
        *out = nullptr;
        return 1;
    }
}
void mhdg_destroy(mhdg_ctx *c) { delete c; }
const char *mhdg_last_error(mhdg_ctx *c) { return c ? c->err.c_str() : g_create_err.c_str(); }
int mhdg_device_info(char *buf, int len)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 1;
    std::snprintf(buf, len, "%s %s CUs=%d L2=%d MB", p.name, p.gcnArchName, p.multiProcessorCount, p.l2CacheSize >> 20);
    return 0;
}
int mhdg_sync(mhdg_ctx *c) { MHDG_TRY(c, hip_check(hipStreamSynchronize(c->stream), "sync")) }
int mhdg_profile(mhdg_ctx *c, int enable)
{
    MHDG_TRY(c, {
        c->prof_flush();
        if (enable < 0) c->prof_acc.clear();
        else c->prof_on = enable != 0;
    })
}
int mhdg_profile_only(mhdg_ctx *c, const char *name)
{
    MHDG_TRY(c, {
        c->prof_flush();
        c->prof_only = name ? name : "";
    })
}
int mhdg_profile_report(mhdg_ctx *c, char *buf, int len)
{
    MHDG_TRY(c, {
        c->prof_flush();
        std::string r;
        char line[160];
        for (const auto &kv : c->prof_acc)
        {
            std::snprintf(line, sizeof(line), "%s %ld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
            r += line;
        }
        std::snprintf(buf, len, "%s", r.c_str());
    })
}

int mhdg_cell_vector(mhdg_ctx *c, const double *flux, double *b)
{
    MHDG_TRY(c, {
        const double *f = c->stage_in(0, flux, c->nf);
        double *o = c->stage(1, 3 * size_t(c->nc));
        hipLaunchKernelGGL(k_cell_vector, dim3(nblk(c->nc)), dim3(256), 0, c->stream, c->d, f, o);
        c->check_launch("cell_vector");
        c->d2h(b, o, 3 * size_t(c->nc) * 8);
    })
}
int mhdg_cell_current(mhdg_ctx *c, const double *B, double *J)
{
    MHDG_TRY(c, {
        const double *b = c->stage_in(0, B, 3 * size_t(c->nc));
        double *o = c->stage(1, 3 * size_t(c->nc));
        SpC none{};
        SpW nw{};
        hipLaunchKernelGGL(k_grad_current<double>, dim3(nblk(c->nc)), dim3(256), 0, c->stream, c->d, c->P, 0, 0, none, none, nw, 0, b, o);
        c->check_launch("cell_current");
        c->d2h(J, o, 3 * size_t(c->nc) * 8);
    })
}
int mhdg_limited_gradients(mhdg_ctx *c, const double *W, int species, const double fs[5], int limit, const double *ghost, double *grad)
{
    MHDG_TRY(c, {
        if (species < 0 || species >= kMaxSpecies) throw std::runtime_error("species out of range");
        Params P = c->P;
        for (int m = 0; m < 5; ++m) P.fs[species][m] = fs[m];
        SpC Ws{}, gh{};
        SpW gr{};
        Ws.p[0] = c->stage_in(0, W, 5 * size_t(c->nc));
        if (ghost) gh.p[0] = c->stage_in(2, ghost, 5 * size_t(c->nf));
        gr.p[0] = c->stage(1, 15 * size_t(c->nc));
        hipLaunchKernelGGL(k_grad_current<double>, dim3(nblk(c->nc)), dim3(256), 0, c->stream, c->d, P, 1, species, Ws, gh, gr, limit,
                           static_cast<const double *>(nullptr), static_cast<double *>(nullptr));
        c->check_launch("limited_gradients");
        c->d2h(grad, gr.p[0], 15 * size_t(c->nc) * 8);
    })
}
int mhdg_flux_rhs(mhdg_ctx *c, int species, const double fs[5], const double *W, const double *grad, const double *ghost, double *rhs)
{
    MHDG_TRY(c, {
        if (species < 0 || species >= kMaxSpecies) throw std::runtime_error("species out of range");
        Params P = c->P;
        for (int m = 0; m < 5; ++m) P.fs[species][m] = fs[m];
        SpC Ws{}, Gs{}, gh{}, none{};
        SpW ro{}, nw{};
        Ws.p[0] = c->stage_in(0, W, 5 * size_t(c->nc));
        Gs.p[0] = c->stage_in(1, grad, 15 * size_t(c->nc));
        if (ghost) gh.p[0] = c->stage_in(2, ghost, 5 * size_t(c->nf));
        double *fa = c->stage(3, 5 * size_t(c->nf));
        ro.p[0] = c->stage(4, 5 * size_t(c->nc));
        hipLaunchKernelGGL(k_face<double>, dim3(nblk(c->nf)), dim3(256), 0, c->stream, c->d, P, 1, species, Ws, Gs, gh, fa,
                           static_cast<const double *>(nullptr), static_cast<const double *>(nullptr), static_cast<const double *>(nullptr),
                           static_cast<const double *>(nullptr), static_cast<const double *>(nullptr), static_cast<double *>(nullptr));
        hipLaunchKernelGGL(k_cell_update<double>, dim3(nblk(c->nc)), dim3(256), 0, c->stream, c->d, P, 1, species, static_cast<const double *>(fa),
                           none, none, static_cast<const double *>(nullptr), static_cast<const double *>(nullptr),
                           static_cast<const double *>(nullptr), 0.0, 0, 0, nw, ro);
        c->check_launch("flux_rhs");
        c->d2h(rhs, ro.p[0], 5 * size_t(c->nc) * 8);
    })
}
int mhdg_ideal_emf(mhdg_ctx *c, const double *u, const double *Bt, const double *Bi, const double *pt, const double *pi, double cuct, double *E)
{
    MHDG_TRY(c, {
        const size_t C3 = 3 * size_t(c->nc);
        const double *du = c->stage_in(0, u, C3), *dbt = c->stage_in(1, Bt, C3), *dbi = c->stage_in(2, Bi, C3);
        const double *dpt = c->stage_in(3, pt, c->nf), *dpi = c->stage_in(4, pi, c->nf);
        double *dE = c->stage_in(5, E, c->ne);
        double *ef = c->stage(6, 3 * size_t(c->nf));
        SpC none{};
        hipLaunchKernelGGL(k_face<double>, dim3(nblk(c->nf)), dim3(256), 0, c->stream, c->d, c->P, 0, 0, none, none, none,
                           static_cast<double *>(nullptr), du, dbt, dbi, dpt, static_cast<const double *>(nullptr), ef);
        c->launch_edge_emf(ef, du, dpi, cuct, dE, dE, 0);
        c->d2h(E, dE, size_t(c->ne) * 8);
    })
}
int mhdg_faraday(mhdg_ctx *c, const double *E, double *Phi, double dt)
{
    MHDG_TRY(c, {
        const double *dE = c->stage_in(0, E, c->ne);
        double *dP = c->stage_in(1, Phi, c->nf);
        hipLaunchKernelGGL(k_faraday<double>, dim3(nblk(c->nf)), dim3(256), 0, c->stream, c->d, dE, static_cast<const double *>(nullptr),
                           static_cast<const double *>(nullptr), dt, dP);
        c->check_launch("faraday");
        c->d2h(Phi, dP, size_t(c->nf) * 8);
    })
}
int mhdg_faraday_edge_dt(mhdg_ctx *c, const double *E, const double *dte, double *Phi)
{
    MHDG_TRY(c, {
        const double *dE = c->stage_in(0, E, c->ne), *dT = c->stage_in(2, dte, c->ne);
        double *dP = c->stage_in(1, Phi, c->nf);
        hipLaunchKernelGGL(k_faraday<double>, dim3(nblk(c->nf)), dim3(256), 0, c->stream, c->d, dE, dT, static_cast<const double *>(nullptr), 0.0, dP);
        c->check_launch("faraday_edge_dt");
        c->d2h(Phi, dP, size_t(c->nf) * 8);
    })
}
int mhdg_div_b(mhdg_ctx *c, const double *B, const double *extra, double *max_abs, double *max_rel, int64_t *worst)
{
    MHDG_TRY(c, {
        const int nb = nblk(c->nc);
        const double *dB = c->stage_in(0, B, c->nf);
        const double *dX = extra ? c->stage_in(1, extra, c->nf) : nullptr;
        double *pa = c->stage(2, nb), *pr = c->stage(3, nb);
        int32_t *pc = reinterpret_cast<int32_t *>(c->stage(4, nb));
        hipLaunchKernelGGL(k_div_b, dim3(nb), dim3(256), 0, c->stream, c->d, dB, dX, pa, pc, pr);
        c->check_launch("div_b");
        std::vector<double> ha(nb), hr(nb);
        std::vector<int32_t> hc(nb);
        c->d2h(ha.data(), pa, nb * 8); c->d2h(hr.data(), pr, nb * 8); c->d2h(hc.data(), pc, nb * 4);
        double ma = 0, mr = 0;
        int32_t wc = INT32_MAX;
        for (int i = 0; i < nb; ++i)
        {
            if (ha[i] > ma || (ha[i] == ma && ha[i] > 0 && hc[i] < wc)) { ma = ha[i]; wc = hc[i]; }
            mr = std::max(mr, hr[i]);
        }
        *max_abs = ma; *max_rel = mr; *worst = wc == INT32_MAX ? -1 : wc;
    })
}
int mhdg_local_dt(mhdg_ctx *c, const double *WH, const double *WNa, const double *Bt, double *dt)
{
    MHDG_TRY(c, {
        if (c->P.ns != 2) throw std::runtime_error("local_dt: the two-species (H+, Na+) entry needs n_species 2");
        const int nb = nblk(c->nc);
        SpC Ws{};
        Ws.p[0] = c->stage_in(0, WH, 5 * size_t(c->nc));
        Ws.p[1] = c->stage_in(1, WNa, 5 * size_t(c->nc));
        const double *bt = c->stage_in(2, Bt, 3 * size_t(c->nc));
        double *part = c->stage(3, nb), *o = c->stage(4, 1);
        hipLaunchKernelGGL(k_local_dt, dim3(nb), dim3(256), 0, c->stream, c->d, c->P, Ws, bt, part);
        hipLaunchKernelGGL(k_dt_final<double>, dim3(1), dim3(256), 0, c->stream, static_cast<const double *>(part), nb, __builtin_inf(), o);
        c->check_launch("local_dt");
        c->d2h(dt, o, 8);
    })
}
int mhdg_add_sources(mhdg_ctx *c, int species, const double *WH, const double *WNa, const double *Bt, const double *Jc, double dt, double *rhs)
{
    MHDG_TRY(c, {
        if (species < 0 || species > 1) throw std::runtime_error("sources: species 0 or 1");
        if (c->P.ns != 2) throw std::runtime_error("add_sources: the two-species (H+, Na+) entry needs n_species 2");
        const size_t C = c->nc;
        SpC Wall{};
        Wall.p[0] = c->stage_in(0, WH, 5 * C);
        Wall.p[1] = c->stage_in(1, WNa, 5 * C);
        const double *bt = c->stage_in(2, Bt, 3 * C), *j = c->stage_in(3, Jc, 3 * C);
        double *r = c->stage_in(4, rhs, 5 * C);
        SpC ri{};
        SpW ro{}, nw{};
        ri.p[0] = r;
        ro.p[0] = r;
        hipLaunchKernelGGL(k_cell_update<double>, dim3(nblk(c->nc)), dim3(256), 0, c->stream, c->d, c->P, 1, species,
                           static_cast<const double *>(nullptr), ri, Wall, bt, j, static_cast<const double *>(nullptr), dt, 1, 0, nw, ro);
        c->check_launch("add_sources");
        c->d2h(rhs, r, 5 * C * 8);
    })
}
int mhdg_ind_set_frozen(mhdg_ctx *c, const double *B0c, const double *ne_eff)
{
    MHDG_TRY(c, {
        const size_t C = c->nc;
        hip_check(hipMemcpyAsync(c->B0, B0c, 3 * C * 8, hipMemcpyHostToDevice, c->stream), "H2D");
        std::vector<double> al(C);
        for (size_t i = 0; i < C; ++i) al[i] = (c->P.hall && ne_eff[i] > 0) ? c->taper_h[i] * c->cfg.hall_coef / ne_eff[i] : 0.0;
        hip_check(hipMemcpyAsync(c->alpha, al.data(), C * 8, hipMemcpyHostToDevice, c->stream), "H2D");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    })
}
int mhdg_ind_apply_F(mhdg_ctx *c, const double *phi, double *out)
{
    MHDG_TRY(c, {
        const double *p = c->stage_in(0, phi, c->nf);
        c->apply_op(p, nullptr, nullptr, nullptr, 0.0, 0, 0, nullptr, c->tmp);
        c->d2h(out, c->tmp, size_t(c->ne) * 8);
    })
}
int mhdg_set_recon_rel_limit(mhdg_ctx *c, double theta)
{
    MHDG_TRY(c, {
        if (!(theta >= 0)) throw std::runtime_error("recon_rel_limit must be >= 0");
        c->P.rel_limit = theta;
        c->cfg.recon_rel_limit = theta;
    })
}
int mhdg_ind_flags(mhdg_ctx *c, int32_t *active, int32_t *has_ideal, int32_t *has_hall)
{
    MHDG_TRY(c, {
        if (active) *active = c->op_res || c->op_hall || c->ideal_cfg;
        if (has_ideal) *has_ideal = c->ideal_cfg;
        if (has_hall) *has_hall = c->op_hall;
    })
}
int mhdg_ind_set_ideal_state(mhdg_ctx *c, const double *u, const int8_t *mask)
{
    MHDG_TRY(c, {
        if (!c->u_ideal)
        {
            c->u_ideal = static_cast<double *>(c->alloc(3 * size_t(c->nc) * 8));
            c->ide = static_cast<double *>(c->alloc(size_t(c->ne) * 8));
            c->Ef_i = static_cast<double *>(c->alloc(3 * size_t(c->nf) * 8));
            c->imask = static_cast<int8_t *>(c->alloc(size_t(c->ne)));
        }
        hip_check(hipMemcpyAsync(c->u_ideal, u, 3 * size_t(c->nc) * 8, hipMemcpyHostToDevice, c->stream), "H2D");
        hip_check(hipMemcpyAsync(c->imask, mask, size_t(c->ne), hipMemcpyHostToDevice, c->stream), "H2D");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    })
}
int mhdg_ind_ideal_linear(mhdg_ctx *c, const double *phi, double *out)
{
    MHDG_TRY(c, {
        const double *p = c->stage_in(0, phi, c->nf);
        const bool h = c->op_hall, r = c->op_res;
        c->op_hall = false;
        c->op_res = false;
        try { c->apply_op(p, nullptr, nullptr, nullptr, 0.0, 0, 0, nullptr, c->tmp); }
        catch (...) { c->op_hall = h; c->op_res = r; throw; }
        c->op_hall = h;
        c->op_res = r;
        c->d2h(out, c->tmp, size_t(c->ne) * 8);
    })
}
int mhdg_ind_solve(mhdg_ctx *c, const double *Phi_n, const double *E_expl, const int8_t *dir, double dt, const double *dt_edge,
                   int precision, double *E, int32_t *iterations, double *rel_res)
{
    MHDG_TRY(c, {
        const double *p = c->stage_in(0, Phi_n, c->nf), *ee = c->stage_in(1, E_expl, c->ne);
        int8_t *dd = reinterpret_cast<int8_t *>(c->stage(2, (size_t(c->ne) + 7) / 8));
        hip_check(hipMemcpyAsync(dd, dir, size_t(c->ne), hipMemcpyHostToDevice, c->stream), "H2D");
        double *eo = c->stage(3, c->ne);
        c->dte_cur = dt_edge ? c->stage_in(4, dt_edge, c->ne) : nullptr;
        double rr = 0;
        int it = 0;
        if (precision < 0 || precision > 3)
            throw std::runtime_error("solve: precision must be 0 (FP64), 1 (mixed), 2 (mixed with multigrid) or 3 (FP64 with classical Gram-Schmidt)");
        try
        {
            it = precision == 0   ? c->solve64(p, ee, dd, dt, eo, &rr)
                 : precision == 3 ? c->solve64_cgs(p, ee, dd, dt, eo, &rr)
                                  : c->solve_mixed(p, ee, dd, dt, eo, &rr, precision == 2, dir);
        }
        catch (...) { c->dte_cur = nullptr; throw; }
        c->dte_cur = nullptr;
        c->d2h(E, eo, size_t(c->ne) * 8);
        *iterations = it;
        if (rel_res) *rel_res = rr;
    })
}
int mhdg_ind_apply_G(mhdg_ctx *c, const double *phi, double *out)
{
    MHDG_TRY(c, {
        const double *p = c->stage_in(0, phi, c->nf);
        c->apply_G(p, c->tmp);
        c->d2h(out, c->tmp, size_t(c->ne) * 8);
    })
}
int mhdg_ind_jvp(mhdg_ctx *c, const double *phi_lin, const double *dphi, double *out)
{
    MHDG_TRY(c, {
        const double *p = c->stage_in(0, phi_lin, c->nf), *q = c->stage_in(1, dphi, c->nf);
        c->jvp(p, q, c->tmp);
        c->d2h(out, c->tmp, size_t(c->ne) * 8);
    })
}
int mhdg_ind_newton(mhdg_ctx *c, const double *Phi_n, const double *E_expl, const int8_t *dir, double dt, const double *E0, double rtol,
                    int max_newton, double eta, double *E, int32_t *newton_its, int32_t *gmres_its, double *rel_res)
{
    MHDG_TRY(c, {
        if (c->ideal_cfg && c->u_ideal) throw std::runtime_error("newton: the implicit ideal part is not supported");
        if (!(rtol > 0) || !(eta > 0) || max_newton < 0) throw std::runtime_error("newton: rtol, eta > 0 and max_newton >= 0 required");
        const double *p = c->stage_in(0, Phi_n, c->nf), *ee = c->stage_in(1, E_expl, c->ne);
        int8_t *dd = reinterpret_cast<int8_t *>(c->stage(2, (size_t(c->ne) + 7) / 8));
        hip_check(hipMemcpyAsync(dd, dir, size_t(c->ne), hipMemcpyHostToDevice, c->stream), "H2D");
        const double *e0 = E0 ? c->stage_in(4, E0, c->ne) : nullptr;
        double *eo = c->stage(3, c->ne);
        int gm = 0;
        double rr = 0;
        const int k = c->newton(p, ee, dd, dt, e0, rtol, max_newton, eta, eo, &gm, &rr);
        c->d2h(E, eo, size_t(c->ne) * 8);
        if (newton_its) *newton_its = k;
        if (gmres_its) *gmres_its = gm;
        if (rel_res) *rel_res = rr;
    })
}
int mhdg_shell_energy(mhdg_ctx *c, double r0, double r1, double *J)
{
    MHDG_TRY(c, {
        c->tangent_alloc();
        c->shell_energy<double>(c->Phi, c->Badd, r0, r1, J);
    })
}
int mhdg_explicit_tangent(mhdg_ctx *c, int n_steps, double dt_max, const double *dU_H, const double *dU_Na, const double *dPhi,
                          const double *dPhiAdd, const double *dBadd, const double *dfs, const double *dE_inf, double r0, double r1,
                          double *J, double *dJ, double *dU_H_out, double *dU_Na_out, double *dPhi_out)
{
    MHDG_TRY(c, {
        if (n_steps < 0) throw std::runtime_error("explicit_tangent: n_steps >= 0");
        if (c->P.ns != 2) throw std::runtime_error("explicit_tangent: two-species (H+, Na+) contexts only");
        if (c->P.induce_coeff > 0) throw std::runtime_error("explicit_tangent: the gyration split has no tangent");
        c->tangent_alloc();
        const size_t C = c->nc, F = c->nf;
        auto seed = [&](const double *val, const double *host_d, size_t n, int slot, Dual *out) {
            const double *dd = host_d ? c->stage_in(slot, host_d, n) : nullptr;
            hipLaunchKernelGGL(k_seed, dim3(nblk(int64_t(n))), dim3(256), 0, c->stream, val, dd, int64_t(n), out);
        };
        seed(c->U[0], dU_H, 5 * C, 0, c->tw.U[0]);
        seed(c->U[1], dU_Na, 5 * C, 1, c->tw.U[1]);
        seed(c->Phi, dPhi, F, 2, c->tw.Phi);
        seed(c->PhiAdd, dPhiAdd, F, 3, c->tw.PhiAdd);
        seed(c->Badd, dBadd, 3 * C, 4, c->tw.Badd);
        Params Pt = c->P;
        for (int s = 0; s < 2; ++s)
            for (int m = 0; m < 5; ++m) Pt.dfs[s][m] = dfs ? dfs[5 * s + m] : 0.0;
        for (int k = 0; k < 3; ++k) Pt.dE_inf[k] = dE_inf ? dE_inf[k] : 0.0;
        c->explicit_tangent(n_steps, dt_max, Pt);
        Dual e;
        c->shell_energy<Dual>(c->tw.Phi, c->tw.Badd, r0, r1, &e);
        if (J) *J = e.v;
        if (dJ) *dJ = e.d;
        auto out = [&](const Dual *x, size_t n, double *host) {
            if (!host) return;
            double *tmpd = c->stage(5, n);
            hipLaunchKernelGGL(k_tangent_of, dim3(nblk(int64_t(n))), dim3(256), 0, c->stream, x, int64_t(n), tmpd);
            c->d2h(host, tmpd, n * 8);
        };
        out(c->tw.U[0], 5 * C, dU_H_out);
        out(c->tw.U[1], 5 * C, dU_Na_out);
        out(c->tw.Phi, F, dPhi_out);
    })
}
int mhdg_step_vjp(mhdg_ctx *c, int n_steps, double dt_max, double r0, double r1, double *J, double *gU_H, double *gU_Na, double *gPhi,
                  double *gPhiAdd, double *gBadd, double *gfs, double *gE_inf)
{
    MHDG_TRY(c, {
        if (n_steps < 0) throw std::runtime_error("step_vjp: n_steps >= 0");
        if (c->P.ns != 2) throw std::runtime_error("step_vjp: two-species (H+, Na+) contexts only");
        if (c->P.induce_coeff > 0) throw std::runtime_error("step_vjp: the gyration split has no adjoint");
        const size_t C = c->nc, F = c->nf;
        const double v = c->step_vjp(n_steps, dt_max, r0, r1, gfs, gE_inf);
        if (J) *J = v;
        if (gU_H) c->d2h(gU_H, c->aw.Ub[0], 5 * C * 8);
        if (gU_Na) c->d2h(gU_Na, c->aw.Ub[1], 5 * C * 8);
        if (gPhi) c->d2h(gPhi, c->aw.Phib, F * 8);
        if (gPhiAdd) c->d2h(gPhiAdd, c->aw.PhiAddb, F * 8);
        if (gBadd) c->d2h(gBadd, c->aw.Baddb, 3 * C * 8);
    })
}
int mhdg_set_state(mhdg_ctx *c, const double *UH, const double *UNa, const double *Ph)
{
    MHDG_TRY(c, {
        hip_check(hipMemcpy(c->U[0], UH, 5 * size_t(c->nc) * 8, hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemcpy(c->U[1], UNa, 5 * size_t(c->nc) * 8, hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemcpy(c->Phi, Ph, size_t(c->nf) * 8, hipMemcpyHostToDevice), "H2D");
        c->have_corr = false;
    })
}
int mhdg_get_state(mhdg_ctx *c, double *UH, double *UNa, double *Ph)
{
    MHDG_TRY(c, {
        hip_check(hipStreamSynchronize(c->stream), "sync");
        hip_check(hipMemcpy(UH, c->U[0], 5 * size_t(c->nc) * 8, hipMemcpyDeviceToHost), "D2H");
        hip_check(hipMemcpy(UNa, c->U[1], 5 * size_t(c->nc) * 8, hipMemcpyDeviceToHost), "D2H");
        hip_check(hipMemcpy(Ph, c->Phi, size_t(c->nf) * 8, hipMemcpyDeviceToHost), "D2H");
    })
}
int mhdg_set_species_state(mhdg_ctx *c, int species, const double *U)
{
    MHDG_TRY(c, {
        if (species < 0 || species >= c->P.ns) throw std::runtime_error("species out of range for this context's n_species");
        hip_check(hipMemcpy(c->U[species], U, 5 * size_t(c->nc) * 8, hipMemcpyHostToDevice), "H2D");
        c->have_corr = false;
    })
}
int mhdg_get_species_state(mhdg_ctx *c, int species, double *U)
{
    MHDG_TRY(c, {
        if (species < 0 || species >= c->P.ns) throw std::runtime_error("species out of range for this context's n_species");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        hip_check(hipMemcpy(U, c->U[species], 5 * size_t(c->nc) * 8, hipMemcpyDeviceToHost), "D2H");
    })
}
int mhdg_get_species_prim(mhdg_ctx *c, int species, double *W)
{
    MHDG_TRY(c, {
        if (species < 0 || species >= c->P.ns) throw std::runtime_error("species out of range for this context's n_species");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        hip_check(hipMemcpy(W, c->W[species], 5 * size_t(c->nc) * 8, hipMemcpyDeviceToHost), "D2H");
    })
}
int mhdg_set_background(mhdg_ctx *c, const double *PA, const double *BA)
{
    MHDG_TRY(c, {
        hip_check(hipMemcpy(c->PhiAdd, PA, size_t(c->nf) * 8, hipMemcpyHostToDevice), "H2D");
        hip_check(hipMemcpy(c->Badd, BA, 3 * size_t(c->nc) * 8, hipMemcpyHostToDevice), "H2D");
    })
}
int mhdg_step(mhdg_ctx *c, double dt_max, int precision, double *dt_used, int32_t *iterations, double *rel_res)
{
    MHDG_TRY(c, {
        int it = 0;
        double rr = 0;
        const double dt = c->step(dt_max, precision, &it, &rr);
        if (dt_used) *dt_used = dt;
        if (iterations) *iterations = it;
        if (rel_res) *rel_res = rr;
    })
}
int mhdg_step_explicit(mhdg_ctx *c, double dt_max, double *dt_used)
{
    MHDG_TRY(c, {
        int it = 0;
        double rr = 0;
        c->skip_implicit = true;
        double dt = 0;
        try { dt = c->step(dt_max, 0, &it, &rr); }
        catch (...) { c->skip_implicit = false; throw; }
        c->skip_implicit = false;
        if (dt_used) *dt_used = dt;
    })
}
int mhdg_step_hall(mhdg_ctx *c, double dt_max, int precision, double *dt_used, int32_t *iterations, double *rel_res)
{
    MHDG_TRY(c, {
        if (!(c->op_hall || c->op_res)) throw std::runtime_error("step_hall: the configuration has no implicit Hall/resistive part");
        int it = 0;
        double rr = 0;
        const double dt = c->step(dt_max, precision, &it, &rr);
        if (dt_used) *dt_used = dt;
        if (iterations) *iterations = it;
        if (rel_res) *rel_res = rr;
    })
}
int mhdg_state_save(mhdg_ctx *c)
{
    MHDG_TRY(c, {
        const size_t C = c->nc, F = c->nf, ns = size_t(c->P.ns);
        if (!c->snapshot) c->snapshot = static_cast<double *>(c->alloc((5 * C * ns + F) * 8));
        for (size_t s = 0; s < ns; ++s)
            hip_check(hipMemcpyAsync(c->snapshot + 5 * C * s, c->U[s], 5 * C * 8, hipMemcpyDeviceToDevice, c->stream), "copy");
        hip_check(hipMemcpyAsync(c->snapshot + 5 * C * ns, c->Phi, F * 8, hipMemcpyDeviceToDevice, c->stream), "copy");
    })
}
int mhdg_state_restore(mhdg_ctx *c)
{
    MHDG_TRY(c, {
        if (!c->snapshot) throw std::runtime_error("state_restore: no saved state");
        const size_t C = c->nc, F = c->nf, ns = size_t(c->P.ns);
        for (size_t s = 0; s < ns; ++s)
            hip_check(hipMemcpyAsync(c->U[s], c->snapshot + 5 * C * s, 5 * C * 8, hipMemcpyDeviceToDevice, c->stream), "copy");
        hip_check(hipMemcpyAsync(c->Phi, c->snapshot + 5 * C * ns, F * 8, hipMemcpyDeviceToDevice, c->stream), "copy");
        c->have_corr = false;
    })
}
int mhdg_prepare(mhdg_ctx *c)
{
    MHDG_TRY(c, {
        const int nb = nblk(c->nc);
        SpC Uc{}, none{};
        SpW Ww{}, noneW{};
        for (int s = 0; s < c->P.ns; ++s) { Uc.p[s] = c->U[s]; Ww.p[s] = c->W[s]; }
        hipLaunchKernelGGL(k_prepare<double>, dim3(nb), dim3(256), 0, c->stream, c->d, c->P, Uc, static_cast<const double *>(c->Phi),
                           static_cast<const double *>(c->Badd), Ww, c->Bind, c->Btot, c->uplus, c->alpha, c->dt_part);
        hipLaunchKernelGGL(k_grad_current<double>, dim3(nb), dim3(256), 0, c->stream, c->d, c->P, 0, 0, none, none, noneW, 1,
                           static_cast<const double *>(c->Bind), c->J);
        c->check_launch("prepare");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    })
}
int mhdg_get_derived(mhdg_ctx *c, int which, double *out)
{
    MHDG_TRY(c, {
        const size_t C = c->nc;
        const double *src = nullptr;
        size_t n = 0;
        switch (which)
        {
        case 0: src = c->W[0]; n = 5 * C; break;
        case 1: src = c->W[1]; n = 5 * C; break;
        case 2: src = c->Bind; n = 3 * C; break;
        case 3: src = c->Btot; n = 3 * C; break;
        case 4: src = c->J; n = 3 * C; break;
        case 5: src = c->uplus; n = 3 * C; break;
        case 6: src = c->E_last; n = c->ne; break;
        default: throw std::runtime_error("unknown derived field");
        }
        c->d2h(out, src, n * 8);
    })
}
}
