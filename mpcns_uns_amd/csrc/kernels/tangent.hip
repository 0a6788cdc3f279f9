// Nonlinear Hall induction operator and its tangent-linear kernel (Newton-Krylov on the implicit induction).
//
// The semi-implicit step of ImplicitInduction (ImplicitInduction.cpp:105-198) freezes the field of the Hall
// term at B0 = B_tot^n.  Made a function of the unknown flux, the Hall EMF of a cell is
//     g_c(Phi) = alpha_c J_c[b(Phi)] x (B_add,c + b_c(Phi)),   b = R Phi (the recon coefficients),
// quadratic in Phi, and the step R(E) = E - E_expl - G(Phi^n - dt d1 E) = 0 is solved by Newton's method with
// matrix-free Jacobian products  J v = v + dt G'(Phi) d1 v,  whose Hall part is the tangent kernel below:
//     g'_c = alpha_c ( J_c[db] x (B_add,c + b_c) + J_c[b] x db_c ).
#include <type_traits>

#include "ctx.h"

namespace mhdg
{
namespace
{
// J_c[b] = (1/V_c) sum_f (s A_f) x b_f, b_f the face average of the two cells or b_c on the boundary
// (ImplicitInduction.cpp:162-172)
__device__ __forceinline__ V3 cell_curl(const DevMesh &d, const double *b, int32_t c)
{
    V3 sum;
    const V3 bc = L3(b, c);
    for (int32_t k = d.c2f_ptr[c]; k < d.c2f_ptr[c + 1]; ++k)
    {
        const int32_t v = d.c2f[k], f = pid(v);
        V3 bf = bc;
        const int32_t l = d.fL[f], r = d.fR[f];
        if (l >= 0 && r >= 0) bf = 0.5 * (bc + L3(b, l == c ? r : l));
        sum += cross(psg(v) * L3(d.farea, f), bf);
    }
    return (1.0 / d.cvol[c]) * sum;
}
} // namespace

// B0 = B_add + b: the total cell field of the iterate
__global__ void __launch_bounds__(256) k_field_of_flux(const double *Badd, const double *b, int64_t n, double *B0)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) B0[i] = Badd[i] + b[i];
}

// g' = alpha (J[db] x (B_add + b) + J[b] x db), zero where alpha = 0
__global__ void __launch_bounds__(256) k_hall_g_tangent(DevMesh d, const double *b, const double *db, const double *alpha,
                                                        const double *Badd, double *g)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    const double a = alpha[c];
    V3 out;
    if (a != 0.0)
    {
        const V3 Jd = cell_curl(d, db, c), Jb = cell_curl(d, b, c);
        const V3 B0 = L3(Badd, c) + L3(b, c);
        out = a * (cross(Jd, B0) + cross(Jb, L3(db, c)));
    }
    st3(g, c, out);
}

// Newton residual R = E - E_expl - G, zero on Dirichlet rows (E = E_expl imposed there)
__global__ void __launch_bounds__(256) k_newton_residual(const double *E, const double *E_expl, const double *G, const int8_t *dir, int32_t n,
                                                         double *R)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) R[e] = dir[e] ? 0.0 : E[e] - E_expl[e] - G[e];
}

// out = s * x
__global__ void __launch_bounds__(256) k_scaled(const double *x, double s, int32_t n, double *out)
{
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = s * x[e];
}
} // namespace mhdg

using namespace mhdg;

namespace
{
inline int nbk(int64_t n) { return static_cast<int>((n + 255) / 256); }
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

void mhdg_ctx::newton_alloc()
{
    if (nt.phik) return;
    auto D = [&](size_t n) { return static_cast<double *>(alloc(n * 8)); };
    nt.phik = D(nf); nt.G = D(ne); nt.R = D(ne); nt.rhs = D(ne); nt.delta = D(ne); nt.E = D(ne);
    nt.lin_b = D(3 * size_t(nc)); nt.B0 = D(3 * size_t(nc)); nt.zero_phi = D(nf); nt.corr_save = D(ne);
    chk(hipMemsetAsync(nt.zero_phi, 0, size_t(nf) * 8, stream), "memset");
}

// G(phi): apply_F with the Hall field B0 = B_add + R phi instead of the frozen one
void mhdg_ctx::apply_G(const double *phi, double *out)
{
    newton_alloc();
    if (op_hall)
    {
        LAUNCH("bcell", k_bcell, dim3(nbk(nc)), dim3(256), 0, stream, d, phi, bcell);
        LAUNCH("field_of_flux", k_field_of_flux, dim3(nbk(3 * int64_t(nc))), dim3(256), 0, stream, Badd, bcell, 3 * int64_t(nc), nt.B0);
        LAUNCH("hall_g", k_hall_g, dim3(nbk(nc)), dim3(256), 0, stream, d, bcell, alpha, nt.B0, gcell);
    }
    LAUNCH("edge_apply", k_edge_apply, dim3(nbk(ne)), dim3(256), 0, stream, d, int32_t(op_res), int32_t(op_hall), phi, gcell,
           static_cast<const double *>(nullptr), static_cast<const int8_t *>(nullptr), static_cast<const double *>(nullptr),
           static_cast<const double *>(nullptr), static_cast<const int8_t *>(nullptr), 0.0, 0, 0, static_cast<const double *>(nullptr), out);
    if (n_heavy)
        LAUNCH("edge_apply_heavy", k_edge_apply_heavy, dim3(n_heavy), dim3(256), 0, stream, d, int32_t(op_res), int32_t(op_hall),
               static_cast<const int32_t *>(heavy), phi, gcell, static_cast<const double *>(nullptr), static_cast<const int8_t *>(nullptr),
               static_cast<const double *>(nullptr), static_cast<const double *>(nullptr), static_cast<const int8_t *>(nullptr), 0.0, 0, 0,
               static_cast<const double *>(nullptr), out);
    check_launch("apply_G");
}

// G'(phi_lin) dphi
void mhdg_ctx::jvp(const double *phi_lin, const double *dphi, double *out)
{
    newton_alloc();
    LAUNCH("bcell", k_bcell, dim3(nbk(nc)), dim3(256), 0, stream, d, phi_lin, nt.lin_b);
    lin_b = nt.lin_b;
    try { apply_op(dphi, nullptr, nullptr, nullptr, 0.0, 0, 0, nullptr, out); }
    catch (...) { lin_b = nullptr; throw; }
    lin_b = nullptr;
}

// Newton-Krylov on R(E) = E - E_expl - G(Phi_n - dt d1 E) = 0: each step solves J delta = -R with the FP64
// GMRES of the semi-implicit solve (Jacobi, MGS) on the tangent operator, relative forcing term eta.
int mhdg_ctx::newton(const double *Phi_n, const double *Eex, const int8_t *dir, double dt, const double *E0, double rtol, int max_newton,
                     double eta, double *Eout, int *gmres_total, double *rel_res)
{
    newton_alloc();
    const size_t E8 = size_t(ne) * 8;
    chk(hipMemcpyAsync(nt.E, E0 ? E0 : Eex, E8, hipMemcpyDeviceToDevice, stream), "copy");
    // scale of the system: |E_expl + G(Phi^n)|
    apply_G(Phi_n, nt.G);
    LAUNCH("axpy", k_add, dim3(nbk(ne)), dim3(256), 0, stream, Eex, nt.G, ne, nt.rhs);
    const double bnorm = std::sqrt(std::max(dot(nt.rhs, nt.rhs, 0, 0), 1e-300));
    // the warm-start correction of the semi-implicit solve is not touched
    const bool had_corr = have_corr;
    if (had_corr) chk(hipMemcpyAsync(nt.corr_save, corr, E8, hipMemcpyDeviceToDevice, stream), "copy");
    int k = 0, gm = 0;
    double rn = 0;
    struct Restore
    {
        mhdg_ctx *c; bool had;
        ~Restore()
        {
            c->lin_b = nullptr; c->rtol_override = 0; c->have_corr = had;
            if (had) (void)hipMemcpyAsync(c->corr, c->nt.corr_save, size_t(c->ne) * 8, hipMemcpyDeviceToDevice, c->stream);
        }
    } restore{this, had_corr};
    for (;; ++k)
    {
        chk(hipMemcpyAsync(nt.phik, Phi_n, size_t(nf) * 8, hipMemcpyDeviceToDevice, stream), "copy");
        LAUNCH("faraday", k_faraday<double>, dim3(nbk(nf)), dim3(256), 0, stream, d, nt.E, static_cast<const double *>(nullptr),
               static_cast<const double *>(nullptr), dt, nt.phik);
        apply_G(nt.phik, nt.G);
        LAUNCH("newton_residual", k_newton_residual, dim3(nbk(ne)), dim3(256), 0, stream, nt.E, Eex, nt.G, dir, ne, nt.R);
        rn = std::sqrt(dot(nt.R, nt.R, 0, 0));
        if (!(rn > rtol * bnorm) || k >= max_newton) break;
        // J delta = -R, J = I + dt G'(Phi_k) d1, Dirichlet rows identity
        LAUNCH("bcell", k_bcell, dim3(nbk(nc)), dim3(256), 0, stream, d, nt.phik, nt.lin_b);
        LAUNCH("scale", k_scaled, dim3(nbk(ne)), dim3(256), 0, stream, nt.R, -1.0, ne, nt.rhs);
        lin_b = nt.lin_b;
        rtol_override = eta;
        have_corr = false;
        gm += solve64(nt.zero_phi, nt.rhs, dir, dt, nt.delta, nullptr);
        lin_b = nullptr;
        rtol_override = 0;
        LAUNCH("axpy", k_add, dim3(nbk(ne)), dim3(256), 0, stream, nt.E, nt.delta, ne, nt.E);
    }
    chk(hipMemcpyAsync(Eout, nt.E, E8, hipMemcpyDeviceToDevice, stream), "copy");
    check_launch("newton");
    if (gmres_total) *gmres_total = gm;
    if (rel_res) *rel_res = rn / bnorm;
    return k;
}

// ------------------------------------------------------------------ tangent-linear of the explicit step
namespace mhdg
{
// value + tangent of a double array: out[i] = {v[i], dv ? dv[i] : 0}
__global__ void __launch_bounds__(256) k_seed(const double *v, const double *dv, int64_t n, Dual *out)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = Dual(v[i], dv ? dv[i] : 0.0);
}
// split a Dual array into its tangent part
__global__ void __launch_bounds__(256) k_tangent_of(const Dual *x, int64_t n, double *d)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) d[i] = x[i].d;
}
} // namespace mhdg

void mhdg_ctx::tangent_alloc()
{
    if (tw.U[0]) return;
    if (n_heavy)
    {
        std::vector<int32_t> hv(n_heavy), ptr(ne + 1);
        chk(hipMemcpy(hv.data(), heavy, sizeof(int32_t) * n_heavy, hipMemcpyDeviceToHost), "D2H");
        chk(hipMemcpy(ptr.data(), d.e2f_ptr, sizeof(int32_t) * (ne + 1), hipMemcpyDeviceToHost), "D2H");
        for (int32_t e : hv)
            if (ptr[e + 1] - ptr[e] > kHeavyShared<Dual>())
                throw std::runtime_error("tangent step: edge valence above " + std::to_string(kHeavyShared<Dual>()));
    }
    auto D = [&](size_t n) { return static_cast<Dual *>(alloc(n * sizeof(Dual))); };
    const size_t C = nc, F = nf, E = ne;
    tw.U[0] = D(5 * C); tw.U[1] = D(5 * C); tw.Phi = D(F); tw.PhiAdd = D(F); tw.Badd = D(3 * C);
    tw.W[0] = D(5 * C); tw.W[1] = D(5 * C); tw.grad[0] = D(15 * C); tw.grad[1] = D(15 * C);
    tw.Bind = D(3 * C); tw.Btot = D(3 * C); tw.uplus = D(3 * C); tw.J = D(3 * C); tw.alpha = D(C);
    tw.FA = D(10 * F); tw.Ef = D(3 * F); tw.E = D(E);
    tw.dt_part = D(size_t(nbk(nc)) + 1); tw.dt_dev = D(1); tw.part = D(size_t(nbk(nc)));
    if (!energy_part) energy_part = static_cast<double *>(alloc(size_t(nbk(nc)) * 8));
}

// n steps of step_first_order's explicit stage (prepare, CFL, gradients, fluxes, sources, update, UCT EMF with
// the boundary EMF, CT) on the Dual state tw, with the seeds of the free-stream states and of the inflow EMF in Pt
void mhdg_ctx::explicit_tangent(int n_steps, double dt_max, const Params &Pt)
{
    const int nb = nbk(nc);
    SpCT<Dual> Ws{}, none{}, gradsC{};
    SpWT<Dual> grads{}, Us{}, noneW{}, Ww{};
    SpCT<Dual> Uc{};
    Ws.p[0] = tw.W[0]; Ws.p[1] = tw.W[1];
    Ww.p[0] = tw.W[0]; Ww.p[1] = tw.W[1];
    Uc.p[0] = tw.U[0]; Uc.p[1] = tw.U[1];
    grads.p[0] = tw.grad[0]; grads.p[1] = tw.grad[1];
    gradsC.p[0] = tw.grad[0]; gradsC.p[1] = tw.grad[1];
    Us.p[0] = tw.U[0]; Us.p[1] = tw.U[1];
    for (int s = 0; s < n_steps; ++s)
    {
        LAUNCH("prepare_t", k_prepare<Dual>, dim3(nb), dim3(256), 0, stream, d, Pt, Uc, static_cast<const Dual *>(tw.Phi),
               static_cast<const Dual *>(tw.Badd), Ww, tw.Bind, tw.Btot, tw.uplus, tw.alpha, tw.dt_part);
        LAUNCH("dt_final_t", k_dt_final<Dual>, dim3(1), dim3(256), 0, stream, tw.dt_part, nb, dt_max, tw.dt_dev);
        if (!Pt.second_order)
        {
            chk(hipMemsetAsync(tw.grad[0], 0, 15 * size_t(nc) * sizeof(Dual), stream), "memset");
            chk(hipMemsetAsync(tw.grad[1], 0, 15 * size_t(nc) * sizeof(Dual), stream), "memset");
        }
        LAUNCH("grad_current_t", k_grad_current<Dual>, dim3(nb), dim3(256), 0, stream, d, Pt, Pt.second_order ? 2 : 0, 0, Ws, none, grads,
               1, static_cast<const Dual *>(tw.Bind), tw.J);
        LAUNCH("face_t", k_face<Dual>, dim3(nbk(nf)), dim3(256), 0, stream, d, Pt, 2, 0, Ws, gradsC, none, tw.FA,
               static_cast<const Dual *>(tw.uplus), static_cast<const Dual *>(tw.Btot), static_cast<const Dual *>(tw.Bind),
               static_cast<const Dual *>(tw.Phi), static_cast<const Dual *>(tw.PhiAdd), tw.Ef);
        LAUNCH("cell_update_t", k_cell_update<Dual>, dim3(nb), dim3(256), 0, stream, d, Pt, 2, 0, static_cast<const Dual *>(tw.FA), none,
               Ws, static_cast<const Dual *>(tw.Btot),
               static_cast<const Dual *>(tw.J), static_cast<const Dual *>(tw.dt_dev), 0.0, 1, 1, Us, noneW);
        LAUNCH("edge_emf_t", k_edge_emf<Dual>, dim3(nbk(ne)), dim3(256), 0, stream, d, Pt, static_cast<const Dual *>(tw.Ef),
               static_cast<const Dual *>(tw.uplus), static_cast<const Dual *>(tw.Phi), Pt.cuct, kLightMax, static_cast<const Dual *>(nullptr),
               tw.E, 1);
        if (n_heavy)
            LAUNCH("edge_emf_heavy_t", k_edge_emf_heavy<Dual>, dim3(n_heavy), dim3(256), 0, stream, d, Pt, static_cast<const Dual *>(tw.Ef),
                   static_cast<const Dual *>(tw.uplus), static_cast<const Dual *>(tw.Phi), Pt.cuct, static_cast<const int32_t *>(heavy),
                   static_cast<const Dual *>(nullptr), tw.E, 1);
        LAUNCH("faraday_t", k_faraday<Dual>, dim3(nbk(nf)), dim3(256), 0, stream, d, static_cast<const Dual *>(tw.E),
               static_cast<const double *>(nullptr), static_cast<const Dual *>(tw.dt_dev), 0.0, tw.Phi);
        check_launch("explicit tangent step");
    }
}

template <class R> void mhdg_ctx::shell_energy(const R *phi, const R *badd, double r0, double r1, R *out)
{
    const int nb = nbk(nc);
    R *part = nullptr;
    if constexpr (std::is_same<R, Dual>::value)
    {
        if (!tw.part) throw std::runtime_error("shell energy: tangent work arrays not allocated");
        part = tw.part;
    }
    else
    {
        // per-block partial sums of the double evaluation (any caller: the energy entry, the tangent, the adjoint)
        if (!energy_part) energy_part = static_cast<double *>(alloc(size_t(nb) * 8));
        part = energy_part;
    }
    hipLaunchKernelGGL(k_shell_energy<R>, dim3(nb), dim3(256), 0, stream, d, phi, badd, r0, r1, part);
    check_launch("shell energy");
    std::vector<R> h(nb);
    chk(hipMemcpyAsync(h.data(), part, sizeof(R) * nb, hipMemcpyDeviceToHost, stream), "D2H");
    chk(hipStreamSynchronize(stream), "sync");
    R s = 0.0;
    for (int i = 0; i < nb; ++i) s += h[i];
    *out = s;
}
template void mhdg_ctx::shell_energy<double>(const double *, const double *, double, double, double *);
template void mhdg_ctx::shell_energy<Dual>(const Dual *, const Dual *, double, double, Dual *);
