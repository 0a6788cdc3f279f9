// MarsSolver's point-implicit sources (MarsUns/driver/MarsSolver.cpp:422-527) on the device, one thread per fluid
// cell over every ion species:
//   1. gyration  dz_s/dt = -i Omega_s (z_s - Z), z = u_x + i u_y in the plane normal to B, Z = sum chi_s z_s:
//      Crank-Nicolson in closed form for the coupled species (Z' = P / (1 - Q)); u_par unchanged; the kinetic
//      energy change goes to E
//   2. ion-neutral drag  dm/dt = -nu m, backward Euler; the lost kinetic energy stays in E as heat
//   3. thermal relaxation and frictional heating  dp/dt = (gamma-1) [k (185 - T) + S_f], backward Euler in p
// Same operations in the same order as the reference; complex arithmetic spelled out on (re, im) pairs, the
// division by Smith's method.
#include <cmath>
#include <stdexcept>

#include "ctx.h"
#include "kernels.h"

namespace mhdg
{
namespace
{
struct Cx
{
    double re, im;
};
__device__ __forceinline__ Cx operator+(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx operator-(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cx operator*(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ Cx operator*(double s, Cx a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ Cx cdiv(Cx a, Cx b)
{
    if (fabs(b.re) >= fabs(b.im))
    {
        const double r = b.im / b.re, den = b.re + b.im * r;
        return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
    }
    const double r = b.re / b.im, den = b.re * r + b.im;
    return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}
__device__ __forceinline__ Cx one_minus(Cx a) { return {1.0 - a.re, -a.im}; }
__device__ __forceinline__ Cx one_plus(Cx a) { return {1.0 + a.re, a.im}; }

struct MarsDev
{
    double a5, qBL, kB, U, n, rho_ref, p_ref, time_ref, rho_floor, p_floor, gamma;
    double ion_mass[kMaxSpecies], neutral_mass[kMaxSpecies];
    int32_t ns, nn, stride;
};
} // namespace

__global__ void __launch_bounds__(256) k_mars_implicit(DevMesh d, MarsDev M, const double *Btot, const double *neutral, double dt, SpW Us)
{
    const int32_t c = gid();
    if (c >= d.nc || d.mat[c] != 0) return;
    const size_t i = size_t(c);
    const double g = M.gamma;
    double nrel[kMaxSpecies], ntot = 0.0;
#pragma unroll
    for (int s = 0; s < kMaxSpecies; ++s)
        if (s < M.ns)
        {
            nrel[s] = mx(Us.p[s][5 * i], 0.0) * M.rho_ref / (M.ion_mass[s] * M.n);
            ntot += nrel[s];
        }
    // ---- 1. gyration
    const V3 B = L3(Btot, c);
    const double Bm = norm(B);
    if (ntot > 1e-300 && Bm > 1e-30)
    {
        const V3 k = (1.0 / Bm) * B;
        const V3 t = fabs(k.x) < 0.9 ? V3(1, 0, 0) : V3(0, 1, 0);
        V3 e1 = cross(k, t);
        e1 = (1.0 / norm(e1)) * e1;
        const V3 e2 = cross(k, e1);
        Cx z[kMaxSpecies], a[kMaxSpecies], Z0{0.0, 0.0};
#pragma unroll
        for (int s = 0; s < kMaxSpecies; ++s)
            if (s < M.ns)
            {
                const double *u = Us.p[s] + 5 * i;
                const double rho = mx(u[0], M.rho_floor);
                const V3 v(u[1] / rho, u[2] / rho, u[3] / rho);
                z[s] = {dot(v, e1), dot(v, e2)};
                // the reference holds al = +i alpha / 2 and uses a = -al
                const double al = M.qBL / (M.ion_mass[s] * M.U) * Bm * dt * 0.5;
                a[s] = {-0.0, -al};
                Z0 = Z0 + (nrel[s] / ntot) * z[s];
            }
        Cx Pn{0.0, 0.0}, Qn{0.0, 0.0};
#pragma unroll
        for (int s = 0; s < kMaxSpecies; ++s)
            if (s < M.ns)
            {
                const double chi = nrel[s] / ntot;
                Pn = Pn + cdiv(chi * (z[s] * one_minus(a[s]) + a[s] * Z0), one_plus(a[s]));
                Qn = Qn + cdiv(chi * a[s], one_plus(a[s]));
            }
        const Cx Zn = cdiv(Pn, one_minus(Qn));
#pragma unroll
        for (int s = 0; s < kMaxSpecies; ++s)
            if (s < M.ns)
            {
                const Cx zn = cdiv(z[s] * one_minus(a[s]) + a[s] * (Z0 + Zn), one_plus(a[s]));
                double *u = Us.p[s] + 5 * i;
                const double rho = mx(u[0], M.rho_floor);
                const V3 v(u[1] / rho, u[2] / rho, u[3] / rho);
                const double vpar = dot(v, k);
                const V3 vn = vpar * k + zn.re * e1 + zn.im * e2;
                u[4] += 0.5 * rho * (dot(vn, vn) - dot(v, v));
                u[1] = rho * vn.x;
                u[2] = rho * vn.y;
                u[3] = rho * vn.z;
            }
    }
    // ---- 2./3. drag and thermal relaxation
    double coll_n[kMaxSpecies], collision = 0.0;
#pragma unroll
    for (int n = 0; n < kMaxSpecies; ++n)
        if (n < M.nn)
        {
            coll_n[n] = mx(neutral[i * size_t(M.stride) + size_t(n)], 0.0) * 2.0e-10;
            collision += coll_n[n];
        }
    const double nu = M.time_ref * collision;
#pragma unroll
    for (int s = 0; s < kMaxSpecies; ++s)
        if (s < M.ns)
        {
            double *u = Us.p[s] + 5 * i;
            const double m = M.ion_mass[s];
            const double rho = mx(u[0], M.rho_floor);
            const V3 v(u[1] / rho, u[2] / rho, u[3] / rho);
            const double ke_old = 0.5 * rho * dot(v, v);
            const double den = 1.0 + nu * dt;
            u[1] /= den;
            u[2] /= den;
            u[3] /= den;
            const double ke = ke_old / (den * den);
            const double p = mx((g - 1.0) * (u[4] - ke), M.p_floor);
            double b1c = 0.0, b2 = 0.0;
#pragma unroll
            for (int n = 0; n < kMaxSpecies; ++n)
                if (n < M.nn)
                {
                    const double mm = m + M.neutral_mass[n];
                    b1c += m * coll_n[n] / mm;
                    b2 += M.neutral_mass[n] * coll_n[n] / mm;
                }
            const double n_m3 = rho * M.rho_ref / m;
            const double ns_cm3 = n_m3 * 1e-6;
            const double kk = M.a5 * ns_cm3 * b1c;
            const double cT = M.p_ref / (mx(n_m3, 1.0) * M.kB);
            const double Sf = M.time_ref * rho * dot(v, v) * b2;
            double pn;
            if (p * cT >= 50.0) pn = (p + dt * (g - 1.0) * (kk * 185.0 + Sf)) / (1.0 + dt * (g - 1.0) * kk * cT);
            else pn = p + dt * (g - 1.0) * (kk * (185.0 - 50.0) + Sf);
            u[4] = pn / (g - 1.0) + ke;
        }
}
} // namespace mhdg

using namespace mhdg;

extern "C" int mhdg_mars_implicit_sources(mhdg_ctx *c, const mhdg_mars_params *prm, const double *neutral, double dt)
{
    try
    {
        if (!prm || (prm->n_neutral > 0 && !neutral)) throw std::runtime_error("mars implicit sources: null argument");
        if (prm->n_neutral < 0 || prm->n_neutral > kMaxSpecies || prm->neutral_stride < prm->n_neutral)
            throw std::runtime_error("mars implicit sources: n_neutral must be 0..10 and neutral_stride >= n_neutral");
        if (!(dt >= 0.0) || !std::isfinite(dt)) throw std::runtime_error("mars implicit sources: invalid dt");
        MarsDev M{};
        M.a5 = 3.0e6 * prm->L * prm->kB / (prm->rho_ref * prm->U * prm->U * prm->U);
        M.qBL = prm->q * prm->B * prm->L;
        M.kB = prm->kB; M.U = prm->U; M.n = prm->n;
        M.rho_ref = prm->rho_ref; M.p_ref = prm->p_ref; M.time_ref = prm->time_ref;
        M.rho_floor = prm->rho_floor; M.p_floor = prm->p_floor;
        M.gamma = c->P.gamma;
        M.ns = c->P.ns;
        M.nn = prm->n_neutral;
        M.stride = prm->neutral_stride;
        for (int s = 0; s < kMaxSpecies; ++s)
        {
            M.ion_mass[s] = prm->ion_mass[s];
            M.neutral_mass[s] = prm->neutral_mass[s];
            if (s < M.ns && !(prm->ion_mass[s] > 0.0)) throw std::runtime_error("mars implicit sources: ion masses must be positive");
        }
        const double *nd = M.nn > 0 ? c->stage_in(0, neutral, size_t(c->nc) * size_t(M.stride)) : nullptr;
        SpW Us{};
        for (int s = 0; s < M.ns; ++s) Us.p[s] = c->U[s];
        hipLaunchKernelGGL(k_mars_implicit, dim3((c->nc + 255) / 256), dim3(256), 0, c->stream, c->d, M,
                           static_cast<const double *>(c->Btot), nd, dt, Us);
        c->check_launch("mars implicit sources");
        return 0;
    }
    catch (const std::exception &e)
    {
        c->err = e.what();
        return 1;
    }
}
