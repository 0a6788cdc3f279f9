#pragma once
// Finite-volume device helpers shared by the explicit-stage kernels and their adjoints: cell vector
// reconstruction, CFL time step, face states, Rusanov flux, ghost states, the UCT edge quantities and the
// boundary EMF.  Templates on the real type R (double, or Dual for forward-mode derivatives).
#include "dev.h"

namespace mhdg
{
namespace
{
constexpr int kGradFast = 4; // cells with up to this many faces keep their face rows in registers (tetrahedra)
// Only the double instantiations do: the Dual (forward-mode) ones would spill the rows to scratch memory
template <class R> constexpr bool kRowsInRegisters = sizeof(R) == sizeof(double);

// src/uns/geometry/Reconstruct.cpp:11-45 for one cell: the face term (c2f entry v, |A|, area vector, flux) and
// the 3x3 solve
template <class R> __device__ __forceinline__ void cell_vector_add(R (&M)[3][3], R (&r)[3], int32_t v, const R &A, const V3 &area, const R &flux)
{
    if (A <= 0) return;
    const V3T<R> n = (psg(v) / A) * area;
    const R nv[3] = {n.x, n.y, n.z};
    const R phi = psg(v) * flux;
    for (int i = 0; i < 3; ++i)
    {
        r[i] += nv[i] * phi;
        for (int j = 0; j < 3; ++j) M[i][j] += A * nv[i] * nv[j];
    }
}
template <class R> __device__ V3T<R> cell_vector_solve(const R (&M)[3][3], const R (&r)[3])
{
    const R det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    const R sc = mx(mx(mx(M[0][0], M[1][1]), M[2][2]), 1e-300);
    if (fabs(det) < 1e-12 * sc * sc * sc) return V3T<R>();
    R I[3][3];
    I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
    I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
    I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
    I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
    I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    return {I[0][0] * r[0] + I[0][1] * r[1] + I[0][2] * r[2], I[1][0] * r[0] + I[1][1] * r[1] + I[1][2] * r[2],
            I[2][0] * r[0] + I[2][1] * r[1] + I[2][2] * r[2]};
}
template <class R> __device__ V3T<R> cell_vector(const DevMesh &d, int32_t c, const R *flux)
{
    R M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, r[3] = {0, 0, 0};
    for (int32_t k = d.c2f_ptr[c]; k < d.c2f_ptr[c + 1]; ++k)
    {
        const int32_t v = d.c2f[k], f = pid(v);
        cell_vector_add(M, r, v, R(d.famag[f]), L3(d.farea, f), flux[f]);
    }
    return cell_vector_solve(M, r);
}

// MercuryUns/fluid/MercurySources.cpp:104-145 for one fluid cell (+inf if no face restricts it)
// (every ion species; the total density of the fast speed summed in species order)
// (the species rows of cell c are read at row `row` of the species arrays Ws: c itself, or 0 for a one-cell copy)
template <class R> __device__ R cell_dt(const DevMesh &d, const Params &P, int32_t c, const R *const *Ws, const V3T<R> &b, size_t row)
{
    R dt = __builtin_inf();
    const R B2 = dot(b, b);
    R rho_sum = Ws[0][5 * row];
    for (int s = 1; s < P.ns; ++s) rho_sum = rho_sum + Ws[s][5 * row];
    const R rho_tot = mx(rho_sum, P.rho_floor_dt);
    const R vA2 = P.inver_MA2 * B2 / rho_tot; // the same for every face and species
    for (int s = 0; s < P.ns; ++s)
    {
        const R *w = Ws[s] + 5 * row;
        const R cs2 = P.gamma * w[4] / mx(w[0], 1e-30);
        R denom = 0;
        for (int32_t k = d.c2f_ptr[c]; k < d.c2f_ptr[c + 1]; ++k)
        {
            const int32_t f = pid(d.c2f[k]);
            const R A = d.famag[f];
            const V3T<R> n = (1.0 / A) * L3(d.farea, f);
            const R vAn2 = P.inver_MA2 * dot(b, n) * dot(b, n) / rho_tot;
            const R term = cs2 + vA2;
            const R cf = sqrt(0.5 * (term + sqrt(mx(0.0, term * term - 4 * cs2 * vAn2))));
            denom += (fabs(w[1] * n.x + w[2] * n.y + w[3] * n.z) + cf) * A;
        }
        if (denom > 0) dt = mn(dt, P.cfl * d.cvol[c] / denom);
    }
    return dt;
}

// cell_dt for a cell whose N or fewer faces are given as (|A|, area vector) rows: the face normal and the normal
// Alfven term, which do not depend on the species, are formed once (same operations as cell_dt)
template <class R, int N>
__device__ R cell_dt_rows(const Params &P, double vol, const R *const *Ws, const V3T<R> &b, size_t row, int nk, const double (&A)[N],
                          const V3 (&area)[N])
{
    R dt = __builtin_inf();
    const R B2 = dot(b, b);
    R rho_sum = Ws[0][5 * row];
    for (int s = 1; s < P.ns; ++s) rho_sum = rho_sum + Ws[s][5 * row];
    const R rho_tot = mx(rho_sum, P.rho_floor_dt);
    const R vA2 = P.inver_MA2 * B2 / rho_tot;
    V3T<R> n[N];
    R vAn2[N];
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (i < nk)
        {
            n[i] = (1.0 / A[i]) * area[i];
            vAn2[i] = P.inver_MA2 * dot(b, n[i]) * dot(b, n[i]) / rho_tot;
        }
    for (int s = 0; s < P.ns; ++s)
    {
        const R *w = Ws[s] + 5 * row;
        const R cs2 = P.gamma * w[4] / mx(w[0], 1e-30);
        R denom = 0;
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (i < nk)
            {
                const R term = cs2 + vA2;
                const R cf = sqrt(0.5 * (term + sqrt(mx(0.0, term * term - 4 * cs2 * vAn2[i]))));
                denom += (fabs(w[1] * n[i].x + w[2] * n[i].y + w[3] * n[i].z) + cf) * A[i];
            }
        if (denom > 0) dt = mn(dt, P.cfl * vol / denom);
    }
    return dt;
}

template <class R> __device__ __forceinline__ R block_min(R v, R *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1)
    {
        if (t < s) sh[t] = mn(sh[t], sh[t + s]);
        __syncthreads();
    }
    return sh[0];
}

// Common/fluid/FluxRHS.cpp:16-27
template <class R> __device__ __forceinline__ PrimT<R> face_state(const R *W, const R *grad, int32_t c, const V3T<R> &dx)
{
    const R *w = W + 5 * size_t(c);
    const R *g = grad + 15 * size_t(c);
    PrimT<R> r;
    r.rho = w[0] + g[0] * dx.x + g[1] * dx.y + g[2] * dx.z;
    r.u = w[1] + g[3] * dx.x + g[4] * dx.y + g[5] * dx.z;
    r.v = w[2] + g[6] * dx.x + g[7] * dx.y + g[8] * dx.z;
    r.w = w[3] + g[9] * dx.x + g[10] * dx.y + g[11] * dx.z;
    r.p = w[4] + g[12] * dx.x + g[13] * dx.y + g[14] * dx.z;
    if (r.rho > 0 && r.p > 0) return r;
    return PrimT<R>{w[0], w[1], w[2], w[3], w[4]};
}

// Common/physics/Euler.h:29-49
template <class R> __device__ void rusanov(const PrimT<R> &L, const PrimT<R> &Rt, const R n[3], double gamma, R *F)
{
    R UL[5], UR[5], FL[5], FR[5];
    to_cons(L, gamma, UL);
    to_cons(Rt, gamma, UR);
    const R unL = L.u * n[0] + L.v * n[1] + L.w * n[2];
    FL[0] = L.rho * unL; FL[1] = L.rho * unL * L.u + L.p * n[0]; FL[2] = L.rho * unL * L.v + L.p * n[1];
    FL[3] = L.rho * unL * L.w + L.p * n[2]; FL[4] = unL * (UL[4] + L.p);
    const R unR = Rt.u * n[0] + Rt.v * n[1] + Rt.w * n[2];
    FR[0] = Rt.rho * unR; FR[1] = Rt.rho * unR * Rt.u + Rt.p * n[0]; FR[2] = Rt.rho * unR * Rt.v + Rt.p * n[1];
    FR[3] = Rt.rho * unR * Rt.w + Rt.p * n[2]; FR[4] = unR * (UR[4] + Rt.p);
    const R sL = fabs(L.u * n[0] + L.v * n[1] + L.w * n[2]) + sqrt(gamma * L.p / L.rho);
    const R sR = fabs(Rt.u * n[0] + Rt.v * n[1] + Rt.w * n[2]) + sqrt(gamma * Rt.p / Rt.rho);
    const R s = mx(sL, sR);
    for (int m = 0; m < 5; ++m) F[m] = 0.5 * (FL[m] + FR[m]) - 0.5 * s * (UR[m] - UL[m]);
}

// ghost_state (Common/fluid/Reconstruction.cpp:9-24)
template <class R> __device__ __forceinline__ PrimT<R> ghost_state(int8_t kind, int species, const PrimT<R> &in, const V3T<R> &n, const R *fs)
{
    if (kind == kWall)
    {
        if (species == 1) return PrimT<R>{1e-12, 0, 0, 0, 1e-12};
        const R un = in.u * n.x + in.v * n.y + in.w * n.z;
        PrimT<R> g = in;
        g.u -= 2 * un * n.x; g.v -= 2 * un * n.y; g.w -= 2 * un * n.z;
        return g;
    }
    if (kind == kInflow) return PrimT<R>{fs[0], fs[1], fs[2], fs[3], fs[4]};
    return in;
}

// The per-face quantities of the UCT edge sum (IdealEMF.cpp:79-88)
__device__ __forceinline__ void face_dirs(const DevMesh &d, int32_t f, int32_t e, const V3 &th, V3 &m, V3 &nrm)
{
    V3 v = L3(d.fcen, f) - L3(d.emid, e);
    v -= dot(v, th) * th;
    const double nv = norm(v);
    m = nv > 0 ? (1.0 / nv) * v : V3();
    nrm = (1.0 / d.famag[f]) * L3(d.farea, f);
}
template <class R> __device__ __forceinline__ R edge_amax(const DevMesh &d, const R *u_cell, int32_t c0, int32_t c1, const V3T<R> &m)
{
    R a = 0.0;
    for (int32_t k = c0; k < c1; ++k) a = mx(a, fabs(dot(L3(u_cell, d.ec_cell[k]), m)));
    return a;
}
__device__ __forceinline__ void plane_basis(const V3 &t, V3 &e1, V3 &e2)
{
    const V3 a = fabs(t.x) < 0.9 ? V3(1, 0, 0) : V3(0, 1, 0);
    e1 = cross(t, a);
    e1 *= 1.0 / norm(e1);
    e2 = cross(t, e1);
}
// eps-regularised 2x2 solve of the omega correction (IdealEMF.cpp:110-118)
template <class R> __device__ __forceinline__ void uct_solve(R M11, R M12, R M22, R g1, R g2, R &l1, R &l2)
{
    const R tr = M11 + M22;
    const R eps = 1e-14 * tr;
    const R det = (M11 + eps) * (M22 + eps) - M12 * M12;
    l1 = 0;
    l2 = 0;
    if (tr > 0 && det > 0)
    {
        l1 = ((M22 + eps) * g1 - M12 * g2) / det;
        l2 = ((M11 + eps) * g2 - M12 * g1) / det;
    }
}
constexpr int kEmfFast = 6; // valence served from registers by k_edge_emf
// boundary EMF of Solver::rhs (MercuryUns/driver/Solver.cpp:162-167)
template <class R> __device__ __forceinline__ R edge_bc(const DevMesh &d, const Params &P, int32_t e, R v)
{
    const int8_t mode = d.emode[e];
    if (mode == 1) return 0.0;
    if (mode == 2)
        return dot(V3T<R>(seeded<R>(P.E_inf[0], P.dE_inf[0]), seeded<R>(P.E_inf[1], P.dE_inf[1]), seeded<R>(P.E_inf[2], P.dE_inf[2])),
                   L3(d.evec, e));
    return v;
}
} // namespace
} // namespace mhdg
