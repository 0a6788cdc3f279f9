// Reverse mode (adjoint) of the explicit stage of the Mercury step (MercuryUns/driver/Solver.cpp:228-277 with
// the Hall solve inactive): the gradient of the shell magnetic energy after n steps with respect to the
// initial state, the background flux and field, the free-stream states and the imposed inflow EMF.
//
// The forward sweep keeps the state (U, Phi) of every step; the reverse sweep recomputes a step's
// intermediates from it (primitives, fields, limited gradients, face fluxes, face and edge EMFs) and runs the
// transposed kernels in reverse order.  Every adjoint kernel works owner-side: an entity adds the parts of
// its own inputs and writes the parts of its neighbours' inputs into slots (one per c2f entry, e2f entry or
// edge-cell entry) that the neighbours then gather in a fixed order, so no atomics and a deterministic
// result.  Local Jacobians of the nonlinear pieces (primitive recovery, floors, limited gradients, Rusanov,
// Lorentz sources, u+, the CFL step, the UCT 2x2 solve) come from the forward-mode Dual instantiation of
// the same device functions, one seeded input at a time; branches (limiters, upwinding, floors, argmax,
// the CFL minimum) follow the values, as in the tangent.
#include <type_traits>

#include "ctx.h"
#include "fv.h"
#include "kernels.h"

namespace mhdg
{
namespace
{
constexpr int kAdjNs = 2;   // species of the adjoint (the two Mercury ions)
constexpr int kRedMax = 16; // components of one block reduction

// per-block sums of K values (fixed tree order); part[blockIdx.x * K + j]
template <int K> __device__ void block_sums(const double (&v)[K], double *part)
{
    __shared__ double sh[256];
    for (int j = 0; j < K; ++j)
    {
        sh[threadIdx.x] = v[j];
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1)
        {
            if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[blockIdx.x * K + j] = sh[0];
        __syncthreads();
    }
}

// I = M^-1 of reconstruct_cell_vector (Reconstruct.cpp:11-45) in the same operation order; false if singular
__device__ bool cell_vector_inverse(const DevMesh &d, int32_t c, double I[3][3])
{
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int32_t k = d.c2f_ptr[c]; k < d.c2f_ptr[c + 1]; ++k)
    {
        const int32_t v = d.c2f[k], f = pid(v);
        const double A = d.famag[f];
        if (A <= 0) continue;
        const V3 n = (psg(v) / A) * L3(d.farea, f);
        const double nv[3] = {n.x, n.y, n.z};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M[i][j] += A * nv[i] * nv[j];
    }
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    const double sc = mx(mx(mx(M[0][0], M[1][1]), M[2][2]), 1e-300);
    if (fabs(det) < 1e-12 * sc * sc * sc) return false;
    I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
    I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
    I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
    I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
    I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    return true;
}

// transpose of b_c = cell_vector(flux): slot k of cell c receives d(bbar . b_c)/d flux_f
__device__ void cell_vector_adjoint(const DevMesh &d, int32_t c, const V3 &bbar, double *slot)
{
    double I[3][3];
    const bool ok = cell_vector_inverse(d, c, I);
    const V3 y = ok ? V3(I[0][0] * bbar.x + I[1][0] * bbar.y + I[2][0] * bbar.z, I[0][1] * bbar.x + I[1][1] * bbar.y + I[2][1] * bbar.z,
                         I[0][2] * bbar.x + I[1][2] * bbar.y + I[2][2] * bbar.z)
                    : V3();
    for (int32_t k = d.c2f_ptr[c]; k < d.c2f_ptr[c + 1]; ++k)
    {
        const int32_t v = d.c2f[k], f = pid(v);
        const double A = d.famag[f];
        double g = 0.0;
        if (ok && A > 0)
        {
            const V3 n = (psg(v) / A) * L3(d.farea, f);
            g = psg(v) * dot(n, y);
        }
        slot[k] = g;
    }
}

// Rusanov flux transposed: Lb, Rb = (dF/dL)^T Fb, (dF/dR)^T Fb (ten forward-mode columns)
__device__ void rusanov_adjoint(const PrimT<double> &L, const PrimT<double> &Rt, const V3 &n, double gamma, const double Fb[5],
                                double Lb[5], double Rb[5])
{
    const Dual nd[3] = {Dual(n.x), Dual(n.y), Dual(n.z)};
    for (int j = 0; j < 10; ++j)
    {
        PrimT<Dual> a{L.rho, L.u, L.v, L.w, L.p}, b{Rt.rho, Rt.u, Rt.v, Rt.w, Rt.p};
        PrimT<Dual> &t = j < 5 ? a : b;
        Dual *q[5] = {&t.rho, &t.u, &t.v, &t.w, &t.p};
        q[j % 5]->d = 1.0;
        Dual F[5];
        rusanov(a, b, nd, gamma, F);
        double g = 0;
        for (int m = 0; m < 5; ++m) g += Fb[m] * F[m].d;
        (j < 5 ? Lb : Rb)[j % 5] = g;
    }
}

// face_state transposed (FluxRHS.cpp:16-27): the reconstructed branch is linear in (w, grad), the fallback
// takes the cell value
__device__ void face_state_adjoint(const double *W, const double *G, int32_t c, const V3 &dx, const double rb[5], double out[20])
{
    const double *w = W + 5 * size_t(c);
    const double *g = G + 15 * size_t(c);
    const double rr = w[0] + g[0] * dx.x + g[1] * dx.y + g[2] * dx.z, pp = w[4] + g[12] * dx.x + g[13] * dx.y + g[14] * dx.z;
    const bool recon = rr > 0 && pp > 0;
    for (int m = 0; m < 5; ++m)
    {
        out[m] = rb[m];
        out[5 + 3 * m] = recon ? rb[m] * dx.x : 0.0;
        out[5 + 3 * m + 1] = recon ? rb[m] * dx.y : 0.0;
        out[5 + 3 * m + 2] = recon ? rb[m] * dx.z : 0.0;
    }
}

// one column of the limited gradient of species sp at cell c (Reconstruction.cpp:80-142 as in k_grad_current):
// the input seeded is W[seed_cell][seed_m], or fs[seed_fs] of the inflow ghost
__device__ void grad_column(const DevMesh &d, const Params &P, int32_t c, int32_t sp, const double *W, int32_t seed_cell, uint32_t seed_m,
                            uint32_t seed_fs, Dual out[15])
{
    // seed_m, seed_fs: bit masks of the primitives seeded (of seed_cell's W, of the free-stream state)
    const int32_t k0 = d.c2f_ptr[c], k1 = d.c2f_ptr[c + 1];
    Dual fs[5];
    for (int m = 0; m < 5; ++m) fs[m] = Dual(P.fs[sp][m], (seed_fs >> m) & 1u ? 1.0 : 0.0);
    const V3 xc = L3(d.ccen, c);
    Dual wc[5], r[5][3], lo[5], hi[5];
    for (int m = 0; m < 5; ++m)
    {
        wc[m] = Dual(W[5 * size_t(c) + m], (c == seed_cell && ((seed_m >> m) & 1u)) ? 1.0 : 0.0);
        lo[m] = hi[m] = wc[m];
        r[m][0] = r[m][1] = r[m][2] = 0.0;
    }
    for (int32_t k = k0; k < k1; ++k)
    {
        const int32_t nb = d.snb[k];
        if (nb == kNoStencil) continue;
        Dual wn[5];
        if (nb >= 0)
            for (int m = 0; m < 5; ++m) wn[m] = Dual(W[5 * size_t(nb) + m], (nb == seed_cell && ((seed_m >> m) & 1u)) ? 1.0 : 0.0);
        else
        {
            const int32_t f = -(nb + 1);
            const int8_t kind = d.kind[f];
            for (int m = 0; m < 5; ++m) wn[m] = wc[m];
            if (kind == kWall)
            {
                if (sp == 1) { wn[0] = 1e-12; wn[1] = 0.0; wn[2] = 0.0; wn[3] = 0.0; wn[4] = 1e-12; }
                else
                {
                    const V3 n = (psg(d.c2f[k]) / d.famag[f]) * L3(d.farea, f);
                    const Dual un = wc[1] * n.x + wc[2] * n.y + wc[3] * n.z;
                    wn[1] -= 2.0 * un * n.x; wn[2] -= 2.0 * un * n.y; wn[3] -= 2.0 * un * n.z;
                }
            }
            else if (kind == kInflow)
                for (int m = 0; m < 5; ++m) wn[m] = fs[m];
        }
        const double w = d.sw[k];
        const V3 dd = L3(d.sd, k);
        for (int m = 0; m < 5; ++m)
        {
            const Dual dq = wn[m] - wc[m];
            r[m][0] += w * dd.x * dq; r[m][1] += w * dd.y * dq; r[m][2] += w * dd.z * dq;
            lo[m] = mn(lo[m], wn[m]);
            hi[m] = mx(hi[m], wn[m]);
        }
    }
    const double *I = d.sinv + 6 * size_t(c);
    const double Im[9] = {I[0], I[1], I[2], I[1], I[3], I[4], I[2], I[4], I[5]};
    for (int m = 0; m < 5; ++m)
    {
        const Dual g[3] = {Im[0] * r[m][0] + Im[1] * r[m][1] + Im[2] * r[m][2], Im[3] * r[m][0] + Im[4] * r[m][1] + Im[5] * r[m][2],
                           Im[6] * r[m][0] + Im[7] * r[m][1] + Im[8] * r[m][2]};
        Dual psi = 1.0;
        for (int32_t k = k0; k < k1; ++k)
        {
            const V3 dx = L3(d.fcen, pid(d.c2f[k])) - xc;
            const Dual dl = g[0] * dx.x + g[1] * dx.y + g[2] * dx.z;
            if (dl > 1e-300) psi = mn(psi, (hi[m] - wc[m]) / dl);
            else if (dl < -1e-300) psi = mn(psi, (lo[m] - wc[m]) / dl);
        }
        if (P.rel_limit > 0 && (m == 0 || m == 4))
            for (int32_t k = k0; k < k1; ++k)
            {
                const V3 dx = L3(d.fcen, pid(d.c2f[k])) - xc;
                const Dual dl = fabs(g[0] * dx.x + g[1] * dx.y + g[2] * dx.z);
                if (dl > 1e-300) psi = mn(psi, P.rel_limit * wc[m] / dl);
            }
        psi = mx(Dual(0.0), mn(Dual(1.0), psi));
        out[3 * m] = psi * g[0]; out[3 * m + 1] = psi * g[1]; out[3 * m + 2] = psi * g[2];
    }
}

// Lorentz source of species s (MercurySources.cpp:20-63 as in k_cell_update) from local values
template <class R> __device__ void lorentz_source(const Params &P, const R (&w)[kAdjNs][5], const V3T<R> &J, const V3T<R> &B, const R &dt, int s,
                                                  R out[4])
{
    R ne = ion_density(P, 0, w[0][0]);
    for (int k = 1; k < kAdjNs; ++k) ne = ne + ion_density(P, k, w[k][0]);
    const R ws = ion_density(P, s, w[s][0]) / sqrt(ne * ne + P.ne_floor * P.ne_floor);
    const R rho = mx(w[s][0], R(0.0));
    const V3T<R> u(w[s][1], w[s][2], w[s][3]);
    V3T<R> S;
    if (P.lorentz) S += (P.inver_MA2 * ws) * cross(J, B);
    out[0] = S.x; out[1] = S.y; out[2] = S.z;
    out[3] = dot(u, S) + (rho > 1e-30 ? 0.5 * dt * dot(S, S) / rho : R(0.0));
}

// forward Euler floors of the update (Solver.cpp:171-191 as in k_cell_update)
template <class R> __device__ void floors(const Params &P, R q[5])
{
    const R ke = q[0] > 0 ? 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0] : R(0.0);
    if (!(q[0] >= P.rho_floor)) to_cons(PrimT<R>{P.rho_floor, 0, 0, 0, P.p_floor}, P.gamma, q);
    else if ((P.gamma - 1.0) * (q[4] - ke) < P.p_floor)
    {
        const PrimT<R> w = to_prim(q, P.gamma, P.rho_floor, P.p_floor);
        to_cons(w, P.gamma, q);
    }
}

// u+ = sum chi_s u_s (Solver.cpp:110-118 as in k_prepare)
template <class R> __device__ V3T<R> ion_velocity(const Params &P, const R (&w)[kAdjNs][5])
{
    R ne = ion_density(P, 0, w[0][0]);
    for (int s = 1; s < kAdjNs; ++s) ne = ne + ion_density(P, s, w[s][0]);
    V3T<R> up;
    for (int s = 0; s < kAdjNs; ++s)
    {
        const R chi = ne > 1e-300 ? ion_density(P, s, w[s][0]) / ne : (s == 0 ? R(1.0) : R(0.0));
        if (s == 0) up = V3T<R>(chi * w[s][1], chi * w[s][2], chi * w[s][3]);
        else up = V3T<R>(up.x + chi * w[s][1], up.y + chi * w[s][2], up.z + chi * w[s][3]);
    }
    return up;
}
} // namespace

// ------------------------------------------------------------------------------------------ kernels
// Faraday transposed: Eb_e = -dt sum over the active faces f of e of sgn(e in f) Phib_f
__global__ void __launch_bounds__(256) k_adj_faraday(DevMesh d, const double *Phib, double dt, double *Eb)
{
    const int32_t e = gid();
    if (e >= d.ne) return;
    double acc = 0.0;
    for (int32_t i = d.e2f_ptr[e]; i < d.e2f_ptr[e + 1]; ++i)
    {
        const int32_t v = d.e2f[i], f = pid(v);
        if (d.fL[f] >= 0 || d.fR[f] >= 0) acc += psg(v) * Phib[f];
    }
    Eb[e] = -dt * acc;
}

// d/d dt of Faraday: per-block sums of -Phib_f (d1 E)_f
__global__ void __launch_bounds__(256) k_adj_faraday_dt(DevMesh d, const double *E, const double *Phib, double *part)
{
    const int32_t f = gid();
    double v[1] = {0.0};
    if (f < d.nf && (d.fL[f] >= 0 || d.fR[f] >= 0))
    {
        double circ = 0.0;
        for (int32_t k = d.f2e_ptr[f]; k < d.f2e_ptr[f + 1]; ++k) circ += psg(d.f2e[k]) * E[pid(d.f2e[k])];
        v[0] = -Phib[f] * circ;
    }
    block_sums<1>(v, part);
}

// UCT edge EMF with the boundary EMF, transposed (IdealEMF.cpp:59-128, Solver.cpp:162-167): per edge, the parts
// of its faces' E_f and induced flux (eF: 4 per e2f entry) and of its cells' u+ (eC: 3 per edge-cell entry);
// per-block sums of the imposed inflow EMF's adjoint (3)
__global__ void __launch_bounds__(256) k_adj_edge_emf(DevMesh d, Params P, const double *u_cell, const double *phi_ind, double cuct,
                                                      const double *Eb, double *eF, double *eC, double *part)
{
    const int32_t e = gid();
    double einf[3] = {0, 0, 0};
    if (e < d.ne)
    {
        const int32_t f0 = d.e2f_ptr[e], f1 = d.e2f_ptr[e + 1], nf = f1 - f0;
        const int32_t c0 = d.ec_ptr[e], c1 = d.ec_ptr[e + 1];
        for (int32_t i = f0; i < f1; ++i)
            for (int q = 0; q < 4; ++q) eF[4 * size_t(i) + q] = 0.0;
        for (int32_t k = c0; k < c1; ++k)
            for (int q = 0; q < 3; ++q) eC[3 * size_t(k) + q] = 0.0;
        double Ebar = Eb[e];
        const int8_t mode = d.emode[e];
        if (mode == 1) Ebar = 0.0;
        else if (mode == 2)
        {
            const V3 t = L3(d.evec, e);
            einf[0] = Ebar * t.x; einf[1] = Ebar * t.y; einf[2] = Ebar * t.z;
            Ebar = 0.0;
        }
        bool any_fluid = false;
        for (int32_t k = c0; k < c1 && !any_fluid; ++k) any_fluid = d.mat[d.ec_cell[k]] == 0;
        // high-valence edges (the pole axes) take their UCT part in k_adj_edge_emf_heavy
        if (Ebar != 0.0 && nf > 0 && nf <= kLightMax && d.elen[e] > 0 && any_fluid)
        {
            const V3 t = L3(d.evec, e);
            const double L = d.elen[e];
            const V3 th = (1.0 / L) * t;
            V3 e1, e2;
            plane_basis(th, e1, e2);
            // face quantities: aef = max_k |u_k . m| (first strict maximum), sig, q = aef sig (nrm.e1, nrm.e2)
            auto face = [&](int32_t i, V3 &m, V3 &nrm, double &aef, int32_t &kmax, double &sig, double &qa, double &qb) {
                const int32_t f = pid(d.e2f[i]);
                face_dirs(d, f, e, th, m, nrm);
                aef = 0.0;
                kmax = -1;
                for (int32_t k = c0; k < c1; ++k)
                {
                    const double a = fabs(dot(L3(u_cell, d.ec_cell[k]), m));
                    if (aef < a) { aef = a; kmax = k; }
                }
                sig = dot(th, cross(m, nrm));
                const V3 qv = (aef * sig) * nrm;
                qa = dot(qv, e1);
                qb = dot(qv, e2);
            };
            double M11 = 0, M12 = 0, M22 = 0, g1 = 0, g2 = 0;
            for (int32_t i = f0; i < f1; ++i)
            {
                V3 m, nrm;
                double aef, sig, qa, qb;
                int32_t km;
                face(i, m, nrm, aef, km, sig, qa, qb);
                const double w = d.efw[i];
                M11 += w * qa * qa; M12 += w * qa * qb; M22 += w * qb * qb; g1 += w * qa; g2 += w * qb;
            }
            double l1, l2;
            uct_solve(M11, M12, M22, g1, g2, l1, l2);
            const double Ec_b = Ebar, Eu_b = 2.0 * cuct * L * Ebar;
            // lambda adjoint from the omega weights
            double l1b = 0, l2b = 0;
            for (int32_t i = f0; i < f1; ++i)
            {
                V3 m, nrm;
                double aef, sig, qa, qb;
                int32_t km;
                face(i, m, nrm, aef, km, sig, qa, qb);
                const int32_t f = pid(d.e2f[i]);
                const double w = d.efw[i], pf = phi_ind[f] / d.famag[f];
                const double omb = Eu_b * aef * sig * pf;
                l1b -= omb * w * qa;
                l2b -= omb * w * qb;
            }
            // (l1, l2) = solve(M11, M12, M22, g1, g2): transposed Jacobian by five forward columns
            double sb[5];
            for (int j = 0; j < 5; ++j)
            {
                Dual in[5] = {Dual(M11), Dual(M12), Dual(M22), Dual(g1), Dual(g2)};
                in[j].d = 1.0;
                Dual a, b;
                uct_solve(in[0], in[1], in[2], in[3], in[4], a, b);
                sb[j] = l1b * a.d + l2b * b.d;
            }
            for (int32_t i = f0; i < f1; ++i)
            {
                V3 m, nrm;
                double aef, sig, qa, qb;
                int32_t km;
                face(i, m, nrm, aef, km, sig, qa, qb);
                const int32_t f = pid(d.e2f[i]);
                const double w = d.efw[i], A = d.famag[f], pf = phi_ind[f] / A;
                const double omega = w * (1.0 - (qa * l1 + qb * l2));
                const double omb = Eu_b * aef * sig * pf;
                double aefb = Eu_b * omega * sig * pf;
                const double pfb = Eu_b * omega * aef * sig;
                const double qab = -omb * w * l1 + 2.0 * w * qa * sb[0] + w * qb * sb[1] + w * sb[3];
                const double qbb = -omb * w * l2 + w * qa * sb[1] + 2.0 * w * qb * sb[2] + w * sb[4];
                const V3 sn = sig * nrm;
                aefb += qab * dot(sn, e1) + qbb * dot(sn, e2);
                eF[4 * size_t(i)] = Ec_b * w * t.x;
                eF[4 * size_t(i) + 1] = Ec_b * w * t.y;
                eF[4 * size_t(i) + 2] = Ec_b * w * t.z;
                eF[4 * size_t(i) + 3] = pfb / A;
                if (km >= 0)
                {
                    const double sg = dot(L3(u_cell, d.ec_cell[km]), m) < 0 ? -1.0 : 1.0;
                    eC[3 * size_t(km)] += aefb * sg * m.x;
                    eC[3 * size_t(km) + 1] += aefb * sg * m.y;
                    eC[3 * size_t(km) + 2] += aefb * sg * m.z;
                }
            }
        }
    }
    block_sums<3>(einf, part);
}

// The UCT part of k_adj_edge_emf for the high-valence edges (valence > kLightMax: the pole axes, n_lon faces and cells),
// one block per edge: the edge's cell velocities staged in LDS, the per-face terms in parallel (each face's max over
// the cells read from LDS), the sums, the 2x2 solve and its adjoint by one thread in face order, and the eC
// accumulation by one thread in face order, as the light kernel does them (which zeroed the edge's eF and eC slots).
__global__ void __launch_bounds__(256) k_adj_edge_emf_heavy(DevMesh d, const double *u_cell, const double *phi_ind, double cuct,
                                                            const double *Eb, const int32_t *heavy, double *eF, double *eC)
{
    constexpr int kH = kHeavyMax;
    constexpr int kPer = kH / 256; // faces per thread
    __shared__ double s_a[kH], s_b[kH], s_c[kH];
    __shared__ int32_t s_k[kH];
    __shared__ double s_r[8];
    const int32_t e = heavy[blockIdx.x];
    const int32_t f0 = d.e2f_ptr[e], nf = d.e2f_ptr[e + 1] - f0;
    const int32_t c0 = d.ec_ptr[e], nc = d.ec_ptr[e + 1] - c0;
    const double Ebar = Eb[e];
    bool any_fluid = false;
    for (int32_t k = 0; k < nc && !any_fluid; ++k) any_fluid = d.mat[d.ec_cell[c0 + k]] == 0;
    // the same condition for every thread of the block: mode 1 and 2 edges carry no UCT part
    if (d.emode[e] != 0 || !(Ebar != 0.0 && nf > kLightMax && nf <= kH && nc <= kH && d.elen[e] > 0 && any_fluid)) return;
    const V3 t = L3(d.evec, e);
    const double L = d.elen[e];
    const V3 th = (1.0 / L) * t;
    V3 e1, e2;
    plane_basis(th, e1, e2);
    for (int32_t k = threadIdx.x; k < nc; k += blockDim.x)
    {
        const V3 u = L3(u_cell, d.ec_cell[c0 + k]);
        s_a[k] = u.x; s_b[k] = u.y; s_c[k] = u.z;
    }
    __syncthreads();
    V3 m_r[kPer], n_r[kPer];
    double aef_r[kPer], sig_r[kPer], qa_r[kPer], qb_r[kPer], sg_r[kPer];
    int32_t km_r[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j)
    {
        const int32_t i = int32_t(threadIdx.x) + j * int32_t(blockDim.x);
        km_r[j] = -1;
        if (i >= nf) continue;
        const int32_t f = pid(d.e2f[f0 + i]);
        face_dirs(d, f, e, th, m_r[j], n_r[j]);
        double aef = 0.0, sg = 1.0;
        for (int32_t k = 0; k < nc; ++k)
        {
            const double as = dot(V3(s_a[k], s_b[k], s_c[k]), m_r[j]), a = fabs(as);
            if (aef < a) { aef = a; km_r[j] = k; sg = as < 0 ? -1.0 : 1.0; }
        }
        aef_r[j] = aef;
        sg_r[j] = sg;
        sig_r[j] = dot(th, cross(m_r[j], n_r[j]));
        const V3 qv = (aef * sig_r[j]) * n_r[j];
        qa_r[j] = dot(qv, e1);
        qb_r[j] = dot(qv, e2);
    }
    __syncthreads(); // every read of the staged velocities is done
#pragma unroll
    for (int j = 0; j < kPer; ++j)
    {
        const int32_t i = int32_t(threadIdx.x) + j * int32_t(blockDim.x);
        if (i < nf) { s_a[i] = qa_r[j]; s_b[i] = qb_r[j]; s_c[i] = d.efw[f0 + i]; }
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double M11 = 0, M12 = 0, M22 = 0, g1 = 0, g2 = 0;
        for (int32_t i = 0; i < nf; ++i)
        {
            const double w = s_c[i], qa = s_a[i], qb = s_b[i];
            M11 += w * qa * qa; M12 += w * qa * qb; M22 += w * qb * qb; g1 += w * qa; g2 += w * qb;
        }
        double l1, l2;
        uct_solve(M11, M12, M22, g1, g2, l1, l2);
        s_r[0] = l1; s_r[1] = l2;
        s_r[2] = M11; s_r[3] = M12; s_r[4] = M22; s_r[5] = g1; s_r[6] = g2;
    }
    __syncthreads();
    const double Eu_b = 2.0 * cuct * L * Ebar, l1 = s_r[0], l2 = s_r[1];
    // the omega weights' parts of the lambda adjoint, per face
    double pf_r[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j)
    {
        const int32_t i = int32_t(threadIdx.x) + j * int32_t(blockDim.x);
        if (i >= nf) continue;
        const int32_t f = pid(d.e2f[f0 + i]);
        pf_r[j] = phi_ind[f] / d.famag[f];
        const double w = d.efw[f0 + i], omb = Eu_b * aef_r[j] * sig_r[j] * pf_r[j];
        s_a[i] = omb * w * qa_r[j];
        s_b[i] = omb * w * qb_r[j];
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double l1b = 0, l2b = 0;
        for (int32_t i = 0; i < nf; ++i) { l1b -= s_a[i]; l2b -= s_b[i]; }
        // (l1, l2) = solve(M11, M12, M22, g1, g2): transposed Jacobian by five forward columns
        double sb[5];
        for (int q = 0; q < 5; ++q)
        {
            Dual in[5] = {Dual(s_r[2]), Dual(s_r[3]), Dual(s_r[4]), Dual(s_r[5]), Dual(s_r[6])};
            in[q].d = 1.0;
            Dual a, b;
            uct_solve(in[0], in[1], in[2], in[3], in[4], a, b);
            sb[q] = l1b * a.d + l2b * b.d;
        }
        for (int q = 0; q < 5; ++q) s_r[2 + q] = sb[q];
    }
    __syncthreads();
    const double sb0 = s_r[2], sb1 = s_r[3], sb2 = s_r[4], sb3 = s_r[5], sb4 = s_r[6];
#pragma unroll
    for (int j = 0; j < kPer; ++j)
    {
        const int32_t i = int32_t(threadIdx.x) + j * int32_t(blockDim.x);
        if (i >= nf) continue;
        const int32_t f = pid(d.e2f[f0 + i]);
        const double w = d.efw[f0 + i], A = d.famag[f], pf = pf_r[j], aef = aef_r[j], sig = sig_r[j], qa = qa_r[j], qb = qb_r[j];
        const double omega = w * (1.0 - (qa * l1 + qb * l2));
        const double omb = Eu_b * aef * sig * pf;
        double aefb = Eu_b * omega * sig * pf;
        const double pfb = Eu_b * omega * aef * sig;
        const double qab = -omb * w * l1 + 2.0 * w * qa * sb0 + w * qb * sb1 + w * sb3;
        const double qbb = -omb * w * l2 + w * qa * sb1 + 2.0 * w * qb * sb2 + w * sb4;
        const V3 sn = sig * n_r[j];
        aefb += qab * dot(sn, e1) + qbb * dot(sn, e2);
        eF[4 * size_t(f0 + i)] = Ebar * w * t.x;
        eF[4 * size_t(f0 + i) + 1] = Ebar * w * t.y;
        eF[4 * size_t(f0 + i) + 2] = Ebar * w * t.z;
        eF[4 * size_t(f0 + i) + 3] = pfb / A;
        s_k[i] = km_r[j];
        const double gsc = aefb * sg_r[j];
        s_a[i] = gsc * m_r[j].x; s_b[i] = gsc * m_r[j].y; s_c[i] = gsc * m_r[j].z;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int32_t i = 0; i < nf; ++i)
            if (s_k[i] >= 0)
            {
                const size_t k = size_t(c0 + s_k[i]);
                eC[3 * k] += s_a[i]; eC[3 * k + 1] += s_b[i]; eC[3 * k + 2] += s_c[i];
            }
}

// face gather of the edge parts: Efb_f = sum over its edges, Phib_f += the induced-flux parts
__global__ void __launch_bounds__(256) k_adj_gather_edge_faces(DevMesh d, const int32_t *f2e_pos, const double *eF, double *Efb, double *Phib)
{
    const int32_t f = gid();
    if (f >= d.nf) return;
    V3 s;
    double p = 0.0;
    for (int32_t k = d.f2e_ptr[f]; k < d.f2e_ptr[f + 1]; ++k)
    {
        const size_t i = size_t(f2e_pos[k]);
        s += V3(eF[4 * i], eF[4 * i + 1], eF[4 * i + 2]);
        p += eF[4 * i + 3];
    }
    st3(Efb, f, s);
    Phib[f] += p;
}

// cell gather of the edge parts of u+
__global__ void __launch_bounds__(256) k_adj_gather_edge_cells(DevMesh d, const int32_t *cec_ptr, const int32_t *cec, const double *eC, double *upb)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    V3 s;
    for (int32_t j = cec_ptr[c]; j < cec_ptr[c + 1]; ++j)
    {
        const size_t i = size_t(cec[j]);
        s += V3(eC[3 * i], eC[3 * i + 1], eC[3 * i + 2]);
    }
    st3(upb, c, s);
}

// Forward Euler update with floors, flux divergence and Lorentz sources, transposed (k_cell_update):
// Ub (in: adjoint of U^{n+1}; out: its part through the update), Wb, Jb, Btb set here; sFA per c2f entry;
// per-block sums of d/d dt
__global__ void __launch_bounds__(256) k_adj_cell_update(DevMesh d, Params P, const double *FA, SpC Uold, SpC Wall, const double *Btot,
                                                         const double *J, double dt, SpW Ub, SpW Wb, double *Jb, double *Btb, SpW sFA,
                                                         double *part)
{
    const int32_t c = gid();
    double dtb[1] = {0.0};
    if (c < d.nc)
    {
        const bool fluid = d.mat[c] == 0;
        const double V = d.cvol[c];
        const int32_t k0 = d.c2f_ptr[c], k1 = d.c2f_ptr[c + 1];
        double rhsb[kAdjNs][5];
        for (int s = 0; s < kAdjNs; ++s)
        {
            double *ub = Ub.p[s] + 5 * size_t(c);
            if (!fluid)
            {
                for (int m = 0; m < 5; ++m) rhsb[s][m] = 0.0;
                for (int32_t k = k0; k < k1; ++k)
                    for (int m = 0; m < 5; ++m) sFA.p[s][5 * size_t(k) + m] = 0.0;
                continue;
            }
            // the step's rhs, as k_cell_update forms it
            double rhs[5] = {0, 0, 0, 0, 0};
            for (int32_t k = k0; k < k1; ++k)
            {
                const int32_t v = d.cfs[k], f = pid(v);
                const double *fa = FA + size_t(5 * kAdjNs) * f + 5 * s;
                if (v >= 0) for (int m = 0; m < 5; ++m) rhs[m] -= fa[m] / V;
                else for (int m = 0; m < 5; ++m) rhs[m] += fa[m] / V;
            }
            {
                double w[kAdjNs][5];
                for (int k = 0; k < kAdjNs; ++k)
                    for (int m = 0; m < 5; ++m) w[k][m] = Wall.p[k][5 * size_t(c) + m];
                double S[4];
                lorentz_source<double>(P, w, L3(J, c), L3(Btot, c), dt, s, S);
                rhs[1] += S[0]; rhs[2] += S[1]; rhs[3] += S[2]; rhs[4] += S[3];
                if (P.src_rate[s] != 0.0 && d.atm)
                {
                    const double q = P.src_rate[s] * d.atm[c];
                    rhs[0] += q;
                    rhs[4] += q * P.src_T / (P.gamma - 1.0);
                }
            }
            // floors: qb = (dU_new/dq)^T ub
            double q0[5], qb[5];
            for (int m = 0; m < 5; ++m) q0[m] = Uold.p[s][5 * size_t(c) + m] + dt * rhs[m];
            for (int j = 0; j < 5; ++j)
            {
                Dual q[5];
                for (int m = 0; m < 5; ++m) q[m] = Dual(q0[m], m == j ? 1.0 : 0.0);
                floors<Dual>(P, q);
                double g = 0;
                for (int m = 0; m < 5; ++m) g += ub[m] * q[m].d;
                qb[j] = g;
            }
            for (int m = 0; m < 5; ++m)
            {
                ub[m] = qb[m];
                rhsb[s][m] = dt * qb[m];
                dtb[0] += rhs[m] * qb[m];
            }
            for (int32_t k = k0; k < k1; ++k)
            {
                const double sg = d.c2f[k] >= 0 ? -1.0 : 1.0;
                for (int m = 0; m < 5; ++m) sFA.p[s][5 * size_t(k) + m] = sg * rhsb[s][m] / V;
            }
        }
        double wb[kAdjNs][5] = {}, jb[3] = {0, 0, 0}, bb[3] = {0, 0, 0};
        if (fluid)
        {
            // sources: columns over (W of every species, J, B_tot, dt)
            constexpr int K = 5 * kAdjNs + 7;
            for (int j = 0; j < K; ++j)
            {
                Dual w[kAdjNs][5];
                for (int k = 0; k < kAdjNs; ++k)
                    for (int m = 0; m < 5; ++m) w[k][m] = Dual(Wall.p[k][5 * size_t(c) + m], j == 5 * k + m ? 1.0 : 0.0);
                const V3 Jv = L3(J, c), Bv = L3(Btot, c);
                const int o = j - 5 * kAdjNs;
                const V3D Jd(Dual(Jv.x, o == 0 ? 1.0 : 0.0), Dual(Jv.y, o == 1 ? 1.0 : 0.0), Dual(Jv.z, o == 2 ? 1.0 : 0.0));
                const V3D Bd(Dual(Bv.x, o == 3 ? 1.0 : 0.0), Dual(Bv.y, o == 4 ? 1.0 : 0.0), Dual(Bv.z, o == 5 ? 1.0 : 0.0));
                const Dual dtd(dt, o == 6 ? 1.0 : 0.0);
                double g = 0;
                for (int s = 0; s < kAdjNs; ++s)
                {
                    Dual S[4];
                    lorentz_source<Dual>(P, w, Jd, Bd, dtd, s, S);
                    for (int m = 0; m < 4; ++m) g += rhsb[s][m + 1] * S[m].d;
                }
                if (o < 0) wb[j / 5][j % 5] = g;
                else if (o < 3) jb[o] = g;
                else if (o < 6) bb[o - 3] = g;
                else dtb[0] += g;
            }
        }
        for (int s = 0; s < kAdjNs; ++s)
            for (int m = 0; m < 5; ++m) Wb.p[s][5 * size_t(c) + m] = wb[s][m];
        for (int q = 0; q < 3; ++q) { Jb[3 * size_t(c) + q] = jb[q]; Btb[3 * size_t(c) + q] = bb[q]; }
    }
    block_sums<1>(dtb, part);
}

// face gather of the flux adjoint: FAb (5 per species per face, the layout of FA) from the two cells' slots
__global__ void __launch_bounds__(256) k_adj_gather_fa(DevMesh d, const int32_t *fsl, SpC sFA, double *FAb)
{
    const int32_t f = gid();
    if (f >= d.nf) return;
    const int32_t a = fsl[2 * f], b = fsl[2 * f + 1];
    for (int s = 0; s < kAdjNs; ++s)
        for (int m = 0; m < 5; ++m)
            FAb[size_t(5 * kAdjNs) * f + 5 * s + m] = (a >= 0 ? sFA.p[s][5 * size_t(a) + m] : 0.0) + (b >= 0 ? sFA.p[s][5 * size_t(b) + m] : 0.0);
}

// Face fluxes and face EMF candidates transposed (k_face): per side cell the parts of its primitives and
// gradients (sF: 20 per c2f entry and species) and of its u+, B_tot, B_ind (sE: 9 per c2f entry); Phib and
// PhiAddb += the face's own parts; per-block sums of the free-stream adjoint (5 per species)
__global__ void __launch_bounds__(256) k_adj_face(DevMesh d, Params P, SpC Ws, SpC Gs, const double *FAb, const double *uplus,
                                                  const double *Btot, const double *Bind, const double *Phi, const double *PhiAdd,
                                                  const double *Efb, const int32_t *fsl, SpW sF, double *sE, double *Phib, double *PhiAddb,
                                                  double *part)
{
    const int32_t f = gid();
    double fsb[5 * kAdjNs] = {};
    if (f < d.nf)
    {
        const int32_t l = d.fL[f], r = d.fR[f];
        const int32_t sl = fsl[2 * f], sr = fsl[2 * f + 1];
        const double A = d.famag[f];
        const V3 n = (1.0 / A) * L3(d.farea, f);
        const int8_t kind = d.kind[f];
        const bool any = (l >= 0 && d.mat[l] == 0) || (r >= 0 && d.mat[r] == 0);
        const V3 xf = L3(d.fcen, f);
        for (int s = 0; s < kAdjNs; ++s)
        {
            double cl[20] = {}, cr[20] = {};
            double Fb[5];
            bool nz = false;
            for (int m = 0; m < 5; ++m) { Fb[m] = FAb[size_t(5 * kAdjNs) * f + 5 * s + m] * A; nz = nz || Fb[m] != 0.0; }
            if (kind != kSkip && any && nz)
            {
                const double *W = Ws.p[s], *G = Gs.p[s];
                if (l >= 0 && r >= 0 && kind == kInterior)
                {
                    const V3 dl = xf - L3(d.ccen, l), dr = xf - L3(d.ccen, r);
                    const PrimT<double> L = face_state(W, G, l, dl), Rt = face_state(W, G, r, dr);
                    double Lb[5], Rb[5];
                    rusanov_adjoint(L, Rt, n, P.gamma, Fb, Lb, Rb);
                    face_state_adjoint(W, G, l, dl, Lb, cl);
                    face_state_adjoint(W, G, r, dr, Rb, cr);
                }
                else
                {
                    int32_t c;
                    if (l >= 0 && r >= 0)
                    {
                        const int32_t first = l < r ? l : r, second = l < r ? r : l;
                        c = d.mat[first] == 1 ? second : first;
                    }
                    else c = l >= 0 ? l : r;
                    const double sg = c == l ? 1.0 : -1.0;
                    const V3 dx = xf - L3(d.ccen, c);
                    const PrimT<double> in = face_state(W, G, c, dx);
                    const double fsv[5] = {P.fs[s][0], P.fs[s][1], P.fs[s][2], P.fs[s][3], P.fs[s][4]};
                    const PrimT<double> gh = ghost_state(kind, s, in, sg * n, fsv);
                    double Lb[5], Rb[5];
                    if (sg > 0) rusanov_adjoint(in, gh, n, P.gamma, Fb, Lb, Rb);
                    else rusanov_adjoint(gh, in, n, P.gamma, Fb, Lb, Rb);
                    double ib[5], gb[5];
                    for (int m = 0; m < 5; ++m) { ib[m] = sg > 0 ? Lb[m] : Rb[m]; gb[m] = sg > 0 ? Rb[m] : Lb[m]; }
                    if (kind == kWall)
                    {
                        if (s != 1)
                        {
                            const V3 nn = sg * n;
                            const double ndg = nn.x * gb[1] + nn.y * gb[2] + nn.z * gb[3];
                            ib[0] += gb[0];
                            ib[1] += gb[1] - 2.0 * ndg * nn.x;
                            ib[2] += gb[2] - 2.0 * ndg * nn.y;
                            ib[3] += gb[3] - 2.0 * ndg * nn.z;
                            ib[4] += gb[4];
                        }
                    }
                    else if (kind == kInflow)
                        for (int m = 0; m < 5; ++m) fsb[5 * s + m] += gb[m];
                    else
                        for (int m = 0; m < 5; ++m) ib[m] += gb[m];
                    face_state_adjoint(W, G, c, dx, ib, c == l ? cl : cr);
                }
            }
            if (sl >= 0)
                for (int q = 0; q < 20; ++q) sF.p[s][20 * size_t(sl) + q] = cl[q];
            if (sr >= 0)
                for (int q = 0; q < 20; ++q) sF.p[s][20 * size_t(sr) + q] = cr[q];
        }
        // E_f = -n x F  (IdealEMF.cpp:32-55 as in k_face)
        double el[9] = {}, er[9] = {};
        const V3 Eb = L3(Efb, f);
        const V3 Fb = cross(n, Eb);
        double Bnb = 0.0;
        const double Bn = (PhiAdd ? Phi[f] + PhiAdd[f] : Phi[f]) / A;
        auto side = [&](int32_t c, const V3 &fxb, double *out) {
            const V3 u = L3(uplus, c), B = L3(Btot, c);
            const V3 ub = dot(B, fxb) * n - Bn * fxb;
            const V3 Bb = dot(u, n) * fxb;
            Bnb -= dot(u, fxb);
            out[0] += ub.x; out[1] += ub.y; out[2] += ub.z;
            out[3] += Bb.x; out[4] += Bb.y; out[5] += Bb.z;
        };
        if (l >= 0 && r >= 0)
        {
            const V3 uL = L3(uplus, l), uR = L3(uplus, r);
            const double aL = fabs(dot(uL, n)), aR = fabs(dot(uR, n));
            const double a = aL < aR ? aR : aL;
            side(l, 0.5 * Fb, el);
            side(r, 0.5 * Fb, er);
            const V3 db = L3(Bind, r) - L3(Bind, l);
            const double ab = -0.5 * dot(Fb, db);
            el[6] += 0.5 * a * Fb.x; el[7] += 0.5 * a * Fb.y; el[8] += 0.5 * a * Fb.z;
            er[6] -= 0.5 * a * Fb.x; er[7] -= 0.5 * a * Fb.y; er[8] -= 0.5 * a * Fb.z;
            const bool right = aL < aR;
            const V3 us = right ? uR : uL;
            const double sgn = dot(us, n) < 0 ? -1.0 : 1.0;
            double *o = right ? er : el;
            o[0] += ab * sgn * n.x; o[1] += ab * sgn * n.y; o[2] += ab * sgn * n.z;
        }
        else
        {
            const int32_t c = l >= 0 ? l : r;
            side(c, Fb, c == l ? el : er);
        }
        Phib[f] += Bnb / A;
        if (PhiAddb) PhiAddb[f] += Bnb / A;
        if (sl >= 0)
            for (int q = 0; q < 9; ++q) sE[9 * size_t(sl) + q] = el[q];
        if (sr >= 0)
            for (int q = 0; q < 9; ++q) sE[9 * size_t(sr) + q] = er[q];
    }
    block_sums<5 * kAdjNs>(fsb, part);
}

// cell gather of the face parts: Wb += primitives, Gb = gradients, upb += u+, Btb += B_tot, Bib = B_ind
__global__ void __launch_bounds__(256) k_adj_gather_faces(DevMesh d, SpC sF, const double *sE, SpW Wb, SpW Gb, double *upb, double *Btb,
                                                          double *Bib)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    const int32_t k0 = d.c2f_ptr[c], k1 = d.c2f_ptr[c + 1];
    for (int s = 0; s < kAdjNs; ++s)
    {
        double w[5] = {0, 0, 0, 0, 0}, g[15] = {};
        for (int32_t k = k0; k < k1; ++k)
        {
            const double *x = sF.p[s] + 20 * size_t(k);
            for (int m = 0; m < 5; ++m) w[m] += x[m];
            for (int q = 0; q < 15; ++q) g[q] += x[5 + q];
        }
        for (int m = 0; m < 5; ++m) Wb.p[s][5 * size_t(c) + m] += w[m];
        for (int q = 0; q < 15; ++q) Gb.p[s][15 * size_t(c) + q] = g[q];
    }
    double e[9] = {};
    for (int32_t k = k0; k < k1; ++k)
        for (int q = 0; q < 9; ++q) e[q] += sE[9 * size_t(k) + q];
    for (int q = 0; q < 3; ++q)
    {
        upb[3 * size_t(c) + q] += e[q];
        Btb[3 * size_t(c) + q] += e[3 + q];
        Bib[3 * size_t(c) + q] = e[6 + q];
    }
}

// Limited gradients transposed (k_grad_current): per cell and species the parts of its own primitives (Wb +=),
// of each stencil neighbour's (sG: 5 per c2f entry, the first entry of that neighbour) and of the inflow
// free stream (per-block sums, 5 per species)
__global__ void __launch_bounds__(256) k_adj_grad(DevMesh d, Params P, SpC Ws, SpC Gb, SpW Wb, SpW sG, double *part)
{
    const int32_t c = gid();
    double fsb[5 * kAdjNs] = {};
    if (c < d.nc)
    {
        const int32_t k0 = d.c2f_ptr[c], k1 = d.c2f_ptr[c + 1];
        for (int s = 0; s < kAdjNs; ++s)
        {
            for (int32_t k = k0; k < k1; ++k)
                for (int m = 0; m < 5; ++m) sG.p[s][5 * size_t(k) + m] = 0.0;
            const double *gb = Gb.p[s] + 15 * size_t(c);
            bool nz = false;
            for (int q = 0; q < 15; ++q) nz = nz || gb[q] != 0.0;
            if (!d.svalid[c] || !nz) continue;
            const double *W = Ws.p[s];
            // Columns of the local Jacobian, seeded several primitives at a time: primitive m's gradient depends on
            // primitive m of the stencil only (LSQ sums, Barth-Jespersen and relative limiters per primitive), except
            // through the wall ghost of the first species, which mirrors the cell's own velocity (components 1-3
            // couple); those three are seeded one at a time, with every output summed, as a single column is.
            auto columns = [&](int32_t cell, uint32_t mm, uint32_t fm, double g[5]) {
                Dual out[15];
                grad_column(d, P, c, s, W, cell, mm, fm, out);
                for (int m = 0; m < 5; ++m) g[m] = gb[3 * m] * out[3 * m].d + gb[3 * m + 1] * out[3 * m + 1].d + gb[3 * m + 2] * out[3 * m + 2].d;
            };
            auto column = [&](int32_t cell, int32_t m) {
                Dual out[15];
                grad_column(d, P, c, s, W, cell, 1u << m, 0u, out);
                double g = 0;
                for (int q = 0; q < 15; ++q) g += gb[q] * out[q].d;
                return g;
            };
            bool wall = false, inflow = false;
            for (int32_t k = k0; k < k1; ++k)
            {
                const int32_t nb = d.snb[k];
                if (nb == kNoStencil || nb >= 0) continue;
                wall = wall || d.kind[-(nb + 1)] == kWall;
                inflow = inflow || d.kind[-(nb + 1)] == kInflow;
            }
            double g[5];
            if (s == 0 && wall)
            {
                columns(c, (1u << 0) | (1u << 4), 0u, g);
                g[1] = column(c, 1); g[2] = column(c, 2); g[3] = column(c, 3);
            }
            else columns(c, 0x1fu, 0u, g);
            for (int m = 0; m < 5; ++m) Wb.p[s][5 * size_t(c) + m] += g[m];
            for (int32_t k = k0; k < k1; ++k)
            {
                const int32_t nb = d.snb[k];
                if (nb == kNoStencil || nb < 0) continue;
                bool first = nb != c;
                for (int32_t j = k0; j < k && first; ++j) first = d.snb[j] != nb;
                if (!first) continue;
                columns(nb, 0x1fu, 0u, g);
                for (int m = 0; m < 5; ++m) sG.p[s][5 * size_t(k) + m] = g[m];
            }
            if (inflow)
            {
                columns(-1, 0u, 0x1fu, g);
                for (int m = 0; m < 5; ++m) fsb[5 * s + m] += g[m];
            }
        }
    }
    block_sums<5 * kAdjNs>(fsb, part);
}

// cell gather of the stencil-neighbour parts: cell j adds slot kc of each face neighbour c whose stencil entry
// across that face is j
__global__ void __launch_bounds__(256) k_adj_gather_grad(DevMesh d, const int32_t *fsl, SpC sG, SpW Wb)
{
    const int32_t j = gid();
    if (j >= d.nc) return;
    for (int32_t k = d.c2f_ptr[j]; k < d.c2f_ptr[j + 1]; ++k)
    {
        const int32_t f = pid(d.c2f[k]);
        const int32_t l = d.fL[f], r = d.fR[f];
        const int32_t c = l == j ? r : l;
        if (c < 0 || c == j) continue;
        const int32_t kc = c == l ? fsl[2 * f] : fsl[2 * f + 1];
        if (kc < 0 || d.snb[kc] != j) continue;
        for (int s = 0; s < kAdjNs; ++s)
            for (int m = 0; m < 5; ++m) Wb.p[s][5 * size_t(j) + m] += sG.p[s][5 * size_t(kc) + m];
    }
}

// cell_current transposed (Current.cpp:9-31 as in k_grad_current): Bib_c += own and neighbour parts
__global__ void __launch_bounds__(256) k_adj_current(DevMesh d, const double *Jb, double *Bib)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    const V3 jc = (1.0 / d.cvol[c]) * L3(Jb, c);
    V3 acc;
    for (int32_t k = d.c2f_ptr[c]; k < d.c2f_ptr[c + 1]; ++k)
    {
        const int32_t v = d.c2f[k], f = pid(v);
        const int32_t l = d.fL[f], r = d.fR[f];
        const V3 a = psg(v) * L3(d.farea, f);
        const V3 own = cross(jc, a);
        if (l >= 0 && r >= 0)
        {
            const int32_t o = l == c ? r : l;
            const V3 ao = (o == l ? 1.0 : -1.0) * L3(d.farea, f);
            const V3 other = cross((1.0 / d.cvol[o]) * L3(Jb, o), ao);
            acc += 0.5 * own + 0.5 * other;
        }
        else acc += own;
    }
    Bib[3 * size_t(c)] += acc.x; Bib[3 * size_t(c) + 1] += acc.y; Bib[3 * size_t(c) + 2] += acc.z;
}

// per-block (smallest cell time step, first cell attaining it) over the fluid cells, as the step's CFL minimum
__global__ void __launch_bounds__(256) k_adj_dt_argmin(DevMesh d, Params P, SpC Ws, const double *Btot, double *pv, int32_t *pc)
{
    __shared__ double sv[256];
    __shared__ int32_t sc[256];
    const int32_t c = gid();
    double v = __builtin_inf();
    int32_t ci = INT32_MAX;
    if (c < d.nc && d.mat[c] == 0) { v = cell_dt(d, P, c, Ws.p, L3(Btot, c), size_t(c)); ci = c; }
    const int t = threadIdx.x;
    sv[t] = v; sc[t] = ci;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1)
    {
        if (t < s && (sv[t + s] < sv[t] || (sv[t + s] == sv[t] && sc[t + s] < sc[t]))) { sv[t] = sv[t + s]; sc[t] = sc[t + s]; }
        __syncthreads();
    }
    if (t == 0) { pv[blockIdx.x] = sv[0]; pc[blockIdx.x] = sc[0]; }
}

// the CFL step of cell c transposed: Wb, Btb of c += dtb d dt_c / d(W, B_tot) (one thread; wd: 5 per species
// of Dual scratch)
__global__ void k_adj_dt_cell(DevMesh d, Params P, int32_t c, SpC Ws, const double *Btot, double dtb, Dual *wd, SpW Wb, double *Btb)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    SpCT<Dual> sp{};
    for (int s = 0; s < kAdjNs; ++s) sp.p[s] = wd + 5 * s;
    const V3 b = L3(Btot, c);
    for (int j = 0; j < 5 * kAdjNs + 3; ++j)
    {
        for (int s = 0; s < kAdjNs; ++s)
            for (int m = 0; m < 5; ++m) wd[5 * s + m] = Dual(Ws.p[s][5 * size_t(c) + m], j == 5 * s + m ? 1.0 : 0.0);
        const int o = j - 5 * kAdjNs;
        const V3D bd(Dual(b.x, o == 0 ? 1.0 : 0.0), Dual(b.y, o == 1 ? 1.0 : 0.0), Dual(b.z, o == 2 ? 1.0 : 0.0));
        const Dual t = cell_dt(d, P, c, sp.p, bd, 0);
        const double g = dtb * t.d;
        if (o < 0) Wb.p[j / 5][5 * size_t(c) + j % 5] += g;
        else Btb[3 * size_t(c) + o] += g;
    }
}

// Solver::prepare transposed (k_prepare): Baddb += Btb; u+ and primitive recovery transposed into Ub +=;
// the reconstructed field's part of the induced flux into sPhi (1 per c2f entry)
__global__ void __launch_bounds__(256) k_adj_prepare(DevMesh d, Params P, SpC Us, SpC Ws, SpC Wb, const double *upb, const double *Btb,
                                                     const double *Bib, SpW Ub, double *Baddb, double *sPhi)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    const V3 btb = L3(Btb, c);
    Baddb[3 * size_t(c)] += btb.x; Baddb[3 * size_t(c) + 1] += btb.y; Baddb[3 * size_t(c) + 2] += btb.z;
    double wb[kAdjNs][5];
    for (int s = 0; s < kAdjNs; ++s)
        for (int m = 0; m < 5; ++m) wb[s][m] = Wb.p[s][5 * size_t(c) + m];
    const V3 ub = L3(upb, c);
    if (d.mat[c] == 0 && (ub.x != 0.0 || ub.y != 0.0 || ub.z != 0.0))
        for (int j = 0; j < 4 * kAdjNs; ++j)
        {
            Dual w[kAdjNs][5];
            for (int s = 0; s < kAdjNs; ++s)
                for (int m = 0; m < 5; ++m) w[s][m] = Dual(Ws.p[s][5 * size_t(c) + m], (j / 4 == s && j % 4 == m) ? 1.0 : 0.0);
            const V3D up = ion_velocity<Dual>(P, w);
            wb[j / 4][j % 4] += ub.x * up.x.d + ub.y * up.y.d + ub.z * up.z.d;
        }
    for (int s = 0; s < kAdjNs; ++s)
        for (int j = 0; j < 5; ++j)
        {
            Dual u[5];
            for (int m = 0; m < 5; ++m) u[m] = Dual(Us.p[s][5 * size_t(c) + m], m == j ? 1.0 : 0.0);
            const PrimT<Dual> w = to_prim(u, P.gamma, P.rho_floor, P.p_floor);
            Ub.p[s][5 * size_t(c) + j] += wb[s][0] * w.rho.d + wb[s][1] * w.u.d + wb[s][2] * w.v.d + wb[s][3] * w.w.d + wb[s][4] * w.p.d;
        }
    cell_vector_adjoint(d, c, L3(Bib, c) + btb, sPhi);
}

// shell energy transposed: J = sum over fluid cells with r0 <= |x| <= r1 of |b + B_add|^2 V
__global__ void __launch_bounds__(256) k_adj_energy(DevMesh d, const double *phi, const double *badd, double r0, double r1, double *Baddb,
                                                    double *sPhi)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    V3 bb;
    if (d.mat[c] == 0)
    {
        const double r = norm(L3(d.ccen, c));
        if (r >= r0 && r <= r1) bb = (2.0 * d.cvol[c]) * (cell_vector(d, c, phi) + L3(badd, c));
    }
    Baddb[3 * size_t(c)] += bb.x; Baddb[3 * size_t(c) + 1] += bb.y; Baddb[3 * size_t(c) + 2] += bb.z;
    cell_vector_adjoint(d, c, bb, sPhi);
}

// face gather of the reconstruction parts: Phib_f += both cells' slots
__global__ void __launch_bounds__(256) k_adj_gather_phi(DevMesh d, const int32_t *fsl, const double *sPhi, double *Phib)
{
    const int32_t f = gid();
    if (f >= d.nf) return;
    const int32_t a = fsl[2 * f], b = fsl[2 * f + 1];
    Phib[f] += (a >= 0 ? sPhi[a] : 0.0) + (b >= 0 ? sPhi[b] : 0.0);
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
// sum of K per-block partials (host, block order)
std::vector<double> sum_parts(mhdg_ctx &c, const double *part, int nb, int K)
{
    std::vector<double> h(size_t(nb) * K), out(K, 0.0);
    c.d2h(h.data(), part, h.size() * 8);
    for (int b = 0; b < nb; ++b)
        for (int j = 0; j < K; ++j) out[j] += h[size_t(b) * K + j];
    return out;
}
} // namespace

#define LAUNCH(NAME, ...)                                                                                                    \
    do                                                                                                                       \
    {                                                                                                                        \
        prof_begin(NAME);                                                                                                    \
        hipLaunchKernelGGL(__VA_ARGS__);                                                                                     \
        prof_end();                                                                                                          \
    } while (0)

// Work arrays and gather maps of the adjoint (first use)
void mhdg_ctx::vjp_alloc(int n_steps)
{
    if (!aw.fsl)
    {
        const size_t C = nc, F = nf, E = ne, K = c2f_nnz;
        // host copies of the incidence tables, for the gather maps
        std::vector<int32_t> c2f_ptr(C + 1), c2f(K), fL(F), fR(F), f2e_ptr(F + 1), f2e(size_t(3) * F), e2f_ptr(E + 1), ec_ptr(E + 1);
        chk(hipMemcpy(c2f_ptr.data(), d.c2f_ptr, 4 * (C + 1), hipMemcpyDeviceToHost), "D2H");
        chk(hipMemcpy(c2f.data(), d.c2f, 4 * K, hipMemcpyDeviceToHost), "D2H");
        chk(hipMemcpy(fL.data(), d.fL, 4 * F, hipMemcpyDeviceToHost), "D2H");
        chk(hipMemcpy(fR.data(), d.fR, 4 * F, hipMemcpyDeviceToHost), "D2H");
        chk(hipMemcpy(f2e_ptr.data(), d.f2e_ptr, 4 * (F + 1), hipMemcpyDeviceToHost), "D2H");
        f2e.resize(size_t(f2e_ptr[F]));
        chk(hipMemcpy(f2e.data(), d.f2e, 4 * f2e.size(), hipMemcpyDeviceToHost), "D2H");
        chk(hipMemcpy(e2f_ptr.data(), d.e2f_ptr, 4 * (E + 1), hipMemcpyDeviceToHost), "D2H");
        std::vector<int32_t> e2f(static_cast<size_t>(e2f_ptr[E]));
        chk(hipMemcpy(e2f.data(), d.e2f, 4 * e2f.size(), hipMemcpyDeviceToHost), "D2H");
        chk(hipMemcpy(ec_ptr.data(), d.ec_ptr, 4 * (E + 1), hipMemcpyDeviceToHost), "D2H");
        std::vector<int32_t> ec_cell(static_cast<size_t>(ec_ptr[E]));
        chk(hipMemcpy(ec_cell.data(), d.ec_cell, 4 * ec_cell.size(), hipMemcpyDeviceToHost), "D2H");
        auto id = [](int32_t v) { return v >= 0 ? v : ~v; };
        // face -> its c2f entry in fL's and in fR's list
        std::vector<int32_t> fsl(2 * F, -1);
        for (size_t c = 0; c < C; ++c)
            for (int32_t k = c2f_ptr[c]; k < c2f_ptr[c + 1]; ++k)
            {
                const int32_t f = id(c2f[k]);
                if (fL[f] == int32_t(c) && fsl[2 * f] < 0) fsl[2 * f] = k;
                else if (fR[f] == int32_t(c) && fsl[2 * f + 1] < 0) fsl[2 * f + 1] = k;
            }
        // f2e entry -> position of the face in the edge's e2f list
        std::vector<int32_t> pos(f2e.size(), -1);
        for (size_t f = 0; f < F; ++f)
            for (int32_t k = f2e_ptr[f]; k < f2e_ptr[f + 1]; ++k)
            {
                const int32_t e = id(f2e[k]);
                for (int32_t i = e2f_ptr[e]; i < e2f_ptr[e + 1]; ++i)
                    if (id(e2f[i]) == int32_t(f)) { pos[k] = i; break; }
                if (pos[k] < 0) throw std::runtime_error("vjp: face missing from its edge's face list");
            }
        // cell -> its edge-cell entries
        std::vector<int32_t> cptr(C + 1, 0), cec(ec_cell.size());
        for (int32_t c : ec_cell) ++cptr[size_t(c) + 1];
        for (size_t c = 0; c < C; ++c) cptr[c + 1] += cptr[c];
        std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
        for (size_t k = 0; k < ec_cell.size(); ++k) cec[size_t(fill[size_t(ec_cell[k])]++)] = int32_t(k);
        auto up32 = [&](const std::vector<int32_t> &v) {
            int32_t *p = static_cast<int32_t *>(alloc(4 * v.size()));
            if (!v.empty()) chk(hipMemcpy(p, v.data(), 4 * v.size(), hipMemcpyHostToDevice), "upload");
            return p;
        };
        aw.fsl = up32(fsl);
        aw.f2e_pos = up32(pos);
        aw.cec_ptr = up32(cptr);
        aw.cec = up32(cec);
        auto D = [&](size_t n) { return static_cast<double *>(alloc(n * sizeof(double))); };
        for (int s = 0; s < 2; ++s)
        {
            aw.U[s] = D(5 * C); aw.W[s] = D(5 * C); aw.G[s] = D(15 * C);
            aw.Ub[s] = D(5 * C); aw.Wb[s] = D(5 * C); aw.Gb[s] = D(15 * C);
            aw.sF[s] = D(20 * K); aw.sG[s] = D(5 * K); aw.sFA[s] = D(5 * K);
        }
        aw.Phi = D(F); aw.Bind = D(3 * C); aw.Btot = D(3 * C); aw.up = D(3 * C); aw.J = D(3 * C); aw.alpha = D(C);
        aw.FA = D(10 * F); aw.Ef = D(3 * F); aw.E = D(E); aw.dt_part = D(size_t(nbk(nc)) + 1); aw.dt_dev = D(1);
        aw.Phib = D(F); aw.PhiAddb = D(F); aw.Baddb = D(3 * C); aw.Eb = D(E); aw.Efb = D(3 * F); aw.FAb = D(10 * F);
        aw.upb = D(3 * C); aw.Btb = D(3 * C); aw.Bib = D(3 * C); aw.Jb = D(3 * C);
        aw.sE = D(9 * K); aw.sPhi = D(K);
        aw.eF = D(4 * e2f.size()); aw.eC = D(3 * ec_cell.size());
        aw.part = D(size_t(kRedMax) * size_t(nbk(std::max<int64_t>({int64_t(C), int64_t(F), int64_t(E)}))));
        aw.argc = static_cast<int32_t *>(alloc(4 * size_t(nbk(nc))));
        aw.wd = static_cast<Dual *>(alloc(sizeof(Dual) * 5 * 2));
    }
    if (int(aw.ck.size()) < n_steps)
    {
        for (double *p : aw.ck) (void)hipFree(p);
        aw.ck.clear();
        const size_t per = 10 * size_t(nc) + size_t(nf);
        for (int n = 0; n < n_steps; ++n)
        {
            void *p = nullptr;
            chk(hipMalloc(&p, per * 8), "hipMalloc");
            aw.ck.push_back(static_cast<double *>(p));
        }
    }
}

// the explicit stage of step_first_order on (U, Phi) of the adjoint work arrays up to the edge EMF; with
// `update`, also the cell update and Faraday (the forward sweep)
double mhdg_ctx::vjp_stage(double dt_max, bool update)
{
    const int nb = nbk(nc);
    SpC Uc{}, Wc{}, Gc{}, none{};
    SpW Ww{}, Gw{}, Uw{}, noneW{};
    for (int s = 0; s < 2; ++s)
    {
        Uc.p[s] = aw.U[s]; Wc.p[s] = aw.W[s]; Gc.p[s] = aw.G[s];
        Ww.p[s] = aw.W[s]; Gw.p[s] = aw.G[s]; Uw.p[s] = aw.U[s];
    }
    LAUNCH("prepare", k_prepare<double>, dim3(nb), dim3(256), 0, stream, d, P, Uc, static_cast<const double *>(aw.Phi),
           static_cast<const double *>(Badd), Ww, aw.Bind, aw.Btot, aw.up, aw.alpha, aw.dt_part);
    LAUNCH("dt_final", k_dt_final<double>, dim3(1), dim3(256), 0, stream, static_cast<const double *>(aw.dt_part), nb, dt_max, aw.dt_dev);
    if (P.second_order)
        LAUNCH("grad_current", k_grad_current<double>, dim3(nb), dim3(256), 0, stream, d, P, 2, 0, Wc, none, Gw, 1,
               static_cast<const double *>(aw.Bind), aw.J);
    else
    {
        for (int s = 0; s < 2; ++s) chk(hipMemsetAsync(aw.G[s], 0, 15 * size_t(nc) * 8, stream), "memset");
        LAUNCH("grad_current", k_grad_current<double>, dim3(nb), dim3(256), 0, stream, d, P, 0, 0, Wc, none, Gw, 1,
               static_cast<const double *>(aw.Bind), aw.J);
    }
    LAUNCH("face", k_face<double>, dim3(nbk(nf)), dim3(256), 0, stream, d, P, 2, 0, Wc, Gc, none, aw.FA,
           static_cast<const double *>(aw.up), static_cast<const double *>(aw.Btot), static_cast<const double *>(aw.Bind),
           static_cast<const double *>(aw.Phi), static_cast<const double *>(PhiAdd), aw.Ef);
    double dt = 0;
    d2h(&dt, aw.dt_dev, 8);
    if (!std::isfinite(dt) || dt <= 0) throw std::runtime_error("invalid time step");
    if (update)
        LAUNCH("cell_update", k_cell_update<double>, dim3(nb), dim3(256), 0, stream, d, P, 2, 0, static_cast<const double *>(aw.FA), none, Wc,
               static_cast<const double *>(aw.Btot), static_cast<const double *>(aw.J), static_cast<const double *>(aw.dt_dev), 0.0, 1, 1,
               Uw, noneW);
    launch_edge_emf(aw.Ef, aw.up, aw.Phi, P.cuct, nullptr, aw.E, 1);
    if (update)
        LAUNCH("faraday", k_faraday<double>, dim3(nbk(nf)), dim3(256), 0, stream, d, static_cast<const double *>(aw.E),
               static_cast<const double *>(nullptr), static_cast<const double *>(aw.dt_dev), 0.0, aw.Phi);
    check_launch("vjp stage");
    return dt;
}

// n explicit steps from the resident state (kept unchanged), J = shell energy of the final state, and the
// reverse sweep; gradients are left in aw.Ub, aw.Phib, aw.PhiAddb, aw.Baddb and fs_b (2x5), einf_b (3)
double mhdg_ctx::step_vjp(int n_steps, double dt_max, double r0, double r1, double *fs_b, double *einf_b)
{
    vjp_alloc(n_steps);
    const size_t C = nc, F = nf, E = ne;
    const int nbc = nbk(nc), nbf = nbk(nf), nbe = nbk(ne);
    for (int s = 0; s < 2; ++s) chk(hipMemcpyAsync(aw.U[s], U[s], 5 * C * 8, hipMemcpyDeviceToDevice, stream), "copy");
    chk(hipMemcpyAsync(aw.Phi, Phi, F * 8, hipMemcpyDeviceToDevice, stream), "copy");
    std::vector<double> dts(static_cast<size_t>(n_steps));
    for (int n = 0; n < n_steps; ++n)
    {
        double *ck = aw.ck[size_t(n)];
        chk(hipMemcpyAsync(ck, aw.U[0], 5 * C * 8, hipMemcpyDeviceToDevice, stream), "copy");
        chk(hipMemcpyAsync(ck + 5 * C, aw.U[1], 5 * C * 8, hipMemcpyDeviceToDevice, stream), "copy");
        chk(hipMemcpyAsync(ck + 10 * C, aw.Phi, F * 8, hipMemcpyDeviceToDevice, stream), "copy");
        dts[size_t(n)] = vjp_stage(dt_max, true);
    }
    double Jv = 0;
    shell_energy<double>(aw.Phi, Badd, r0, r1, &Jv);
    // adjoint of the objective
    for (int s = 0; s < 2; ++s) chk(hipMemsetAsync(aw.Ub[s], 0, 5 * C * 8, stream), "memset");
    chk(hipMemsetAsync(aw.Phib, 0, F * 8, stream), "memset");
    chk(hipMemsetAsync(aw.PhiAddb, 0, F * 8, stream), "memset");
    chk(hipMemsetAsync(aw.Baddb, 0, 3 * C * 8, stream), "memset");
    LAUNCH("adj_energy", k_adj_energy, dim3(nbc), dim3(256), 0, stream, d, static_cast<const double *>(aw.Phi),
           static_cast<const double *>(Badd), r0, r1, aw.Baddb, aw.sPhi);
    LAUNCH("adj_gather_phi", k_adj_gather_phi, dim3(nbf), dim3(256), 0, stream, d, static_cast<const int32_t *>(aw.fsl),
           static_cast<const double *>(aw.sPhi), aw.Phib);
    double fsb[10] = {}, eib[3] = {};
    SpC Wc{}, Gc{}, Uck{}, Wbc{}, Gbc{}, sFc{}, sGc{}, sFAc{};
    SpW Ubw{}, Wbw{}, Gbw{}, sFw{}, sGw{}, sFAw{};
    for (int s = 0; s < 2; ++s)
    {
        Wc.p[s] = aw.W[s]; Gc.p[s] = aw.G[s];
        Ubw.p[s] = aw.Ub[s]; Wbw.p[s] = aw.Wb[s]; Gbw.p[s] = aw.Gb[s]; Wbc.p[s] = aw.Wb[s]; Gbc.p[s] = aw.Gb[s];
        sFw.p[s] = aw.sF[s]; sFc.p[s] = aw.sF[s]; sGw.p[s] = aw.sG[s]; sGc.p[s] = aw.sG[s]; sFAw.p[s] = aw.sFA[s]; sFAc.p[s] = aw.sFA[s];
    }
    for (int n = n_steps - 1; n >= 0; --n)
    {
        // recompute the step's intermediates from its checkpoint
        const double *ck = aw.ck[size_t(n)];
        chk(hipMemcpyAsync(aw.U[0], ck, 5 * C * 8, hipMemcpyDeviceToDevice, stream), "copy");
        chk(hipMemcpyAsync(aw.U[1], ck + 5 * C, 5 * C * 8, hipMemcpyDeviceToDevice, stream), "copy");
        chk(hipMemcpyAsync(aw.Phi, ck + 10 * C, F * 8, hipMemcpyDeviceToDevice, stream), "copy");
        const double dt = vjp_stage(dt_max, false);
        if (dt != dts[size_t(n)]) throw std::runtime_error("vjp: recomputed time step differs from the forward sweep");
        for (int s = 0; s < 2; ++s) Uck.p[s] = aw.U[s];
        // Faraday
        LAUNCH("adj_faraday_dt", k_adj_faraday_dt, dim3(nbf), dim3(256), 0, stream, d, static_cast<const double *>(aw.E),
               static_cast<const double *>(aw.Phib), aw.part);
        double dtb = sum_parts(*this, aw.part, nbf, 1)[0];
        LAUNCH("adj_faraday", k_adj_faraday, dim3(nbe), dim3(256), 0, stream, d, static_cast<const double *>(aw.Phib), dt, aw.Eb);
        // edge EMF
        LAUNCH("adj_edge_emf", k_adj_edge_emf, dim3(nbe), dim3(256), 0, stream, d, P, static_cast<const double *>(aw.up),
               static_cast<const double *>(aw.Phi), P.cuct, static_cast<const double *>(aw.Eb), aw.eF, aw.eC, aw.part);
        if (n_heavy > 0)
            LAUNCH("adj_edge_emf_heavy", k_adj_edge_emf_heavy, dim3(n_heavy), dim3(256), 0, stream, d, static_cast<const double *>(aw.up),
                   static_cast<const double *>(aw.Phi), P.cuct, static_cast<const double *>(aw.Eb), static_cast<const int32_t *>(heavy), aw.eF,
                   aw.eC);
        const std::vector<double> ei = sum_parts(*this, aw.part, nbe, 3);
        for (int k = 0; k < 3; ++k) eib[k] += ei[size_t(k)];
        LAUNCH("adj_gather_edge_faces", k_adj_gather_edge_faces, dim3(nbf), dim3(256), 0, stream, d, static_cast<const int32_t *>(aw.f2e_pos),
               static_cast<const double *>(aw.eF), aw.Efb, aw.Phib);
        LAUNCH("adj_gather_edge_cells", k_adj_gather_edge_cells, dim3(nbc), dim3(256), 0, stream, d, static_cast<const int32_t *>(aw.cec_ptr),
               static_cast<const int32_t *>(aw.cec), static_cast<const double *>(aw.eC), aw.upb);
        // cell update and sources
        LAUNCH("adj_cell_update", k_adj_cell_update, dim3(nbc), dim3(256), 0, stream, d, P, static_cast<const double *>(aw.FA), Uck, Wc,
               static_cast<const double *>(aw.Btot), static_cast<const double *>(aw.J), dt, Ubw, Wbw, aw.Jb, aw.Btb, sFAw, aw.part);
        dtb += sum_parts(*this, aw.part, nbc, 1)[0];
        LAUNCH("adj_gather_fa", k_adj_gather_fa, dim3(nbf), dim3(256), 0, stream, d, static_cast<const int32_t *>(aw.fsl), sFAc, aw.FAb);
        // face fluxes and face EMF candidates
        LAUNCH("adj_face", k_adj_face, dim3(nbf), dim3(256), 0, stream, d, P, Wc, Gc, static_cast<const double *>(aw.FAb),
               static_cast<const double *>(aw.up), static_cast<const double *>(aw.Btot), static_cast<const double *>(aw.Bind),
               static_cast<const double *>(aw.Phi), static_cast<const double *>(PhiAdd), static_cast<const double *>(aw.Efb),
               static_cast<const int32_t *>(aw.fsl), sFw, aw.sE, aw.Phib, aw.PhiAddb, aw.part);
        {
            const std::vector<double> f = sum_parts(*this, aw.part, nbf, 10);
            for (int k = 0; k < 10; ++k) fsb[k] += f[size_t(k)];
        }
        LAUNCH("adj_gather_faces", k_adj_gather_faces, dim3(nbc), dim3(256), 0, stream, d, sFc, static_cast<const double *>(aw.sE), Wbw, Gbw,
               aw.upb, aw.Btb, aw.Bib);
        // limited gradients and cell current
        if (P.second_order)
        {
            LAUNCH("adj_grad", k_adj_grad, dim3(nbc), dim3(256), 0, stream, d, P, Wc, Gbc, Wbw, sGw, aw.part);
            const std::vector<double> f = sum_parts(*this, aw.part, nbc, 10);
            for (int k = 0; k < 10; ++k) fsb[k] += f[size_t(k)];
            LAUNCH("adj_gather_grad", k_adj_gather_grad, dim3(nbc), dim3(256), 0, stream, d, static_cast<const int32_t *>(aw.fsl), sGc, Wbw);
        }
        LAUNCH("adj_current", k_adj_current, dim3(nbc), dim3(256), 0, stream, d, static_cast<const double *>(aw.Jb), aw.Bib);
        // the CFL minimum: its cell's primitives and field
        if (dtb != 0.0 && dt < dt_max)
        {
            LAUNCH("adj_dt_argmin", k_adj_dt_argmin, dim3(nbc), dim3(256), 0, stream, d, P, Wc, static_cast<const double *>(aw.Btot), aw.part,
                   aw.argc);
            std::vector<double> pv(static_cast<size_t>(nbc));
            std::vector<int32_t> pc(static_cast<size_t>(nbc));
            d2h(pv.data(), aw.part, 8 * pv.size());
            d2h(pc.data(), aw.argc, 4 * pc.size());
            double best = __builtin_inf();
            int32_t cell = -1;
            for (int b = 0; b < nbc; ++b)
                if (pv[size_t(b)] < best || (pv[size_t(b)] == best && pc[size_t(b)] < cell)) { best = pv[size_t(b)]; cell = pc[size_t(b)]; }
            if (cell < 0 || best != dt) throw std::runtime_error("vjp: CFL minimum cell not found");
            LAUNCH("adj_dt_cell", k_adj_dt_cell, dim3(1), dim3(1), 0, stream, d, P, cell, Wc, static_cast<const double *>(aw.Btot), dtb, aw.wd,
                   Wbw, aw.Btb);
        }
        // prepare
        LAUNCH("adj_prepare", k_adj_prepare, dim3(nbc), dim3(256), 0, stream, d, P, Uck, Wc, Wbc, static_cast<const double *>(aw.upb),
               static_cast<const double *>(aw.Btb), static_cast<const double *>(aw.Bib), Ubw, aw.Baddb, aw.sPhi);
        LAUNCH("adj_gather_phi", k_adj_gather_phi, dim3(nbf), dim3(256), 0, stream, d, static_cast<const int32_t *>(aw.fsl),
               static_cast<const double *>(aw.sPhi), aw.Phib);
        check_launch("vjp reverse step");
    }
    chk(hipStreamSynchronize(stream), "sync");
    if (fs_b)
        for (int k = 0; k < 10; ++k) fs_b[k] = fsb[k];
    if (einf_b)
        for (int k = 0; k < 3; ++k) einf_b[k] = eib[k];
    (void)E;
    return Jv;
}
