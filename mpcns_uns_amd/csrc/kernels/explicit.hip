// Explicit stage of the Hall-MHD step: cell preparation + CFL, limited gradients + cell current, face
// fluxes + face EMF candidates, deterministic cell owner-gather + sources + forward Euler + floors, edge UCT
// EMF, constrained transport and div B.  Arithmetic mirrors the reference line by line (cited per kernel)
// so that results agree bit for bit with the CPU code built without FP contraction.
#include "fv.h"
#include "kernels.h"

namespace mhdg
{

// Solver::prepare (MercuryUns/driver/Solver.cpp:94-125) + cell_time_steps + n_e,eff / alpha of the
// frozen Hall state (Solver.cpp:257-260, ImplicitInduction.cpp:105-111), for P.ns ion species (u+ and n_e summed
// in species order: for the two Mercury ions the reference's expressions).  One thread per cell.
template <class R> __global__ void __launch_bounds__(256) k_prepare(DevMesh d, Params P, SpCT<R> Us, const R *Phi, const R *Badd, SpWT<R> Ws,
                                                                    R *Bind, R *Btot, R *uplus, R *alpha, R *dt_part)
{
    __shared__ R sh[256];
    const int32_t c = gid();
    R dtc = __builtin_inf();
    if (c < d.nc)
    {
        for (int s = 0; s < P.ns; ++s)
        {
            const PrimT<R> a = to_prim(Us.p[s] + 5 * size_t(c), P.gamma, P.rho_floor, P.p_floor);
            R *w = Ws.p[s] + 5 * size_t(c);
            w[0] = a.rho; w[1] = a.u; w[2] = a.v; w[3] = a.w; w[4] = a.p;
        }
        // the species rows through the kernel argument's pointer table itself (a local copy of the table, indexed by
        // a run-time species number, went to scratch memory)
        const R *const *Wc = Ws.p;
        // up to kGradFast faces: the face rows are loaded once, together, for the cell vector and the time step
        const int32_t k0 = d.c2f_ptr[c], k1 = d.c2f_ptr[c + 1], nk = k1 - k0;
        const bool fast = kRowsInRegisters<R> && nk <= kGradFast;
        double Ak[kGradFast];
        V3 ak[kGradFast];
        V3T<R> bi;
        if (fast)
        {
            int32_t vk[kGradFast];
            R fl[kGradFast];
#pragma unroll
            for (int i = 0; i < kGradFast; ++i)
                if (i < nk) vk[i] = d.c2f[k0 + i];
#pragma unroll
            for (int i = 0; i < kGradFast; ++i)
                if (i < nk)
                {
                    const int32_t f = pid(vk[i]);
                    Ak[i] = d.famag[f]; ak[i] = L3(d.farea, f); fl[i] = Phi[f];
                }
            R M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, r[3] = {0, 0, 0};
#pragma unroll
            for (int i = 0; i < kGradFast; ++i)
                if (i < nk) cell_vector_add(M, r, vk[i], R(Ak[i]), ak[i], fl[i]);
            bi = cell_vector_solve(M, r);
        }
        else
            bi = cell_vector(d, c, Phi);
        const V3T<R> ba = L3(Badd, c);
        const V3T<R> bt(bi.x + ba.x, bi.y + ba.y, bi.z + ba.z);
        st3(Bind, c, bi);
        st3(Btot, c, bt);
        V3T<R> up;
        R al = 0.0;
        if (d.mat[c] == 0)
        {
            const R ne = electron_density(P, Wc, c);
            const R ne_eff = sqrt(ne * ne + P.ne_floor * P.ne_floor);
            for (int s = 0; s < P.ns; ++s)
            {
                const R *w = Wc[s] + 5 * size_t(c);
                const R chi = ne > 1e-300 ? ion_density(P, s, w[0]) / ne : (s == 0 ? R(1.0) : R(0.0));
                if (s == 0) up = V3T<R>(chi * w[1], chi * w[2], chi * w[3]);
                else up = V3T<R>(up.x + chi * w[1], up.y + chi * w[2], up.z + chi * w[3]);
            }
            al = (P.hall && ne_eff > 0) ? d.taper[c] * P.hall_coef / ne_eff : 0.0;
            dtc = fast ? cell_dt_rows<R, kGradFast>(P, d.cvol[c], Wc, bt, size_t(c), nk, Ak, ak) : cell_dt(d, P, c, Wc, bt, size_t(c));
        }
        st3(uplus, c, up);
        alpha[c] = al;
    }
    const R m = block_min(dtc, sh);
    if (threadIdx.x == 0) dt_part[blockIdx.x] = m;
}

// Final min of the block minima; out[0] = min(dt, dt_max).
template <class R> __global__ void __launch_bounds__(256) k_dt_final(const R *part, int32_t n, double dt_max, R *out)
{
    __shared__ R sh[256];
    R v = __builtin_inf();
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x) v = mn(v, part[i]);
    v = block_min(v, sh);
    if (threadIdx.x == 0) out[0] = mn(v, dt_max);
}

// Stencil neighbour state of a boundary entry (nb < 0, face f = -(nb + 1), c2f entry v): the ghost table row, or
// ghost_state (Common/fluid/Reconstruction.cpp:9-24) with the cell value wc inside.
template <class R>
__device__ __forceinline__ void stencil_ghost(const DevMesh &d, int32_t sp, const R *gh, const R *wc, const R *fs, int32_t v, int32_t nb,
                                              R *wn)
{
    const int32_t f = -(nb + 1);
    if (gh)
    {
        for (int m = 0; m < 5; ++m) wn[m] = gh[5 * size_t(f) + m];
        return;
    }
    const int8_t kind = d.kind[f];
    for (int m = 0; m < 5; ++m) wn[m] = wc[m];
    if (kind == kWall)
    {
        if (sp == 1) { wn[0] = 1e-12; wn[1] = 0; wn[2] = 0; wn[3] = 0; wn[4] = 1e-12; }
        else
        {
            const V3T<R> n = (psg(v) / d.famag[f]) * L3(d.farea, f);
            const R un = wc[1] * n.x + wc[2] * n.y + wc[3] * n.z;
            wn[1] -= 2 * un * n.x; wn[2] -= 2 * un * n.y; wn[3] -= 2 * un * n.z;
        }
    }
    else if (kind == kInflow)
        for (int m = 0; m < 5; ++m) wn[m] = fs[m];
}

// Stencil::limited_gradients (Common/fluid/Reconstruction.cpp:80-142) for `ns` species and cell_current
// (Common/induction/Current.cpp:9-31, by the blocks of grid row 0).  ghost: optional per-face table of 5
// primitives per species.
// Each block stages its cells' 15-value gradient rows in LDS and writes them as one contiguous range (the
// per-thread 120-byte rows written directly left partially written cache lines behind: 2.2x write traffic).
// Cells of at most kGradFast faces load their face and stencil rows once, together, before the gathers, and
// the limiter's face offsets once for every species; the operations and their order are those of the general
// path.
template <class R, bool HOIST> __global__ void __launch_bounds__(256, 3) k_grad_current(DevMesh d, Params P, int32_t ns, int32_t s0, SpCT<R> Ws, SpCT<R> ghosts, SpWT<R> grads,
                                                      int32_t limit, const R *Bind, R *J)
{
    __shared__ R stage[256 * 15];
    const int32_t c = gid();
    const int32_t c_first = blockIdx.x * blockDim.x;
    const int32_t n_here = min(int32_t(blockDim.x), d.nc - c_first);
    const bool act = c < d.nc;
    const int32_t k0 = act ? d.c2f_ptr[c] : 0, k1 = act ? d.c2f_ptr[c + 1] : 0;
    const int32_t nk = k1 - k0;
    const bool fast = kRowsInRegisters<R> && act && nk <= kGradFast;
    int32_t vk[kGradFast];
    if (fast)
#pragma unroll
        for (int i = 0; i < kGradFast; ++i)
            if (i < nk) vk[i] = d.c2f[k0 + i];
    if (act && J && blockIdx.y == 0)
    {
        V3T<R> sum;
        const V3T<R> bc = L3(Bind, c);
        if (fast)
        {
            int32_t o[kGradFast];
            V3 A[kGradFast];
            V3T<R> bo[kGradFast];
#pragma unroll
            for (int i = 0; i < kGradFast; ++i)
                if (i < nk)
                {
                    o[i] = d.c2n[k0 + i];
                    A[i] = L3(d.farea, pid(vk[i]));
                }
#pragma unroll
            for (int i = 0; i < kGradFast; ++i)
                if (i < nk) bo[i] = L3(Bind, o[i] >= 0 ? o[i] : c);
#pragma unroll
            for (int i = 0; i < kGradFast; ++i)
                if (i < nk)
                {
                    V3T<R> bf = bc;
                    if (o[i] >= 0) bf = 0.5 * (bc + bo[i]);
                    sum += cross(psg(vk[i]) * A[i], bf);
                }
        }
        else
            for (int32_t k = k0; k < k1; ++k)
            {
                const int32_t v = d.c2f[k], f = pid(v);
                V3T<R> bf = bc;
                const int32_t l = d.fL[f], r = d.fR[f];
                if (l >= 0 && r >= 0) bf = 0.5 * (bc + L3(Bind, l == c ? r : l));
                sum += cross(psg(v) * L3(d.farea, f), bf);
            }
        const R invV = 1.0 / d.cvol[c];
        st3(J, c, V3T<R>(sum.x * invV, sum.y * invV, sum.z * invV));
    }
    const V3T<R> xc = act ? L3(d.ccen, c) : V3T<R>();
    V3T<R> dxk[kGradFast];
    if (fast && limit)
#pragma unroll
        for (int i = 0; i < kGradFast; ++i)
            if (i < nk) dxk[i] = L3(d.fcen, pid(vk[i])) - xc;
    // HOIST (many species): the stencil rows too are loaded once for every species
    int32_t nbh[kGradFast];
    double swh[kGradFast];
    V3 sdh[kGradFast];
    if (HOIST && fast)
#pragma unroll
        for (int i = 0; i < kGradFast; ++i)
            if (i < nk) { nbh[i] = d.snb[k0 + i]; swh[i] = d.sw[k0 + i]; sdh[i] = L3(d.sd, k0 + i); }
    // species s of the block: blockIdx.y, blockIdx.y + gridDim.y, ... (one grid row, every species, measured
    // faster than one row per species)
    for (int32_t s = int32_t(blockIdx.y); s < ns; s += int32_t(gridDim.y))
    {
        const int32_t sp = s0 + s;
        R *out = stage + 15 * threadIdx.x;
        if (!act || !d.svalid[c])
            for (int q = 0; q < 15; ++q) out[q] = 0.0;
        else
        {
        const R *W = Ws.p[s];
        const R *gh = ghosts.p[s];
        R fs[5];
        for (int m = 0; m < 5; ++m) fs[m] = seeded<R>(P.fs[sp][m], P.dfs[sp][m]);
        R wc[5], r[5][3], lo[5], hi[5];
        for (int m = 0; m < 5; ++m) { wc[m] = W[5 * size_t(c) + m]; lo[m] = hi[m] = wc[m]; r[m][0] = r[m][1] = r[m][2] = 0.0; }
        auto accumulate = [&](int32_t k, const R *wn) {
            const R w = HOIST && fast ? R(swh[k - k0]) : R(d.sw[k]);
            const V3T<R> dd = HOIST && fast ? V3T<R>(sdh[k - k0]) : V3T<R>(L3(d.sd, k));
            for (int m = 0; m < 5; ++m)
            {
                const R dq = wn[m] - wc[m];
                r[m][0] += w * dd.x * dq; r[m][1] += w * dd.y * dq; r[m][2] += w * dd.z * dq;
                lo[m] = mn(lo[m], wn[m]);
                hi[m] = mx(hi[m], wn[m]);
            }
        };
        if (fast)
        {
            int32_t nbk[kGradFast];
#pragma unroll
            for (int i = 0; i < kGradFast; ++i)
                if (i < nk) nbk[i] = HOIST ? nbh[i] : d.snb[k0 + i];
#pragma unroll
            for (int i = 0; i < kGradFast; ++i)
                if (i < nk && nbk[i] != kNoStencil)
                {
                    R wn[5];
                    if (nbk[i] >= 0)
                        for (int m = 0; m < 5; ++m) wn[m] = W[5 * size_t(nbk[i]) + m];
                    else
                        stencil_ghost(d, sp, gh, wc, fs, vk[i], nbk[i], wn);
                    accumulate(k0 + i, wn);
                }
        }
        else
            for (int32_t k = k0; k < k1; ++k)
            {
                const int32_t nb = d.snb[k];
                if (nb == kNoStencil) continue;
                R wn[5];
                if (nb >= 0)
                    for (int m = 0; m < 5; ++m) wn[m] = W[5 * size_t(nb) + m];
                else
                    stencil_ghost(d, sp, gh, wc, fs, d.c2f[k], nb, wn);
                accumulate(k, wn);
            }
        const double *I6 = d.sinv + 6 * size_t(c);
        const R I[9] = {I6[0], I6[1], I6[2], I6[1], I6[3], I6[4], I6[2], I6[4], I6[5]};
        for (int m = 0; m < 5; ++m)
        {
            const R g[3] = {I[0] * r[m][0] + I[1] * r[m][1] + I[2] * r[m][2], I[3] * r[m][0] + I[4] * r[m][1] + I[5] * r[m][2],
                                 I[6] * r[m][0] + I[7] * r[m][1] + I[8] * r[m][2]};
            const bool rel = limit && P.rel_limit > 0 && (m == 0 || m == 4);
            R psi = 1.0;
            if (fast)
            {
                if (limit)
#pragma unroll
                    for (int i = 0; i < kGradFast; ++i)
                        if (i < nk)
                        {
                            const R dl = g[0] * dxk[i].x + g[1] * dxk[i].y + g[2] * dxk[i].z;
                            if (dl > 1e-300) psi = mn(psi, (hi[m] - wc[m]) / dl);
                            else if (dl < -1e-300) psi = mn(psi, (lo[m] - wc[m]) / dl);
                        }
                if (rel)
#pragma unroll
                    for (int i = 0; i < kGradFast; ++i)
                        if (i < nk)
                        {
                            const R dl = fabs(g[0] * dxk[i].x + g[1] * dxk[i].y + g[2] * dxk[i].z);
                            if (dl > 1e-300) psi = mn(psi, P.rel_limit * wc[m] / dl);
                        }
            }
            else
            {
                if (limit)
                    for (int32_t k = k0; k < k1; ++k)
                    {
                        const V3T<R> dx = L3(d.fcen, pid(d.c2f[k])) - xc;
                        const R dl = g[0] * dx.x + g[1] * dx.y + g[2] * dx.z;
                        if (dl > 1e-300) psi = mn(psi, (hi[m] - wc[m]) / dl);
                        else if (dl < -1e-300) psi = mn(psi, (lo[m] - wc[m]) / dl);
                    }
                if (rel)
                    for (int32_t k = k0; k < k1; ++k)
                    {
                        const V3T<R> dx = L3(d.fcen, pid(d.c2f[k])) - xc;
                        const R dl = fabs(g[0] * dx.x + g[1] * dx.y + g[2] * dx.z);
                        if (dl > 1e-300) psi = mn(psi, P.rel_limit * wc[m] / dl);
                    }
            }
            psi = mx(0.0, mn(1.0, psi));
            out[3 * m] = psi * g[0]; out[3 * m + 1] = psi * g[1]; out[3 * m + 2] = psi * g[2];
        }
        }
        __syncthreads();
        R *dst = grads.p[s] + 15 * size_t(c_first);
        for (int32_t i = threadIdx.x; i < 15 * n_here; i += blockDim.x) dst[i] = stage[i];
        __syncthreads();
    }
}


// mercury::gyrate_n (MercuryUns/physics/Gyration.cpp:17-51) for the two ions over dt = scale * dt_dev[0]: the relative
// velocity's component normal to B rotates by theta = (a_H chi_Na + a_Na chi_H) |B| dt, a_s = induce_coeff M_ref / M_s,
// shared by the ions in proportion to a_H chi_Na and a_Na chi_H; momentum and kinetic energy follow, thermal energy
// stays.  The operation order is the reference's; cos and sin are the device library's.
__global__ void __launch_bounds__(256) k_gyrate(DevMesh d, Params P, double *UH, double *UNa, const double *Btot, const double *dt_dev,
                                                double scale)
{
    const int32_t c = gid();
    if (c >= d.nc || d.mat[c] != 0) return;
    const double aH = P.induce_coeff * P.M_ref / P.M_s[0], aN = P.induce_coeff * P.M_ref / P.M_s[1];
    double *uh = UH + 5 * size_t(c), *un = UNa + 5 * size_t(c);
    const Prim wh = to_prim(uh, P.gamma, P.rho_floor, P.p_floor), wn = to_prim(un, P.gamma, P.rho_floor, P.p_floor);
    // Scales::numinfo on the conservative densities
    const double nH = mx(uh[0], 0.0) / P.M_s[0] * P.M_ref, nN = mx(un[0], 0.0) / P.M_s[1] * P.M_ref;
    const double ne = nH + nN;
    const double chiH = ne > 1e-300 ? nH / ne : 1.0, chiN = ne > 1e-300 ? nN / ne : 0.0;
    const V3 b = L3(Btot, c);
    const double bm = norm(b);
    const double denom = aH * chiN + aN * chiH;
    if (bm <= 0 || denom <= 0) return;
    const V3 bh = (1.0 / bm) * b;
    const V3 uH(wh.u, wh.v, wh.w), uN(wn.u, wn.v, wn.w);
    const V3 w = uH - uN;
    const V3 w_par = dot(w, bh) * bh, w_perp = w - w_par;
    const double th = denom * bm * (scale * dt_dev[0]);
    const V3 dw = (cos(th) - 1.0) * w_perp + sin(th) * cross(w_perp, bh);
    const double fH = aH * chiN / denom, fN = aN * chiH / denom;
    const V3 dH = fH * dw, dN = -fN * dw;
    const V3 uH2 = uH + dH, uN2 = uN + dN;
    auto set = [](double *u, const V3 &v0, const V3 &v1) {
        const double rho = u[0];
        u[1] = rho * v1.x; u[2] = rho * v1.y; u[3] = rho * v1.z;
        u[4] += 0.5 * rho * (dot(v1, v1) - dot(v0, v0));
    };
    set(uh, uH, uH2);
    set(un, uN, uN2);
}

namespace
{
// Rusanov flux F*|A| of species slot s (species sp) through non-skipped face f with at least one fluid cell
// (Common/fluid/FluxRHS.cpp:43-70)
template <class R>
__device__ __forceinline__ void face_species_flux(const DevMesh &d, const Params &P, int32_t f, int32_t l, int32_t r, int8_t kind, const R &A,
                                                  const V3T<R> &n, int32_t s, int32_t sp, const SpCT<R> &Ws, const SpCT<R> &grads,
                                                  const SpCT<R> &ghosts, R *out)
{
    const R nv[3] = {n.x, n.y, n.z};
    const V3T<R> xf = L3(d.fcen, f);
    const R *W = Ws.p[s], *G = grads.p[s];
    PrimT<R> L, Rt;
    if (l >= 0 && r >= 0 && kind == kInterior)
    {
        L = face_state(W, G, l, xf - L3(d.ccen, l));
        Rt = face_state(W, G, r, xf - L3(d.ccen, r));
    }
    else
    {
        // the single fluid cell; f2c order puts the lower cell id first (FluxRHS.cpp:57-62)
        int32_t c;
        R sg;
        if (l >= 0 && r >= 0)
        {
            const int32_t first = l < r ? l : r, second = l < r ? r : l;
            c = d.mat[first] == 1 ? second : first;
        }
        else c = l >= 0 ? l : r;
        sg = c == l ? 1.0 : -1.0;
        const PrimT<R> in = face_state(W, G, c, xf - L3(d.ccen, c));
        PrimT<R> gh;
        if (ghosts.p[s])
        {
            const R *q = ghosts.p[s] + 5 * size_t(f);
            gh = PrimT<R>{q[0], q[1], q[2], q[3], q[4]};
        }
        else
        {
            R fsv[5];
            for (int m = 0; m < 5; ++m) fsv[m] = seeded<R>(P.fs[sp][m], P.dfs[sp][m]);
            gh = ghost_state(kind, sp, in, sg * n, fsv);
        }
        if (sg > 0) { L = in; Rt = gh; } else { Rt = in; L = gh; }
    }
    R F[5];
    rusanov(L, Rt, nv, P.gamma, F);
    for (int m = 0; m < 5; ++m) out[m] = F[m] * A;
}
} // namespace

// Face fluxes of many species, one thread per (face, species) with the species fastest: a block's threads
// write its faces' 5*ns flux values as one contiguous range (FA layout of k_face).  The face EMF candidate
// is left to k_face.
template <class R> __global__ void __launch_bounds__(256) k_face_species(DevMesh d, Params P, int32_t ns, SpCT<R> Ws, SpCT<R> grads, R *FA)
{
    const int32_t per = int32_t(blockDim.x) / ns;
    const int32_t t = int32_t(threadIdx.x);
    const int32_t f = int32_t(blockIdx.x) * per + t / ns, s = t % ns;
    if (t >= per * ns || f >= d.nf) return;
    const int32_t l = d.fL[f], r = d.fR[f];
    const int8_t kind = d.kind[f];
    const bool any = (l >= 0 && d.mat[l] == 0) || (r >= 0 && d.mat[r] == 0);
    R *out = FA + size_t(5 * ns) * f + 5 * s;
    if (kind == kSkip || !any)
    {
        for (int m = 0; m < 5; ++m) out[m] = 0.0;
        return;
    }
    const R A = d.famag[f];
    const V3T<R> n = (1.0 / A) * L3(d.farea, f);
    SpCT<R> none{};
    face_species_flux(d, P, f, l, r, kind, A, n, s, s, Ws, grads, none, out);
}

// Face kernel: fluid_flux_rhs's face body (Common/fluid/FluxRHS.cpp:36-73) for `ns` species, storing
// F*|A| per face and species (FA, 5*ns per face, zero for skipped faces), and optionally the ideal EMF
// face candidate E_f (Common/induction/IdealEMF.cpp:32-55, 3 per face).
template <class R> __global__ void __launch_bounds__(256) k_face(DevMesh d, Params P, int32_t ns, int32_t s0, SpCT<R> Ws, SpCT<R> grads, SpCT<R> ghosts, R *FA,
                                              const R *uplus, const R *Btot, const R *Bind, const R *Phi,
                                              const R *PhiAdd, R *Ef)
{
    // up to two species the block's flux rows are staged in LDS and written as one contiguous range
    __shared__ R stage[256 * 10];
    const int32_t f = gid();
    const bool act = f < d.nf;
    const bool staged = FA && ns <= 2;
    const int32_t l = act ? d.fL[f] : -1, r = act ? d.fR[f] : -1;
    const R A = act ? d.famag[f] : 1.0;
    const V3T<R> n = act ? (1.0 / A) * L3(d.farea, f) : V3T<R>();
    // the EMF candidate first: its gathers need only l and r
    if (Ef && act)
    {
        const R Bn_tot = (PhiAdd ? Phi[f] + PhiAdd[f] : Phi[f]) / A;
        V3T<R> F;
        if (l >= 0 && r >= 0)
        {
            const V3T<R> uL = L3(uplus, l), uR = L3(uplus, r);
            const V3T<R> fLx = dot(uL, n) * L3(Btot, l) - Bn_tot * uL;
            const V3T<R> fRx = dot(uR, n) * L3(Btot, r) - Bn_tot * uR;
            const R a = mx(fabs(dot(uL, n)), fabs(dot(uR, n)));
            F = 0.5 * (fLx + fRx) - (0.5 * a) * (L3(Bind, r) - L3(Bind, l));
        }
        else
        {
            const int32_t c = l >= 0 ? l : r;
            const V3T<R> u = L3(uplus, c);
            F = dot(u, n) * L3(Btot, c) - Bn_tot * u;
        }
        st3(Ef, f, -1.0 * cross(n, F));
    }
    if (FA && act)
    {
        const int8_t kind = d.kind[f];
        const bool any = (l >= 0 && d.mat[l] == 0) || (r >= 0 && d.mat[r] == 0);
        R *out = staged ? stage + 5 * ns * threadIdx.x : FA + size_t(5 * ns) * f;
        if (kind == kSkip || !any)
            for (int q = 0; q < 5 * ns; ++q) out[q] = 0.0;
        else
            for (int32_t s = 0; s < ns; ++s) face_species_flux(d, P, f, l, r, kind, A, n, s, s0 + s, Ws, grads, ghosts, out + 5 * s);
    }
    if (staged)
    {
        __syncthreads();
        const int32_t f_first = blockIdx.x * blockDim.x, n_here = min(int32_t(blockDim.x), d.nf - f_first);
        R *dst = FA + size_t(5 * ns) * f_first;
        for (int32_t i = threadIdx.x; i < 5 * ns * n_here; i += blockDim.x) dst[i] = stage[i];
    }
}

// Cell owner-gather of the face fluxes in ascending face order (the reference's scatter order,
// FluxRHS.cpp:71-72), Lorentz sources (MercurySources.cpp:20-63), and - when dt != 0 - the forward Euler
// update with floors (Solver.cpp:246-252, 171-191).  rhs_out (nullable) receives the rhs.
// FA nullable (no flux term), rhs_in nullable (start from zero); species of slot s is s0 + s.
template <class R> __global__ void __launch_bounds__(256) k_cell_update(DevMesh d, Params P, int32_t ns, int32_t s0, const R *FA, SpCT<R> rhs_in,
                                                     SpCT<R> Wall, const R *Btot, const R *J,
                                                     const R *dt_dev, double dt_val, int32_t sources, int32_t update, SpWT<R> Us,
                                                     SpWT<R> rhs_out)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    const bool fluid = d.mat[c] == 0;
    const R dt = dt_dev ? dt_dev[0] : dt_val;
    const R V = d.cvol[c];
    // n_e of the cell's ions, once for every species' source
    const R ne_c = (sources && fluid) ? electron_density(P, Wall.p, c) : R(0.0);
    // up to kGradFast faces: the face rows are loaded once, and each species' flux rows together before the sums
    const int32_t k0 = d.c2f_ptr[c], k1 = d.c2f_ptr[c + 1], nk = k1 - k0;
    const bool fast = kRowsInRegisters<R> && nk <= kGradFast;
    int32_t vk[kGradFast];
    if (fast && fluid && FA)
#pragma unroll
        for (int i = 0; i < kGradFast; ++i)
            if (i < nk) vk[i] = d.cfs[k0 + i];
    for (int32_t s = 0; s < ns; ++s)
    {
        const int32_t sp = s0 + s;
        R rhs[5] = {0, 0, 0, 0, 0};
        if (rhs_in.p[s])
            for (int m = 0; m < 5; ++m) rhs[m] = rhs_in.p[s][5 * size_t(c) + m];
        if (fluid)
        {
            if (FA && fast)
            {
                R fa[kGradFast][5];
#pragma unroll
                for (int i = 0; i < kGradFast; ++i)
                    if (i < nk)
                        for (int m = 0; m < 5; ++m) fa[i][m] = FA[size_t(5 * ns) * pid(vk[i]) + 5 * s + m];
#pragma unroll
                for (int i = 0; i < kGradFast; ++i)
                    if (i < nk)
                    {
                        if (vk[i] >= 0) for (int m = 0; m < 5; ++m) rhs[m] -= fa[i][m] / V;
                        else for (int m = 0; m < 5; ++m) rhs[m] += fa[i][m] / V;
                    }
            }
            else if (FA)
            for (int32_t k = k0; k < k1; ++k)
            {
                const int32_t v = d.cfs[k], f = pid(v);
                const R *fa = FA + size_t(5 * ns) * f + 5 * s;
                if (v >= 0) for (int m = 0; m < 5; ++m) rhs[m] -= fa[m] / V;
                else for (int m = 0; m < 5; ++m) rhs[m] += fa[m] / V;
            }
            if (sources)
            {
                // mercury::add_sources for ion sp: w_s = n_s / n_e,eff over all P.ns ions (MercurySources.cpp:20-63)
                const R *W = Wall.p[sp];
                const R ne = ne_c;
                const R ws = ion_density(P, sp, W[5 * size_t(c)]) / sqrt(ne * ne + P.ne_floor * P.ne_floor);
                const R rho = mx(W[5 * size_t(c)], 0.0);
                const V3T<R> u(W[5 * size_t(c) + 1], W[5 * size_t(c) + 2], W[5 * size_t(c) + 3]);
                V3T<R> S;
                if (P.lorentz) S += (P.inver_MA2 * ws) * cross(L3(J, c), L3(Btot, c));
                rhs[1] += S.x; rhs[2] += S.y; rhs[3] += S.z;
                rhs[4] += dot(u, S) + (rho > 1e-30 ? 0.5 * dt * dot(S, S) / rho : 0.0);
                if (P.src_rate[sp] != 0.0 && d.atm)
                {
                    // exponential atmosphere (profile tabulated at set-up): ions produced at rest with temperature src_T
                    const double q = P.src_rate[sp] * d.atm[c];
                    rhs[0] += q;
                    rhs[4] += q * P.src_T / (P.gamma - 1.0);
                }
            }
        }
        if (rhs_out.p[s])
            for (int m = 0; m < 5; ++m) rhs_out.p[s][5 * size_t(c) + m] = rhs[m];
        if (update && fluid)
        {
            R *u = Us.p[s] + 5 * size_t(c);
            R q[5];
            for (int m = 0; m < 5; ++m) q[m] = u[m] + dt * rhs[m];
            const R ke = q[0] > 0 ? 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0] : 0.0;
            if (!(q[0] >= P.rho_floor)) to_cons(PrimT<R>{P.rho_floor, 0, 0, 0, P.p_floor}, P.gamma, q);
            else if ((P.gamma - 1.0) * (q[4] - ke) < P.p_floor)
            {
                const PrimT<R> w = to_prim(q, P.gamma, P.rho_floor, P.p_floor);
                to_cons(w, P.gamma, q);
            }
            for (int m = 0; m < 5; ++m) u[m] = q[m];
        }
    }
}


// k_cell_update of the fused step for more than two ions (rhs_in, rhs_out none, s0 = 0, sources and update on):
// every species' flux sum is gathered in one pass over the cell's faces, so each face row (5 ns values) is read
// once; with one pass per species the rows' lines were fetched again for every species (3x the algorithmic bytes
// for ten ions).  Each species' sum runs over the faces in the same ascending order, so the results are those of
// k_cell_update bit for bit.
template <class R> __global__ void __launch_bounds__(256) k_cell_update_many(DevMesh d, Params P, int32_t ns, const R *FA, SpCT<R> Wall,
                                                                            const R *Btot, const R *J, const R *dt_dev, SpWT<R> Us)
{
    const int32_t c = gid();
    if (c >= d.nc) return;
    const bool fluid = d.mat[c] == 0;
    const R dt = dt_dev[0];
    const R V = d.cvol[c];
    const R ne_c = fluid ? electron_density(P, Wall.p, c) : R(0.0);
    R rhs[kMaxSpecies][5];
#pragma unroll
    for (int s = 0; s < kMaxSpecies; ++s)
        for (int m = 0; m < 5; ++m) rhs[s][m] = 0;
    if (fluid)
        for (int32_t k = d.c2f_ptr[c]; k < d.c2f_ptr[c + 1]; ++k)
        {
            const int32_t v = d.cfs[k];
            const R *fa = FA + size_t(5 * ns) * pid(v);
#pragma unroll
            for (int s = 0; s < kMaxSpecies; ++s)
                if (s < ns)
                {
                    if (v >= 0) for (int m = 0; m < 5; ++m) rhs[s][m] -= fa[5 * s + m] / V;
                    else for (int m = 0; m < 5; ++m) rhs[s][m] += fa[5 * s + m] / V;
                }
        }
#pragma unroll
    for (int s = 0; s < kMaxSpecies; ++s)
        if (s < ns && fluid)
        {
            R *r = rhs[s];
            // mercury::add_sources for ion s (MercurySources.cpp:20-63), as in k_cell_update
            const R *W = Wall.p[s];
            const R ws = ion_density(P, s, W[5 * size_t(c)]) / sqrt(ne_c * ne_c + P.ne_floor * P.ne_floor);
            const R rho = mx(W[5 * size_t(c)], 0.0);
            const V3T<R> u(W[5 * size_t(c) + 1], W[5 * size_t(c) + 2], W[5 * size_t(c) + 3]);
            V3T<R> S;
            if (P.lorentz) S += (P.inver_MA2 * ws) * cross(L3(J, c), L3(Btot, c));
            r[1] += S.x; r[2] += S.y; r[3] += S.z;
            r[4] += dot(u, S) + (rho > 1e-30 ? 0.5 * dt * dot(S, S) / rho : 0.0);
            if (P.src_rate[s] != 0.0 && d.atm)
            {
                const double q = P.src_rate[s] * d.atm[c];
                r[0] += q;
                r[4] += q * P.src_T / (P.gamma - 1.0);
            }
            // forward Euler with floors (Solver.cpp:246-252, 171-191), as in k_cell_update
            R *u0 = Us.p[s] + 5 * size_t(c);
            R q[5];
            for (int m = 0; m < 5; ++m) q[m] = u0[m] + dt * r[m];
            const R ke = q[0] > 0 ? 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0] : 0.0;
            if (!(q[0] >= P.rho_floor)) to_cons(PrimT<R>{P.rho_floor, 0, 0, 0, P.p_floor}, P.gamma, q);
            else if ((P.gamma - 1.0) * (q[4] - ke) < P.p_floor)
            {
                const PrimT<R> w = to_prim(q, P.gamma, P.rho_floor, P.p_floor);
                to_cons(w, P.gamma, q);
            }
            for (int m = 0; m < 5; ++m) u0[m] = q[m];
        }
}
template __global__ void k_cell_update_many<double>(DevMesh d, Params P, int32_t ns, const double *FA, SpCT<double> Wall, const double *Btot,
                                                    const double *J, const double *dt_dev, SpWT<double> Us);

// add_ideal_emf edge loop (Common/induction/IdealEMF.cpp:59-128), one thread per edge with valence at most
// `vmax`; larger edges are left to k_edge_emf_heavy.  E_out[e] = E_in[e] + E_ideal (E_in nullable = 0);
// apply_bc additionally applies the boundary EMF of Solver::rhs.
template <class R> __global__ void __launch_bounds__(256) k_edge_emf(DevMesh d, Params P, const R *Ef, const R *u_cell, const R *phi_ind,
                                                  double cuct, int32_t vmax, const R *E_in, R *E_out, int32_t apply_bc)
{
    const int32_t e = gid();
    if (e >= d.ne) return;
    const int32_t f0 = d.e2f_ptr[e], f1 = d.e2f_ptr[e + 1];
    const int32_t nf = f1 - f0;
    if (nf > vmax) return;
    R out = E_in ? E_in[e] : 0.0;
    const int32_t c0 = d.ec_ptr[e], c1 = d.ec_ptr[e + 1];
    const int32_t nc = c1 - c0;
    const bool fast = nf <= kEmfFast && nc <= kEmfFast;
    // common valence: the row indices are loaded once, together, before their gathers
    int32_t fid[kEmfFast], cid[kEmfFast];
    bool any_fluid = false;
    if (fast)
    {
#pragma unroll
        for (int i = 0; i < kEmfFast; ++i)
        {
            if (i < nf) fid[i] = pid(d.e2f[f0 + i]);
            if (i < nc) cid[i] = d.ec_cell[c0 + i];
        }
#pragma unroll
        for (int k = 0; k < kEmfFast; ++k)
            if (k < nc) any_fluid |= d.mat[cid[k]] == 0;
    }
    else
        for (int32_t k = c0; k < c1 && !any_fluid; ++k) any_fluid = d.mat[d.ec_cell[k]] == 0;
    if (nf > 0 && d.elen[e] > 0 && any_fluid)
    {
        const V3 t = L3(d.evec, e);
        const double L = d.elen[e];
        const V3 th = (1.0 / L) * t;
        R Ec = 0;
        if (fast)
        {
            V3T<R> ef[kEmfFast];
#pragma unroll
            for (int i = 0; i < kEmfFast; ++i)
                if (i < nf) ef[i] = L3(Ef, fid[i]);
#pragma unroll
            for (int i = 0; i < kEmfFast; ++i)
                if (i < nf) Ec += d.efw[f0 + i] * dot(ef[i], t);
        }
        else
            for (int32_t i = f0; i < f1; ++i) Ec += d.efw[i] * dot(L3(Ef, pid(d.e2f[i])), t);
        V3 e1, e2;
        plane_basis(th, e1, e2);
        R M11 = 0, M12 = 0, M22 = 0, g1 = 0, g2 = 0, Eu = 0;
        if (fast)
        {
            // the edge's cell velocities and the per-face terms stay in registers, so the second pass re-reads
            // nothing (same operations in the same order as the general path)
            V3T<R> uc[kEmfFast];
#pragma unroll
            for (int k = 0; k < kEmfFast; ++k)
                if (k < nc) uc[k] = L3(u_cell, cid[k]);
            R fqa[kEmfFast], fqb[kEmfFast], faef[kEmfFast], fsig[kEmfFast];
#pragma unroll
            for (int i = 0; i < kEmfFast; ++i)
                if (i < nf)
                {
                    const int32_t f = fid[i];
                    V3 m, nrm;
                    face_dirs(d, f, e, th, m, nrm);
                    R aef = 0.0;
#pragma unroll
                    for (int k = 0; k < kEmfFast; ++k)
                        if (k < nc) aef = mx(aef, fabs(dot(uc[k], m)));
                    const R sig = dot(th, cross(m, nrm));
                    const V3T<R> qv = (aef * sig) * nrm;
                    const R qa = dot(qv, e1), qb = dot(qv, e2), w = d.efw[f0 + i];
                    M11 += w * qa * qa;
                    M12 += w * qa * qb;
                    M22 += w * qb * qb;
                    g1 += w * qa;
                    g2 += w * qb;
                    fqa[i] = qa; fqb[i] = qb; faef[i] = aef; fsig[i] = sig;
                }
            R l1, l2;
            uct_solve(M11, M12, M22, g1, g2, l1, l2);
            // the face weight and flux are loaded again (cache hits) rather than held across the solve
#pragma unroll
            for (int i = 0; i < kEmfFast; ++i)
                if (i < nf)
                {
                    const R omega = d.efw[f0 + i] * (1.0 - (fqa[i] * l1 + fqb[i] * l2));
                    Eu += omega * faef[i] * fsig[i] * (phi_ind[fid[i]] / d.famag[fid[i]]);
                }
        }
        else
        {
            for (int32_t i = f0; i < f1; ++i)
            {
                const int32_t f = pid(d.e2f[i]);
                V3 m, nrm;
                face_dirs(d, f, e, th, m, nrm);
                const R aef = edge_amax(d, u_cell, c0, c1, m);
                const V3T<R> qv = (aef * dot(th, cross(m, nrm))) * nrm;
                const R qa = dot(qv, e1), qb = dot(qv, e2), w = d.efw[i];
                M11 += w * qa * qa;
                M12 += w * qa * qb;
                M22 += w * qb * qb;
                g1 += w * qa;
                g2 += w * qb;
            }
            R l1, l2;
            uct_solve(M11, M12, M22, g1, g2, l1, l2);
            for (int32_t i = f0; i < f1; ++i)
            {
                const int32_t f = pid(d.e2f[i]);
                V3 m, nrm;
                face_dirs(d, f, e, th, m, nrm);
                const R aef = edge_amax(d, u_cell, c0, c1, m);
                const V3T<R> qv = (aef * dot(th, cross(m, nrm))) * nrm;
                const R qa = dot(qv, e1), qb = dot(qv, e2), w = d.efw[i];
                const R omega = w * (1.0 - (qa * l1 + qb * l2));
                const R sig = dot(th, cross(m, nrm));
                Eu += omega * aef * sig * (phi_ind[f] / d.famag[f]);
            }
        }
        out += Ec + 2.0 * cuct * L * Eu;
    }
    E_out[e] = apply_bc ? edge_bc(d, P, e, out) : out;
}

// The same for one high-valence edge per block (pole axes of a lat-lon net): per-face terms in parallel,
// every sum sequential in the reference's order.
template <class R> __global__ void __launch_bounds__(256) k_edge_emf_heavy(DevMesh d, Params P, const R *Ef, const R *u_cell,
                                                        const R *phi_ind, double cuct, const int32_t *heavy,
                                                        const R *E_in, R *E_out, int32_t apply_bc)
{
    constexpr int kHS = kHeavyShared<R>();
    __shared__ R s_w[kHS], s_ec[kHS], s_qa[kHS], s_qb[kHS], s_sig[kHS], s_aef[kHS];
    __shared__ R s_l[2];
    const int32_t e = heavy[blockIdx.x];
    const int32_t f0 = d.e2f_ptr[e], f1 = d.e2f_ptr[e + 1], nf = f1 - f0;
    const int32_t c0 = d.ec_ptr[e], c1 = d.ec_ptr[e + 1];
    bool any_fluid = false;
    for (int32_t k = c0; k < c1 && !any_fluid; ++k) any_fluid = d.mat[d.ec_cell[k]] == 0;
    const bool active = nf > 0 && d.elen[e] > 0 && any_fluid;
    const V3 t = L3(d.evec, e);
    const double L = d.elen[e];
    const V3 th = (1.0 / L) * t;
    V3 e1, e2;
    plane_basis(th, e1, e2);
    // The edge's cell velocities are staged in LDS (in s_qa, s_qb, s_sig, before those take the per-face terms):
    // every face's max over the cells then reads LDS, where a dependent global gather per (face, cell) pair made
    // this kernel latency-bound (n_lon^2 pairs per pole-axis edge).  Same values, same order.
    const int32_t nc = c1 - c0;
    const bool staged = active && nc <= kHS;
    if (staged)
        for (int32_t k = threadIdx.x; k < nc; k += blockDim.x)
        {
            const V3T<R> u = L3(u_cell, d.ec_cell[c0 + k]);
            s_qa[k] = u.x; s_qb[k] = u.y; s_sig[k] = u.z;
        }
    __syncthreads();
    constexpr int kPer = (kHS + 255) / 256; // faces per thread (nf <= kHS, 256 threads)
    R qa_r[kPer], qb_r[kPer], sig_r[kPer];
    if (active)
#pragma unroll
        for (int j = 0; j < kPer; ++j)
        {
            const int32_t i = int32_t(threadIdx.x) + j * int32_t(blockDim.x);
            if (i >= nf) break;
            const int32_t f = pid(d.e2f[f0 + i]);
            s_w[i] = d.efw[f0 + i];
            s_ec[i] = dot(L3(Ef, f), t);
            V3 m, nrm;
            face_dirs(d, f, e, th, m, nrm);
            R aef = 0.0;
            if (staged)
                for (int32_t k = 0; k < nc; ++k) aef = mx(aef, fabs(dot(V3T<R>(s_qa[k], s_qb[k], s_sig[k]), m)));
            else
                aef = edge_amax(d, u_cell, c0, c1, m);
            const R sig = dot(th, cross(m, nrm));
            const V3T<R> qv = (aef * sig) * nrm;
            qa_r[j] = dot(qv, e1);
            qb_r[j] = dot(qv, e2);
            sig_r[j] = sig;
            s_aef[i] = aef;
        }
    __syncthreads(); // every read of the staged velocities is done
    if (active)
#pragma unroll
        for (int j = 0; j < kPer; ++j)
        {
            const int32_t i = int32_t(threadIdx.x) + j * int32_t(blockDim.x);
            if (i >= nf) break;
            s_qa[i] = qa_r[j]; s_qb[i] = qb_r[j]; s_sig[i] = sig_r[j];
        }
    __syncthreads();
    R Ec = 0;
    if (threadIdx.x == 0 && active)
    {
        R M11 = 0, M12 = 0, M22 = 0, g1 = 0, g2 = 0;
        for (int32_t i = 0; i < nf; ++i)
        {
            Ec += s_w[i] * s_ec[i];
            M11 += s_w[i] * s_qa[i] * s_qa[i];
            M12 += s_w[i] * s_qa[i] * s_qb[i];
            M22 += s_w[i] * s_qb[i] * s_qb[i];
            g1 += s_w[i] * s_qa[i];
            g2 += s_w[i] * s_qb[i];
        }
        const R tr = M11 + M22;
        const R eps = 1e-14 * tr;
        const R det = (M11 + eps) * (M22 + eps) - M12 * M12;
        R l1 = 0, l2 = 0;
        if (tr > 0 && det > 0)
        {
            l1 = ((M22 + eps) * g1 - M12 * g2) / det;
            l2 = ((M11 + eps) * g2 - M12 * g1) / det;
        }
        s_l[0] = l1; s_l[1] = l2;
    }
    __syncthreads();
    if (active)
        for (int32_t i = threadIdx.x; i < nf; i += blockDim.x)
        {
            const int32_t f = pid(d.e2f[f0 + i]);
            const R omega = s_w[i] * (1.0 - (s_qa[i] * s_l[0] + s_qb[i] * s_l[1]));
            s_ec[i] = omega * s_aef[i] * s_sig[i] * (phi_ind[f] / d.famag[f]);
        }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        R out = E_in ? E_in[e] : 0.0;
        if (active)
        {
            R Eu = 0;
            for (int32_t i = 0; i < nf; ++i) Eu += s_ec[i];
            out += Ec + 2.0 * cuct * L * Eu;
        }
        E_out[e] = apply_bc ? edge_bc(d, P, e, out) : out;
    }
}

// faraday_update / faraday_update_edge_dt (Common/induction/Faraday.cpp:15-36): B -= dt d1 E.
template <class R> __global__ void __launch_bounds__(256) k_faraday(DevMesh d, const R *E, const double *dte, const R *dt_dev, double dt_val,
                                                 R *B)
{
    const int32_t f = gid();
    if (f >= d.nf) return;
    if (d.fL[f] < 0 && d.fR[f] < 0) return;
    R circ = 0.0;
    if (dte)
    {
        for (int32_t k = d.f2e_ptr[f]; k < d.f2e_ptr[f + 1]; ++k)
        {
            const int32_t v = d.f2e[k], e = pid(v);
            circ += psg(v) * dte[e] * E[e];
        }
        B[f] -= circ;
    }
    else
    {
        for (int32_t k = d.f2e_ptr[f]; k < d.f2e_ptr[f + 1]; ++k)
        {
            const int32_t v = d.f2e[k];
            circ += psg(v) * E[pid(v)];
        }
        const R dt = dt_dev ? dt_dev[0] : dt_val;
        B[f] -= dt * circ;
    }
}

// div_b (Common/diagnostics/DivB.cpp:8-24): per-block (max |s|, first cell attaining it, max rel).
__global__ void __launch_bounds__(256) k_div_b(DevMesh d, const double *B, const double *extra, double *part_abs, int32_t *part_cell,
                                               double *part_rel)
{
    __shared__ double sa[256], sr[256];
    __shared__ int32_t sc[256];
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    double va = 0, vr = 0;
    int32_t vc = INT32_MAX;
    if (c < d.nc)
    {
        double s = 0, a = 0;
        for (int32_t k = d.c2f_ptr[c]; k < d.c2f_ptr[c + 1]; ++k)
        {
            const int32_t v = d.c2f[k], f = pid(v);
            s += psg(v) * B[f];
            a += fabs(B[f]) + (extra ? fabs(extra[f]) : 0.0);
        }
        vr = a > 0 ? fabs(s) / a : 0.0;
        va = fabs(s);
        vc = va > 0 ? c : INT32_MAX;
    }
    const int t = threadIdx.x;
    sa[t] = va; sr[t] = vr; sc[t] = vc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1)
    {
        if (t < s)
        {
            if (sa[t + s] > sa[t] || (sa[t + s] == sa[t] && sc[t + s] < sc[t])) { sa[t] = sa[t + s]; sc[t] = sc[t + s]; }
            sr[t] = mx(sr[t], sr[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) { part_abs[blockIdx.x] = sa[0]; part_cell[blockIdx.x] = sc[0]; part_rel[blockIdx.x] = sr[0]; }
}
// reconstruct_cell_vector for every cell
__global__ void __launch_bounds__(256) k_cell_vector(DevMesh d, const double *flux, double *b)
{
    const int32_t c = gid();
    if (c < d.nc) st3(b, c, cell_vector(d, c, flux));
}

// cell_time_steps + block minima from given primitives (local_time_step, MercurySources.cpp:104-116)
__global__ void __launch_bounds__(256) k_local_dt(DevMesh d, Params P, SpC Ws, const double *Btot, double *dt_part)
{
    __shared__ double sh[256];
    const int32_t c = gid();
    double dtc = __builtin_inf();
    if (c < d.nc && d.mat[c] == 0) dtc = cell_dt(d, P, c, Ws.p, L3(Btot, c), size_t(c));
    const double m = block_min(dtc, sh);
    if (threadIdx.x == 0) dt_part[blockIdx.x] = m;
}

// |B|^2 V summed over the fluid cells with r0 <= |x_c| <= r1, B = reconstructed cell field of phi + B_add:
// per-block partial sums in a fixed tree order
template <class R>
__global__ void __launch_bounds__(256) k_shell_energy(DevMesh d, const R *phi, const R *badd, double r0, double r1, R *part)
{
    __shared__ R sh[256];
    const int32_t c = gid();
    R e = 0.0;
    if (c < d.nc && d.mat[c] == 0)
    {
        const V3 x = L3(d.ccen, c);
        const double r = norm(x);
        if (r >= r0 && r <= r1)
        {
            const V3T<R> b = cell_vector(d, c, phi) + L3(badd, c);
            e = dot(b, b) * d.cvol[c];
        }
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1)
    {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// the step (double) and its tangent-linear (Dual)
template __global__ void k_shell_energy<double>(DevMesh, const double *, const double *, double, double, double *);
template __global__ void k_shell_energy<Dual>(DevMesh, const Dual *, const Dual *, double, double, Dual *);
template __global__ void k_prepare<double>(DevMesh d, Params P, SpCT<double> Us, const double *Phi, const double *Badd, SpWT<double> Ws, double *Bind, double *Btot, double *uplus, double *alpha, double *dt_part);
template __global__ void k_dt_final<double>(const double *part, int32_t n, double dt_max, double *out);
template __global__ void k_grad_current<double, false>(DevMesh d, Params P, int32_t ns, int32_t s0, SpCT<double> Ws, SpCT<double> ghosts, SpWT<double> grads, int32_t limit, const double *Bind, double *J);
template __global__ void k_grad_current<double, true>(DevMesh d, Params P, int32_t ns, int32_t s0, SpCT<double> Ws, SpCT<double> ghosts, SpWT<double> grads, int32_t limit, const double *Bind, double *J);
template __global__ void k_face_species<double>(DevMesh d, Params P, int32_t ns, SpCT<double> Ws, SpCT<double> grads, double *FA);
template __global__ void k_face<double>(DevMesh d, Params P, int32_t ns, int32_t s0, SpCT<double> Ws, SpCT<double> grads, SpCT<double> ghosts, double *FA, const double *uplus, const double *Btot, const double *Bind, const double *Phi, const double *PhiAdd, double *Ef);
template __global__ void k_cell_update<double>(DevMesh d, Params P, int32_t ns, int32_t s0, const double *FA, SpCT<double> rhs_in, SpCT<double> Wall, const double *Btot, const double *J, const double *dt_dev, double dt_val, int32_t sources, int32_t update, SpWT<double> Us, SpWT<double> rhs_out);
template __global__ void k_edge_emf<double>(DevMesh d, Params P, const double *Ef, const double *u_cell, const double *phi_ind, double cuct, int32_t vmax, const double *E_in, double *E_out, int32_t apply_bc);
template __global__ void k_edge_emf_heavy<double>(DevMesh d, Params P, const double *Ef, const double *u_cell, const double *phi_ind, double cuct, const int32_t *heavy, const double *E_in, double *E_out, int32_t apply_bc);
template __global__ void k_faraday<double>(DevMesh d, const double *E, const double *dte, const double *dt_dev, double dt_val, double *B);
template __global__ void k_prepare<Dual>(DevMesh d, Params P, SpCT<Dual> Us, const Dual *Phi, const Dual *Badd, SpWT<Dual> Ws, Dual *Bind, Dual *Btot, Dual *uplus, Dual *alpha, Dual *dt_part);
template __global__ void k_dt_final<Dual>(const Dual *part, int32_t n, double dt_max, Dual *out);
template __global__ void k_grad_current<Dual, false>(DevMesh d, Params P, int32_t ns, int32_t s0, SpCT<Dual> Ws, SpCT<Dual> ghosts, SpWT<Dual> grads, int32_t limit, const Dual *Bind, Dual *J);
template __global__ void k_face<Dual>(DevMesh d, Params P, int32_t ns, int32_t s0, SpCT<Dual> Ws, SpCT<Dual> grads, SpCT<Dual> ghosts, Dual *FA, const Dual *uplus, const Dual *Btot, const Dual *Bind, const Dual *Phi, const Dual *PhiAdd, Dual *Ef);
template __global__ void k_cell_update<Dual>(DevMesh d, Params P, int32_t ns, int32_t s0, const Dual *FA, SpCT<Dual> rhs_in, SpCT<Dual> Wall, const Dual *Btot, const Dual *J, const Dual *dt_dev, double dt_val, int32_t sources, int32_t update, SpWT<Dual> Us, SpWT<Dual> rhs_out);
template __global__ void k_edge_emf<Dual>(DevMesh d, Params P, const Dual *Ef, const Dual *u_cell, const Dual *phi_ind, double cuct, int32_t vmax, const Dual *E_in, Dual *E_out, int32_t apply_bc);
template __global__ void k_edge_emf_heavy<Dual>(DevMesh d, Params P, const Dual *Ef, const Dual *u_cell, const Dual *phi_ind, double cuct, const int32_t *heavy, const Dual *E_in, Dual *E_out, int32_t apply_bc);
template __global__ void k_faraday<Dual>(DevMesh d, const Dual *E, const double *dte, const Dual *dt_dev, double dt_val, Dual *B);
} // namespace mhdg
