// ORACLE — test infrastructure only.  CPU FP64 restatement of the MPCNS_Uns Hall-MHD step (one rank:
// every entity owned, no halo), used as the parity checker of the device path and as the CPU baseline.
// Never linked into or called by the product.
//
// Each function cites the reference file:line it restates and evaluates in the reference's order
// (entities ascending, CSR rows in stored order, left-to-right arithmetic, no contraction: build with
// -ffp-contract=off), so that with identical inputs it reproduces the reference bit for bit.  Loops over
// independent entities run under OpenMP; every sum and every reduction with an order stays serial.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/mhdg.h"

namespace
{
struct V3
{
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3 &operator-=(const V3 &o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    V3 &operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
};
V3 operator+(V3 a, const V3 &b) { return a += b; }
V3 operator-(V3 a, const V3 &b) { return a -= b; }
V3 operator*(V3 a, double s) { return a *= s; }
V3 operator*(double s, V3 a) { return a *= s; }
double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(const V3 &a, const V3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(const V3 &a) { return std::sqrt(dot(a, a)); }
V3 L3(const double *p, long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
V3 L3(const std::vector<double> &p, long i) { return L3(p.data(), i); }

struct Prim { double rho = 1, u = 0, v = 0, w = 0, p = 1; };

// Common/physics/Euler.h:13-49
Prim to_prim(const double *U, double gamma, double rho_floor, double p_floor)
{
    Prim W;
    W.rho = std::max(U[0], rho_floor);
    W.u = U[1] / W.rho; W.v = U[2] / W.rho; W.w = U[3] / W.rho;
    const double ke = 0.5 * W.rho * (W.u * W.u + W.v * W.v + W.w * W.w);
    W.p = std::max((gamma - 1.0) * (U[4] - ke), p_floor);
    return W;
}
void to_cons(const Prim &W, double gamma, double *U)
{
    U[0] = W.rho; U[1] = W.rho * W.u; U[2] = W.rho * W.v; U[3] = W.rho * W.w;
    U[4] = W.p / (gamma - 1.0) + 0.5 * W.rho * (W.u * W.u + W.v * W.v + W.w * W.w);
}
double sound(const Prim &W, double gamma) { return std::sqrt(gamma * W.p / W.rho); }
void rusanov(const Prim &L, const Prim &R, const double n[3], double gamma, double *F)
{
    double UL[5], UR[5], FL[5], FR[5];
    to_cons(L, gamma, UL);
    to_cons(R, gamma, UR);
    auto phys = [&](const Prim &W, const double *U, double *Fx) {
        const double un = W.u * n[0] + W.v * n[1] + W.w * n[2];
        Fx[0] = W.rho * un;
        Fx[1] = W.rho * un * W.u + W.p * n[0];
        Fx[2] = W.rho * un * W.v + W.p * n[1];
        Fx[3] = W.rho * un * W.w + W.p * n[2];
        Fx[4] = un * (U[4] + W.p);
    };
    phys(L, UL, FL);
    phys(R, UR, FR);
    const double sL = std::abs(L.u * n[0] + L.v * n[1] + L.w * n[2]) + sound(L, gamma);
    const double sR = std::abs(R.u * n[0] + R.v * n[1] + R.w * n[2]) + sound(R, gamma);
    const double s = std::max(sL, sR);
    for (int m = 0; m < 5; ++m) F[m] = 0.5 * (FL[m] + FR[m]) - 0.5 * s * (UR[m] - UL[m]);
}

enum Kind : int8_t { Interior = 0, Wall = 1, Inflow = 2, Outflow = 3, Skip = 4 };

// Common/fluid/Reconstruction.cpp:9-24
Prim ghost_state(int kind, int species, const Prim &in, const V3 &n, const Prim &fs)
{
    switch (kind)
    {
    case Wall:
    {
        if (species == 1) return Prim{1e-12, 0, 0, 0, 1e-12};
        const double un = in.u * n.x + in.v * n.y + in.w * n.z;
        Prim g = in;
        g.u -= 2 * un * n.x; g.v -= 2 * un * n.y; g.w -= 2 * un * n.z;
        return g;
    }
    case Inflow: return fs;
    default: return in;
    }
}

struct Csr
{
    std::vector<int64_t> ptr;
    std::vector<int32_t> idx;
    std::vector<int8_t> sgn;
    int cnt(long r) const { return static_cast<int>(ptr[r + 1] - ptr[r]); }
};
Csr copy_csr(const mhdg_incidence &a, long rows)
{
    Csr c;
    c.ptr.assign(a.ptr, a.ptr + rows + 1);
    c.idx.assign(a.idx, a.idx + c.ptr[rows]);
    c.sgn.assign(a.sgn, a.sgn + c.ptr[rows]);
    return c;
}
template <class T> std::vector<T> cp(const T *p, size_t n) { return std::vector<T>(p, p + n); }
} // namespace

struct orc_ctx
{
    int32_t ne, nf, nc;
    Csr f2e, c2f, e2f, f2c;
    std::vector<int32_t> mat;
    std::vector<int8_t> kind, emode;
    std::vector<double> farea, fcen, famag, ccen, cvol, evec, emid, elen, ecw, efw, star1, star2;
    std::vector<int64_t> ecp;
    std::vector<int32_t> ecc;
    mhdg_config cfg;
    // Stencil (Reconstruction.h:34-41)
    std::vector<int64_t> sptr;
    std::vector<int32_t> snb;
    std::vector<double> sw;
    std::vector<V3> sd;
    std::vector<double> sinv;
    std::vector<char> svalid;
    // ImplicitInduction (ImplicitInduction.h:61-93)
    bool resistive = false, hall = false, ideal = false;
    std::vector<double> m_, taper_, alpha_, B0_, recon_, d2_, corr_;
    std::vector<double> u_ideal;      // set_ideal_state (ImplicitInduction.cpp:113-117), empty until set
    std::vector<char> ideal_mask;
    std::vector<int64_t> ce_ptr;
    std::vector<int32_t> ce_edge;
    std::vector<V3> ce_wt;
    // ion species of the step: ns = 2 (H+, Na+: M_H, M_Na, fs) or n_species (ion_mass, ion_fs)
    int ns = 2;
    double M_s[10] = {}, fs_s[10][5] = {};
    std::vector<double> atm; // exp(-(|x_c| - r0) / H) per cell, empty without ion production
    // step state (mercury::Solver members, Solver.h:128-131), one row per species
    std::vector<double> U[10], Phi, PhiAdd, BaddCell, W[10], grad[10], rhs[10], Bind, Btot, uplus, J, PhiTot, E;
    std::string err;
};

namespace
{
// Common/fluid/Reconstruction.cpp:26-78
void build_stencil(orc_ctx &o)
{
    const int32_t nc = o.nc;
    o.sptr.assign(nc + 1, 0);
    o.sinv.assign(9 * size_t(nc), 0.0);
    o.svalid.assign(nc, 0);
    for (int32_t c = 0; c < nc; ++c)
    {
        if (o.mat[c] == 1) { o.sptr[c + 1] = o.sptr[c]; continue; }
        double M[3][3] = {{0}};
        for (int64_t k = o.c2f.ptr[c]; k < o.c2f.ptr[c + 1]; ++k)
        {
            const int32_t f = o.c2f.idx[k];
            V3 d;
            int32_t nb;
            const int kind = o.kind[f];
            if (kind == Skip) continue;
            if (o.f2c.cnt(f) == 2 && kind == Interior)
            {
                nb = o.f2c.idx[o.f2c.ptr[f]] == c ? o.f2c.idx[o.f2c.ptr[f] + 1] : o.f2c.idx[o.f2c.ptr[f]];
                d = L3(o.ccen, nb) - L3(o.ccen, c);
            }
            else
            {
                const V3 n = (o.c2f.sgn[k] / o.famag[f]) * L3(o.farea, f);
                d = (2.0 * dot(L3(o.fcen, f) - L3(o.ccen, c), n)) * n;
                nb = -(f + 1);
            }
            const double w = 1.0 / dot(d, d);
            o.snb.push_back(nb); o.sw.push_back(w); o.sd.push_back(d);
            const double dv[3] = {d.x, d.y, d.z};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) M[a][b] += w * dv[a] * dv[b];
        }
        o.sptr[c + 1] = static_cast<int64_t>(o.snb.size());
        const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        const double sc = std::max({M[0][0], M[1][1], M[2][2], 1e-300});
        if (std::abs(det) < 1e-9 * sc * sc * sc) continue;
        double *I = &o.sinv[9 * size_t(c)];
        I[0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
        I[1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
        I[2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
        I[3] = I[1];
        I[4] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
        I[5] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
        I[6] = I[2];
        I[7] = I[5];
        I[8] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
        o.svalid[c] = 1;
    }
}

// Common/induction/ImplicitInduction.cpp:10-103 (one rank: assemble_edges is the identity)
void build_induction(orc_ctx &o)
{
    const mhdg_config &cfg = o.cfg;
    const int32_t ne = o.ne, nc = o.nc;
    o.resistive = cfg.resistive && cfg.inver_Rem > 0;
    o.hall = cfg.hall && cfg.hall_coef > 0;
    o.ideal = cfg.implicit_ideal != 0;
    o.m_.assign(ne, 0.0);
    if (o.resistive)
        for (int32_t e = 0; e < ne; ++e)
        {
            const double r = norm(L3(o.emid, e));
            const double shape = 0.5 * (std::tanh((r - cfg.eta_r_inner) / cfg.eta_width) - std::tanh((r - cfg.eta_r_outer) / cfg.eta_width));
            if (std::isfinite(shape) && shape > 1e-10) o.m_[e] = cfg.inver_Rem * std::min(1.0, shape) / o.star1[e];
        }
    o.taper_.assign(nc, 0.0);
    o.alpha_.assign(nc, 0.0);
    o.B0_.assign(3 * size_t(nc), 0.0);
    const double mid = 0.5 * (cfg.hall_r_min + cfg.hall_r_max), width = cfg.hall_r_max - cfg.hall_r_min;
    for (int32_t c = 0; c < nc; ++c)
    {
        if (o.mat[c] == 1) continue;
        const double r = norm(L3(o.ccen, c));
        o.taper_[c] = width <= 0 ? (r >= cfg.hall_r_min ? 1.0 : 0.0) : 0.5 * (1.0 + std::tanh(2.0 * (r - mid) / width));
    }
    o.recon_.assign(3 * o.c2f.idx.size(), 0.0);
    for (int32_t c = 0; c < nc; ++c)
    {
        double M[3][3] = {{0}};
        for (int64_t k = o.c2f.ptr[c]; k < o.c2f.ptr[c + 1]; ++k)
        {
            const int32_t f = o.c2f.idx[k];
            const double A = o.famag[f];
            if (A <= 0) continue;
            const V3 n = (1.0 / A) * L3(o.farea, f);
            const double nv[3] = {n.x, n.y, n.z};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) M[i][j] += A * nv[i] * nv[j];
        }
        const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        const double sc = std::max({M[0][0], M[1][1], M[2][2], 1e-300});
        if (std::abs(det) < 1e-12 * sc * sc * sc) continue;
        double I[3][3];
        I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det; I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det; I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
        I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det; I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det; I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
        I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det; I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det; I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
        for (int64_t k = o.c2f.ptr[c]; k < o.c2f.ptr[c + 1]; ++k)
        {
            const int32_t f = o.c2f.idx[k];
            const double A = o.famag[f];
            if (A <= 0) continue;
            const V3 n = (1.0 / A) * L3(o.farea, f);
            const double nv[3] = {n.x, n.y, n.z};
            for (int i = 0; i < 3; ++i) o.recon_[3 * size_t(k) + i] = I[i][0] * nv[0] + I[i][1] * nv[1] + I[i][2] * nv[2];
        }
    }
    o.ce_ptr.assign(nc + 1, 0);
    std::vector<int32_t> es;
    for (int32_t c = 0; c < nc; ++c)
    {
        es.clear();
        if (o.mat[c] == 0)
            for (int64_t kf = o.c2f.ptr[c]; kf < o.c2f.ptr[c + 1]; ++kf)
                for (int64_t ke = o.f2e.ptr[o.c2f.idx[kf]]; ke < o.f2e.ptr[o.c2f.idx[kf] + 1]; ++ke) es.push_back(o.f2e.idx[ke]);
        std::sort(es.begin(), es.end());
        es.erase(std::unique(es.begin(), es.end()), es.end());
        for (int32_t e : es)
        {
            const auto lo = o.ecc.begin() + o.ecp[e], hi = o.ecc.begin() + o.ecp[e + 1];
            const auto it = std::lower_bound(lo, hi, c);
            if (it == hi || *it != c) continue;
            o.ce_edge.push_back(e);
            o.ce_wt.push_back(o.ecw[it - o.ecc.begin()] * L3(o.evec, e));
        }
        o.ce_ptr[c + 1] = static_cast<int64_t>(o.ce_edge.size());
    }
    o.d2_.assign(ne, 0.0);
    for (int32_t f = 0; f < o.nf; ++f)
        for (int64_t k = o.f2e.ptr[f]; k < o.f2e.ptr[f + 1]; ++k) o.d2_[o.f2e.idx[k]] += o.star2[f];
}

// src/uns/geometry/Reconstruct.cpp:8-46
void cell_vector(const orc_ctx &o, const double *flux, double *b)
{
#pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < o.nc; ++c)
    {
        double M[3][3] = {{0}}, r[3] = {0};
        for (int64_t k = o.c2f.ptr[c]; k < o.c2f.ptr[c + 1]; ++k)
        {
            const int32_t f = o.c2f.idx[k];
            const double A = o.famag[f];
            if (A <= 0) continue;
            const V3 n = (o.c2f.sgn[k] / A) * L3(o.farea, f);
            const double nv[3] = {n.x, n.y, n.z};
            const double phi = o.c2f.sgn[k] * flux[f];
            for (int i = 0; i < 3; ++i)
            {
                r[i] += nv[i] * phi;
                for (int j = 0; j < 3; ++j) M[i][j] += A * nv[i] * nv[j];
            }
        }
        const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        const double sc = std::max({M[0][0], M[1][1], M[2][2], 1e-300});
        double out[3] = {0, 0, 0};
        if (!(std::abs(det) < 1e-12 * sc * sc * sc))
        {
            double inv[3][3];
            inv[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
            inv[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
            inv[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
            inv[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
            inv[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
            inv[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
            inv[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
            inv[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
            inv[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
            for (int i = 0; i < 3; ++i) out[i] = inv[i][0] * r[0] + inv[i][1] * r[1] + inv[i][2] * r[2];
        }
        b[3 * size_t(c)] = out[0]; b[3 * size_t(c) + 1] = out[1]; b[3 * size_t(c) + 2] = out[2];
    }
}

// Common/induction/Current.cpp:9-31
void cell_current(const orc_ctx &o, const double *B, double *J)
{
#pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < o.nc; ++c)
    {
        V3 sum;
        const V3 bc(B[3 * size_t(c)], B[3 * size_t(c) + 1], B[3 * size_t(c) + 2]);
        for (int64_t k = o.c2f.ptr[c]; k < o.c2f.ptr[c + 1]; ++k)
        {
            const int32_t f = o.c2f.idx[k];
            V3 bf = bc;
            if (o.f2c.cnt(f) == 2)
            {
                const int32_t ot = o.f2c.idx[o.f2c.ptr[f]] == c ? o.f2c.idx[o.f2c.ptr[f] + 1] : o.f2c.idx[o.f2c.ptr[f]];
                bf = 0.5 * (bc + L3(B, ot));
            }
            sum += cross(static_cast<double>(o.c2f.sgn[k]) * L3(o.farea, f), bf);
        }
        const double invV = 1.0 / o.cvol[c];
        J[3 * size_t(c)] = sum.x * invV; J[3 * size_t(c) + 1] = sum.y * invV; J[3 * size_t(c) + 2] = sum.z * invV;
    }
}

// Common/fluid/Reconstruction.cpp:80-142
void limited_gradients(const orc_ctx &o, const double *W, int species, const Prim &fs, bool limit, const double *ghost, double *grad)
{
    std::fill(grad, grad + 15 * size_t(o.nc), 0.0);
    const double rel = o.cfg.recon_rel_limit;
    auto prim_at = [&](long c) { return Prim{W[5 * c], W[5 * c + 1], W[5 * c + 2], W[5 * c + 3], W[5 * c + 4]}; };
#pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < o.nc; ++c)
    {
        if (!o.svalid[c]) continue;
        const Prim wc = prim_at(c);
        const double wcv[5] = {wc.rho, wc.u, wc.v, wc.w, wc.p};
        double r[5][3] = {{0}}, lo[5], hi[5];
        for (int m = 0; m < 5; ++m) lo[m] = hi[m] = wcv[m];
        for (int64_t k = o.sptr[c]; k < o.sptr[c + 1]; ++k)
        {
            Prim wn;
            if (o.snb[k] >= 0) wn = prim_at(o.snb[k]);
            else
            {
                const int32_t f = -(o.snb[k] + 1);
                int64_t kk = o.c2f.ptr[c];
                while (o.c2f.idx[kk] != f) ++kk;
                const V3 n = (o.c2f.sgn[kk] / o.famag[f]) * L3(o.farea, f);
                wn = ghost ? Prim{ghost[5 * size_t(f)], ghost[5 * size_t(f) + 1], ghost[5 * size_t(f) + 2], ghost[5 * size_t(f) + 3], ghost[5 * size_t(f) + 4]}
                           : ghost_state(o.kind[f], species, wc, n, fs);
            }
            const double wnv[5] = {wn.rho, wn.u, wn.v, wn.w, wn.p};
            for (int m = 0; m < 5; ++m)
            {
                const double dq = wnv[m] - wcv[m];
                r[m][0] += o.sw[k] * o.sd[k].x * dq; r[m][1] += o.sw[k] * o.sd[k].y * dq; r[m][2] += o.sw[k] * o.sd[k].z * dq;
                lo[m] = std::min(lo[m], wnv[m]);
                hi[m] = std::max(hi[m], wnv[m]);
            }
        }
        const double *I = &o.sinv[9 * size_t(c)];
        for (int m = 0; m < 5; ++m)
        {
            double g[3] = {I[0] * r[m][0] + I[1] * r[m][1] + I[2] * r[m][2], I[3] * r[m][0] + I[4] * r[m][1] + I[5] * r[m][2],
                           I[6] * r[m][0] + I[7] * r[m][1] + I[8] * r[m][2]};
            double psi = 1.0;
            if (limit)
                for (int64_t kf = o.c2f.ptr[c]; kf < o.c2f.ptr[c + 1]; ++kf)
                {
                    const V3 dx = L3(o.fcen, o.c2f.idx[kf]) - L3(o.ccen, c);
                    const double dl = g[0] * dx.x + g[1] * dx.y + g[2] * dx.z;
                    if (dl > 1e-300) psi = std::min(psi, (hi[m] - wcv[m]) / dl);
                    else if (dl < -1e-300) psi = std::min(psi, (lo[m] - wcv[m]) / dl);
                }
            if (limit && rel > 0 && (m == 0 || m == 4))
                for (int64_t kf = o.c2f.ptr[c]; kf < o.c2f.ptr[c + 1]; ++kf)
                {
                    const V3 dx = L3(o.fcen, o.c2f.idx[kf]) - L3(o.ccen, c);
                    const double dl = std::abs(g[0] * dx.x + g[1] * dx.y + g[2] * dx.z);
                    if (dl > 1e-300) psi = std::min(psi, rel * wcv[m] / dl);
                }
            psi = std::max(0.0, std::min(1.0, psi));
            double *out = grad + 15 * size_t(c) + 3 * m;
            out[0] = psi * g[0]; out[1] = psi * g[1]; out[2] = psi * g[2];
        }
    }
}

// Common/fluid/FluxRHS.cpp:16-27
Prim face_state(const double *W, const double *grad, long c, const V3 &dx)
{
    const Prim w{W[5 * c], W[5 * c + 1], W[5 * c + 2], W[5 * c + 3], W[5 * c + 4]};
    const double *g = grad + 15 * c;
    Prim r;
    r.rho = w.rho + g[0] * dx.x + g[1] * dx.y + g[2] * dx.z;
    r.u = w.u + g[3] * dx.x + g[4] * dx.y + g[5] * dx.z;
    r.v = w.v + g[6] * dx.x + g[7] * dx.y + g[8] * dx.z;
    r.w = w.w + g[9] * dx.x + g[10] * dx.y + g[11] * dx.z;
    r.p = w.p + g[12] * dx.x + g[13] * dx.y + g[14] * dx.z;
    return (r.rho > 0 && r.p > 0) ? r : w;
}

// Common/fluid/FluxRHS.cpp:30-74
void flux_rhs(const orc_ctx &o, int species, const Prim &fs, const double *W, const double *grad, const double *ghost, double *rhs)
{
    std::fill(rhs, rhs + 5 * size_t(o.nc), 0.0);
    for (int32_t f = 0; f < o.nf; ++f)
    {
        const int kind = o.kind[f];
        if (kind == Skip) continue;
        const int64_t k0 = o.f2c.ptr[f];
        const int nc = o.f2c.cnt(f);
        bool any = false;
        for (int q = 0; q < nc; ++q) any |= o.mat[o.f2c.idx[k0 + q]] == 0;
        if (!any) continue;
        const double A = o.famag[f];
        const V3 n = (1.0 / A) * L3(o.farea, f);
        const double nv[3] = {n.x, n.y, n.z};
        Prim L, R;
        int32_t cL = -1, cR = -1;
        const V3 xf = L3(o.fcen, f);
        if (nc == 2 && kind == Interior)
        {
            cL = o.f2c.idx[k0]; cR = o.f2c.idx[k0 + 1];
            if (o.f2c.sgn[k0] < 0) std::swap(cL, cR);
            L = face_state(W, grad, cL, xf - L3(o.ccen, cL));
            R = face_state(W, grad, cR, xf - L3(o.ccen, cR));
        }
        else
        {
            int32_t c = o.f2c.idx[k0];
            int sgn = o.f2c.sgn[k0];
            if (nc == 2 && o.mat[c] == 1) { c = o.f2c.idx[k0 + 1]; sgn = o.f2c.sgn[k0 + 1]; }
            const Prim in = face_state(W, grad, c, xf - L3(o.ccen, c));
            const Prim gh = ghost ? Prim{ghost[5 * size_t(f)], ghost[5 * size_t(f) + 1], ghost[5 * size_t(f) + 2], ghost[5 * size_t(f) + 3], ghost[5 * size_t(f) + 4]}
                                  : ghost_state(kind, species, in, (double(sgn)) * n, fs);
            if (sgn > 0) { cL = c; L = in; R = gh; } else { cR = c; R = in; L = gh; }
        }
        double F[5];
        rusanov(L, R, nv, o.cfg.gamma, F);
        for (int m = 0; m < 5; ++m) F[m] *= A;
        if (cL >= 0) for (int m = 0; m < 5; ++m) rhs[5 * size_t(cL) + m] -= F[m] / o.cvol[cL];
        if (cR >= 0) for (int m = 0; m < 5; ++m) rhs[5 * size_t(cR) + m] += F[m] / o.cvol[cR];
    }
}

// n_s = max(rho_s, 0) / M_s * M_ref and n_e = sum of n_s in species order (Common/physics/Scales.cpp:50-62 for
// the two Mercury ions, the same expressions for any number of ions)
double ion_n(const orc_ctx &o, int s, double rho) { return std::max(rho, 0.0) / o.M_s[s] * o.cfg.M_ref; }
double electron_n(const orc_ctx &o, const double *const *Wall, size_t i)
{
    double ne = ion_n(o, 0, Wall[0][5 * i]);
    for (int s = 1; s < o.ns; ++s) ne = ne + ion_n(o, s, Wall[s][5 * i]);
    return ne;
}

// MercuryUns/fluid/MercurySources.cpp:20-63 (Lorentz part; chemistry and ambipolar are off) for ion s of the
// o.ns ions Wall, w_s = n_s / n_e,eff; with `production`, the exponential-atmosphere ion source of the step
void add_sources(const orc_ctx &o, int s, const double *const *Wall, const double *B, const double *J, double dt, double *rhs,
                 bool production)
{
    const double *W = Wall[s];
    const double src = production ? o.cfg.src_rate[s] : 0.0;
#pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < o.nc; ++c)
    {
        if (o.mat[c] == 1) continue;
        const size_t i = c;
        const double ne = electron_n(o, Wall, i);
        const double ws = ion_n(o, s, W[5 * i]) / std::sqrt(ne * ne + o.cfg.ne_floor * o.cfg.ne_floor);
        const double rho = std::max(W[5 * i], 0.0);
        const V3 u(W[5 * i + 1], W[5 * i + 2], W[5 * i + 3]);
        double *r = rhs + 5 * i;
        V3 S;
        if (o.cfg.lorentz_force)
        {
            const V3 b(B[3 * i], B[3 * i + 1], B[3 * i + 2]), j(J[3 * i], J[3 * i + 1], J[3 * i + 2]);
            S += (o.cfg.inver_MA2 * ws) * cross(j, b);
        }
        r[1] += S.x; r[2] += S.y; r[3] += S.z;
        r[4] += dot(u, S) + (rho > 1e-30 ? 0.5 * dt * dot(S, S) / rho : 0.0);
        if (src != 0.0 && !o.atm.empty())
        {
            // ions produced at rest with temperature src_T
            const double q = src * o.atm[i];
            r[0] += q;
            r[4] += q * o.cfg.src_T / (o.cfg.gamma - 1.0);
        }
    }
}

// MercuryUns/fluid/MercurySources.cpp:65-116
double local_dt(const orc_ctx &o, const double *const *Wall, const double *B)
{
    const double inf = std::numeric_limits<double>::infinity();
    std::vector<double> d(o.nc, 0.0);
    const mhdg_config &cfg = o.cfg;
#pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < o.nc; ++c)
    {
        if (o.mat[c] == 1) continue;
        double dt = inf;
        const V3 b(B[3 * size_t(c)], B[3 * size_t(c) + 1], B[3 * size_t(c) + 2]);
        const double B2 = dot(b, b);
        double rho_sum = Wall[0][5 * size_t(c)];
        for (int s = 1; s < o.ns; ++s) rho_sum = rho_sum + Wall[s][5 * size_t(c)];
        const double rho_tot = std::max(rho_sum, cfg.rho_floor_dt);
        for (int s = 0; s < o.ns; ++s)
        {
            const double *w = Wall[s] + 5 * size_t(c);
            const double cs2 = cfg.gamma * w[4] / std::max(w[0], 1e-30);
            double denom = 0;
            for (int64_t k = o.c2f.ptr[c]; k < o.c2f.ptr[c + 1]; ++k)
            {
                const int32_t f = o.c2f.idx[k];
                const double A = o.famag[f];
                const V3 n = (1.0 / A) * L3(o.farea, f);
                const double vA2 = cfg.inver_MA2 * B2 / rho_tot;
                const double vAn2 = cfg.inver_MA2 * dot(b, n) * dot(b, n) / rho_tot;
                const double term = cs2 + vA2;
                const double cf = std::sqrt(0.5 * (term + std::sqrt(std::max(0.0, term * term - 4 * cs2 * vAn2))));
                denom += (std::abs(w[1] * n.x + w[2] * n.y + w[3] * n.z) + cf) * A;
            }
            if (denom > 0) dt = std::min(dt, cfg.cfl * o.cvol[c] / denom);
        }
        d[c] = dt;
    }
    double dt = inf;
    for (int32_t c = 0; c < o.nc; ++c)
        if (o.mat[c] != 1) dt = std::min(dt, d[c]);
    return dt;
}

// Common/induction/IdealEMF.cpp:16-22
void plane_basis(const V3 &t, V3 &e1, V3 &e2)
{
    const V3 a = std::abs(t.x) < 0.9 ? V3(1, 0, 0) : V3(0, 1, 0);
    e1 = cross(t, a);
    e1 *= 1.0 / norm(e1);
    e2 = cross(t, e1);
}

// Common/induction/IdealEMF.cpp:25-129
void ideal_emf(const orc_ctx &o, const double *u_cell, const double *Btot, const double *Bind, const double *phi_tot, const double *phi_ind,
               double cuct, double *E)
{
    std::vector<V3> Ef(o.nf);
#pragma omp parallel for schedule(static)
    for (int32_t f = 0; f < o.nf; ++f)
    {
        const double A = o.famag[f];
        const V3 n = (1.0 / A) * L3(o.farea, f);
        const double Bn_tot = phi_tot[f] / A;
        auto flux = [&](int32_t c) {
            const V3 u = L3(u_cell, c);
            return dot(u, n) * L3(Btot, c) - Bn_tot * u;
        };
        V3 F;
        if (o.f2c.cnt(f) == 2)
        {
            int32_t Lc = o.f2c.idx[o.f2c.ptr[f]], Rc = o.f2c.idx[o.f2c.ptr[f] + 1];
            if (o.f2c.sgn[o.f2c.ptr[f]] < 0) std::swap(Lc, Rc);
            const double a = std::max(std::abs(dot(L3(u_cell, Lc), n)), std::abs(dot(L3(u_cell, Rc), n)));
            F = 0.5 * (flux(Lc) + flux(Rc)) - (0.5 * a) * (L3(Bind, Rc) - L3(Bind, Lc));
        }
        else
            F = flux(o.f2c.idx[o.f2c.ptr[f]]);
        Ef[f] = -1.0 * cross(n, F);
    }
#pragma omp parallel
    {
        std::vector<double> w, aef, qa, qb;
        std::vector<V3> m, nrm;
#pragma omp for schedule(dynamic, 256)
        for (int32_t e = 0; e < o.ne; ++e)
        {
            const int64_t f0 = o.e2f.ptr[e], f1 = o.e2f.ptr[e + 1];
            const int nf = static_cast<int>(f1 - f0);
            if (nf == 0 || o.elen[e] <= 0) continue;
            bool any_fluid = false;
            for (int64_t k = o.ecp[e]; k < o.ecp[e + 1] && !any_fluid; ++k) any_fluid = o.mat[o.ecc[k]] == 0;
            if (!any_fluid) continue;
            const V3 t = L3(o.evec, e);
            const double L = o.elen[e];
            const V3 th = (1.0 / L) * t;
            double Ec = 0;
            w.assign(nf, 0.0); m.assign(nf, V3()); nrm.assign(nf, V3()); aef.assign(nf, 0.0);
            for (int i = 0; i < nf; ++i)
            {
                const int32_t f = o.e2f.idx[f0 + i];
                w[i] = o.efw[f0 + i];
                Ec += w[i] * dot(Ef[f], t);
                V3 v = L3(o.fcen, f) - L3(o.emid, e);
                v -= dot(v, th) * th;
                const double nv = norm(v);
                m[i] = nv > 0 ? (1.0 / nv) * v : V3();
                nrm[i] = (1.0 / o.famag[f]) * L3(o.farea, f);
            }
            for (int64_t k = o.ecp[e]; k < o.ecp[e + 1]; ++k)
            {
                const V3 u = L3(u_cell, o.ecc[k]);
                for (int i = 0; i < nf; ++i) aef[i] = std::max(aef[i], std::abs(dot(u, m[i])));
            }
            V3 e1, e2;
            plane_basis(th, e1, e2);
            double M11 = 0, M12 = 0, M22 = 0, g1 = 0, g2 = 0;
            qa.assign(nf, 0.0); qb.assign(nf, 0.0);
            for (int i = 0; i < nf; ++i)
            {
                const V3 qv = (aef[i] * dot(th, cross(m[i], nrm[i]))) * nrm[i];
                qa[i] = dot(qv, e1);
                qb[i] = dot(qv, e2);
                M11 += w[i] * qa[i] * qa[i];
                M12 += w[i] * qa[i] * qb[i];
                M22 += w[i] * qb[i] * qb[i];
                g1 += w[i] * qa[i];
                g2 += w[i] * qb[i];
            }
            const double tr = M11 + M22;
            const double eps = 1e-14 * tr;
            const double det = (M11 + eps) * (M22 + eps) - M12 * M12;
            double l1 = 0, l2 = 0;
            if (tr > 0 && det > 0)
            {
                l1 = ((M22 + eps) * g1 - M12 * g2) / det;
                l2 = ((M11 + eps) * g2 - M12 * g1) / det;
            }
            double Eu = 0;
            for (int i = 0; i < nf; ++i)
            {
                const int32_t f = o.e2f.idx[f0 + i];
                const double omega = w[i] * (1.0 - (qa[i] * l1 + qb[i] * l2));
                const double sig = dot(th, cross(m[i], nrm[i]));
                Eu += omega * aef[i] * sig * (phi_ind[f] / o.famag[f]);
            }
            E[e] += Ec + 2.0 * cuct * L * Eu;
        }
    }
}

// Common/induction/Faraday.cpp:15-36 (one rank: every face belongs to an owned cell)
void faraday(const orc_ctx &o, const double *E, const double *dte, double *B, double dt)
{
#pragma omp parallel for schedule(static)
    for (int32_t f = 0; f < o.nf; ++f)
    {
        if (o.f2c.cnt(f) == 0) continue;
        double circ = 0.0;
        if (dte)
        {
            for (int64_t k = o.f2e.ptr[f]; k < o.f2e.ptr[f + 1]; ++k) circ += o.f2e.sgn[k] * dte[o.f2e.idx[k]] * E[o.f2e.idx[k]];
            B[f] -= circ;
        }
        else
        {
            for (int64_t k = o.f2e.ptr[f]; k < o.f2e.ptr[f + 1]; ++k) circ += o.f2e.sgn[k] * E[o.f2e.idx[k]];
            B[f] -= dt * circ;
        }
    }
}

// Common/induction/ImplicitInduction.cpp:127-183
void apply_F(orc_ctx &o, const double *phi, double *out, std::vector<double> &part, std::vector<double> &bcell)
{
    const int32_t ne = o.ne, nc = o.nc;
    std::fill(part.begin(), part.end(), 0.0);
    if (o.resistive)
        for (int32_t f = 0; f < o.nf; ++f)
        {
            const double v = o.star2[f] * phi[f];
            for (int64_t k = o.f2e.ptr[f]; k < o.f2e.ptr[f + 1]; ++k) part[3 * size_t(o.f2e.idx[k])] += o.f2e.sgn[k] * v;
        }
    if (o.hall || o.ideal)
#pragma omp parallel for schedule(static)
        for (int32_t c = 0; c < nc; ++c)
        {
            V3 b;
            for (int64_t k = o.c2f.ptr[c]; k < o.c2f.ptr[c + 1]; ++k)
            {
                const double v = phi[o.c2f.idx[k]];
                b.x += o.recon_[3 * size_t(k)] * v; b.y += o.recon_[3 * size_t(k) + 1] * v; b.z += o.recon_[3 * size_t(k) + 2] * v;
            }
            bcell[3 * size_t(c)] = b.x; bcell[3 * size_t(c) + 1] = b.y; bcell[3 * size_t(c) + 2] = b.z;
        }
    if (o.ideal && !o.u_ideal.empty())
    {
        // ImplicitInduction.cpp:150-157: ideal EMF of (u, b, phi) with b as total and induced field
        std::vector<double> ide(ne, 0.0);
        ideal_emf(o, o.u_ideal.data(), bcell.data(), bcell.data(), phi, phi, o.cfg.cuct, ide.data());
        for (int32_t e = 0; e < ne; ++e) part[3 * size_t(e) + 2] = o.ideal_mask[e] ? ide[e] : 0.0;
    }
    if (o.hall)
    {
        for (int32_t c = 0; c < nc; ++c)
        {
            if (o.alpha_[c] == 0.0) continue;
            V3 sum;
            const V3 bc = L3(bcell, c);
            for (int64_t k = o.c2f.ptr[c]; k < o.c2f.ptr[c + 1]; ++k)
            {
                const int32_t f = o.c2f.idx[k];
                V3 bf = bc;
                if (o.f2c.cnt(f) == 2)
                {
                    const int32_t ot = o.f2c.idx[o.f2c.ptr[f]] == c ? o.f2c.idx[o.f2c.ptr[f] + 1] : o.f2c.idx[o.f2c.ptr[f]];
                    bf = 0.5 * (bc + L3(bcell, ot));
                }
                sum += cross(static_cast<double>(o.c2f.sgn[k]) * L3(o.farea, f), bf);
            }
            const V3 J = (1.0 / o.cvol[c]) * sum;
            const V3 g = o.alpha_[c] * cross(J, L3(o.B0_, c));
            for (int64_t q = o.ce_ptr[c]; q < o.ce_ptr[c + 1]; ++q) part[3 * size_t(o.ce_edge[q]) + 1] += dot(g, o.ce_wt[q]);
        }
    }
    for (int32_t e = 0; e < ne; ++e) out[e] = o.m_[e] * part[3 * size_t(e)] + part[3 * size_t(e) + 1] + part[3 * size_t(e) + 2];
}

struct IndWork
{
    std::vector<double> part, bcell, phi, tmp;
    explicit IndWork(const orc_ctx &o) : part(3 * size_t(o.ne)), bcell(3 * size_t(o.nc)), phi(o.nf), tmp(o.ne) {}
};

// Common/induction/ImplicitInduction.cpp:185-198 (dte: per-edge pseudo-time step, null for the scalar dt)
void apply_A(orc_ctx &o, IndWork &wk, const std::vector<double> &x, std::vector<double> &y, double dt, const double *dte)
{
    for (int32_t f = 0; f < o.nf; ++f)
    {
        double s = 0;
        if (dte) for (int64_t k = o.f2e.ptr[f]; k < o.f2e.ptr[f + 1]; ++k) s += o.f2e.sgn[k] * dte[o.f2e.idx[k]] * x[o.f2e.idx[k]];
        else for (int64_t k = o.f2e.ptr[f]; k < o.f2e.ptr[f + 1]; ++k) s += o.f2e.sgn[k] * x[o.f2e.idx[k]];
        wk.phi[f] = s;
    }
    apply_F(o, wk.phi.data(), wk.tmp.data(), wk.part, wk.bcell);
    if (dte) for (size_t e = 0; e < y.size(); ++e) y[e] = x[e] + wk.tmp[e];
    else for (size_t e = 0; e < y.size(); ++e) y[e] = x[e] + dt * wk.tmp[e];
}

double dot_all(const std::vector<double> &a, const std::vector<double> &b)
{
    double s = 0;
    for (size_t e = 0; e < a.size(); ++e) s += a[e] * b[e];
    return s;
}

// Common/induction/ImplicitInduction.cpp:299-390 (Jacobi right preconditioner, MGS Arnoldi)
int ind_solve(orc_ctx &o, const double *Phi_n, const double *E_expl, const int8_t *dir, double dt, double *E, double *rel_res,
              const double *dte = nullptr)
{
    const int32_t ne = o.ne;
    using Vecd = std::vector<double>;
    IndWork wk(o);
    const double rtol = o.cfg.ind_rtol;
    const int max_it = o.cfg.ind_max_it, m = std::max(2, o.cfg.ind_restart);
    Vecd x0(E_expl, E_expl + ne), b(ne), r(ne), Ax(ne);
    if (o.corr_.size() == size_t(ne))
        for (int32_t e = 0; e < ne; ++e) if (!dir[e]) x0[e] += o.corr_[e];
    Vecd Fn(ne);
    apply_F(o, Phi_n, Fn.data(), wk.part, wk.bcell);
    for (int32_t e = 0; e < ne; ++e) b[e] = E_expl[e] + Fn[e];
    apply_A(o, wk, x0, Ax, dt, dte);
    for (int32_t e = 0; e < ne; ++e) r[e] = dir[e] ? 0.0 : b[e] - Ax[e];
    Vecd inv_diag(ne);
    for (int32_t e = 0; e < ne; ++e) inv_diag[e] = 1.0 / (1.0 + (dte ? dte[e] : dt) * o.m_[e] * o.d2_[e]);
    const double bnorm = std::sqrt(std::max(dot_all(b, b), 1e-300));
    double beta = std::sqrt(dot_all(r, r));
    int total = 0;
    Vecd delta(ne, 0.0);
    std::vector<Vecd> V(m + 1, Vecd(ne));
    std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
    std::vector<double> cs(m), sn(m), g(m + 1), yv(m);
    Vecd w(ne), z(ne);
    while (beta > rtol * bnorm && total < max_it)
    {
        for (int32_t e = 0; e < ne; ++e) V[0][e] = r[e] / beta;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m && total < max_it; ++j, ++total)
        {
            for (int32_t e = 0; e < ne; ++e) z[e] = dir[e] ? 0.0 : V[j][e] * inv_diag[e];
            apply_A(o, wk, z, w, dt, dte);
            for (int32_t e = 0; e < ne; ++e) if (dir[e]) w[e] = 0.0;
            for (int i = 0; i <= j; ++i)
            {
                H[i][j] = dot_all(w, V[i]);
                for (int32_t e = 0; e < ne; ++e) w[e] -= H[i][j] * V[i][e];
            }
            H[j + 1][j] = std::sqrt(dot_all(w, w));
            if (H[j + 1][j] > 1e-300) for (int32_t e = 0; e < ne; ++e) V[j + 1][e] = w[e] / H[j + 1][j];
            for (int i = 0; i < j; ++i)
            {
                const double t = cs[i] * H[i][j] + sn[i] * H[i + 1][j];
                H[i + 1][j] = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
                H[i][j] = t;
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
        for (int32_t e = 0; e < ne; ++e)
        {
            double s = 0;
            for (int i = 0; i < j; ++i) s += yv[i] * V[i][e];
            delta[e] += dir[e] ? 0.0 : s * inv_diag[e];
        }
        Vecd xt(ne);
        for (int32_t e = 0; e < ne; ++e) xt[e] = x0[e] + delta[e];
        apply_A(o, wk, xt, Ax, dt, dte);
        for (int32_t e = 0; e < ne; ++e) r[e] = dir[e] ? 0.0 : b[e] - Ax[e];
        beta = std::sqrt(dot_all(r, r));
    }
    for (int32_t e = 0; e < ne; ++e) E[e] = x0[e] + delta[e];
    o.corr_.assign(ne, 0.0);
    for (int32_t e = 0; e < ne; ++e) o.corr_[e] = dir[e] ? 0.0 : E[e] - E_expl[e];
    if (rel_res) *rel_res = beta / bnorm;
    return total;
}

// MercuryUns/driver/Solver.cpp:94-125
void prepare(orc_ctx &o)
{
    const int32_t nc = o.nc;
    const mhdg_config &c = o.cfg;
    for (int k = 0; k < o.ns; ++k)
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < nc; ++i)
        {
            const Prim w = to_prim(&o.U[k][5 * size_t(i)], c.gamma, c.rho_floor, c.p_floor);
            double *q = &o.W[k][5 * size_t(i)];
            q[0] = w.rho; q[1] = w.u; q[2] = w.v; q[3] = w.w; q[4] = w.p;
        }
    std::vector<double> bi(3 * size_t(nc));
    cell_vector(o, o.Phi.data(), bi.data());
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < nc; ++i)
    {
        for (int a = 0; a < 3; ++a) { o.Bind[3 * size_t(i) + a] = bi[3 * size_t(i) + a]; o.Btot[3 * size_t(i) + a] = bi[3 * size_t(i) + a] + o.BaddCell[3 * size_t(i) + a]; }
        double up[3] = {0, 0, 0};
        if (o.mat[i] == 0)
        {
            // u+ = sum of chi_s u_s, chi_s = n_s / n_e (chi = 1, 0, ... for n_e <= 1e-300)
            const double *Wall[10];
            for (int s = 0; s < o.ns; ++s) Wall[s] = o.W[s].data();
            const double ne = electron_n(o, Wall, i);
            for (int s = 0; s < o.ns; ++s)
            {
                const double chi = ne > 1e-300 ? ion_n(o, s, o.W[s][5 * size_t(i)]) / ne : (s == 0 ? 1.0 : 0.0);
                for (int a = 0; a < 3; ++a) up[a] = s == 0 ? chi * o.W[0][5 * size_t(i) + 1 + a] : up[a] + chi * o.W[s][5 * size_t(i) + 1 + a];
            }
        }
        for (int a = 0; a < 3; ++a) o.uplus[3 * size_t(i) + a] = up[a];
    }
    for (int32_t f = 0; f < o.nf; ++f) o.PhiTot[f] = o.Phi[f] + o.PhiAdd[f];
    cell_current(o, o.Bind.data(), o.J.data());
}

// MercuryUns/driver/Solver.cpp:171-191
void enforce_floors(orc_ctx &o)
{
    const mhdg_config &c = o.cfg;
    for (int k = 0; k < o.ns; ++k)
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < o.nc; ++i)
        {
            if (o.mat[i] == 1) continue;
            double *u = &o.U[k][5 * size_t(i)];
            const double ke = u[0] > 0 ? 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0] : 0.0;
            if (!(u[0] >= c.rho_floor)) to_cons(Prim{c.rho_floor, 0, 0, 0, c.p_floor}, c.gamma, u);
            else if ((c.gamma - 1.0) * (u[4] - ke) < c.p_floor)
            {
                const Prim w = to_prim(u, c.gamma, c.rho_floor, c.p_floor);
                to_cons(w, c.gamma, u);
            }
        }
}

Prim fs_of(const double *p) { return Prim{p[0], p[1], p[2], p[3], p[4]}; }

// MercuryUns/physics/Gyration.cpp:17-51 (gyrate_n, dt_cell = null) for the two ions
void gyrate(orc_ctx &o, double dt)
{
    const mhdg_config &c = o.cfg;
    const double aH = c.induce_coeff * c.M_ref / o.M_s[0], aN = c.induce_coeff * c.M_ref / o.M_s[1];
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < o.nc; ++i)
    {
        if (o.mat[i] == 1) continue;
        double *uh = &o.U[0][5 * size_t(i)], *un = &o.U[1][5 * size_t(i)];
        const Prim wh = to_prim(uh, c.gamma, c.rho_floor, c.p_floor), wn = to_prim(un, c.gamma, c.rho_floor, c.p_floor);
        const double nH = std::max(uh[0], 0.0) / o.M_s[0] * c.M_ref, nN = std::max(un[0], 0.0) / o.M_s[1] * c.M_ref;
        const double ne = nH + nN;
        const double chiH = ne > 1e-300 ? nH / ne : 1.0, chiN = ne > 1e-300 ? nN / ne : 0.0;
        const V3 b = L3(o.Btot, i);
        const double bm = norm(b);
        const double denom = aH * chiN + aN * chiH;
        if (bm <= 0 || denom <= 0) continue;
        const V3 bh = (1.0 / bm) * b;
        const V3 uH(wh.u, wh.v, wh.w), uN(wn.u, wn.v, wn.w);
        const V3 w = uH - uN;
        const V3 w_par = dot(w, bh) * bh, w_perp = w - w_par;
        const double th = denom * bm * dt;
        const V3 dw = (std::cos(th) - 1.0) * w_perp + std::sin(th) * cross(w_perp, bh);
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
}

// MercuryUns/driver/Solver.cpp:228-277 (gyration split inactive: induce_coeff = 0)
double step(orc_ctx &o, double dt_max, int *iters, double *res)
{
    const mhdg_config &c = o.cfg;
    const int32_t nc = o.nc, ne = o.ne;
    prepare(o);
    const double *Wall[10];
    for (int s = 0; s < o.ns; ++s) Wall[s] = o.W[s].data();
    double gdt = local_dt(o, Wall, o.Btot.data());
    gdt = std::min(gdt, dt_max);
    if (!std::isfinite(gdt) || gdt <= 0) throw std::runtime_error("invalid time step");
    // Strang split of the gyration (Solver.cpp:237-241, 272)
    const bool split = c.induce_coeff > 0 && c.lorentz_force;
    if (split)
    {
        gyrate(o, 0.5 * gdt);
        prepare(o);
    }
    for (int k = 0; k < o.ns; ++k)
    {
        if (c.second_order) limited_gradients(o, o.W[k].data(), k, fs_of(o.fs_s[k]), true, nullptr, o.grad[k].data());
        else std::fill(o.grad[k].begin(), o.grad[k].end(), 0.0);
        flux_rhs(o, k, fs_of(o.fs_s[k]), o.W[k].data(), o.grad[k].data(), nullptr, o.rhs[k].data());
    }
    for (int k = 0; k < o.ns; ++k) add_sources(o, k, Wall, o.Btot.data(), o.J.data(), gdt, o.rhs[k].data(), true);
    std::fill(o.E.begin(), o.E.end(), 0.0);
    ideal_emf(o, o.uplus.data(), o.Btot.data(), o.Bind.data(), o.PhiTot.data(), o.Phi.data(), c.cuct, o.E.data());
    const V3 u_inf(c.u_inf[0], c.u_inf[1], c.u_inf[2]), B_imf(c.B_imf[0], c.B_imf[1], c.B_imf[2]);
    const V3 E_inf = -1.0 * cross(u_inf, B_imf);
    for (int32_t e = 0; e < ne; ++e)
    {
        if (o.emode[e] == 1) o.E[e] = 0.0;
        else if (o.emode[e] == 2) o.E[e] = dot(E_inf, L3(o.evec, e));
    }
    for (int k = 0; k < o.ns; ++k)
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < nc; ++i)
        {
            if (o.mat[i] == 1) continue;
            for (int q = 0; q < 5; ++q) o.U[k][5 * size_t(i) + q] += gdt * o.rhs[k][5 * size_t(i) + q];
        }
    enforce_floors(o);
    std::vector<double> Eout = o.E;
    *iters = 0;
    *res = 0;
    if (o.resistive || o.hall || o.ideal) // ImplicitInduction::active()
    {
        std::vector<double> ne_eff(nc, 0.0);
        for (int32_t i = 0; i < nc; ++i)
            if (o.mat[i] == 0)
            {
                const double n = electron_n(o, Wall, i);
                ne_eff[i] = std::sqrt(n * n + c.ne_floor * c.ne_floor);
            }
        for (size_t i = 0; i < o.B0_.size(); ++i) o.B0_[i] = o.Btot[i];
        for (int32_t i = 0; i < nc; ++i) o.alpha_[i] = (o.hall && ne_eff[i] > 0) ? o.taper_[i] * c.hall_coef / ne_eff[i] : 0.0;
        std::vector<int8_t> dir(ne);
        for (int32_t e = 0; e < ne; ++e) dir[e] = c.wall_dirichlet ? o.emode[e] != 0 : o.emode[e] == 2;
        *iters = ind_solve(o, o.Phi.data(), o.E.data(), dir.data(), gdt, Eout.data(), res);
    }
    faraday(o, Eout.data(), nullptr, o.Phi.data(), gdt);
    if (split) gyrate(o, 0.5 * gdt);
    o.E = Eout;
    return gdt;
}
} // namespace

// ------------------------------------------------------------------------------------------- C-ABI
extern "C" {
int orc_create(const mhdg_mesh_desc *d, const mhdg_config *cfg, orc_ctx **out)
{
    try
    {
        auto *o = new orc_ctx();
        o->ne = d->n_edges; o->nf = d->n_faces; o->nc = d->n_cells;
        const size_t E = o->ne, F = o->nf, C = o->nc;
        o->f2e = copy_csr(d->f2e, F); o->c2f = copy_csr(d->c2f, C); o->e2f = copy_csr(d->e2f, E); o->f2c = copy_csr(d->f2c, F);
        o->mat = cp(d->cell_material, C); o->kind = cp(d->face_kind, F); o->emode = cp(d->edge_mode, E);
        o->farea = cp(d->face_area, 3 * F); o->fcen = cp(d->face_center, 3 * F); o->famag = cp(d->face_area_mag, F);
        o->ccen = cp(d->cell_center, 3 * C); o->cvol = cp(d->cell_volume, C);
        o->evec = cp(d->edge_vec, 3 * E); o->emid = cp(d->edge_mid, 3 * E); o->elen = cp(d->edge_len, E);
        o->ecp = cp(d->ec_ptr, E + 1); o->ecc = cp(d->ec_cell, size_t(o->ecp[E])); o->ecw = cp(d->ec_weight, size_t(o->ecp[E]));
        o->efw = cp(d->ef_weight, o->e2f.idx.size()); o->star1 = cp(d->star1, E); o->star2 = cp(d->star2, F);
        o->cfg = *cfg;
        if (cfg->n_species > 10 || cfg->n_species < 0 || cfg->n_species == 1) throw std::runtime_error("n_species must be 0 or 2..10");
        o->ns = cfg->n_species > 2 ? cfg->n_species : 2;
        for (int s = 0; s < o->ns; ++s)
        {
            o->M_s[s] = o->ns == 2 ? (s == 0 ? cfg->M_H : cfg->M_Na) : cfg->ion_mass[s];
            for (int m = 0; m < 5; ++m) o->fs_s[s][m] = o->ns == 2 ? cfg->fs[s][m] : cfg->ion_fs[s][m];
        }
        bool production = false;
        for (int s = 0; s < o->ns; ++s) production = production || cfg->src_rate[s] != 0.0;
        if (production)
        {
            o->atm.resize(C);
            for (size_t i = 0; i < C; ++i)
            {
                const double *x = &o->ccen[3 * i];
                o->atm[i] = std::exp(-(std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) - cfg->src_r0) / cfg->src_scale_height);
            }
        }
        build_stencil(*o);
        build_induction(*o);
        for (int k = 0; k < o->ns; ++k) { o->U[k].assign(5 * C, 0.0); o->W[k].assign(5 * C, 0.0); o->grad[k].assign(15 * C, 0.0); o->rhs[k].assign(5 * C, 0.0); }
        o->Phi.assign(F, 0.0); o->PhiAdd.assign(F, 0.0); o->BaddCell.assign(3 * C, 0.0); o->PhiTot.assign(F, 0.0);
        o->Bind.assign(3 * C, 0.0); o->Btot.assign(3 * C, 0.0); o->uplus.assign(3 * C, 0.0); o->J.assign(3 * C, 0.0); o->E.assign(E, 0.0);
        *out = o;
        return 0;
    }
    catch (const std::exception &) { return 1; }
}
void orc_destroy(orc_ctx *o) { delete o; }
const char *orc_last_error(orc_ctx *o) { return o ? o->err.c_str() : ""; }

int orc_cell_vector(orc_ctx *o, const double *flux, double *b) { cell_vector(*o, flux, b); return 0; }
int orc_cell_current(orc_ctx *o, const double *B, double *J) { cell_current(*o, B, J); return 0; }
int orc_limited_gradients(orc_ctx *o, const double *W, int species, const double fs[5], int limit, const double *ghost, double *grad)
{
    limited_gradients(*o, W, species, fs_of(fs), limit != 0, ghost, grad);
    return 0;
}
int orc_flux_rhs(orc_ctx *o, int species, const double fs[5], const double *W, const double *grad, const double *ghost, double *rhs)
{
    flux_rhs(*o, species, fs_of(fs), W, grad, ghost, rhs);
    return 0;
}
int orc_ideal_emf(orc_ctx *o, const double *u, const double *Bt, const double *Bi, const double *pt, const double *pi, double cuct, double *E)
{
    ideal_emf(*o, u, Bt, Bi, pt, pi, cuct, E);
    return 0;
}
int orc_faraday(orc_ctx *o, const double *E, double *Phi, double dt) { faraday(*o, E, nullptr, Phi, dt); return 0; }
int orc_faraday_edge_dt(orc_ctx *o, const double *E, const double *dte, double *Phi) { faraday(*o, E, dte, Phi, 0.0); return 0; }
// Common/diagnostics/DivB.cpp:8-24
int orc_div_b(orc_ctx *o, const double *B, const double *extra, double *max_abs, double *max_rel, int64_t *worst)
{
    double ma = 0, mr = 0;
    int64_t wc = -1;
    for (int32_t c = 0; c < o->nc; ++c)
    {
        double s = 0, a = 0;
        for (int64_t k = o->c2f.ptr[c]; k < o->c2f.ptr[c + 1]; ++k)
        {
            s += o->c2f.sgn[k] * B[o->c2f.idx[k]];
            a += std::abs(B[o->c2f.idx[k]]) + (extra ? std::abs(extra[o->c2f.idx[k]]) : 0.0);
        }
        const double rel = a > 0 ? std::abs(s) / a : 0.0;
        if (std::abs(s) > ma) { ma = std::abs(s); wc = c; }
        mr = std::max(mr, rel);
    }
    *max_abs = ma; *max_rel = mr; *worst = wc;
    return 0;
}
int orc_local_dt(orc_ctx *o, const double *WH, const double *WNa, const double *B, double *dt)
{
    if (o->ns != 2) return 1;
    const double *Wall[2] = {WH, WNa};
    *dt = local_dt(*o, Wall, B);
    return 0;
}
int orc_add_sources(orc_ctx *o, int s, const double *WH, const double *WNa, const double *B, const double *J, double dt, double *rhs)
{
    if (o->ns != 2 || s < 0 || s > 1) return 1;
    const double *Wall[2] = {WH, WNa};
    add_sources(*o, s, Wall, B, J, dt, rhs, false);
    return 0;
}
int orc_ind_set_frozen(orc_ctx *o, const double *B0, const double *ne)
{
    for (size_t i = 0; i < o->B0_.size(); ++i) o->B0_[i] = B0[i];
    for (int32_t c = 0; c < o->nc; ++c) o->alpha_[c] = (o->hall && ne[c] > 0) ? o->taper_[c] * o->cfg.hall_coef / ne[c] : 0.0;
    return 0;
}
int orc_ind_apply_F(orc_ctx *o, const double *phi, double *out)
{
    IndWork wk(*o);
    apply_F(*o, phi, out, wk.part, wk.bcell);
    return 0;
}
int orc_ind_solve(orc_ctx *o, const double *Phi_n, const double *E_expl, const int8_t *dir, double dt, const double *dte, double *E,
                  int32_t *iters, double *rel_res)
{
    *iters = ind_solve(*o, Phi_n, E_expl, dir, dt, E, rel_res, dte);
    return 0;
}
// ImplicitInduction::set_ideal_state / ideal_linear (ImplicitInduction.cpp:113-125)
int orc_ind_set_ideal_state(orc_ctx *o, const double *u, const int8_t *mask)
{
    o->u_ideal.assign(u, u + 3 * size_t(o->nc));
    o->ideal_mask.assign(mask, mask + o->ne);
    return 0;
}
int orc_ind_ideal_linear(orc_ctx *o, const double *phi, double *out)
{
    const bool h = o->hall, r = o->resistive;
    o->hall = false;
    o->resistive = false;
    IndWork wk(*o);
    apply_F(*o, phi, out, wk.part, wk.bcell);
    o->hall = h;
    o->resistive = r;
    return 0;
}
int orc_set_state(orc_ctx *o, const double *UH, const double *UNa, const double *Phi)
{
    std::copy(UH, UH + o->U[0].size(), o->U[0].begin());
    std::copy(UNa, UNa + o->U[1].size(), o->U[1].begin());
    std::copy(Phi, Phi + o->Phi.size(), o->Phi.begin());
    return 0;
}
int orc_get_state(orc_ctx *o, double *UH, double *UNa, double *Phi)
{
    std::copy(o->U[0].begin(), o->U[0].end(), UH);
    std::copy(o->U[1].begin(), o->U[1].end(), UNa);
    std::copy(o->Phi.begin(), o->Phi.end(), Phi);
    return 0;
}
int orc_set_species_state(orc_ctx *o, int s, const double *U)
{
    if (s < 0 || s >= o->ns) return 1;
    std::copy(U, U + o->U[s].size(), o->U[s].begin());
    return 0;
}
int orc_get_species_state(orc_ctx *o, int s, double *U)
{
    if (s < 0 || s >= o->ns) return 1;
    std::copy(o->U[s].begin(), o->U[s].end(), U);
    return 0;
}
int orc_get_species_prim(orc_ctx *o, int s, double *W)
{
    if (s < 0 || s >= o->ns) return 1;
    std::copy(o->W[s].begin(), o->W[s].end(), W);
    return 0;
}
int orc_set_background(orc_ctx *o, const double *PhiAdd, const double *Badd)
{
    std::copy(PhiAdd, PhiAdd + o->PhiAdd.size(), o->PhiAdd.begin());
    std::copy(Badd, Badd + o->BaddCell.size(), o->BaddCell.begin());
    return 0;
}
int orc_step(orc_ctx *o, double dt_max, double *dt_used, int32_t *iters, double *res)
{
    try
    {
        int it = 0;
        *dt_used = step(*o, dt_max, &it, res);
        *iters = it;
        return 0;
    }
    catch (const std::exception &e) { o->err = e.what(); return 1; }
}
// MARS::MarsSolver::implicit_sources_ (MarsUns/driver/MarsSolver.cpp:422-527) on the current state of every ion
// species with B_tot of the last prepare: Crank-Nicolson gyration of the coupled ions (closed form, std::complex as
// the reference), backward-Euler drag and thermal relaxation, per fluid cell in ascending order.
int orc_mars_implicit_sources(orc_ctx *o, const mhdg_mars_params *prm, const double *neutral, double dt)
{
    using cd = std::complex<double>;
    const double g = o->cfg.gamma;
    const double a5 = 3.0e6 * prm->L * prm->kB / (prm->rho_ref * prm->U * prm->U * prm->U);
    const int NS = o->ns, NN = prm->n_neutral;
    for (int32_t c = 0; c < o->nc; ++c)
    {
        if (o->mat[size_t(c)] != 0) continue;
        const size_t i = size_t(c);
        double nrel[10], ntot = 0;
        for (int s = 0; s < NS; ++s)
        {
            nrel[s] = std::max(o->U[s][5 * i], 0.0) * prm->rho_ref / (prm->ion_mass[s] * prm->n);
            ntot += nrel[s];
        }
        const V3 B(o->Btot[3 * i], o->Btot[3 * i + 1], o->Btot[3 * i + 2]);
        const double Bm = norm(B);
        if (ntot > 1e-300 && Bm > 1e-30)
        {
            const V3 k = (1.0 / Bm) * B;
            const V3 t = std::abs(k.x) < 0.9 ? V3(1, 0, 0) : V3(0, 1, 0);
            V3 e1 = cross(k, t);
            e1 = (1.0 / norm(e1)) * e1;
            const V3 e2 = cross(k, e1);
            cd z[10], al[10], Z0 = 0;
            for (int s = 0; s < NS; ++s)
            {
                const double *u = &o->U[s][5 * i];
                const double rho = std::max(u[0], prm->rho_floor);
                const V3 v(u[1] / rho, u[2] / rho, u[3] / rho);
                z[s] = cd(dot(v, e1), dot(v, e2));
                al[s] = cd(0.0, prm->q * prm->B * prm->L / (prm->ion_mass[s] * prm->U) * Bm * dt * 0.5);
                Z0 += (nrel[s] / ntot) * z[s];
            }
            cd Pn = 0, Qn = 0;
            for (int s = 0; s < NS; ++s)
            {
                const double chi = nrel[s] / ntot;
                const cd a = -al[s];
                Pn += chi * (z[s] * (1.0 - a) + a * Z0) / (1.0 + a);
                Qn += chi * a / (1.0 + a);
            }
            const cd Zn = Pn / (1.0 - Qn);
            for (int s = 0; s < NS; ++s)
            {
                const cd a = -al[s];
                const cd zn = (z[s] * (1.0 - a) + a * (Z0 + Zn)) / (1.0 + a);
                double *u = &o->U[s][5 * i];
                const double rho = std::max(u[0], prm->rho_floor);
                const V3 v(u[1] / rho, u[2] / rho, u[3] / rho);
                const double vpar = dot(v, k);
                const V3 vn = vpar * k + zn.real() * e1 + zn.imag() * e2;
                u[4] += 0.5 * rho * (dot(vn, vn) - dot(v, v));
                u[1] = rho * vn.x; u[2] = rho * vn.y; u[3] = rho * vn.z;
            }
        }
        double coll_n[10], collision = 0;
        for (int n = 0; n < NN; ++n)
        {
            coll_n[n] = std::max(neutral[i * size_t(prm->neutral_stride) + size_t(n)], 0.0) * 2.0e-10;
            collision += coll_n[n];
        }
        const double nu = prm->time_ref * collision;
        for (int s = 0; s < NS; ++s)
        {
            double *u = &o->U[s][5 * i];
            const double m = prm->ion_mass[s];
            const double rho = std::max(u[0], prm->rho_floor);
            const V3 v(u[1] / rho, u[2] / rho, u[3] / rho);
            const double ke_old = 0.5 * rho * dot(v, v);
            const double den = 1.0 + nu * dt;
            u[1] /= den; u[2] /= den; u[3] /= den;
            const double ke = ke_old / (den * den);
            const double p = std::max((g - 1.0) * (u[4] - ke), prm->p_floor);
            double b1c = 0, b2 = 0;
            for (int n = 0; n < NN; ++n)
            {
                const double mm = m + prm->neutral_mass[n];
                b1c += m * coll_n[n] / mm;
                b2 += prm->neutral_mass[n] * coll_n[n] / mm;
            }
            const double n_m3 = rho * prm->rho_ref / m;
            const double ns_cm3 = n_m3 * 1e-6;
            const double kk = a5 * ns_cm3 * b1c;
            const double cT = prm->p_ref / (std::max(n_m3, 1.0) * prm->kB);
            const double Sf = prm->time_ref * rho * dot(v, v) * b2;
            double pn;
            if (p * cT >= 50.0) pn = (p + dt * (g - 1.0) * (kk * 185.0 + Sf)) / (1.0 + dt * (g - 1.0) * kk * cT);
            else pn = p + dt * (g - 1.0) * (kk * (185.0 - 50.0) + Sf);
            u[4] = pn / (g - 1.0) + ke;
        }
    }
    return 0;
}
// Solver::prepare (Solver.cpp:94-125) on the current state
int orc_prepare(orc_ctx *o)
{
    prepare(*o);
    return 0;
}
int orc_get_derived(orc_ctx *o, int which, double *out)
{
    const std::vector<double> *v = nullptr;
    switch (which)
    {
    case 0: v = &o->W[0]; break;
    case 1: v = &o->W[1]; break;
    case 2: v = &o->Bind; break;
    case 3: v = &o->Btot; break;
    case 4: v = &o->J; break;
    case 5: v = &o->uplus; break;
    case 6: v = &o->E; break;
    default: return 1;
    }
    std::copy(v->begin(), v->end(), out);
    return 0;
}
}
