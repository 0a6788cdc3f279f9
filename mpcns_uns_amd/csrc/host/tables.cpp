#include "tables.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

#include "vec3.h"

namespace mhdg
{
namespace
{
inline int32_t pack(int32_t idx, int sgn) { return sgn > 0 ? idx : ~idx; }

void pack_csr(const mhdg_incidence &a, int32_t rows, std::vector<int32_t> &ptr, std::vector<int32_t> &out)
{
    const int64_t nnz = a.ptr[rows];
    if (nnz >= (int64_t(1) << 31)) throw std::runtime_error("incidence table too large for 32-bit offsets");
    ptr.resize(size_t(rows) + 1);
    for (int32_t r = 0; r <= rows; ++r) ptr[r] = static_cast<int32_t>(a.ptr[r]);
    out.resize(size_t(nnz));
    for (int64_t k = 0; k < nnz; ++k)
    {
        if (a.sgn[k] != 1 && a.sgn[k] != -1) throw std::runtime_error("incidence sign must be +1 or -1");
        out[k] = pack(a.idx[k], a.sgn[k]);
    }
}
} // namespace

// A descriptor is checked completely on the host before anything reaches the device: every index a kernel
// follows must stay inside its table.
static void check_csr(const mhdg_incidence &t, int64_t rows, int64_t cols, const char *name, bool need_sgn = true)
{
    const std::string n(name);
    if (!t.ptr || !t.idx || (need_sgn && !t.sgn)) throw std::runtime_error("mesh descriptor: " + n + " missing");
    if (t.ptr[0] != 0) throw std::runtime_error("mesh descriptor: " + n + ".ptr[0] != 0");
    for (int64_t r = 0; r < rows; ++r)
        if (t.ptr[r + 1] < t.ptr[r]) throw std::runtime_error("mesh descriptor: " + n + ".ptr not monotonic");
    if (t.ptr[rows] > int64_t(INT32_MAX)) throw std::runtime_error("mesh descriptor: " + n + " has more than 2^31 entries");
    for (int64_t k = 0; k < t.ptr[rows]; ++k)
    {
        if (t.idx[k] < 0 || t.idx[k] >= cols) throw std::runtime_error("mesh descriptor: " + n + ".idx out of range");
        if (need_sgn && t.sgn[k] != 1 && t.sgn[k] != -1) throw std::runtime_error("mesh descriptor: " + n + ".sgn not +-1");
    }
}
void check_desc(const mhdg_mesh_desc &d)
{
    if (d.n_cells <= 0 || d.n_faces <= 0 || d.n_edges <= 0) throw std::runtime_error("mesh descriptor: empty mesh");
    check_csr(d.c2f, d.n_cells, d.n_faces, "c2f");
    check_csr(d.f2e, d.n_faces, d.n_edges, "f2e");
    check_csr(d.e2f, d.n_edges, d.n_faces, "e2f");
    check_csr(d.f2c, d.n_faces, d.n_cells, "f2c");
    if (d.f2e.ptr[d.n_faces] != d.e2f.ptr[d.n_edges] || d.c2f.ptr[d.n_cells] != d.f2c.ptr[d.n_faces])
        throw std::runtime_error("mesh descriptor: e2f / f2c are not the transposes of f2e / c2f");
    const void *need[] = {d.cell_material, d.face_kind, d.edge_mode, d.face_area, d.face_center, d.face_area_mag, d.cell_center,
                          d.cell_volume, d.edge_vec, d.edge_mid, d.edge_len, d.ec_ptr, d.ec_cell, d.ec_weight, d.ef_weight, d.star1, d.star2};
    for (const void *p : need)
        if (!p) throw std::runtime_error("mesh descriptor: a geometry or boundary array is missing");
    if (d.ec_ptr[0] != 0) throw std::runtime_error("mesh descriptor: ec_ptr[0] != 0");
    for (int64_t e = 0; e < d.n_edges; ++e)
        if (d.ec_ptr[e + 1] < d.ec_ptr[e]) throw std::runtime_error("mesh descriptor: ec_ptr not monotonic");
    for (int64_t k = 0; k < d.ec_ptr[d.n_edges]; ++k)
        if (d.ec_cell[k] < 0 || d.ec_cell[k] >= d.n_cells) throw std::runtime_error("mesh descriptor: ec_cell out of range");
    for (int64_t c = 0; c < d.n_cells; ++c)
        if (d.cell_material[c] != 0 && d.cell_material[c] != 1) throw std::runtime_error("mesh descriptor: cell_material not 0/1");
    for (int64_t f = 0; f < d.n_faces; ++f)
        if (d.face_kind[f] < 0 || d.face_kind[f] > 4) throw std::runtime_error("mesh descriptor: face_kind not 0..4");
}

Tables build_tables(const mhdg_mesh_desc &d, const mhdg_config &cfg, int32_t light_max)
{
    check_desc(d);
    Tables t;
    t.nc = d.n_cells; t.nf = d.n_faces; t.ne = d.n_edges;
    const int32_t nc = t.nc, nf = t.nf, ne = t.ne;
    pack_csr(d.c2f, nc, t.c2f_ptr, t.c2f);
    pack_csr(d.f2e, nf, t.f2e_ptr, t.f2e);
    pack_csr(d.e2f, ne, t.e2f_ptr, t.e2f);
    t.mat8.resize(nc);
    for (int32_t c = 0; c < nc; ++c) t.mat8[c] = static_cast<int8_t>(d.cell_material[c]);

    // face -> (outward cell, inward cell)
    t.fL.assign(nf, -1);
    t.fR.assign(nf, -1);
    for (int32_t f = 0; f < nf; ++f)
    {
        const int64_t k0 = d.f2c.ptr[f], k1 = d.f2c.ptr[f + 1];
        if (k1 - k0 > 2) throw std::runtime_error("a face has more than two cells");
        for (int64_t k = k0; k < k1; ++k)
        {
            int32_t &slot = d.f2c.sgn[k] > 0 ? t.fL[f] : t.fR[f];
            if (slot >= 0) throw std::runtime_error("interior face with equal orientation in both cells");
            slot = d.f2c.idx[k];
        }
    }
    // cell rows sorted by face id (owner-gather order of fluid_flux_rhs, FluxRHS.cpp:36-72)
    t.cfs = t.c2f;
    for (int32_t c = 0; c < nc; ++c)
        std::sort(t.cfs.begin() + t.c2f_ptr[c], t.cfs.begin() + t.c2f_ptr[c + 1],
                  [](int32_t a, int32_t b) { return (a >= 0 ? a : ~a) < (b >= 0 ? b : ~b); });
    t.ec_ptr.resize(size_t(ne) + 1);
    if (d.ec_ptr[ne] >= (int64_t(1) << 31)) throw std::runtime_error("edge-cell table too large");
    for (int32_t e = 0; e <= ne; ++e) t.ec_ptr[e] = static_cast<int32_t>(d.ec_ptr[e]);
    t.ec_cell.assign(d.ec_cell, d.ec_cell + d.ec_ptr[ne]);
    for (int32_t e = 0; e < ne; ++e)
    {
        const int32_t v = t.e2f_ptr[e + 1] - t.e2f_ptr[e];
        t.max_valence = std::max(t.max_valence, v);
        if (v > light_max) t.heavy.push_back(e);
    }

    auto A3 = [&](const double *p, long i) { return V3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); };
    // Stencil constructor (Common/fluid/Reconstruction.cpp:26-78), entries aligned with c2f
    t.snb.assign(t.c2f.size(), std::numeric_limits<int32_t>::min());
    t.sw.assign(t.c2f.size(), 0.0);
    t.sd.assign(3 * t.c2f.size(), 0.0);
    t.sinv.assign(6 * size_t(nc), 0.0);
    t.svalid.assign(nc, 0);
    for (int32_t c = 0; c < nc; ++c)
    {
        if (d.cell_material[c] == 1) continue;
        double M[3][3] = {{0}};
        for (int32_t k = t.c2f_ptr[c]; k < t.c2f_ptr[c + 1]; ++k)
        {
            const int32_t f = d.c2f.idx[k];
            const int kind = d.face_kind[f];
            if (kind == 4) continue;
            V3 dv;
            int32_t nb;
            if (t.fL[f] >= 0 && t.fR[f] >= 0 && kind == 0)
            {
                nb = t.fL[f] == c ? t.fR[f] : t.fL[f];
                dv = A3(d.cell_center, nb) - A3(d.cell_center, c);
            }
            else
            {
                const V3 n = (d.c2f.sgn[k] / d.face_area_mag[f]) * A3(d.face_area, f);
                dv = (2.0 * dot(A3(d.face_center, f) - A3(d.cell_center, c), n)) * n;
                nb = -(f + 1);
            }
            const double w = 1.0 / dot(dv, dv);
            t.snb[k] = nb; t.sw[k] = w;
            t.sd[3 * size_t(k)] = dv.x; t.sd[3 * size_t(k) + 1] = dv.y; t.sd[3 * size_t(k) + 2] = dv.z;
            const double v3[3] = {dv.x, dv.y, dv.z};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) M[a][b] += w * v3[a] * v3[b];
        }
        const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        const double sc = std::max({M[0][0], M[1][1], M[2][2], 1e-300});
        if (std::abs(det) < 1e-9 * sc * sc * sc) continue;
        double *I = &t.sinv[6 * size_t(c)];
        I[0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
        I[1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
        I[2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
        I[3] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
        I[4] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
        I[5] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
        t.svalid[c] = 1;
    }

    // ImplicitInduction constructor (Common/induction/ImplicitInduction.cpp:10-103)
    const bool resistive = cfg.resistive && cfg.inver_Rem > 0;
    t.m_e.assign(ne, 0.0);
    if (resistive)
        for (int32_t e = 0; e < ne; ++e)
        {
            const double r = norm(A3(d.edge_mid, e));
            const double shape = 0.5 * (std::tanh((r - cfg.eta_r_inner) / cfg.eta_width) - std::tanh((r - cfg.eta_r_outer) / cfg.eta_width));
            if (std::isfinite(shape) && shape > 1e-10) t.m_e[e] = cfg.inver_Rem * std::min(1.0, shape) / d.star1[e];
        }
    t.taper.assign(nc, 0.0);
    const double mid = 0.5 * (cfg.hall_r_min + cfg.hall_r_max), width = cfg.hall_r_max - cfg.hall_r_min;
    for (int32_t c = 0; c < nc; ++c)
    {
        if (d.cell_material[c] == 1) continue;
        const double r = norm(A3(d.cell_center, c));
        t.taper[c] = width <= 0 ? (r >= cfg.hall_r_min ? 1.0 : 0.0) : 0.5 * (1.0 + std::tanh(2.0 * (r - mid) / width));
    }
    t.recon.assign(3 * t.c2f.size(), 0.0);
    for (int32_t c = 0; c < nc; ++c)
    {
        double M[3][3] = {{0}};
        for (int32_t k = t.c2f_ptr[c]; k < t.c2f_ptr[c + 1]; ++k)
        {
            const int32_t f = d.c2f.idx[k];
            const double A = d.face_area_mag[f];
            if (A <= 0) continue;
            const V3 n = (1.0 / A) * A3(d.face_area, f);
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
        for (int32_t k = t.c2f_ptr[c]; k < t.c2f_ptr[c + 1]; ++k)
        {
            const int32_t f = d.c2f.idx[k];
            const double A = d.face_area_mag[f];
            if (A <= 0) continue;
            const V3 n = (1.0 / A) * A3(d.face_area, f);
            const double nv[3] = {n.x, n.y, n.z};
            for (int i = 0; i < 3; ++i) t.recon[3 * size_t(k) + i] = I[i][0] * nv[0] + I[i][1] * nv[1] + I[i][2] * nv[2];
        }
    }
    // d2_e = sum over the faces of e (ascending) of star2_f
    t.d2.assign(ne, 0.0);
    for (int32_t e = 0; e < ne; ++e)
    {
        double s = 0.0;
        for (int32_t k = t.e2f_ptr[e]; k < t.e2f_ptr[e + 1]; ++k) s += d.star2[d.e2f.idx[k]];
        t.d2[e] = s;
    }
    return t;
}
} // namespace mhdg
