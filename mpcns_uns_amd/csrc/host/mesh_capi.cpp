// Host C-ABI for mesh setup (include/mhdg_mesh.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>

#include "../../../include/mhdg_mesh.h"
#include "complex.h"
#include "geometry.h"
#include "multigrid.h"
#include "sfc.h"
#include "tables.h"
#include "shell.h"
#include "unsbin.h"
#include "vtu.h"
#include "vec3.h"

struct mhdg_mesh
{
    mhdg::Complex cx;
    mhdg::Geometry g;
    std::vector<double> xyz;
    std::vector<int32_t> tets;
    std::vector<int8_t> face_kind, edge_mode;
    bool have_bc = false;
    std::vector<std::string> tag_names;
};
struct mhdg_fields
{
    std::vector<mhdg::unsbin::Field> f;
};

namespace mhdg
{
// Crustal field of point dipoles (MarsUns/physics/MarsDipoles.cpp:63-107): vector potential
// A(x) = sum_i m_i x (x - c_i) / (|x - c_i|^2 + core_i^2)^{3/2}, summed over the dipoles in order with a
// compensated (Kahan) sum per component, dipoles beyond the cutoff distance (cut2 > 0) skipped; the edge
// integral of A by the symmetric two-point Gauss rule on the straight edge.
static void dipole_potential(int32_t n, const double *pos, const double *mom, const double *core, double cut2, const double x[3],
                             double out[3])
{
    double sum[3] = {0, 0, 0}, corr[3] = {0, 0, 0};
    for (int32_t i = 0; i < n; ++i)
    {
        double r[3], d2 = 0;
        for (int k = 0; k < 3; ++k) { r[k] = x[k] - pos[3 * i + k]; d2 += r[k] * r[k]; }
        if (!std::isfinite(d2)) throw std::runtime_error("dipoles: non-finite evaluation distance");
        if (cut2 > 0 && d2 > cut2) continue;
        const double r2 = d2 + core[i] * core[i];
        if (!(r2 > 0) || !std::isfinite(r2)) throw std::runtime_error("dipoles: singular evaluation point");
        const double inv = 1.0 / (r2 * std::sqrt(r2));
        for (int k = 0; k < 3; ++k)
        {
            const int j = (k + 1) % 3, l = (k + 2) % 3;
            const double term = (mom[3 * i + j] * r[l] - mom[3 * i + l] * r[j]) * inv;
            const double y = term - corr[k], t = sum[k] + y;
            corr[k] = (t - sum[k]) - y;
            sum[k] = t;
        }
    }
    for (int k = 0; k < 3; ++k)
    {
        if (!std::isfinite(sum[k])) throw std::runtime_error("dipoles: non-finite vector potential");
        out[k] = sum[k];
    }
}
double dipole_edge_integral(int32_t n, const double *pos, const double *mom, const double *core, double cut2, const double *a,
                            const double *b)
{
    double mid[3], dr[3];
    for (int k = 0; k < 3; ++k) { mid[k] = 0.5 * (a[k] + b[k]); dr[k] = b[k] - a[k]; }
    if (dr[0] == 0 && dr[1] == 0 && dr[2] == 0) return 0.0;
    const double q = 0.5 / std::sqrt(3.0);
    double p[3], mq[3], ap[3], am[3];
    for (int k = 0; k < 3; ++k) { p[k] = mid[k] + q * dr[k]; mq[k] = mid[k] - q * dr[k]; }
    dipole_potential(n, pos, mom, core, cut2, p, ap);
    dipole_potential(n, pos, mom, core, cut2, mq, am);
    double v = 0;
    for (int k = 0; k < 3; ++k) v += 0.5 * (ap[k] + am[k]) * dr[k];
    if (!std::isfinite(v)) throw std::runtime_error("dipoles: non-finite edge integral");
    return v;
}
} // namespace mhdg

namespace
{
thread_local std::string g_err;

mhdg_mesh *finish_mesh(mhdg_mesh *m)
{
    m->g = mhdg::compute_geometry(m->cx, m->xyz.data());
    return m;
}
} // namespace

namespace
{
std::vector<mhdg::vtk::Array> vtk_arrays(int32_t n, const mhdg_vtk_array *a)
{
    std::vector<mhdg::vtk::Array> out;
    for (int32_t i = 0; i < n; ++i)
    {
        if (!a[i].name || a[i].ncomp < 1 || !a[i].values) throw std::runtime_error("vtk array: name, ncomp >= 1 and values required");
        out.push_back({a[i].name, a[i].ncomp, a[i].values});
    }
    return out;
}
} // namespace

extern "C" {
const char *mhdg_mesh_error(void) { return g_err.c_str(); }

int mhdg_mesh_read_unsbin(const char *path, mhdg_mesh **out)
{
    try
    {
        auto m = std::make_unique<mhdg_mesh>();
        mhdg::unsbin::read_complex(path, m->cx, m->xyz, m->tag_names);
        *out = finish_mesh(m.release());
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}
int mhdg_mesh_read_unsbin_parts(const char *const *paths, int32_t n_parts, mhdg_mesh **out)
{
    try
    {
        if (!paths || n_parts < 1) throw std::runtime_error("mhdg_mesh_read_unsbin_parts: no part files");
        std::vector<std::string> v;
        for (int32_t k = 0; k < n_parts; ++k)
        {
            if (!paths[k]) throw std::runtime_error("mhdg_mesh_read_unsbin_parts: null path");
            v.emplace_back(paths[k]);
        }
        auto m = std::make_unique<mhdg_mesh>();
        mhdg::unsbin::read_partition(v, m->cx, m->xyz, m->tag_names);
        *out = finish_mesh(m.release());
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}
int mhdg_mesh_write_unsbin(const mhdg_mesh *m, const char *path)
{
    try
    {
        mhdg::unsbin::write_complex(path, m->cx, m->xyz, m->tag_names);
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}
int mhdg_mesh_topology_hash(const mhdg_mesh *m, uint64_t *hash)
{
    *hash = mhdg::unsbin::topology_hash(m->cx);
    return 0;
}
int mhdg_mesh_tag_names(const mhdg_mesh *m, char *buf, int len)
{
    std::string s;
    for (const auto &t : m->tag_names) s += t + "\n";
    if (len <= 0) return 1;
    std::snprintf(buf, static_cast<size_t>(len), "%s", s.c_str());
    return static_cast<int>(s.size()) < len ? 0 : 1;
}
int mhdg_mesh_set_tag_names(mhdg_mesh *m, const char *joined)
{
    m->tag_names.clear();
    std::string cur;
    for (const char *p = joined; *p; ++p)
    {
        if (*p == '\n') { m->tag_names.push_back(cur); cur.clear(); }
        else cur += *p;
    }
    if (!cur.empty()) m->tag_names.push_back(cur);
    return 0;
}
int mhdg_fields_write(const char *path, uint64_t hash, const mhdg_field *fields, int32_t n)
{
    try
    {
        std::vector<mhdg::unsbin::Field> v(n);
        for (int32_t i = 0; i < n; ++i)
        {
            v[i].name = fields[i].name;
            v[i].dim = fields[i].dim;
            v[i].ncomp = fields[i].ncomp;
            v[i].values.assign(fields[i].values, fields[i].values + fields[i].count);
        }
        mhdg::unsbin::write_fields(path, hash, v);
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}
int mhdg_fields_read(const char *path, mhdg_fields **out, uint64_t *hash, int32_t *n)
{
    try
    {
        auto f = std::make_unique<mhdg_fields>();
        f->f = mhdg::unsbin::read_fields(path, hash);
        *n = static_cast<int32_t>(f->f.size());
        *out = f.release();
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}
int mhdg_fields_get(const mhdg_fields *f, int32_t i, mhdg_field *out)
{
    if (i < 0 || i >= static_cast<int32_t>(f->f.size())) { g_err = "field index out of range"; return 1; }
    const auto &x = f->f[i];
    out->name = x.name.c_str();
    out->dim = x.dim;
    out->ncomp = x.ncomp;
    out->count = static_cast<int64_t>(x.values.size());
    out->values = x.values.data();
    return 0;
}
void mhdg_fields_free(mhdg_fields *f) { delete f; }

int mhdg_mesh_shell(int n_lat, int n_lon, int n_r, double r_in, double r_out, double ratio, double jitter, uint64_t seed, mhdg_mesh **out)
{
    try
    {
        mhdg::ShellSpec s;
        s.n_lat = n_lat; s.n_lon = n_lon; s.n_r = n_r; s.r_in = r_in; s.r_out = r_out; s.ratio = ratio; s.jitter = jitter; s.seed = seed;
        mhdg::ShellMesh sm = mhdg::make_shell(s);
        auto *m = new mhdg_mesh();
        m->xyz = std::move(sm.xyz);
        m->tets = std::move(sm.tets);
        const int32_t nn = static_cast<int32_t>(m->xyz.size() / 3);
        m->cx = mhdg::build_tet_complex(nn, m->tets.data(), static_cast<int64_t>(m->tets.size() / 4), sm.tags.data(),
                                        static_cast<int64_t>(sm.tags.size() / 4));
        m->tag_names = {"", "inner", "outer"}; // tag ids 1 (r = r_in) and 2 (r = r_out)
        *out = finish_mesh(m);
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}

int mhdg_mesh_from_cells(int32_t n_nodes, const double *xyz, int64_t n_cells, const int32_t *cell_ptr, const int32_t *loop_ptr,
                         const int32_t *loop_nodes, const int32_t *material, int64_t n_tags, const int32_t *tag_ptr,
                         const int32_t *tag_nodes, const int32_t *tag_id, mhdg_mesh **out)
{
    try
    {
        auto *m = new mhdg_mesh();
        m->xyz.assign(xyz, xyz + 3 * size_t(n_nodes));
        mhdg::Builder b(n_nodes, n_cells);
        std::vector<int32_t> lp;
        for (int64_t c = 0; c < n_cells; ++c)
        {
            const int32_t l0 = cell_ptr[c], l1 = cell_ptr[c + 1];
            lp.assign(size_t(l1 - l0) + 1, 0);
            for (int32_t l = l0; l <= l1; ++l) lp[l - l0] = loop_ptr[l] - loop_ptr[l0];
            b.add_cell(lp.data(), loop_nodes + loop_ptr[l0], l1 - l0, material ? material[c] : 0);
        }
        for (int64_t t = 0; t < n_tags; ++t) b.tag_face(tag_nodes + tag_ptr[t], tag_ptr[t + 1] - tag_ptr[t], tag_id[t]);
        m->cx = b.finish();
        *out = finish_mesh(m);
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}

void mhdg_mesh_free(mhdg_mesh *m) { delete m; }

int mhdg_mesh_hilbert_order(const mhdg_mesh *m, int32_t *order)
{
    try
    {
        const std::vector<int32_t> o = mhdg::hilbert_order(m->g.cell_center.data(), m->cx.nc);
        std::copy(o.begin(), o.end(), order);
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}

int mhdg_mesh_renumber_cells(const mhdg_mesh *m, const int32_t *order, mhdg_mesh **out)
{
    try
    {
        const int64_t nc = m->cx.nc;
        if (m->tets.size() != size_t(4 * nc)) throw std::runtime_error("renumber: only tetrahedral meshes built by mhdg_mesh_shell");
        std::vector<char> seen(size_t(nc), 0);
        auto r = std::make_unique<mhdg_mesh>();
        r->xyz = m->xyz;
        r->tets.resize(m->tets.size());
        for (int64_t k = 0; k < nc; ++k)
        {
            const int32_t c = order[k];
            if (c < 0 || c >= nc || seen[size_t(c)]) throw std::runtime_error("renumber: order is not a permutation of the cells");
            seen[size_t(c)] = 1;
            std::copy(m->tets.begin() + 4 * int64_t(c), m->tets.begin() + 4 * int64_t(c) + 4, r->tets.begin() + 4 * k);
        }
        // tagged faces by their nodes: the builder finds them again in the new numbering
        std::vector<int32_t> tags;
        for (int32_t f = 0; f < m->cx.nf; ++f)
        {
            if (!m->cx.face_tag[size_t(f)]) continue;
            int32_t nodes[8];
            int n = 0;
            m->cx.face_loop(f, nodes, n);
            if (n != 3) throw std::runtime_error("renumber: tagged face is not a triangle");
            tags.insert(tags.end(), {nodes[0], nodes[1], nodes[2], m->cx.face_tag[size_t(f)]});
        }
        r->cx = mhdg::build_tet_complex(m->cx.nn, r->tets.data(), nc, tags.data(), static_cast<int64_t>(tags.size() / 4));
        r->tag_names = m->tag_names;
        *out = finish_mesh(r.release());
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}

int mhdg_mesh_counts(const mhdg_mesh *m, int64_t c[8])
{
    c[0] = m->cx.nn; c[1] = m->cx.ne; c[2] = m->cx.nf; c[3] = m->cx.nc;
    c[4] = m->cx.f2e.nnz(); c[5] = m->cx.c2f.nnz(); c[6] = m->cx.e2f.nnz(); c[7] = static_cast<int64_t>(m->g.ec_cell.size());
    return 0;
}

int mhdg_mesh_array(const mhdg_mesh *m, const char *name, const void **ptr, int64_t *count, int32_t *elem)
{
    const std::string n(name);
    auto give = [&](const auto &v) {
        *ptr = v.data();
        *count = static_cast<int64_t>(v.size());
        *elem = static_cast<int32_t>(sizeof(v[0]));
        return 0;
    };
    const mhdg::Complex &cx = m->cx;
    const mhdg::Geometry &g = m->g;
    auto csr = [&](const mhdg::Csr &t, const std::string &part) {
        if (part == "ptr") return give(t.ptr);
        if (part == "idx") return give(t.idx);
        return give(t.sgn);
    };
    const auto dotp = n.find('.');
    if (dotp != std::string::npos)
    {
        const std::string tb = n.substr(0, dotp), part = n.substr(dotp + 1);
        if (tb == "e2n") return csr(cx.e2n, part);
        if (tb == "f2e") return csr(cx.f2e, part);
        if (tb == "c2f") return csr(cx.c2f, part);
        if (tb == "e2f") return csr(cx.e2f, part);
        if (tb == "f2c") return csr(cx.f2c, part);
        if (tb == "n2e") return csr(cx.n2e, part);
        return 1;
    }
    if (n == "xyz") return give(m->xyz);
    if (n == "tets") return give(m->tets);
    if (n == "face_tag") return give(cx.face_tag);
    if (n == "material") return give(cx.material);
    if (n == "face_kind") return give(m->face_kind);
    if (n == "edge_mode") return give(m->edge_mode);
    if (n == "edge_vec") return give(g.edge_vec);
    if (n == "edge_mid") return give(g.edge_mid);
    if (n == "edge_len") return give(g.edge_len);
    if (n == "face_area") return give(g.face_area);
    if (n == "face_center") return give(g.face_center);
    if (n == "face_area_mag") return give(g.face_area_mag);
    if (n == "cell_center") return give(g.cell_center);
    if (n == "cell_volume") return give(g.cell_volume);
    if (n == "cell_min_tet") return give(g.cell_min_tet);
    if (n == "edge_dual_area") return give(g.edge_dual_area);
    if (n == "ec_ptr") return give(g.ec_ptr);
    if (n == "ec_cell") return give(g.ec_cell);
    if (n == "ec_weight") return give(g.ec_weight);
    if (n == "ef_weight") return give(g.ef_weight);
    if (n == "ec_misoriented") return give(g.ec_misoriented);
    if (n == "star1") return give(g.star1);
    if (n == "star2") return give(g.star2);
    if (n == "star3") return give(g.star3);
    return 1;
}

// Common/boundary/BoundaryInfo.cpp:9-49
int mhdg_mesh_boundary(mhdg_mesh *m, const int32_t *tag_type, int32_t n_types, int32_t default_type, const double u_inf[3])
{
    using namespace mhdg;
    const Complex &cx = m->cx;
    m->face_kind.assign(cx.nf, 0);
    m->edge_mode.assign(cx.ne, 0);
    const V3 u(u_inf[0], u_inf[1], u_inf[2]);
    for (int32_t f = 0; f < cx.nf; ++f)
    {
        const int64_t k0 = cx.f2c.ptr[f];
        const int nc = static_cast<int>(cx.f2c.ptr[f + 1] - k0);
        if (nc == 2)
        {
            const bool s0 = cx.material[cx.f2c.idx[k0]] == 1, s1 = cx.material[cx.f2c.idx[k0 + 1]] == 1;
            m->face_kind[f] = (s0 && s1) ? 4 : (s0 || s1) ? 1 : 0;
            continue;
        }
        if (cx.material[cx.f2c.idx[k0]] == 1) { m->face_kind[f] = 4; continue; }
        const int tag = cx.face_tag[f];
        const int type = (tag > 0 && tag < n_types && tag_type[tag] >= 0) ? tag_type[tag] : default_type;
        if (type == 0) m->face_kind[f] = 1;
        else if (type == 1) m->face_kind[f] = 2;
        else if (type == 2) m->face_kind[f] = 3;
        else if (type == 3)
        {
            const V3 n_out = (cx.f2c.sgn[k0] / m->g.face_area_mag[f]) * ld3(m->g.face_area.data(), f);
            m->face_kind[f] = dot(u, n_out) < 0 ? 2 : 3;
        }
        else { g_err = "unknown boundary type"; return 1; }
    }
    for (int32_t f = 0; f < cx.nf; ++f)
        if (m->face_kind[f] == 2)
            for (int64_t k = cx.f2e.ptr[f]; k < cx.f2e.ptr[f + 1]; ++k) m->edge_mode[cx.f2e.idx[k]] = 2;
    for (int32_t f = 0; f < cx.nf; ++f)
        if (m->face_kind[f] == 1)
            for (int64_t k = cx.f2e.ptr[f]; k < cx.f2e.ptr[f + 1]; ++k) m->edge_mode[cx.f2e.idx[k]] = 1;
    m->have_bc = true;
    return 0;
}

int mhdg_mesh_desc_fill(const mhdg_mesh *m, mhdg_mesh_desc *d)
{
    if (!m->have_bc) { g_err = "mhdg_mesh_boundary has not been called"; return 1; }
    const mhdg::Complex &cx = m->cx;
    const mhdg::Geometry &g = m->g;
    std::memset(d, 0, sizeof(*d));
    d->n_edges = cx.ne; d->n_faces = cx.nf; d->n_cells = cx.nc;
    d->f2e = {cx.f2e.ptr.data(), cx.f2e.idx.data(), cx.f2e.sgn.data()};
    d->c2f = {cx.c2f.ptr.data(), cx.c2f.idx.data(), cx.c2f.sgn.data()};
    d->e2f = {cx.e2f.ptr.data(), cx.e2f.idx.data(), cx.e2f.sgn.data()};
    d->f2c = {cx.f2c.ptr.data(), cx.f2c.idx.data(), cx.f2c.sgn.data()};
    d->cell_material = cx.material.data();
    d->face_kind = m->face_kind.data();
    d->edge_mode = m->edge_mode.data();
    d->face_area = g.face_area.data(); d->face_center = g.face_center.data(); d->face_area_mag = g.face_area_mag.data();
    d->cell_center = g.cell_center.data(); d->cell_volume = g.cell_volume.data();
    d->edge_vec = g.edge_vec.data(); d->edge_mid = g.edge_mid.data(); d->edge_len = g.edge_len.data();
    d->ec_ptr = g.ec_ptr.data(); d->ec_cell = g.ec_cell.data(); d->ec_weight = g.ec_weight.data(); d->ef_weight = g.ef_weight.data();
    d->star1 = g.star1.data(); d->star2 = g.star2.data();
    d->e2n = {cx.e2n.ptr.data(), cx.e2n.idx.data(), cx.e2n.sgn.data()};
    d->n_nodes = cx.nn;
    return 0;
}

// MercuryUns/initial/Initial.cpp:66-84
int mhdg_mesh_background(const mhdg_mesh *m, const double B_imf[3], int dipole, const double dm[3], const double dc[3], double soft,
                         double *PhiAdd, double *BaddCell)
{
    using namespace mhdg;
    const Complex &cx = m->cx;
    const Geometry &g = m->g;
    const V3 B(B_imf[0], B_imf[1], B_imf[2]);
    for (int32_t f = 0; f < cx.nf; ++f) PhiAdd[f] = dot(B, ld3(g.face_area.data(), f));
    if (dipole)
    {
        const V3 mm(dm[0], dm[1], dm[2]), cc(dc[0], dc[1], dc[2]);
        std::vector<double> A(cx.ne, 0.0);
        for (int32_t e = 0; e < cx.ne; ++e)
        {
            const V3 r = ld3(g.edge_mid.data(), e) - cc;
            const double d2 = dot(r, r) + soft * soft;
            const V3 a = (1.0 / (d2 * std::sqrt(d2))) * cross(mm, r);
            A[e] = dot(a, ld3(g.edge_vec.data(), e));
        }
        for (int32_t f = 0; f < cx.nf; ++f)
            for (int64_t k = cx.f2e.ptr[f]; k < cx.f2e.ptr[f + 1]; ++k) PhiAdd[f] += cx.f2e.sgn[k] * A[cx.f2e.idx[k]];
    }
    reconstruct_cell_vector(cx, g, PhiAdd, BaddCell);
    return 0;
}

int mhdg_dipole_edge_integrals(int32_t n, const double *pos, const double *mom, const double *core, double cutoff, int64_t n_edges,
                               const double *xa, const double *xb, double *out)
{
    try
    {
        const double cut2 = cutoff * cutoff;
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < n_edges; ++e) out[e] = mhdg::dipole_edge_integral(n, pos, mom, core, cut2, xa + 3 * e, xb + 3 * e);
        return 0;
    }
    catch (const std::exception &ex) { g_err = ex.what(); return 1; }
}

int mhdg_mesh_background_dipoles(const mhdg_mesh *m, const double B_imf[3], int32_t n, const double *pos, const double *mom,
                                 const double *core, double cutoff, double *PhiAdd, double *BaddCell)
{
    try
    {
        using namespace mhdg;
        const Complex &cx = m->cx;
        const Geometry &g = m->g;
        const V3 B(B_imf[0], B_imf[1], B_imf[2]);
        for (int32_t f = 0; f < cx.nf; ++f) PhiAdd[f] = dot(B, ld3(g.face_area.data(), f));
        if (n > 0)
        {
            // A.dl along every edge from its start node to its end node, then the flux d1 A
            std::vector<double> A(cx.ne, 0.0);
            const double cut2 = cutoff * cutoff;
            std::exception_ptr err;
#pragma omp parallel for schedule(static)
            for (int32_t e = 0; e < cx.ne; ++e)
            {
                int32_t a, b;
                cx.edge_nodes(e, a, b);
                try { A[e] = dipole_edge_integral(n, pos, mom, core, cut2, &m->xyz[3 * size_t(a)], &m->xyz[3 * size_t(b)]); }
                catch (...)
                {
#pragma omp critical
                    err = std::current_exception();
                }
            }
            if (err) std::rethrow_exception(err);
            for (int32_t f = 0; f < cx.nf; ++f)
                for (int64_t k = cx.f2e.ptr[f]; k < cx.f2e.ptr[f + 1]; ++k) PhiAdd[f] += cx.f2e.sgn[k] * A[cx.f2e.idx[k]];
        }
        reconstruct_cell_vector(cx, g, PhiAdd, BaddCell);
        return 0;
    }
    catch (const std::exception &ex) { g_err = ex.what(); return 1; }
}

int mhdg_mesh_cell_vector(const mhdg_mesh *m, const double *flux, double *b)
{
    mhdg::reconstruct_cell_vector(m->cx, m->g, flux, b);
    return 0;
}

int mhdg_mesh_seeded_state(const mhdg_mesh *m, const mhdg_config *cfg, double noise, uint64_t seed, const double *Badd, double *UH,
                           double *UNa, double *Phi)
{
    const mhdg::Complex &cx = m->cx;
    const mhdg::Geometry &g = m->g;
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> N(0.0, 1.0);
    const double gm1 = cfg->gamma - 1.0;
    for (int32_t c = 0; c < cx.nc; ++c)
        for (int s = 0; s < 2; ++s)
        {
            const double *fs = cfg->fs[s];
            const double umag = std::sqrt(fs[1] * fs[1] + fs[2] * fs[2] + fs[3] * fs[3]);
            const double rho = fs[0] * (1.0 + noise * N(rng));
            const double u = fs[1] + noise * umag * N(rng), v = fs[2] + noise * umag * N(rng), w = fs[3] + noise * umag * N(rng);
            const double p = fs[4] * (1.0 + noise * N(rng));
            double *U = (s == 0 ? UH : UNa) + 5 * size_t(c);
            U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = rho * w;
            U[4] = p / gm1 + 0.5 * rho * (u * u + v * v + w * w);
        }
    std::vector<double> A(cx.ne);
    for (int32_t e = 0; e < cx.ne; ++e) A[e] = N(rng);
    double s_phi = 0, s_b = 0;
    for (int32_t f = 0; f < cx.nf; ++f)
    {
        double v = 0;
        for (int64_t k = cx.f2e.ptr[f]; k < cx.f2e.ptr[f + 1]; ++k) v += cx.f2e.sgn[k] * A[cx.f2e.idx[k]];
        Phi[f] = v;
        const double bn = v / g.face_area_mag[f];
        s_phi += bn * bn;
    }
    for (int32_t c = 0; c < cx.nc; ++c)
        s_b += Badd[3 * size_t(c)] * Badd[3 * size_t(c)] + Badd[3 * size_t(c) + 1] * Badd[3 * size_t(c) + 1] + Badd[3 * size_t(c) + 2] * Badd[3 * size_t(c) + 2];
    const double rms_phi = std::sqrt(s_phi / cx.nf), rms_b = std::sqrt(s_b / cx.nc);
    const double scale = rms_phi > 0 ? noise * rms_b / rms_phi : 0.0;
    for (int32_t f = 0; f < cx.nf; ++f) Phi[f] *= scale;
    return 0;
}

int mhdg_mesh_write_vtu(const mhdg_mesh *m, const char *path, int32_t n_cell_arrays, const mhdg_vtk_array *cell_arrays,
                        int32_t n_point_arrays, const mhdg_vtk_array *point_arrays)
{
    try
    {
        mhdg::vtk::write_vtu(path, m->cx, m->xyz, m->cx.nc, vtk_arrays(n_cell_arrays, cell_arrays), vtk_arrays(n_point_arrays, point_arrays));
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}
int mhdg_write_pvtu(const char *path, int32_t n_pieces, const char *const *pieces, int32_t n_cell_arrays, const mhdg_vtk_array *cell_arrays,
                    int32_t n_point_arrays, const mhdg_vtk_array *point_arrays)
{
    try
    {
        // master files carry names only: values may be NULL here
        auto names = [](int32_t n, const mhdg_vtk_array *a) {
            std::vector<mhdg::vtk::Array> out;
            for (int32_t i = 0; i < n; ++i) out.push_back({a[i].name ? a[i].name : "", a[i].ncomp, nullptr});
            return out;
        };
        std::vector<std::string> p(pieces, pieces + n_pieces);
        mhdg::vtk::write_pvtu(path, p, names(n_cell_arrays, cell_arrays), names(n_point_arrays, point_arrays));
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}
int mhdg_write_pvd(const char *path, int32_t n, const double *times, const char *const *files)
{
    try
    {
        std::vector<std::pair<double, std::string>> e;
        for (int32_t i = 0; i < n; ++i) e.emplace_back(times[i], files[i]);
        mhdg::vtk::write_pvd(path, e);
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}
}

// ---------------------------------------------------------------------------- multigrid hierarchy
struct mhdg_mg
{
    mhdg::MgHierarchy h;
};

extern "C" {
int mhdg_mg_build(const mhdg_mesh_desc *mesh, const mhdg_config *cfg, const int8_t *dirichlet, mhdg_mg **out)
{
    try
    {
        using namespace mhdg;
        check_desc(*mesh);
        const Tables t = build_tables(*mesh, *cfg, 32);
        std::vector<int32_t> tail, head;
        edge_nodes(*mesh, tail, head);
        std::vector<int8_t> dir(size_t(mesh->n_edges), 0);
        if (dirichlet) dir.assign(dirichlet, dirichlet + mesh->n_edges);
        auto m = std::make_unique<mhdg_mg>();
        m->h = build_mg_hierarchy(mesh->n_nodes, mesh->n_edges, tail.data(), head.data(), mesh->star1, mesh->n_faces, t.f2e_ptr.data(),
                                  t.f2e.data(), mesh->star2, t.m_e.data(), dir.data(), cfg->mg_levels > 0 ? cfg->mg_levels : 4,
                                  cfg->mg_agglom_passes > 0 ? cfg->mg_agglom_passes : 3);
        *out = m.release();
        return 0;
    }
    catch (const std::exception &e) { g_err = e.what(); return 1; }
}
void mhdg_mg_free(mhdg_mg *mg) { delete mg; }
int mhdg_mg_levels(const mhdg_mg *mg) { return static_cast<int>(mg->h.lv.size()); }
int mhdg_mg_level_size(const mhdg_mg *mg, int level, int32_t *n, int32_t *n_nodes, int64_t *k_nnz)
{
    if (level < 0 || level >= static_cast<int>(mg->h.lv.size())) { g_err = "multigrid level out of range"; return 1; }
    const mhdg::MgLevel &L = mg->h.lv[size_t(level)];
    if (n) *n = L.n;
    if (n_nodes) *n_nodes = L.n_nodes;
    if (k_nnz) *k_nnz = static_cast<int64_t>(L.k_col.size());
    return 0;
}
int mhdg_mg_level_nodes(const mhdg_mg *mg, int level, int32_t *tail, int32_t *head, int32_t *agg)
{
    if (level < 0 || level >= static_cast<int>(mg->h.lv.size())) { g_err = "multigrid level out of range"; return 1; }
    const mhdg::MgLevel &L = mg->h.lv[size_t(level)];
    auto put = [](const auto &v, auto *o) { if (o) std::copy(v.begin(), v.end(), o); };
    put(L.tail, tail); put(L.head, head); put(L.agg, agg);
    return 0;
}
int mhdg_mg_level_arrays(const mhdg_mg *mg, int level, int32_t *map, double *mass, int32_t *k_ptr, int32_t *k_col, double *k_val)
{
    if (level < 0 || level >= static_cast<int>(mg->h.lv.size())) { g_err = "multigrid level out of range"; return 1; }
    const mhdg::MgLevel &L = mg->h.lv[size_t(level)];
    auto put = [](const auto &v, auto *o) { if (o) std::copy(v.begin(), v.end(), o); };
    put(L.map, map); put(L.mass, mass); put(L.k_ptr, k_ptr); put(L.k_col, k_col); put(L.k_val, k_val);
    return 0;
}
}
