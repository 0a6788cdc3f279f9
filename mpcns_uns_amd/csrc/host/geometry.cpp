#include "geometry.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <limits>

#include "vec3.h"

namespace mhdg
{
namespace
{
// unique sorted node list of a cell (all its face loops)
int cell_nodes(const Complex &cx, int32_t c, int32_t *out)
{
    int n = 0;
    int32_t loop[8];
    int m;
    for (int64_t k = cx.c2f.ptr[c]; k < cx.c2f.ptr[c + 1]; ++k)
    {
        cx.face_loop(cx.c2f.idx[k], loop, m);
        for (int i = 0; i < m; ++i) out[n++] = loop[i];
    }
    std::sort(out, out + n);
    return static_cast<int>(std::unique(out, out + n) - out);
}

// dual piece of the i-th loop edge of face f (c2f sign s) inside cell c with centre xc
inline V3 dual_piece(const Complex &cx, const Geometry &g, int64_t fe, int s, const V3 &xc, const V3 &cf)
{
    const int32_t e = cx.f2e.idx[fe];
    const int se = s * cx.f2e.sgn[fe];
    const V3 mid = ld3(g.edge_mid.data(), e);
    return (0.5 * se) * cross(xc - mid, cf - mid);
}
} // namespace

Geometry compute_geometry(const Complex &cx, const double *xyz)
{
    const bool tm = std::getenv("MHDG_SETUP_TIMING") != nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto lap = [&](const char *what) {
        if (!tm) return;
        const auto t = std::chrono::steady_clock::now();
        std::fprintf(stderr, "  geometry %-12s %.2fs\n", what, std::chrono::duration<double>(t - t0).count());
        t0 = t;
    };
    const int32_t ne = cx.ne, nf = cx.nf, nc = cx.nc;
    Geometry g;
    g.edge_vec.resize(3 * size_t(ne)); g.edge_mid.resize(3 * size_t(ne)); g.edge_len.resize(ne);
#pragma omp parallel for schedule(static)
    for (int32_t e = 0; e < ne; ++e)
    {
        int32_t a, b;
        cx.edge_nodes(e, a, b);
        const V3 xa = ld3(xyz, a), xb = ld3(xyz, b);
        const V3 t = xb - xa;
        st3(g.edge_vec.data(), e, t);
        st3(g.edge_mid.data(), e, 0.5 * (xa + xb));
        g.edge_len[e] = norm(t);
    }

    g.face_area.resize(3 * size_t(nf)); g.face_center.resize(3 * size_t(nf)); g.face_area_mag.resize(nf);
#pragma omp parallel for schedule(static)
    for (int32_t f = 0; f < nf; ++f)
    {
        int32_t loop[8];
        int k;
        cx.face_loop(f, loop, k);
        V3 p0;
        for (int i = 0; i < k; ++i) p0 += ld3(xyz, loop[i]);
        p0 *= 1.0 / static_cast<double>(k);
        V3 cw;
        double wsum = 0;
        for (int i = 0; i < k; ++i)
        {
            const V3 a = ld3(xyz, loop[i]), b = ld3(xyz, loop[(i + 1) % k]);
            const double w = 0.5 * norm(cross(a - p0, b - p0));
            cw += w * ((p0 + a + b) * (1.0 / 3.0));
            wsum += w;
        }
        const V3 c = wsum > 0 ? cw * (1.0 / wsum) : p0;
        V3 A;
        for (int i = 0; i < k; ++i) A += 0.5 * cross(ld3(xyz, loop[i]) - c, ld3(xyz, loop[(i + 1) % k]) - c);
        st3(g.face_center.data(), f, c);
        st3(g.face_area.data(), f, A);
        g.face_area_mag[f] = norm(A);
    }

    lap("edges+faces");
    // edge -> incident cells (sorted unique)
    g.ec_ptr.assign(size_t(ne) + 1, 0);
    {
        auto gather = [&](int32_t e, std::vector<int32_t> &tmp) {
            tmp.clear();
            for (int64_t k = cx.e2f.ptr[e]; k < cx.e2f.ptr[e + 1]; ++k)
            {
                const int32_t f = cx.e2f.idx[k];
                for (int64_t q = cx.f2c.ptr[f]; q < cx.f2c.ptr[f + 1]; ++q) tmp.push_back(cx.f2c.idx[q]);
            }
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        };
        std::vector<int32_t> cnt(ne);
#pragma omp parallel
        {
            std::vector<int32_t> tmp;
#pragma omp for schedule(static)
            for (int32_t e = 0; e < ne; ++e) { gather(e, tmp); cnt[e] = static_cast<int32_t>(tmp.size()); }
        }
        for (int32_t e = 0; e < ne; ++e) g.ec_ptr[e + 1] = g.ec_ptr[e] + cnt[e];
        g.ec_cell.resize(g.ec_ptr[ne]);
#pragma omp parallel
        {
            std::vector<int32_t> tmp;
#pragma omp for schedule(static)
            for (int32_t e = 0; e < ne; ++e) { gather(e, tmp); std::copy(tmp.begin(), tmp.end(), g.ec_cell.begin() + g.ec_ptr[e]); }
        }
    }

    lap("ec lists");
    // cell centres (pass 1) and volumes (pass 2): independent per cell
    g.cell_center.resize(3 * size_t(nc)); g.cell_volume.assign(nc, 0.0); g.cell_min_tet.assign(nc, 0.0);
    g.star3.assign(nc, 0.0);
#pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < nc; ++c)
    {
        int32_t cn[64], loop[8];
        const int nn = cell_nodes(cx, c, cn);
        V3 xc0;
        for (int i = 0; i < nn; ++i) xc0 += ld3(xyz, cn[i]);
        xc0 *= 1.0 / static_cast<double>(nn);
        double vsum = 0;
        V3 vc;
        for (int64_t kf = cx.c2f.ptr[c]; kf < cx.c2f.ptr[c + 1]; ++kf)
        {
            const int32_t f = cx.c2f.idx[kf];
            const int s = cx.c2f.sgn[kf];
            int k;
            cx.face_loop(f, loop, k);
            const V3 cf = ld3(g.face_center.data(), f);
            for (int i = 0; i < k; ++i)
            {
                const V3 a = ld3(xyz, loop[i]), b = ld3(xyz, loop[(i + 1) % k]);
                const V3 at = (0.5 * s) * cross(a - cf, b - cf);
                const double vt = dot(at, cf - xc0) / 3.0;
                vc += vt * ((xc0 + cf + a + b) * 0.25);
                vsum += vt;
            }
        }
        const V3 xc = vsum > 0 ? vc * (1.0 / vsum) : xc0;
        st3(g.cell_center.data(), c, xc);
        double vol = 0, mintet = std::numeric_limits<double>::max();
        for (int64_t kf = cx.c2f.ptr[c]; kf < cx.c2f.ptr[c + 1]; ++kf)
        {
            const int32_t f = cx.c2f.idx[kf];
            const int s = cx.c2f.sgn[kf];
            int k;
            cx.face_loop(f, loop, k);
            const V3 cf = ld3(g.face_center.data(), f);
            for (int i = 0; i < k; ++i)
            {
                const V3 a = ld3(xyz, loop[i]), b = ld3(xyz, loop[(i + 1) % k]);
                const V3 at = (0.5 * s) * cross(a - cf, b - cf);
                const double vt = dot(at, cf - xc) / 3.0;
                vol += vt;
                mintet = std::min(mintet, vt);
            }
        }
        g.cell_volume[c] = vol;
        g.cell_min_tet[c] = mintet;
        g.star3[c] = vol > 0 ? 1.0 / vol : 0.0;
    }

    lap("cells");
    // w_{e,c}: the pieces of cell c at edge e, in (c2f, loop position) order
    g.ec_weight.assign(g.ec_cell.size(), 0.0);
    g.edge_dual_area.assign(3 * size_t(ne), 0.0);
    g.ec_misoriented.assign(ne, 0);
#pragma omp parallel for schedule(static)
    for (int32_t e = 0; e < ne; ++e)
    {
        V3 dual;
        int mis = 0;
        const V3 t = ld3(g.edge_vec.data(), e);
        for (int64_t q = g.ec_ptr[e]; q < g.ec_ptr[e + 1]; ++q)
        {
            const int32_t c = g.ec_cell[q];
            const V3 xc = ld3(g.cell_center.data(), c);
            double w = 0;
            for (int64_t kf = cx.c2f.ptr[c]; kf < cx.c2f.ptr[c + 1]; ++kf)
            {
                const int32_t f = cx.c2f.idx[kf];
                const V3 cf = ld3(g.face_center.data(), f);
                for (int64_t fe = cx.f2e.ptr[f]; fe < cx.f2e.ptr[f + 1]; ++fe)
                {
                    if (cx.f2e.idx[fe] != e) continue;
                    const V3 piece = dual_piece(cx, g, fe, cx.c2f.sgn[kf], xc, cf);
                    dual += piece;
                    if (dot(piece, t) <= 0) ++mis;
                    w += norm(piece);
                }
            }
            g.ec_weight[q] = w;
        }
        st3(g.edge_dual_area.data(), e, dual);
        g.ec_misoriented[e] = mis;
    }

    lap("w_ec");
    // w_{e,f}: pieces of the cells of f (ascending) through f at e; dual length of faces
    g.ef_weight.assign(cx.e2f.idx.size(), 0.0);
#pragma omp parallel for schedule(static)
    for (int32_t e = 0; e < ne; ++e)
        for (int64_t k = cx.e2f.ptr[e]; k < cx.e2f.ptr[e + 1]; ++k)
        {
            const int32_t f = cx.e2f.idx[k];
            const V3 cf = ld3(g.face_center.data(), f);
            int64_t fe = cx.f2e.ptr[f];
            while (cx.f2e.idx[fe] != e) ++fe;
            double w = 0;
            for (int64_t q = cx.f2c.ptr[f]; q < cx.f2c.ptr[f + 1]; ++q)
            {
                const int32_t c = cx.f2c.idx[q];
                // c2f sign of f in c equals the f2c sign
                w += norm(dual_piece(cx, g, fe, cx.f2c.sgn[q], ld3(g.cell_center.data(), c), cf));
            }
            g.ef_weight[k] = w;
        }

    lap("w_ef");
    g.star1.assign(ne, 0.0);
    g.star2.assign(nf, 0.0);
#pragma omp parallel for schedule(static)
    for (int32_t e = 0; e < ne; ++e)
    {
        double s = 0;
        for (int64_t k = g.ec_ptr[e]; k < g.ec_ptr[e + 1]; ++k) s += g.ec_weight[k];
        if (s > 0)
        {
            for (int64_t k = g.ec_ptr[e]; k < g.ec_ptr[e + 1]; ++k) g.ec_weight[k] /= s;
            for (int64_t k = cx.e2f.ptr[e]; k < cx.e2f.ptr[e + 1]; ++k) g.ef_weight[k] /= s;
        }
        g.star1[e] = g.edge_len[e] > 0 ? norm(ld3(g.edge_dual_area.data(), e)) / g.edge_len[e] : 0.0;
    }
#pragma omp parallel for schedule(static)
    for (int32_t f = 0; f < nf; ++f)
    {
        double dl = 0;
        const V3 cf = ld3(g.face_center.data(), f);
        for (int64_t q = cx.f2c.ptr[f]; q < cx.f2c.ptr[f + 1]; ++q) dl += norm(cf - ld3(g.cell_center.data(), cx.f2c.idx[q]));
        g.star2[f] = g.face_area_mag[f] > 0 ? dl / g.face_area_mag[f] : 0.0;
    }
    lap("stars");
    return g;
}

void reconstruct_cell_vector(const Complex &cx, const Geometry &g, const double *flux, double *b3)
{
#pragma omp parallel for schedule(static)
    for (int32_t c = 0; c < cx.nc; ++c)
    {
        double M[3][3] = {{0}}, r[3] = {0};
        for (int64_t k = cx.c2f.ptr[c]; k < cx.c2f.ptr[c + 1]; ++k)
        {
            const int32_t f = cx.c2f.idx[k];
            const double A = g.face_area_mag[f];
            if (A <= 0) continue;
            const V3 n = (cx.c2f.sgn[k] / A) * ld3(g.face_area.data(), f);
            const double nv[3] = {n.x, n.y, n.z};
            const double phi = cx.c2f.sgn[k] * flux[f];
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
            double I[3][3];
            I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
            I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
            I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
            I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
            I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
            I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
            I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
            I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
            I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
            for (int i = 0; i < 3; ++i) out[i] = I[i][0] * r[0] + I[i][1] * r[1] + I[i][2] * r[2];
        }
        b3[3 * size_t(c)] = out[0]; b3[3 * size_t(c) + 1] = out[1]; b3[3 * size_t(c) + 2] = out[2];
    }
}
} // namespace mhdg
