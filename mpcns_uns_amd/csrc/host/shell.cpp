#include "shell.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>

namespace mhdg
{
void tet_loops(const int32_t *t, int32_t L[12])
{
    const int32_t a = t[0], b = t[1], c = t[2], d = t[3];
    const int32_t v[12] = {a, c, b, a, b, d, a, d, c, b, c, d};
    for (int i = 0; i < 12; ++i) L[i] = v[i];
}

ShellMesh make_shell(const ShellSpec &s)
{
    if (s.n_lat < 2 || s.n_lon < 3 || s.n_r < 1) throw std::runtime_error("shell: n_lat >= 2, n_lon >= 3, n_r >= 1");
    const int nlat = s.n_lat, nlon = s.n_lon, nr = s.n_r;
    const double pi = 3.14159265358979323846;
    // radial levels
    std::vector<double> r(nr + 1), dr(nr);
    {
        double w = 1, sum = 0;
        for (int k = 0; k < nr; ++k) { dr[k] = w; sum += w; w *= s.ratio; }
        r[0] = s.r_in;
        for (int k = 0; k < nr; ++k) { dr[k] *= (s.r_out - s.r_in) / sum; r[k + 1] = r[k] + dr[k]; }
        r[nr] = s.r_out;
    }
    const int per_level = 2 + (nlat - 1) * nlon;
    auto nid = [&](int i, int j, int k) -> int32_t { // j = 0 north pole, j = nlat south pole
        if (j == 0) return k * per_level;
        if (j == nlat) return k * per_level + 1;
        return k * per_level + 2 + (j - 1) * nlon + ((i % nlon) + nlon) % nlon;
    };
    ShellMesh m;
    m.xyz.resize(3 * size_t(per_level) * (nr + 1));
    std::mt19937_64 rng(s.seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double dth = pi / nlat, dph = 2 * pi / nlon;
    auto place = [&](int32_t id, double rr, double th, double ph) {
        m.xyz[3 * size_t(id)] = rr * std::sin(th) * std::cos(ph);
        m.xyz[3 * size_t(id) + 1] = rr * std::sin(th) * std::sin(ph);
        m.xyz[3 * size_t(id) + 2] = rr * std::cos(th);
    };
    for (int k = 0; k <= nr; ++k)
    {
        const bool interior = k > 0 && k < nr;
        const double hr = interior ? std::min(dr[k - 1], dr[k]) : 0.0;
        for (int pole = 0; pole < 2; ++pole)
        {
            const int32_t id = nid(0, pole ? nlat : 0, k);
            double rr = r[k];
            double jx = 0, jy = 0;
            if (interior)
            {
                rr += s.jitter * hr * U(rng);
                jx = s.jitter * r[k] * dth * U(rng);
                jy = s.jitter * r[k] * dth * U(rng);
            }
            m.xyz[3 * size_t(id)] = jx;
            m.xyz[3 * size_t(id) + 1] = jy;
            m.xyz[3 * size_t(id) + 2] = pole ? -rr : rr;
        }
        for (int j = 1; j < nlat; ++j)
            for (int i = 0; i < nlon; ++i)
            {
                double rr = r[k], th = j * dth, ph = i * dph;
                if (interior)
                {
                    rr += s.jitter * hr * U(rng);
                    th += s.jitter * dth * U(rng);
                    ph += s.jitter * dph * U(rng);
                }
                place(nid(i, j, k), rr, th, ph);
            }
    }
    // Kuhn split of hexahedron (j, i, k); corner q at (i + (q&1), j + ((q>>1)&1), k + ((q>>2)&1))
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    m.tets.reserve(size_t(nlat) * nlon * nr * 24);
    for (int j = 0; j < nlat; ++j)
        for (int i = 0; i < nlon; ++i)
            for (int k = 0; k < nr; ++k)
            {
                int32_t v[8];
                for (int q = 0; q < 8; ++q) v[q] = nid(i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1));
                for (const auto &p : perm)
                {
                    const int q1 = 1 << p[0], q2 = q1 | (1 << p[1]);
                    int32_t t[4] = {v[0], v[q1], v[q2], v[7]};
                    if (t[0] == t[1] || t[0] == t[2] || t[0] == t[3] || t[1] == t[2] || t[1] == t[3] || t[2] == t[3]) continue;
                    auto X = [&](int32_t n, int a) { return m.xyz[3 * size_t(n) + a]; };
                    double e[3][3];
                    for (int a = 1; a < 4; ++a)
                        for (int c = 0; c < 3; ++c) e[a - 1][c] = X(t[a], c) - X(t[0], c);
                    const double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) - e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                                       e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
                    if (det < 0) std::swap(t[2], t[3]);
                    m.tets.insert(m.tets.end(), t, t + 4);
                    // boundary triangles: faces with all three nodes on the inner / outer level
                    int32_t L[12];
                    tet_loops(t, L);
                    for (int fc = 0; fc < 4; ++fc)
                    {
                        int lv[3];
                        for (int a = 0; a < 3; ++a) lv[a] = L[3 * fc + a] / per_level;
                        if (lv[0] == lv[1] && lv[1] == lv[2] && (lv[0] == 0 || lv[0] == nr))
                        {
                            m.tags.insert(m.tags.end(), {L[3 * fc], L[3 * fc + 1], L[3 * fc + 2], lv[0] == 0 ? 1 : 2});
                        }
                    }
                }
            }
    return m;
}

Complex build_tet_complex(int32_t n_nodes, const int32_t *tets, int64_t n_cells, const int32_t *tags, int64_t n_tags, bool sorted_builder)
{
    if (sorted_builder && n_nodes < (1 << 21))
    {
        std::vector<int32_t> loops(12 * size_t(n_cells));
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < n_cells; ++c) tet_loops(tets + 4 * c, &loops[12 * c]);
        return build_triangle_complex(n_nodes, loops.data(), n_cells, 4, tags, n_tags);
    }
    Builder b(n_nodes, n_cells);
    const int32_t lptr[5] = {0, 3, 6, 9, 12};
    int32_t L[12];
    for (int64_t c = 0; c < n_cells; ++c)
    {
        tet_loops(tets + 4 * c, L);
        b.add_cell(lptr, L, 4, 0);
    }
    for (int64_t t = 0; t < n_tags; ++t) b.tag_face(tags + 4 * t, 3, tags[4 * t + 3]);
    return b.finish();
}
} // namespace mhdg
