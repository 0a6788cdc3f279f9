// Agglomeration hierarchy of the edge complex (multigrid.h).
#include "multigrid.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>

namespace mhdg
{
namespace
{
inline int32_t unfold(int32_t v) { return v >= 0 ? v : ~v; }
inline double fsign(int32_t v) { return v >= 0 ? 1.0 : -1.0; }

// Heavy-edge matching of the node graph (n nodes, edges tail/head with weights), `passes` times over the groups
// formed so far: each unmatched group, in group order, pairs with its unmatched neighbour of largest total weight;
// groups left unmatched join the neighbouring pair they are most strongly connected to; groups are renumbered in
// order of their representative.  The steps of MercuryUns/multigrid/Multigrid.cpp:219-282 on cells, here on nodes.
std::vector<int32_t> match_nodes(int32_t n, const std::vector<int32_t> &tail, const std::vector<int32_t> &head,
                                 const std::vector<double> &w, int passes, int32_t &n_groups)
{
    std::vector<int32_t> grp(n);
    std::iota(grp.begin(), grp.end(), 0);
    int32_t ng = n;
    const size_t ne = tail.size();
    for (int pass = 0; pass < passes && ng > 1; ++pass)
    {
        // group adjacency: (neighbour, summed weight), neighbours ascending
        std::vector<int64_t> cnt(size_t(ng) + 1, 0);
        for (size_t e = 0; e < ne; ++e)
        {
            const int32_t a = grp[tail[e]], b = grp[head[e]];
            if (a == b) continue;
            ++cnt[size_t(a) + 1];
            ++cnt[size_t(b) + 1];
        }
        for (int32_t g = 0; g < ng; ++g) cnt[size_t(g) + 1] += cnt[size_t(g)];
        std::vector<std::pair<int32_t, double>> adj(size_t(cnt[size_t(ng)]));
        {
            std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
            for (size_t e = 0; e < ne; ++e)
            {
                const int32_t a = grp[tail[e]], b = grp[head[e]];
                if (a == b) continue;
                const double x = std::abs(w[e]);
                adj[size_t(pos[size_t(a)]++)] = {b, x};
                adj[size_t(pos[size_t(b)]++)] = {a, x};
            }
        }
        std::vector<int64_t> ptr(size_t(ng) + 1, 0);
        std::vector<int32_t> nb;
        std::vector<double> nw;
        nb.reserve(adj.size());
        nw.reserve(adj.size());
        for (int32_t g = 0; g < ng; ++g)
        {
            auto b0 = adj.begin() + cnt[size_t(g)], b1 = adj.begin() + cnt[size_t(g) + 1];
            std::sort(b0, b1, [](const auto &x, const auto &y) { return x.first < y.first; });
            for (auto it = b0; it != b1;)
            {
                auto jt = it;
                double s = 0;
                for (; jt != b1 && jt->first == it->first; ++jt) s += jt->second;
                nb.push_back(it->first);
                nw.push_back(s);
                it = jt;
            }
            ptr[size_t(g) + 1] = int64_t(nb.size());
        }
        std::vector<int32_t> mate(size_t(ng), -1), rep(size_t(ng), -1);
        for (int32_t g = 0; g < ng; ++g)
        {
            if (mate[size_t(g)] >= 0) continue;
            int32_t best = -1;
            double bw = 0;
            for (int64_t k = ptr[size_t(g)]; k < ptr[size_t(g) + 1]; ++k)
                if (mate[size_t(nb[size_t(k)])] < 0 && nw[size_t(k)] > bw) { bw = nw[size_t(k)]; best = nb[size_t(k)]; }
            if (best >= 0)
            {
                mate[size_t(g)] = best;
                mate[size_t(best)] = g;
                rep[size_t(g)] = rep[size_t(best)] = std::min(g, best);
            }
        }
        for (int32_t g = 0; g < ng; ++g)
        {
            if (rep[size_t(g)] >= 0) continue;
            int32_t best = -1;
            double bw = 0;
            for (int64_t k = ptr[size_t(g)]; k < ptr[size_t(g) + 1]; ++k)
                if (rep[size_t(nb[size_t(k)])] >= 0 && nw[size_t(k)] > bw) { bw = nw[size_t(k)]; best = nb[size_t(k)]; }
            rep[size_t(g)] = best >= 0 ? rep[size_t(best)] : g;
        }
        std::vector<int32_t> label(size_t(ng), -1);
        int32_t nn = 0;
        for (int32_t g = 0; g < ng; ++g)
            if (rep[size_t(g)] == g) label[size_t(g)] = nn++;
        for (int32_t g = 0; g < ng; ++g) label[size_t(g)] = label[size_t(rep[size_t(g)])];
        if (nn >= ng) break;
        for (auto &x : grp) x = label[size_t(x)];
        ng = nn;
    }
    n_groups = ng;
    return grp;
}

// Rows of a sparse matrix accumulated as (column, value) pairs, one row per task: columns sorted, duplicates summed
void compress_row(std::vector<std::pair<int32_t, double>> &row, std::vector<int32_t> &col, std::vector<double> &val)
{
    std::sort(row.begin(), row.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    for (size_t i = 0; i < row.size();)
    {
        size_t j = i;
        double s = 0;
        for (; j < row.size() && row[j].first == row[i].first; ++j) s += row[j].second;
        col.push_back(row[i].first);
        val.push_back(s);
        i = j;
    }
}

// CSR from per-row (column, value) lists built in parallel
template <class RowFn> void assemble(int32_t n, RowFn fn, std::vector<int32_t> &ptr, std::vector<int32_t> &col, std::vector<double> &val)
{
    std::vector<std::vector<int32_t>> cols(static_cast<size_t>(n));
    std::vector<std::vector<double>> vals(static_cast<size_t>(n));
#pragma omp parallel
    {
        std::vector<std::pair<int32_t, double>> row;
#pragma omp for schedule(dynamic, 256)
        for (int32_t I = 0; I < n; ++I)
        {
            row.clear();
            fn(I, row);
            compress_row(row, cols[size_t(I)], vals[size_t(I)]);
        }
    }
    ptr.assign(size_t(n) + 1, 0);
    for (int32_t I = 0; I < n; ++I) ptr[size_t(I) + 1] = ptr[size_t(I)] + int32_t(cols[size_t(I)].size());
    col.resize(size_t(ptr[size_t(n)]));
    val.resize(size_t(ptr[size_t(n)]));
    for (int32_t I = 0; I < n; ++I)
    {
        std::copy(cols[size_t(I)].begin(), cols[size_t(I)].end(), col.begin() + ptr[size_t(I)]);
        std::copy(vals[size_t(I)].begin(), vals[size_t(I)].end(), val.begin() + ptr[size_t(I)]);
    }
}

// Coarse unknowns of level `c` from the unknowns of its finer level `f` and the node aggregation agg
void coarsen(const MgLevel &f, const std::vector<int32_t> &agg, int32_t n_agg, const int8_t *dir, MgLevel &c)
{
    const int32_t n = f.n;
    std::vector<int64_t> key;
    key.reserve(size_t(n));
    for (int32_t i = 0; i < n; ++i)
    {
        if (dir && dir[i]) continue;
        const int64_t a = agg[size_t(f.tail[size_t(i)])], b = agg[size_t(f.head[size_t(i)])];
        if (a != b) key.push_back(std::min(a, b) * int64_t(n_agg) + std::max(a, b));
    }
    std::sort(key.begin(), key.end());
    key.erase(std::unique(key.begin(), key.end()), key.end());
    if (key.size() > size_t(INT32_MAX)) throw std::runtime_error("multigrid: too many coarse unknowns");
    c.n = int32_t(key.size());
    c.n_nodes = n_agg;
    c.tail.resize(key.size());
    c.head.resize(key.size());
    c.weight.assign(key.size(), 0.0);
    for (size_t k = 0; k < key.size(); ++k)
    {
        c.tail[k] = int32_t(key[k] / n_agg);
        c.head[k] = int32_t(key[k] % n_agg);
    }
    c.map.assign(size_t(n), kMgNone);
    std::vector<int32_t> count(key.size() + 1, 0);
    for (int32_t i = 0; i < n; ++i)
    {
        if (dir && dir[i]) continue;
        const int64_t a = agg[size_t(f.tail[size_t(i)])], b = agg[size_t(f.head[size_t(i)])];
        if (a == b) continue;
        const int64_t k = std::lower_bound(key.begin(), key.end(), std::min(a, b) * int64_t(n_agg) + std::max(a, b)) - key.begin();
        c.map[size_t(i)] = a < b ? int32_t(k) : ~int32_t(k);
        c.weight[size_t(k)] += std::abs(f.weight[size_t(i)]);
        ++count[size_t(k) + 1];
    }
    c.pt_ptr.assign(key.size() + 1, 0);
    for (size_t k = 0; k < key.size(); ++k) c.pt_ptr[k + 1] = c.pt_ptr[k] + count[k + 1];
    c.pt.resize(size_t(c.pt_ptr.back()));
    std::vector<int32_t> pos(c.pt_ptr.begin(), c.pt_ptr.end() - 1);
    for (int32_t i = 0; i < n; ++i) // ascending finer unknowns in every row
    {
        const int32_t v = c.map[size_t(i)];
        if (v == kMgNone) continue;
        const int32_t k = unfold(v);
        c.pt[size_t(pos[size_t(k)]++)] = v >= 0 ? i : ~i;
    }
    c.mass.assign(key.size(), 0.0);
    for (size_t k = 0; k < key.size(); ++k)
        for (int32_t q = c.pt_ptr[k]; q < c.pt_ptr[k + 1]; ++q) c.mass[k] += f.mass[size_t(unfold(c.pt[size_t(q)]))];
}
} // namespace

void edge_nodes(const mhdg_mesh_desc &d, std::vector<int32_t> &tail, std::vector<int32_t> &head)
{
    if (!d.e2n.ptr || !d.e2n.idx || !d.e2n.sgn || d.n_nodes <= 0) throw std::runtime_error("the mesh descriptor carries no e2n incidence");
    tail.assign(size_t(d.n_edges), -1);
    head.assign(size_t(d.n_edges), -1);
    for (int32_t e = 0; e < d.n_edges; ++e)
    {
        if (d.e2n.ptr[e + 1] - d.e2n.ptr[e] != 2) throw std::runtime_error("e2n: an edge must have two nodes");
        for (int64_t k = d.e2n.ptr[e]; k < d.e2n.ptr[e + 1]; ++k)
        {
            if (d.e2n.idx[k] < 0 || d.e2n.idx[k] >= d.n_nodes) throw std::runtime_error("e2n: node out of range");
            (d.e2n.sgn[k] < 0 ? tail : head)[size_t(e)] = d.e2n.idx[k];
        }
        if (tail[size_t(e)] < 0 || head[size_t(e)] < 0) throw std::runtime_error("e2n: an edge needs a start and an end node");
    }
}

MgHierarchy build_mg_hierarchy(int32_t n_nodes, int32_t ne, const int32_t *e_tail, const int32_t *e_head, const double *star1,
                               int32_t nf, const int32_t *f2e_ptr, const int32_t *f2e, const double *star2, const double *m_e,
                               const int8_t *dir, int levels, int passes)
{
    if (levels < 1 || passes < 1) throw std::runtime_error("multigrid: levels and agglomeration passes must be at least 1");
    if (n_nodes <= 0 || !e_tail || !e_head) throw std::runtime_error("multigrid: the mesh carries no edge -> node incidence");
    MgHierarchy H;
    H.lv.resize(1);
    MgLevel &L0 = H.lv[0];
    L0.n = ne;
    L0.n_nodes = n_nodes;
    L0.tail.assign(e_tail, e_tail + ne);
    L0.head.assign(e_head, e_head + ne);
    L0.weight.assign(star1, star1 + ne);
    L0.mass.assign(size_t(ne), 1.0);
    for (int32_t e = 0; e < ne; ++e)
    {
        if (L0.tail[size_t(e)] < 0 || L0.tail[size_t(e)] >= n_nodes || L0.head[size_t(e)] < 0 || L0.head[size_t(e)] >= n_nodes)
            throw std::runtime_error("multigrid: edge node out of range");
        if (dir[e]) L0.mass[size_t(e)] = 0.0;
    }
    // edge -> (face, sign of the edge in the face's loop) for the first Galerkin product
    std::vector<int32_t> e2f_ptr(size_t(ne) + 1, 0), e2f;
    for (int32_t k = 0; k < f2e_ptr[nf]; ++k) ++e2f_ptr[size_t(unfold(f2e[k])) + 1];
    for (int32_t e = 0; e < ne; ++e) e2f_ptr[size_t(e) + 1] += e2f_ptr[size_t(e)];
    e2f.resize(size_t(e2f_ptr[size_t(ne)]));
    {
        std::vector<int32_t> pos(e2f_ptr.begin(), e2f_ptr.end() - 1);
        for (int32_t f = 0; f < nf; ++f)
            for (int32_t k = f2e_ptr[f]; k < f2e_ptr[f + 1]; ++k) e2f[size_t(pos[size_t(unfold(f2e[k]))]++)] = f2e[k] >= 0 ? f : ~f;
    }
    for (int l = 1; l < levels; ++l)
    {
        const MgLevel &F = H.lv[size_t(l - 1)];
        int32_t n_agg = 0;
        const std::vector<int32_t> agg = match_nodes(F.n_nodes, F.tail, F.head, F.weight, passes, n_agg);
        MgLevel C;
        coarsen(F, agg, n_agg, l == 1 ? dir : nullptr, C);
        C.agg = agg;
        if (C.n == 0 || C.n >= F.n) break; // nothing left to coarsen
        if (l == 1)
        {
            // K_1 = P^T diag(m) d1^T diag(star2) d1 P: row I gathers its fine edges e (sign p_e), their faces f (sign s_fe) and
            // each face's edges e' mapped to coarse J (signs s_fe', p_e')
            const std::vector<int32_t> &map = C.map;
            assemble(C.n, [&](int32_t I, std::vector<std::pair<int32_t, double>> &row) {
                for (int32_t q = C.pt_ptr[size_t(I)]; q < C.pt_ptr[size_t(I) + 1]; ++q)
                {
                    const int32_t ev = C.pt[size_t(q)], e = unfold(ev);
                    const double pe = fsign(ev);
                    for (int32_t kf = e2f_ptr[size_t(e)]; kf < e2f_ptr[size_t(e) + 1]; ++kf)
                    {
                        const int32_t fv = e2f[size_t(kf)], f = unfold(fv);
                        const double a = pe * m_e[e] * fsign(fv) * star2[f];
                        for (int32_t k = f2e_ptr[f]; k < f2e_ptr[f + 1]; ++k)
                        {
                            const int32_t Jv = map[size_t(unfold(f2e[k]))];
                            if (Jv == kMgNone) continue;
                            row.emplace_back(unfold(Jv), a * fsign(f2e[k]) * fsign(Jv));
                        }
                    }
                }
            }, C.k_ptr, C.k_col, C.k_val);
        }
        else
        {
            // K_l = P^T K_{l-1} P
            assemble(C.n, [&](int32_t I, std::vector<std::pair<int32_t, double>> &row) {
                for (int32_t q = C.pt_ptr[size_t(I)]; q < C.pt_ptr[size_t(I) + 1]; ++q)
                {
                    const int32_t iv = C.pt[size_t(q)], i = unfold(iv);
                    const double pi = fsign(iv);
                    for (int32_t k = F.k_ptr[size_t(i)]; k < F.k_ptr[size_t(i) + 1]; ++k)
                    {
                        const int32_t Jv = C.map[size_t(F.k_col[size_t(k)])];
                        if (Jv == kMgNone) continue;
                        row.emplace_back(unfold(Jv), pi * fsign(Jv) * F.k_val[size_t(k)]);
                    }
                }
            }, C.k_ptr, C.k_col, C.k_val);
        }
        H.lv.push_back(std::move(C));
    }
    return H;
}
} // namespace mhdg
