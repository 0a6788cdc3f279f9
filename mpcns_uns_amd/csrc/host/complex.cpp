#include "complex.h"

#include <algorithm>
#include <parallel/algorithm>
#include <stdexcept>
#include <unordered_map>
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include <omp.h>

namespace mhdg
{
Csr transpose(const Csr &a, int32_t ncols)
{
    Csr t;
    t.ptr.assign(static_cast<size_t>(ncols) + 1, 0);
    for (int32_t c : a.idx) ++t.ptr[static_cast<size_t>(c) + 1];
    for (int32_t c = 0; c < ncols; ++c) t.ptr[c + 1] += t.ptr[c];
    t.idx.resize(a.idx.size());
    t.sgn.resize(a.sgn.size());
    std::vector<int64_t> fill(t.ptr.begin(), t.ptr.end() - 1);
    for (int32_t r = 0; r < a.rows(); ++r)
        for (int64_t k = a.ptr[r]; k < a.ptr[r + 1]; ++k)
        {
            const int64_t q = fill[a.idx[k]]++;
            t.idx[q] = r;
            t.sgn[q] = a.sgn[k];
        }
    return t;
}

void Complex::build_up()
{
    n2e = transpose(e2n, nn);
    e2f = transpose(f2e, ne);
    f2c = transpose(c2f, nf);
}

void Complex::edge_nodes(int32_t e, int32_t &a, int32_t &b) const
{
    a = b = -1;
    for (int64_t k = e2n.ptr[e]; k < e2n.ptr[e + 1]; ++k) (e2n.sgn[k] < 0 ? a : b) = e2n.idx[k];
}

void Complex::face_loop(int32_t f, int32_t *nodes, int &n) const
{
    n = 0;
    for (int64_t k = f2e.ptr[f]; k < f2e.ptr[f + 1]; ++k)
    {
        int32_t a, b;
        edge_nodes(f2e.idx[k], a, b);
        nodes[n++] = f2e.sgn[k] > 0 ? a : b;
    }
}

namespace
{
// rows of upper*lower that are not identically zero (integer arithmetic)
long nonzero_products(const Csr &upper, const Csr &lower, int32_t ncols)
{
    long bad = 0;
    std::vector<long> acc(static_cast<size_t>(ncols), 0);
    std::vector<int32_t> touched;
    for (int32_t r = 0; r < upper.rows(); ++r)
    {
        touched.clear();
        for (int64_t k = upper.ptr[r]; k < upper.ptr[r + 1]; ++k)
        {
            const int32_t m = upper.idx[k];
            for (int64_t q = lower.ptr[m]; q < lower.ptr[m + 1]; ++q)
            {
                acc[lower.idx[q]] += static_cast<long>(upper.sgn[k]) * lower.sgn[q];
                touched.push_back(lower.idx[q]);
            }
        }
        bool nz = false;
        for (int32_t c : touched) { nz |= acc[c] != 0; acc[c] = 0; }
        bad += nz;
    }
    return bad;
}
} // namespace

TopoReport validate(const Complex &cx)
{
    TopoReport r;
    r.euler = static_cast<long>(cx.nn) - cx.ne + cx.nf - cx.nc;
    for (int32_t e = 0; e < cx.ne; ++e)
    {
        int s = 0, t = 0;
        int32_t a = -1, b = -1;
        for (int64_t k = cx.e2n.ptr[e]; k < cx.e2n.ptr[e + 1]; ++k)
        {
            if (cx.e2n.sgn[k] < 0) { ++s; a = cx.e2n.idx[k]; }
            else { ++t; b = cx.e2n.idx[k]; }
        }
        if (s != 1 || t != 1 || a == b || cx.e2n.ptr[e + 1] - cx.e2n.ptr[e] != 2) ++r.bad_edges;
    }
    for (int32_t f = 0; f < cx.nf; ++f)
    {
        const int m = static_cast<int>(cx.f2e.ptr[f + 1] - cx.f2e.ptr[f]);
        bool ok = m >= 3;
        for (int i = 0; ok && i < m; ++i)
        {
            const int64_t k0 = cx.f2e.ptr[f] + i, k1 = cx.f2e.ptr[f] + (i + 1) % m;
            int32_t a0, b0, a1, b1;
            cx.edge_nodes(cx.f2e.idx[k0], a0, b0);
            cx.edge_nodes(cx.f2e.idx[k1], a1, b1);
            const int32_t leave = cx.f2e.sgn[k0] > 0 ? b0 : a0;
            const int32_t enter = cx.f2e.sgn[k1] > 0 ? a1 : b1;
            if (leave != enter) ok = false;
        }
        if (!ok) ++r.bad_loops;
        const int64_t c = cx.f2c.ptr[f + 1] - cx.f2c.ptr[f];
        if (c == 1) ++r.boundary_faces;
        else if (c == 2) { if (cx.f2c.sgn[cx.f2c.ptr[f]] + cx.f2c.sgn[cx.f2c.ptr[f] + 1] != 0) ++r.bad_face_cells; }
        else ++r.bad_face_cells;
    }
    r.d1d0_bad = nonzero_products(cx.f2e, cx.e2n, cx.nn);
    r.d2d1_bad = nonzero_products(cx.c2f, cx.f2e, cx.ne);
    return r;
}

// ------------------------------------------------------------------------------------------ builder
namespace
{
struct Key2
{
    uint64_t a, b;
    bool operator==(const Key2 &o) const { return a == o.a && b == o.b; }
};
inline uint64_t mix64(uint64_t x)
{
    x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
    return x;
}

// Open-addressing map Key2 -> int32 (value -1 = empty slot).
class KeyMap
{
public:
    explicit KeyMap(size_t expect) { rehash(std::max<size_t>(1024, expect * 2)); }
    // returns the stored value, inserting `v` if absent
    int32_t find_or_insert(const Key2 &k, int32_t v, bool &inserted)
    {
        if (2 * (size_ + 1) > cap_) rehash(cap_ * 2);
        size_t h = mix64(k.a * 0x9E3779B97F4A7C15ull ^ k.b) & (cap_ - 1);
        while (val_[h] >= 0)
        {
            if (key_[h] == k) { inserted = false; return val_[h]; }
            h = (h + 1) & (cap_ - 1);
        }
        key_[h] = k; val_[h] = v; ++size_;
        inserted = true;
        return v;
    }
    int32_t find(const Key2 &k) const
    {
        size_t h = mix64(k.a * 0x9E3779B97F4A7C15ull ^ k.b) & (cap_ - 1);
        while (val_[h] >= 0)
        {
            if (key_[h] == k) return val_[h];
            h = (h + 1) & (cap_ - 1);
        }
        return -1;
    }

private:
    void rehash(size_t ncap)
    {
        size_t c = 1;
        while (c < ncap) c <<= 1;
        std::vector<Key2> ok(std::move(key_));
        std::vector<int32_t> ov(std::move(val_));
        key_.assign(c, Key2{0, 0});
        val_.assign(c, -1);
        cap_ = c;
        size_ = 0;
        for (size_t i = 0; i < ov.size(); ++i)
            if (ov[i] >= 0) { bool ins; find_or_insert(ok[i], ov[i], ins); }
    }
    std::vector<Key2> key_;
    std::vector<int32_t> val_;
    size_t cap_ = 0, size_ = 0;
};

Key2 face_key(const int32_t *nodes, int n)
{
    uint32_t s[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    for (int i = 0; i < n; ++i) s[i] = static_cast<uint32_t>(nodes[i]);
    std::sort(s, s + n);
    return {(static_cast<uint64_t>(s[0]) << 32) | s[1], (static_cast<uint64_t>(s[2]) << 32) | s[3]};
}
} // namespace

struct Builder::Impl
{
    int32_t nn;
    KeyMap edges, faces;
    std::vector<int32_t> e_lo, e_hi;
    Csr f2e, c2f;
    std::vector<int32_t> material;
    std::vector<std::pair<Key2, int>> tags;
    Impl(int32_t n, int64_t expect) : nn(n), edges(static_cast<size_t>(expect) * 2), faces(static_cast<size_t>(expect) * 3) {}

    int32_t edge_id(int32_t a, int32_t b)
    {
        const int32_t lo = std::min(a, b), hi = std::max(a, b);
        bool ins;
        const int32_t id = edges.find_or_insert(Key2{(static_cast<uint64_t>(lo) << 32) | static_cast<uint32_t>(hi), 0},
                                                static_cast<int32_t>(e_lo.size()), ins);
        if (ins) { e_lo.push_back(lo); e_hi.push_back(hi); }
        return id;
    }
};

Builder::Builder(int32_t n_nodes, int64_t expect_cells) : p_(new Impl(n_nodes, expect_cells)) {}
Builder::~Builder() { delete p_; }

int32_t Builder::add_cell(const int32_t *loop_ptr, const int32_t *loop_nodes, int n_loops, int material)
{
    if (n_loops < 4) throw std::runtime_error("cell with fewer than 4 faces");
    for (int l = 0; l < n_loops; ++l)
    {
        const int m = loop_ptr[l + 1] - loop_ptr[l];
        if (m < 3 || m > 4) throw std::runtime_error("face loops must have 3 or 4 nodes");
        int32_t loop[4];
        for (int i = 0; i < m; ++i) loop[i] = loop_nodes[loop_ptr[l] + i];
        // canonical loop: start at the smallest node, second node smaller than the last
        const int s = static_cast<int>(std::min_element(loop, loop + m) - loop);
        std::rotate(loop, loop + s, loop + m);
        const bool flipped = loop[1] > loop[m - 1];
        if (flipped) std::reverse(loop + 1, loop + m);
        bool ins;
        const int32_t fid = p_->faces.find_or_insert(face_key(loop, m), p_->f2e.rows(), ins);
        if (ins)
        {
            for (int i = 0; i < m; ++i)
            {
                const int32_t a = loop[i], b = loop[(i + 1) % m];
                p_->f2e.push(p_->edge_id(a, b), a < b ? +1 : -1);
            }
            p_->f2e.close_row();
        }
        p_->c2f.push(fid, flipped ? -1 : +1);
    }
    p_->c2f.close_row();
    p_->material.push_back(material);
    return static_cast<int32_t>(p_->material.size()) - 1;
}

void Builder::tag_face(const int32_t *nodes, int n, int tag) { p_->tags.emplace_back(face_key(nodes, n), tag); }

Complex Builder::finish()
{
    Complex cx;
    cx.nn = p_->nn;
    cx.ne = static_cast<int32_t>(p_->e_lo.size());
    cx.e2n.ptr.reserve(cx.ne + 1);
    for (int32_t e = 0; e < cx.ne; ++e)
    {
        cx.e2n.push(p_->e_lo[e], -1);
        cx.e2n.push(p_->e_hi[e], +1);
        cx.e2n.close_row();
    }
    cx.f2e = std::move(p_->f2e);
    cx.c2f = std::move(p_->c2f);
    cx.nf = cx.f2e.rows();
    cx.nc = cx.c2f.rows();
    cx.material = std::move(p_->material);
    cx.face_tag.assign(cx.nf, 0);
    for (const auto &t : p_->tags)
    {
        const int32_t f = p_->faces.find(t.first);
        if (f >= 0) cx.face_tag[f] = t.second; // later tags win, as with a map assignment
    }
    cx.build_up();
    return cx;
}
// ------------------------------------------------------------------------------- sort-based builder
// Same numbering as Builder (faces by first appearance in cell/loop order, edges by first appearance in
// face/loop-position order of newly created faces), computed with parallel sorts and scans instead of a
// hash map: the occurrence of each face (edge) is keyed by its sorted node set, sorted by (key,
// occurrence), and the first occurrence of every key, ranked in occurrence order, is its number.
namespace
{
struct KeyOcc
{
    uint64_t key;
    int64_t occ;
    bool operator<(const KeyOcc &o) const { return key < o.key || (key == o.key && occ < o.occ); }
};

// ids[i] = number of the key group of entry i, groups numbered by their first occurrence.
// Returns the number of groups; first[g] = first occurrence of group g.
int64_t number_by_first(std::vector<KeyOcc> &v, int64_t n_occ, std::vector<int32_t> &id_of_occ, std::vector<int64_t> &first)
{
    __gnu_parallel::sort(v.begin(), v.end());
    const int64_t n = static_cast<int64_t>(v.size());
    std::vector<uint8_t> is_first(n_occ, 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        if (i == 0 || v[i].key != v[i - 1].key) is_first[v[i].occ] = 1;
    // exclusive scan of is_first over occurrences
    std::vector<int32_t> rank(n_occ);
    int64_t total = 0;
    {
        const int nt = std::max(1, omp_get_max_threads());
        std::vector<int64_t> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            const int64_t lo = n_occ * t / nt, hi = n_occ * (t + 1) / nt;
            int64_t s = 0;
            for (int64_t i = lo; i < hi; ++i) s += is_first[i];
            part[t + 1] = s;
#pragma omp barrier
#pragma omp single
            for (int k = 0; k < nt; ++k) part[k + 1] += part[k];
            int64_t acc = part[t];
            for (int64_t i = lo; i < hi; ++i) { rank[i] = static_cast<int32_t>(acc); acc += is_first[i]; }
        }
        total = part[nt];
    }
    first.assign(total, 0);
    id_of_occ.assign(n_occ, -1);
    // group start positions
    std::vector<int64_t> starts;
    starts.reserve(total + 1);
    for (int64_t i = 0; i < n; ++i)
        if (i == 0 || v[i].key != v[i - 1].key) starts.push_back(i);
    starts.push_back(n);
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < static_cast<int64_t>(starts.size()) - 1; ++g)
    {
        const int64_t f0 = v[starts[g]].occ;
        const int32_t id = rank[f0];
        first[id] = f0;
        for (int64_t i = starts[g]; i < starts[g + 1]; ++i) id_of_occ[v[i].occ] = id;
    }
    return total;
}
} // namespace

Complex build_triangle_complex(int32_t n_nodes, const int32_t *tri_loops, int64_t n_cells, int faces_per_cell, const int32_t *tags,
                               int64_t n_tags)
{
    if (n_nodes >= (1 << 21)) throw std::runtime_error("sort-based builder: node ids must fit in 21 bits");
    const bool tm = std::getenv("MHDG_SETUP_TIMING") != nullptr;
    auto t0 = std::chrono::steady_clock::now();
    auto lap = [&](const char *what) {
        if (!tm) return;
        const auto t = std::chrono::steady_clock::now();
        std::fprintf(stderr, "  complex %-12s %.2fs\n", what, std::chrono::duration<double>(t - t0).count());
        t0 = t;
    };
    const int64_t L = n_cells * faces_per_cell;
    auto pack3 = [](uint32_t a, uint32_t b, uint32_t c) {
        uint32_t s[3] = {a, b, c};
        std::sort(s, s + 3);
        return (uint64_t(s[0]) << 42) | (uint64_t(s[1]) << 21) | uint64_t(s[2]);
    };
    // faces
    std::vector<KeyOcc> occ(L);
    std::vector<uint8_t> flipped(L);
    std::vector<int32_t> canon(3 * L);
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < L; ++l)
    {
        int32_t lp[3] = {tri_loops[3 * l], tri_loops[3 * l + 1], tri_loops[3 * l + 2]};
        const int sft = static_cast<int>(std::min_element(lp, lp + 3) - lp);
        std::rotate(lp, lp + sft, lp + 3);
        const bool fl = lp[1] > lp[2];
        if (fl) std::swap(lp[1], lp[2]);
        flipped[l] = fl;
        canon[3 * l] = lp[0]; canon[3 * l + 1] = lp[1]; canon[3 * l + 2] = lp[2];
        occ[l] = {pack3(lp[0], lp[1], lp[2]), l};
    }
    lap("face keys");
    std::vector<int32_t> face_of;
    std::vector<int64_t> face_first;
    const int64_t nf = number_by_first(occ, L, face_of, face_first);
    lap("face numbers");
    Complex cx;
    cx.nn = n_nodes;
    cx.nc = static_cast<int32_t>(n_cells);
    cx.nf = static_cast<int32_t>(nf);
    cx.c2f.ptr.resize(n_cells + 1);
    cx.c2f.idx.resize(L);
    cx.c2f.sgn.resize(L);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c <= n_cells; ++c) cx.c2f.ptr[c] = c * faces_per_cell;
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < L; ++l) { cx.c2f.idx[l] = face_of[l]; cx.c2f.sgn[l] = flipped[l] ? -1 : 1; }
    // edges: occurrences (face, loop position) of the canonical loops of the faces
    std::vector<KeyOcc> eocc(3 * nf);
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nf; ++f)
    {
        const int32_t *lp = &canon[3 * face_first[f]];
        for (int i = 0; i < 3; ++i)
        {
            const uint32_t a = lp[i], b = lp[(i + 1) % 3];
            const uint32_t lo = std::min(a, b), hi = std::max(a, b);
            eocc[3 * f + i] = {(uint64_t(lo) << 21) | hi, 3 * f + i};
        }
    }
    std::vector<int32_t> edge_of;
    std::vector<int64_t> edge_first;
    lap("edge keys");
    const int64_t ne = number_by_first(eocc, 3 * nf, edge_of, edge_first);
    lap("edge numbers");
    cx.ne = static_cast<int32_t>(ne);
    cx.f2e.ptr.resize(nf + 1);
    cx.f2e.idx.resize(3 * nf);
    cx.f2e.sgn.resize(3 * nf);
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f <= nf; ++f) cx.f2e.ptr[f] = 3 * f;
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nf; ++f)
    {
        const int32_t *lp = &canon[3 * face_first[f]];
        for (int i = 0; i < 3; ++i)
        {
            cx.f2e.idx[3 * f + i] = edge_of[3 * f + i];
            cx.f2e.sgn[3 * f + i] = lp[i] < lp[(i + 1) % 3] ? 1 : -1;
        }
    }
    cx.e2n.ptr.resize(ne + 1);
    cx.e2n.idx.resize(2 * ne);
    cx.e2n.sgn.resize(2 * ne);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e)
    {
        const int64_t o = edge_first[e];
        const int32_t *lp = &canon[3 * face_first[o / 3]];
        const int i = static_cast<int>(o % 3);
        const int32_t a = lp[i], b = lp[(i + 1) % 3];
        cx.e2n.ptr[e] = 2 * e;
        cx.e2n.idx[2 * e] = std::min(a, b); cx.e2n.sgn[2 * e] = -1;
        cx.e2n.idx[2 * e + 1] = std::max(a, b); cx.e2n.sgn[2 * e + 1] = 1;
    }
    cx.e2n.ptr[ne] = 2 * ne;
    cx.material.assign(n_cells, 0);
    cx.face_tag.assign(nf, 0);
    // tags: look the sorted key up among the first occurrences (occ is sorted by key)
    for (int64_t t = 0; t < n_tags; ++t)
    {
        const uint64_t k = pack3(tags[4 * t], tags[4 * t + 1], tags[4 * t + 2]);
        const auto it = std::lower_bound(occ.begin(), occ.end(), KeyOcc{k, -1});
        if (it != occ.end() && it->key == k) cx.face_tag[face_of[it->occ]] = tags[4 * t + 3];
    }
    lap("tables");
    cx.build_up();
    lap("transposes");
    return cx;
}
} // namespace mhdg
