#include "complex.h"

#include <algorithm>
#include <stdexcept>
#include <unordered_map>

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
} // namespace mhdg
