#include "unsbin.h"

#include <climits>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <stdexcept>

namespace mhdg::unsbin
{
namespace
{
constexpr char kMagic[8] = {'U', 'N', 'S', 'B', 'I', 'N', 0, 1};
constexpr int kDims = 4;

size_t elem_bytes(uint8_t t)
{
    switch (t)
    {
    case kI8: return 1;
    case kI32: case kF32: return 4;
    case kI64: case kF64: return 8;
    default: throw std::runtime_error("UNSBIN: unknown data type " + std::to_string(int(t)));
    }
}

struct FileHandle
{
    std::FILE *f = nullptr;
    FileHandle(const std::string &path, const char *mode) : f(std::fopen(path.c_str(), mode))
    {
        if (!f) throw std::runtime_error("cannot open '" + path + "'");
    }
    ~FileHandle() { if (f) std::fclose(f); }
    void put(const void *p, size_t n)
    {
        if (n && std::fwrite(p, 1, n, f) != n) throw std::runtime_error("UNSBIN: write failed");
    }
    void take(void *p, size_t n)
    {
        if (n && std::fread(p, 1, n, f) != n) throw std::runtime_error("UNSBIN: truncated file");
    }
};

std::string join_lines(const std::vector<std::string> &v)
{
    std::string s;
    for (const auto &x : v) s += x + "\n";
    return s;
}
std::vector<std::string> split_lines(const std::vector<int8_t> &raw)
{
    std::istringstream is(std::string(raw.begin(), raw.end()));
    std::vector<std::string> out;
    for (std::string s; std::getline(is, s);) out.push_back(s);
    return out;
}
} // namespace

void Writer::add_raw(const std::string &name, uint8_t type, uint32_t cols, uint64_t rows, const void *data, size_t nbytes)
{
    if (rows * cols * elem_bytes(type) != nbytes) throw std::runtime_error("UNSBIN: record '" + name + "' size mismatch");
    Record r;
    r.type = type;
    r.cols = cols;
    r.rows = rows;
    r.bytes.resize(nbytes);
    if (nbytes) std::memcpy(r.bytes.data(), data, nbytes);
    recs_.emplace_back(name, std::move(r));
}

void Writer::write(const std::string &path) const
{
    FileHandle fh(path, "wb");
    fh.put(kMagic, 8);
    const uint64_t n = recs_.size();
    fh.put(&n, 8);
    const char zero[8] = {0};
    for (const auto &[name, r] : recs_)
    {
        const uint32_t len = static_cast<uint32_t>(name.size());
        const uint8_t tag[4] = {r.type, 0, 0, 0};
        fh.put(&len, 4);
        fh.put(name.data(), len);
        fh.put(tag, 4);
        fh.put(&r.cols, 4);
        fh.put(&r.rows, 8);
        fh.put(r.bytes.data(), r.bytes.size());
        fh.put(zero, (8 - r.bytes.size() % 8) % 8);
    }
}

Reader::Reader(const std::string &path) : path_(path)
{
    FileHandle fh(path, "rb");
    char magic[8];
    fh.take(magic, 8);
    if (std::memcmp(magic, kMagic, 8) != 0) throw std::runtime_error("'" + path + "' is not an UNSBIN file");
    uint64_t n = 0;
    fh.take(&n, 8);
    for (uint64_t i = 0; i < n; ++i)
    {
        uint32_t len = 0;
        fh.take(&len, 4);
        if (len > (1u << 20)) throw std::runtime_error("UNSBIN: corrupt record name in '" + path + "'");
        std::string name(len, '\0');
        fh.take(name.data(), len);
        uint8_t tag[4];
        fh.take(tag, 4);
        Record r;
        r.type = tag[0];
        fh.take(&r.cols, 4);
        fh.take(&r.rows, 8);
        r.bytes.resize(r.rows * r.cols * elem_bytes(r.type));
        fh.take(r.bytes.data(), r.bytes.size());
        char pad[8];
        fh.take(pad, (8 - r.bytes.size() % 8) % 8);
        recs_[name] = std::move(r);
    }
}

const Record &Reader::get(const std::string &name, uint8_t type) const
{
    auto it = recs_.find(name);
    if (it == recs_.end()) throw std::runtime_error("UNSBIN: '" + path_ + "' has no record '" + name + "'");
    if (it->second.type != type) throw std::runtime_error("UNSBIN: record '" + name + "' has another data type");
    return it->second;
}

std::vector<std::string> Reader::names() const
{
    std::vector<std::string> v;
    for (const auto &kv : recs_) v.push_back(kv.first);
    return v;
}

uint64_t topology_hash(const Complex &cx)
{
    // FNV-1a with the 64-bit FNV prime; the start value is the reference's (CellComplex.cpp:58), one digit shorter
    // than the published offset basis, and is kept so that hashes agree with files the reference writes
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](const void *p, size_t n) {
        const auto *b = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    };
    for (const Csr *t : {&cx.e2n, &cx.f2e, &cx.c2f})
    {
        mix(t->ptr.data(), t->ptr.size() * sizeof(int64_t));
        mix(t->idx.data(), t->idx.size() * sizeof(int32_t));
        mix(t->sgn.data(), t->sgn.size());
    }
    return h;
}

void write_complex(const std::string &path, const Complex &cx, const std::vector<double> &xyz, const std::vector<std::string> &tag_names)
{
    const int64_t counts[kDims] = {cx.nn, cx.ne, cx.nf, cx.nc};
    std::vector<int64_t> meta = {1, 0, 1};
    for (int d = 0; d < kDims; ++d) meta.push_back(counts[d]);
    for (int d = 0; d < kDims; ++d) meta.push_back(counts[d]);
    meta.push_back(static_cast<int64_t>(topology_hash(cx)));
    Writer w;
    w.add("meta", meta);
    const std::pair<const char *, const Csr *> inc[] = {{"e2n", &cx.e2n}, {"f2e", &cx.f2e}, {"c2f", &cx.c2f}};
    for (const auto &[n, t] : inc)
    {
        w.add(std::string(n) + ".ptr", t->ptr);
        w.add(std::string(n) + ".idx", t->idx);
        w.add(std::string(n) + ".sgn", t->sgn);
    }
    for (int d = 0; d < kDims; ++d)
    {
        std::vector<int64_t> gid(counts[d]);
        for (int64_t i = 0; i < counts[d]; ++i) gid[i] = i;
        const std::string p = "ent" + std::to_string(d);
        w.add(p + ".gid", gid);
        w.add(p + ".owner", std::vector<int32_t>(counts[d], 0));
        w.add(p + ".alias", std::vector<int8_t>(counts[d], 1));
    }
    w.add("face_tag", cx.face_tag);
    w.add("cell_material", cx.material);
    w.add("x", xyz, 3);
    const std::string names = join_lines(tag_names);
    w.add("tag_names", std::vector<int8_t>(names.begin(), names.end()));
    const std::vector<int32_t> no_peers;
    const std::vector<int64_t> ptr0{0};
    const std::vector<int32_t> no_idx;
    auto lists = [&](const std::string &p) {
        w.add(p + ".peers", no_peers);
        w.add(p + ".ptr", ptr0);
        w.add(p + ".idx", no_idx);
    };
    for (int d = 0; d < kDims; ++d)
    {
        lists("plan.send" + std::to_string(d));
        lists("plan.recv" + std::to_string(d));
    }
    lists("plan.reduce_send");
    lists("plan.reduce_recv");
    w.write(path);
}

void read_complex(const std::string &path, Complex &cx, std::vector<double> &xyz, std::vector<std::string> &tag_names)
{
    Reader r(path);
    const auto meta = r.vec<int64_t>("meta");
    if (meta.size() != 12 || meta[0] != 1) throw std::runtime_error("unsupported complex file version: " + path);
    if (meta[2] != 1)
        throw std::runtime_error("'" + path + "' is one part of a " + std::to_string(meta[2]) +
                                 "-rank partition; one GPU reads a complex written for one rank");
    cx = Complex();
    const std::pair<const char *, Csr *> inc[] = {{"e2n", &cx.e2n}, {"f2e", &cx.f2e}, {"c2f", &cx.c2f}};
    for (const auto &[n, t] : inc)
    {
        t->ptr = r.vec<int64_t>(std::string(n) + ".ptr");
        t->idx = r.vec<int32_t>(std::string(n) + ".idx");
        t->sgn = r.vec<int8_t>(std::string(n) + ".sgn");
        if (t->ptr.empty() || t->idx.size() != t->sgn.size() || t->ptr.back() != static_cast<int64_t>(t->idx.size()))
            throw std::runtime_error("corrupt incidence '" + std::string(n) + "' in " + path);
    }
    cx.nn = static_cast<int32_t>(meta[3]);
    cx.ne = static_cast<int32_t>(meta[4]);
    cx.nf = static_cast<int32_t>(meta[5]);
    cx.nc = static_cast<int32_t>(meta[6]);
    if (cx.e2n.rows() != cx.ne || cx.f2e.rows() != cx.nf || cx.c2f.rows() != cx.nc) throw std::runtime_error("corrupt complex file: " + path);
    cx.face_tag = r.vec<int32_t>("face_tag");
    cx.material = r.vec<int32_t>("cell_material");
    if (static_cast<int32_t>(cx.face_tag.size()) != cx.nf || static_cast<int32_t>(cx.material.size()) != cx.nc)
        throw std::runtime_error("corrupt complex file: " + path);
    if (static_cast<int64_t>(topology_hash(cx)) != meta[11]) throw std::runtime_error("topology hash mismatch: " + path);
    cx.build_up();
    uint32_t cols = 0;
    xyz = r.vec<double>("x", &cols);
    if (cols != 3 || static_cast<int64_t>(xyz.size()) != 3 * meta[3]) throw std::runtime_error("corrupt node coordinates in " + path);
    tag_names = split_lines(r.vec<int8_t>("tag_names"));
}

namespace
{
struct Part
{
    int64_t rank = 0, nranks = 0, counts[kDims]{}, owned[kDims]{};
    Complex cx;                       // the part's local complex (no upward tables)
    std::vector<int64_t> gid[kDims];  // global id of every local entity
    std::vector<double> x;
    std::vector<std::string> names;
};

Part read_part(const std::string &path)
{
    Reader r(path);
    const auto meta = r.vec<int64_t>("meta");
    if (meta.size() != 12 || meta[0] != 1) throw std::runtime_error("unsupported complex file version: " + path);
    Part p;
    p.rank = meta[1];
    p.nranks = meta[2];
    for (int d = 0; d < kDims; ++d) { p.counts[d] = meta[3 + d]; p.owned[d] = meta[7 + d]; }
    const std::pair<const char *, Csr *> inc[] = {{"e2n", &p.cx.e2n}, {"f2e", &p.cx.f2e}, {"c2f", &p.cx.c2f}};
    for (const auto &[n, t] : inc)
    {
        t->ptr = r.vec<int64_t>(std::string(n) + ".ptr");
        t->idx = r.vec<int32_t>(std::string(n) + ".idx");
        t->sgn = r.vec<int8_t>(std::string(n) + ".sgn");
        if (t->ptr.empty() || t->idx.size() != t->sgn.size() || t->ptr.back() != static_cast<int64_t>(t->idx.size()))
            throw std::runtime_error("corrupt incidence '" + std::string(n) + "' in " + path);
    }
    p.cx.nn = static_cast<int32_t>(p.counts[0]);
    p.cx.ne = static_cast<int32_t>(p.counts[1]);
    p.cx.nf = static_cast<int32_t>(p.counts[2]);
    p.cx.nc = static_cast<int32_t>(p.counts[3]);
    if (p.cx.e2n.rows() != p.cx.ne || p.cx.f2e.rows() != p.cx.nf || p.cx.c2f.rows() != p.cx.nc) throw std::runtime_error("corrupt complex file: " + path);
    if (static_cast<int64_t>(topology_hash(p.cx)) != meta[11]) throw std::runtime_error("topology hash mismatch: " + path);
    for (int d = 0; d < kDims; ++d)
    {
        const std::string q = "ent" + std::to_string(d);
        p.gid[d] = r.vec<int64_t>(q + ".gid");
        const auto alias = r.vec<int8_t>(q + ".alias");
        if (static_cast<int64_t>(p.gid[d].size()) != p.counts[d] || alias.size() != p.gid[d].size() || p.owned[d] < 0 || p.owned[d] > p.counts[d])
            throw std::runtime_error("corrupt entity records '" + q + "' in " + path);
        for (int8_t a : alias)
            if (a != 1) throw std::runtime_error("'" + path + "': an entity copy oriented against its owner's is not supported");
    }
    p.cx.face_tag = r.vec<int32_t>("face_tag");
    p.cx.material = r.vec<int32_t>("cell_material");
    if (static_cast<int64_t>(p.cx.face_tag.size()) != p.counts[2] || static_cast<int64_t>(p.cx.material.size()) != p.counts[3])
        throw std::runtime_error("corrupt complex file: " + path);
    uint32_t cols = 0;
    p.x = r.vec<double>("x", &cols);
    if (cols != 3 || static_cast<int64_t>(p.x.size()) != 3 * p.counts[0]) throw std::runtime_error("corrupt node coordinates in " + path);
    p.names = split_lines(r.vec<int8_t>("tag_names"));
    return p;
}
} // namespace

void read_partition(const std::vector<std::string> &paths, Complex &cx, std::vector<double> &xyz, std::vector<std::string> &tag_names)
{
    const int64_t n = static_cast<int64_t>(paths.size());
    if (n < 1) throw std::runtime_error("no part files given");
    std::vector<Part> parts(static_cast<size_t>(n));
    std::vector<int> seen(static_cast<size_t>(n), 0);
    for (int64_t k = 0; k < n; ++k)
    {
        parts[size_t(k)] = read_part(paths[size_t(k)]);
        const Part &p = parts[size_t(k)];
        if (p.nranks != n) throw std::runtime_error("'" + paths[size_t(k)] + "' is one part of a " + std::to_string(p.nranks) +
                                                    "-rank partition, " + std::to_string(n) + " files given");
        if (p.rank < 0 || p.rank >= n || seen[size_t(p.rank)]++) throw std::runtime_error("part ranks are not 0 .. n-1 once each");
        if (p.names != parts[0].names) throw std::runtime_error("the parts' tag names differ");
    }
    // owner (part, local index) of every global id, from the owned ranges
    int64_t G[kDims] = {0, 0, 0, 0};
    for (const Part &p : parts)
        for (int d = 0; d < kDims; ++d) G[d] += p.owned[d];
    for (int d = 0; d < kDims; ++d)
        if (G[d] > INT32_MAX) throw std::runtime_error("partition too large for 32-bit ids");
    std::vector<std::pair<int32_t, int32_t>> own[kDims];
    for (int d = 0; d < kDims; ++d) own[d].assign(size_t(G[d]), {-1, -1});
    for (int32_t k = 0; k < n; ++k)
        for (int d = 0; d < kDims; ++d)
            for (int64_t i = 0; i < parts[size_t(k)].owned[d]; ++i)
            {
                const int64_t g = parts[size_t(k)].gid[d][size_t(i)];
                if (g < 0 || g >= G[d] || own[d][size_t(g)].first >= 0)
                    throw std::runtime_error("global ids of dimension " + std::to_string(d) + " are not owned once each");
                own[d][size_t(g)] = {k, static_cast<int32_t>(i)};
            }
    // global rows: the owner's row with its local column ids mapped through the owner's gids (signs as stored)
    auto merge = [&](Csr Complex::*tab, int row_dim, int col_dim) {
        Csr out;
        for (int64_t g = 0; g < G[row_dim]; ++g)
        {
            const auto [k, i] = own[row_dim][size_t(g)];
            const Part &p = parts[size_t(k)];
            const Csr &t = p.cx.*tab;
            for (int64_t q = t.ptr[size_t(i)]; q < t.ptr[size_t(i) + 1]; ++q)
            {
                const int32_t l = t.idx[size_t(q)];
                if (l < 0 || l >= p.counts[col_dim]) throw std::runtime_error("corrupt incidence in a part file");
                out.push(static_cast<int32_t>(p.gid[col_dim][size_t(l)]), t.sgn[size_t(q)]);
            }
            out.close_row();
        }
        return out;
    };
    cx = Complex();
    cx.nn = static_cast<int32_t>(G[0]);
    cx.ne = static_cast<int32_t>(G[1]);
    cx.nf = static_cast<int32_t>(G[2]);
    cx.nc = static_cast<int32_t>(G[3]);
    cx.e2n = merge(&Complex::e2n, 1, 0);
    cx.f2e = merge(&Complex::f2e, 2, 1);
    cx.c2f = merge(&Complex::c2f, 3, 2);
    xyz.assign(size_t(3 * G[0]), 0.0);
    for (int64_t g = 0; g < G[0]; ++g)
    {
        const auto [k, i] = own[0][size_t(g)];
        for (int c = 0; c < 3; ++c) xyz[size_t(3 * g + c)] = parts[size_t(k)].x[size_t(3 * i + c)];
    }
    cx.face_tag.assign(size_t(G[2]), 0);
    for (int64_t g = 0; g < G[2]; ++g) cx.face_tag[size_t(g)] = parts[size_t(own[2][size_t(g)].first)].cx.face_tag[size_t(own[2][size_t(g)].second)];
    cx.material.assign(size_t(G[3]), 0);
    for (int64_t g = 0; g < G[3]; ++g) cx.material[size_t(g)] = parts[size_t(own[3][size_t(g)].first)].cx.material[size_t(own[3][size_t(g)].second)];
    // every copy (ghosts included) must agree with the merged complex
    const std::pair<Csr Complex::*, int> tabs[] = {{&Complex::e2n, 1}, {&Complex::f2e, 2}, {&Complex::c2f, 3}};
    for (const Part &p : parts)
    {
        for (const auto &[tab, d] : tabs)
        {
            const Csr &t = p.cx.*tab, &gt = cx.*tab;
            for (int64_t i = 0; i < p.counts[d]; ++i)
            {
                const int64_t g = p.gid[d][size_t(i)];
                if (g < 0 || g >= G[d]) throw std::runtime_error("a part's ghost has no owner");
                const int64_t a0 = t.ptr[size_t(i)], a1 = t.ptr[size_t(i) + 1], b0 = gt.ptr[size_t(g)];
                bool same = a1 - a0 == gt.ptr[size_t(g) + 1] - b0;
                for (int64_t q = 0; same && q < a1 - a0; ++q)
                {
                    const int32_t l = t.idx[size_t(a0 + q)];
                    if (l < 0 || l >= p.counts[d - 1]) throw std::runtime_error("corrupt incidence in a part file");
                    same = p.gid[d - 1][size_t(l)] == gt.idx[size_t(b0 + q)] && t.sgn[size_t(a0 + q)] == gt.sgn[size_t(b0 + q)];
                }
                if (!same) throw std::runtime_error("the parts disagree on the rows of dimension " + std::to_string(d));
            }
        }
        for (int64_t i = 0; i < p.counts[0]; ++i)
            for (int c = 0; c < 3; ++c)
                if (p.x[size_t(3 * i + c)] != xyz[size_t(3 * p.gid[0][size_t(i)] + c)]) throw std::runtime_error("the parts disagree on node coordinates");
        for (int64_t i = 0; i < p.counts[2]; ++i)
            if (p.cx.face_tag[size_t(i)] != cx.face_tag[size_t(p.gid[2][size_t(i)])]) throw std::runtime_error("the parts disagree on face tags");
        for (int64_t i = 0; i < p.counts[3]; ++i)
            if (p.cx.material[size_t(i)] != cx.material[size_t(p.gid[3][size_t(i)])]) throw std::runtime_error("the parts disagree on cell materials");
    }
    cx.build_up();
    tag_names = parts[0].names;
}

void write_fields(const std::string &path, uint64_t hash, const std::vector<Field> &fields)
{
    Writer w;
    w.add("hash", std::vector<int64_t>{static_cast<int64_t>(hash)});
    std::vector<std::string> index;
    for (const Field &f : fields)
    {
        if (f.ncomp < 1 || f.values.size() % f.ncomp) throw std::runtime_error("field '" + f.name + "': size not a multiple of ncomp");
        index.push_back(f.name);
        w.add("f." + f.name + ".info", std::vector<int32_t>{f.dim, f.ncomp});
        w.add("f." + f.name, f.values, static_cast<uint32_t>(f.ncomp));
    }
    const std::string idx = join_lines(index);
    w.add("index", std::vector<int8_t>(idx.begin(), idx.end()));
    w.write(path);
}

std::vector<Field> read_fields(const std::string &path, uint64_t *hash)
{
    Reader r(path);
    const auto h = r.vec<int64_t>("hash");
    if (h.size() != 1) throw std::runtime_error("corrupt field file: " + path);
    if (hash) *hash = static_cast<uint64_t>(h[0]);
    std::vector<Field> out;
    for (const std::string &name : split_lines(r.vec<int8_t>("index")))
    {
        Field f;
        f.name = name;
        const auto info = r.vec<int32_t>("f." + name + ".info");
        if (info.size() != 2) throw std::runtime_error("corrupt field info '" + name + "' in " + path);
        f.dim = info[0];
        f.ncomp = info[1];
        f.values = r.vec<double>("f." + name);
        out.push_back(std::move(f));
    }
    return out;
}
} // namespace mhdg::unsbin
