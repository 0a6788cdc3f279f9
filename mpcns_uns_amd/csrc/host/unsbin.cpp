#include "unsbin.h"

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
