#pragma once
// UNSBIN containers (the reference's mesh, sub-complex and field files, include/uns/io/BinContainer.h:1-6,
// docs/FileFormats.md:3-9): little-endian
//     magic[8] = "UNSBIN\0\1" | u64 n_records | records
//     record   = u32 name_len | name | u8 dtype | u8[3] 0 | u32 cols | u64 rows | payload, zero-padded to 8 bytes
// with dtype 0 int8, 1 int32, 2 int64, 3 float32, 4 float64.  Complex files (src/uns/io/MeshIO.cpp:38-74) carry
// meta = {version 1, rank, nranks, counts[4], n_owned[4], topology hash}, the e2n/f2e/c2f incidences, the
// per-dimension gid/owner/alias arrays, face_tag, cell_material, x (3 columns), the newline-joined tag names
// and the communication plans; field files (MeshIO.cpp:117-130) carry hash, one "f.<name>.info" = {dim, ncomp}
// and one "f.<name>" (ncomp columns) per field, and the newline-joined index of the names.
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "complex.h"

namespace mhdg::unsbin
{
enum DType : uint8_t { kI8 = 0, kI32 = 1, kI64 = 2, kF32 = 3, kF64 = 4 };

struct Record
{
    uint8_t type = kF64;
    uint32_t cols = 1;
    uint64_t rows = 0;
    std::vector<char> bytes;
};

template <class T> constexpr uint8_t dtype_of();
template <> constexpr uint8_t dtype_of<int8_t>() { return kI8; }
template <> constexpr uint8_t dtype_of<int32_t>() { return kI32; }
template <> constexpr uint8_t dtype_of<int64_t>() { return kI64; }
template <> constexpr uint8_t dtype_of<float>() { return kF32; }
template <> constexpr uint8_t dtype_of<double>() { return kF64; }

class Writer
{
public:
    template <class T> void add(const std::string &name, const std::vector<T> &v, uint32_t cols = 1)
    {
        add_raw(name, dtype_of<T>(), cols, cols ? v.size() / cols : 0, v.data(), v.size() * sizeof(T));
    }
    void add_raw(const std::string &name, uint8_t type, uint32_t cols, uint64_t rows, const void *data, size_t nbytes);
    void write(const std::string &path) const;

private:
    std::vector<std::pair<std::string, Record>> recs_; // file order = insertion order
};

class Reader
{
public:
    explicit Reader(const std::string &path);
    bool has(const std::string &name) const { return recs_.count(name) != 0; }
    const Record &get(const std::string &name, uint8_t type) const;
    template <class T> std::vector<T> vec(const std::string &name, uint32_t *cols = nullptr) const
    {
        const Record &r = get(name, dtype_of<T>());
        if (cols) *cols = r.cols;
        std::vector<T> v(r.bytes.size() / sizeof(T));
        if (!v.empty()) std::memcpy(v.data(), r.bytes.data(), r.bytes.size());
        return v;
    }
    std::vector<std::string> names() const;

private:
    std::map<std::string, Record> recs_;
    std::string path_;
};

// FNV-1a style 64-bit hash over the bytes of e2n, f2e, c2f (ptr, idx, sgn each) - the reference's topology hash
// (src/uns/complex/CellComplex.cpp:56-70), which ties field files to the complex they belong to.
uint64_t topology_hash(const Complex &cx);

// Complex file of one rank holding the whole mesh (every entity owned, empty plans).
void write_complex(const std::string &path, const Complex &cx, const std::vector<double> &xyz, const std::vector<std::string> &tag_names);
// Reads a complex file written for one rank (nranks = 1); verifies the stored topology hash.
void read_complex(const std::string &path, Complex &cx, std::vector<double> &xyz, std::vector<std::string> &tag_names);
// The n part files of an n-rank partition (uns::Distributor; src/uns/parallel/Distributor.cpp:115-221) merged into
// the global complex they were cut from, in its global numbering: each entity's row from its owner's file, column
// ids through that file's global ids; each file's topology hash and every ghost copy are checked.
void read_partition(const std::vector<std::string> &paths, Complex &cx, std::vector<double> &xyz, std::vector<std::string> &tag_names);

struct Field
{
    std::string name;
    int32_t dim = 3;   // 0 node, 1 edge, 2 face, 3 cell
    int32_t ncomp = 1;
    std::vector<double> values;
};
void write_fields(const std::string &path, uint64_t hash, const std::vector<Field> &fields);
std::vector<Field> read_fields(const std::string &path, uint64_t *hash);
} // namespace mhdg::unsbin
