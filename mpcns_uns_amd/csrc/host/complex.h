#pragma once
// Signed cell complex cell -> face -> edge -> node, host side.
//
// Orientation and numbering follow the reference's rules (src/uns/complex/ComplexBuilder.cpp:20-134,
// include/uns/complex/CellComplex.h:4-13):
//   * an edge runs from its smaller node id to its larger one (e2n sign -1 = start, +1 = end);
//   * a face loop is stored rotated to start at its smallest node, and reversed if its second node is
//     larger than its last; f2e sign is +1 where the loop runs along the edge;
//   * c2f sign is +1 when the cell's outward loop was stored unreversed;
//   * faces and edges are numbered by first appearance while cells are added in order.
// Transposed tables (n2e, e2f, f2c) have rows sorted by ascending column (Incidence.cpp:8-26).
#include <cstdint>
#include <string>
#include <vector>

namespace mhdg
{
struct Csr
{
    std::vector<int64_t> ptr{0};
    std::vector<int32_t> idx;
    std::vector<int8_t> sgn;
    int32_t rows() const { return static_cast<int32_t>(ptr.size()) - 1; }
    int64_t nnz() const { return static_cast<int64_t>(idx.size()); }
    void push(int32_t col, int s) { idx.push_back(col); sgn.push_back(static_cast<int8_t>(s)); }
    void close_row() { ptr.push_back(static_cast<int64_t>(idx.size())); }
};

// Column-sorted transpose (each new row lists the old rows in ascending order).
Csr transpose(const Csr &a, int32_t ncols);

struct Complex
{
    int32_t nn = 0, ne = 0, nf = 0, nc = 0;
    Csr e2n, f2e, c2f; // downward
    Csr n2e, e2f, f2c; // upward
    std::vector<int32_t> face_tag; // 0 = untagged
    std::vector<int32_t> material; // 0 fluid, 1 solid
    void build_up();
    void edge_nodes(int32_t e, int32_t &a, int32_t &b) const;
    void face_loop(int32_t f, int32_t *nodes, int &n) const; // node at which each loop edge is entered
};

// Combinatorial checks (reference Validate.cpp:18-69): number of violations of each invariant.
struct TopoReport
{
    long euler = 0, boundary_faces = 0, bad_edges = 0, bad_loops = 0, bad_face_cells = 0;
    long d1d0_bad = 0, d2d1_bad = 0;
    bool ok() const { return bad_edges == 0 && bad_loops == 0 && bad_face_cells == 0 && d1d0_bad == 0 && d2d1_bad == 0; }
};
TopoReport validate(const Complex &cx);

// Same complex as Builder for cells whose faces are all triangles, `faces_per_cell` loops of 3 nodes per
// cell in tri_loops (nodes < 2^21), built with parallel sorts.  tags: (n0, n1, n2, tag) per tagged face.
Complex build_triangle_complex(int32_t n_nodes, const int32_t *tri_loops, int64_t n_cells, int faces_per_cell, const int32_t *tags,
                               int64_t n_tags);

// Builds a Complex from cells given as outward (counter-clockwise seen from outside) node loops.
// Loops of 3 or 4 nodes (tetrahedra, prisms, hexahedra).
class Builder
{
public:
    explicit Builder(int32_t n_nodes, int64_t expect_cells = 0);
    ~Builder();
    int32_t add_cell(const int32_t *loop_ptr, const int32_t *loop_nodes, int n_loops, int material);
    void tag_face(const int32_t *nodes, int n, int tag);
    Complex finish();

private:
    struct Impl;
    Impl *p_;
};
} // namespace mhdg
