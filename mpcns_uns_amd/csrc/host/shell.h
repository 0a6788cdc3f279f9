#pragma once
// Seeded synthetic spherical-shell tetrahedral mesh (lat-lon net, geometric radial layers, Kuhn split).
//
// Nodes: n_lat rows of colatitude (the two poles are single nodes per radial level) x n_lon longitudes x
// (n_r + 1) radial levels; layer thickness grows by `ratio`; interior levels are jittered by a uniform
// `jitter` fraction of the local spacing (seeded mt19937_64).  Each hexahedron (j, i, k) is split into the
// six Kuhn tetrahedra along its (0,0,0)-(1,1,1) diagonal; next to a pole the hexahedron is a wedge and the
// three degenerate tetrahedra are dropped.  Cells are ordered (j, i, k) with k innermost so that the
// radial column of a net cell is contiguous.  Boundary faces are tagged 1 ("inner", r = r_in) and 2
// ("outer", r = r_out).
#include <cstdint>
#include <vector>

#include "complex.h"

namespace mhdg
{
struct ShellSpec
{
    int n_lat = 24, n_lon = 48, n_r = 9;
    double r_in = 1.0, r_out = 4.0, ratio = 1.05, jitter = 0.05;
    uint64_t seed = 20261202;
};

struct ShellMesh
{
    std::vector<double> xyz;      // 3 per node
    std::vector<int32_t> tets;    // 4 per cell, positive orientation
    std::vector<int32_t> tags;    // (n0, n1, n2, tag) per tagged boundary triangle
};

ShellMesh make_shell(const ShellSpec &s);

// Outward loops of a positively oriented tetrahedron (a, b, c, d): [a c b], [a b d], [a d c], [b c d].
void tet_loops(const int32_t *t, int32_t loops[12]);

// Complex of a tetrahedral mesh (cells in order, faces/edges by first appearance), tags applied.  The
// sort-based parallel builder is used when `sorted_builder` is set and node ids fit in 21 bits; both
// builders give identical complexes.
Complex build_tet_complex(int32_t n_nodes, const int32_t *tets, int64_t n_cells, const int32_t *tags, int64_t n_tags,
                          bool sorted_builder = true);
} // namespace mhdg
