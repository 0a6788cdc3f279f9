#pragma once
// Metric arrays of a complex (face area vectors and centres, cell centres and volumes, edge vectors,
// dual-face weights w_{e,c} and w_{e,f}, Hodge stars star1/star2).
//
// Definitions follow reference src/uns/geometry/Geometry.cpp:17-169 (face centre = area-weighted centroid
// of the triangle fan around the vertex mean; cell centre = volume-weighted centroid of the tetrahedra
// (x_c0, c_f, v_i, v_i+1); dual piece of edge e in cell c through face f = 1/2 (x_c - m_e) x (c_f - m_e)).
// Every accumulation is evaluated in the reference's order (cells ascending, faces in c2f order, loop
// positions ascending) but organised as gathers so that the loops run in parallel and the result is
// bit-identical to a serial evaluation.
#include "complex.h"

namespace mhdg
{
struct Geometry
{
    std::vector<double> edge_vec, edge_mid, edge_len;         // 3E, 3E, E
    std::vector<double> face_area, face_center, face_area_mag; // 3F, 3F, F
    std::vector<double> cell_center, cell_volume, cell_min_tet; // 3C, C, C
    std::vector<double> edge_dual_area;                       // 3E
    std::vector<int64_t> ec_ptr;                              // E+1
    std::vector<int32_t> ec_cell;                             // sorted unique cells at each edge
    std::vector<double> ec_weight;                            // w_{e,c}
    std::vector<double> ef_weight;                            // w_{e,f}, aligned with e2f
    std::vector<int32_t> ec_misoriented;                      // E
    std::vector<double> star1, star2, star3;                  // E, F, C
};

Geometry compute_geometry(const Complex &cx, const double *xyz);

// Area-weighted least-squares cell vector from face fluxes (reference Reconstruct.cpp:8-46).
void reconstruct_cell_vector(const Complex &cx, const Geometry &g, const double *flux, double *b3);
} // namespace mhdg
