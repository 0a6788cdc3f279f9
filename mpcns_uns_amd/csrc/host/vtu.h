#pragma once
// XML VTK output for visual checks of device results: unstructured-grid files with every cell a
// polyhedron (VTK type 42) and raw appended binary payloads, the parallel master file and the time-series
// collection, in the layout of the reference's writer (src/uns/io/VtkWriter.cpp) so that both programs
// produce the same bytes for the same mesh and fields.
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "complex.h"

namespace mhdg::vtk
{
struct Array
{
    std::string name;
    int32_t ncomp = 1;
    const double *values = nullptr; // ncomp per cell or per node
};

// the first n_cells cells of the complex, nodes xyz (3 per node)
void write_vtu(const std::string &path, const Complex &cx, const std::vector<double> &xyz, int32_t n_cells,
               const std::vector<Array> &cell_data, const std::vector<Array> &point_data);
void write_pvtu(const std::string &path, const std::vector<std::string> &pieces, const std::vector<Array> &cell_data,
                const std::vector<Array> &point_data);
void write_pvd(const std::string &path, const std::vector<std::pair<double, std::string>> &entries);
} // namespace mhdg::vtk
