#pragma once
// Space-filling-curve order of cells: the Hilbert key of the reference's partitioner
// (src/uns/partition/Partitioner.cpp:37-108) on cell centres, which the one-GPU build uses to renumber the
// whole mesh (cells in key order, faces and edges by first appearance) instead of cutting it into ranks.
#include <cstdint>
#include <vector>

namespace mhdg
{
// 3 x 20-bit Hilbert index of integer coordinates (J. Skilling, "Programming the Hilbert curve", AIP Conf. Proc.
// 707, 381 (2004): axes -> transpose, then bit interleave x, y, z from the top bit).
uint64_t hilbert_key(uint32_t x, uint32_t y, uint32_t z);
// Cells sorted by the Hilbert key of their centres scaled into the bounding box of all centres (ties by id):
// order[k] = the cell at position k.
std::vector<int32_t> hilbert_order(const double *centers, int32_t n);
} // namespace mhdg
