#pragma once
// Agglomeration hierarchy of the edge complex for the multigrid preconditioner of the implicit induction solve.
//
// Level 0 is the mesh's edges.  Level l+1 agglomerates the nodes of level l by heavy-edge matching passes (pairs,
// then pairs of pairs, ...: aggregates of ~2^passes nodes, the cell matching of the reference's agglomeration
// multigrid, MercuryUns/multigrid/Multigrid.cpp:219-282, applied to the node graph); its unknowns are the pairs of
// adjacent aggregates ("coarse edges", oriented from the lower aggregate number to the higher).  The prolongation
// of a level-l edge is +-1 onto the coarse edge joining its two end aggregates, with the sign of the orientations,
// and 0 for an edge inside one aggregate or a Dirichlet edge: coarse gradients then prolong to fine gradients, the
// de Rham structure of the complex carried to every level.
//
// Coarse operators are Galerkin products of the state-independent part of A = I + dt F d1, the identity and the
// resistive curl-curl K = diag(m_e) d1^T diag(star2) d1 (what the reference's own ILU preconditioner factors,
// ImplicitInduction.cpp:209-244): A_l = M_l + dt K_l with M_l = P^T M_{l-1} P (diagonal, as every P row has at most
// one entry) and K_l = P^T K_{l-1} P, built once per mesh and Dirichlet set; dt enters per solve.
#include <cstdint>
#include <vector>

#include "../../../include/mhdg.h"

namespace mhdg
{
constexpr int32_t kMgNone = INT32_MIN; // prolongation row without an entry

struct MgLevel
{
    int32_t n = 0;                        // unknowns of this level
    int32_t n_nodes = 0;                  // aggregates (nodes) of this level
    std::vector<int32_t> tail, head;      // end nodes of each unknown (tail < head on coarse levels)
    std::vector<double> weight;           // matching weight of each unknown (sum of the fine star1 it gathers)
    // prolongation from this level to the finer one (levels >= 1): per finer unknown, the signed index of its
    // coarse unknown (v >= 0: +1 on v, v < 0: -1 on ~v) or kMgNone; and its transpose by rows of this level
    std::vector<int32_t> map, pt_ptr, pt;
    std::vector<int32_t> agg;             // per node of the finer level: its aggregate, the node of this level
    std::vector<double> mass;             // diagonal of M_l
    std::vector<int32_t> k_ptr, k_col;    // K_l, columns ascending within a row
    std::vector<double> k_val;
};

struct MgHierarchy
{
    std::vector<MgLevel> lv; // lv[0]: the fine edges (n, tail/head/weight; mass 1 on free edges, K not stored)
};

// Fine complex: n_nodes nodes, ne edges with end nodes e_tail/e_head and weights star1; f2e rows (folded signs:
// v >= 0 is +1 on edge v, v < 0 is -1 on edge ~v) with star2 per face; m_e per edge; dir per edge (Dirichlet:
// excluded from every coarse space).  levels >= 1 (1: no coarse level); passes >= 1.  Throws on bad input.
MgHierarchy build_mg_hierarchy(int32_t n_nodes, int32_t ne, const int32_t *e_tail, const int32_t *e_head, const double *star1,
                               int32_t nf, const int32_t *f2e_ptr, const int32_t *f2e, const double *star2, const double *m_e,
                               const int8_t *dir, int levels, int passes);
// Start and end node of every edge from a descriptor's e2n (sign -1 the start, +1 the end); throws without e2n.
void edge_nodes(const mhdg_mesh_desc &d, std::vector<int32_t> &tail, std::vector<int32_t> &head);
} // namespace mhdg
