#pragma once
// Mixed-precision implicit induction solve: the operator A = I + dt F d1 and the Krylov basis in FP32, dot
// products accumulated in FP64, Hessenberg least squares and Givens rotations in FP64 on the host, and an
// FP64 iterative-refinement loop whose residual b - A x uses the FP64 operator.
#include <vector>

#include "induction.h"

namespace mhdg
{
// FP32 copies of the metric arrays the operator reads
struct DevMesh32
{
    const float *star2, *m_e, *recon, *ecw, *evec, *invdiag;
    float *aB0;               // per cell alpha_c B0_c of the frozen state (3)
    const float *aV;          // per (cell, face) entry: s A_f / V_c (3)
    const int32_t *nbr;       // per (cell, face) entry: the cell across the face, -1 on the boundary
    // The edge kernel's rows in slices of 64 edges (one wavefront).  Entries 2p and 2p+1 of the row of edge 64 s + i
    // form pair p of lane i; a slice holds as many pairs per edge as its longest row needs, the pairs of one p side by
    // side, so each pair load of a wavefront reads one contiguous range.  A slice whose entries' ids span less than
    // 2^16 is compact: 12-byte pairs (the two ids as 16-bit offsets from the slice's base id in one word, then the
    // two FP32 weights); any other slice is wide: 16-byte pairs (id, weight, id, weight).  Slots past a row, and the
    // rows of high-valence edges (the heavy blocks' edges), hold offset or id 0 with weight +0, so the kernel gathers
    // and adds them without a branch (a zero term leaves the value of every sum as it was).
    // Descriptor of slice s: {first word, pairs per edge, base id, wide}.
    const int4 *fdesc;        // edge-face rows, entries (face, signed star2 of the face): slice descriptors
    const int32_t *fwords;    //   and the slices' words
    const int4 *cdesc;        // edge-cell rows, entries (cell, edge-cell weight)
    const int32_t *cwords;
    const int8_t *eheavy;     // 1 on the edges of valence above kLightMax (the edge kernel's heavy blocks)
    // Tetrahedra in the four-lane cell-vector kernel (one wavefront = 16 cells): the face of each (cell, face) entry as
    // a 16-bit offset from its wavefront's smallest face id; null when some wavefront's faces span 2^16 or more, and
    // the kernel reads c2f instead.
    const uint16_t *c2f16;
    const int32_t *c2fbase;
    int32_t cdeg, fdeg;       // faces per cell / edges per face when uniform (tets: 4, 3), else 0 (ptr arrays)
};

constexpr int kEdgeBlock = 256; // threads per block of the FP32 edge kernel (a multiple of 64: whole slices)
constexpr int kDotChunk = 16; // basis vectors per multi-dot pass
constexpr int kOrthoBlocks = 1024; // grid of the FP32 orthogonalisation kernels (512 and 2048 measured no better)

__global__ void k_d1_f32(DevMesh d, DevMesh32 s, const float *pv, const float *dte, float *phi);
constexpr int kCellLanes = 4; // lanes per cell of the FP32 cell kernels
template <int G> __global__ void k_bcell_f32_g(DevMesh d, DevMesh32 s, const float *phi, float *b);
template <int G> __global__ void k_hall_g_f32_g(DevMesh d, DevMesh32 s, const float *b, float *g);
enum : int { kOpShifted = 0, kOpApply = 1, kOpSmooth = 2, kOpResid = 3 }; // output modes of k_edge_apply_f32 (mixed.hip)
template <int MODE>
__global__ void k_edge_apply_f32(DevMesh d, DevMesh32 s, int32_t resistive, int32_t hall, const float *phi, const float *g, const float *v,
                                 const float *pv, const int8_t *dir, float dt, float *out, const int32_t *heavy, int32_t n_light, const float *b,
                                 float omega);
__global__ void k_tables_f32(DevMesh d, float *aV, int32_t *nbr);
__global__ void k_multiaxpy_f32(const float *const *V, int32_t nv, const double *h, float *w, int32_t n, RedScratch rs, double *nrm);
__global__ void k_to_f32_scaled(const double *x, const double *s, const float *invdiag, const int8_t *dir, int32_t n, float *out, float *pv);
__global__ void k_scale_f32(const float *w, const double *h, const float *invdiag, const int8_t *dir, int32_t n, float *out, float *pv);
__global__ void k_update_mixed(const float *const *V, const double *y, int32_t j, const float *invdiag, const int8_t *dir, int32_t n,
                               double *delta);
__global__ void k_f64_to_f32(const double *x, int64_t n, float *out);
__global__ void k_alpha_f32(const double *alpha, const double *B0, int32_t nc, float *aB0);
// multigrid.hip
__global__ void k_update_flex(const float *const *Z, const double *y, int32_t j, const int8_t *dir, int32_t n, double *delta);

struct MgDevice; // multigrid.hip

struct MixedSolver
{
    DevMesh32 s{};
    float *phi = nullptr, *b = nullptr, *g = nullptr, *w = nullptr, *dte32 = nullptr, *pv = nullptr;
    std::vector<float *> V, Z;      // Krylov basis; Z_j = M V_j of the flexible (multigrid) inner solve
    float **Vptrs = nullptr, **Zptrs = nullptr;
    double *h = nullptr, *part = nullptr, *part_axpy = nullptr;
    int m = 30;
    bool built = false;
};
} // namespace mhdg
