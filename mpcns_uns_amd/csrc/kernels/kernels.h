#pragma once
// Kernel declarations shared between the kernel files and the context code.
#include "dev.h"

namespace mhdg
{
constexpr int32_t kHeavyMax = 1024; // largest edge valence handled by k_edge_emf_heavy (LDS arrays)
constexpr int32_t kLightMax = 32;   // edges up to this valence use one thread each
// Gather loops issue the index loads of this many table entries together, then their gathers, then the sums in
// entry order (a loop of one entry at a time waits on each index load before its gather)
constexpr int kEdgeUnroll = 8; // entries of an edge row (faces or cells around the edge)
constexpr int kRowUnroll = 4;  // entries of a face or cell row (edges of a face, faces of a cell)
// Block index that gives each XCD one contiguous range of blocks (the hardware runs block b on XCD b % 8), so that
// neighbouring edges share an L2 for their gathers: 4 % faster for the FP32 edge kernel, slower for the cell kernels
__device__ __forceinline__ int32_t xcd_block(uint32_t nb)
{
    const uint32_t b = blockIdx.x, per = nb / 8;
    if (b >= 8 * per) return int32_t(b);
    return int32_t((b % 8) * per + b / 8);
}
// LDS arrays of k_edge_emf_heavy: half the valence for the 16-byte Dual
template <class R> constexpr int32_t kHeavyShared() { return sizeof(R) == sizeof(double) ? kHeavyMax : kHeavyMax / 2; }

// explicit.hip: templates on the real type R (double: the step; Dual: its tangent-linear)
template <class R> __global__ void k_prepare(DevMesh d, Params P, SpCT<R> Us, const R *Phi, const R *Badd, SpWT<R> Ws, R *Bind, R *Btot,
                                             R *uplus, R *alpha, R *dt_part);
template <class R> __global__ void k_dt_final(const R *part, int32_t n, double dt_max, R *out);
template <class R, bool HOIST = false> __global__ void k_grad_current(DevMesh d, Params P, int32_t ns, int32_t s0, SpCT<R> Ws, SpCT<R> ghosts, SpWT<R> grads, int32_t limit,
                               const R *Bind, R *J);
__global__ void k_gyrate(DevMesh d, Params P, double *UH, double *UNa, const double *Btot, const double *dt_dev, double scale);
template <class R> __global__ void k_face_species(DevMesh d, Params P, int32_t ns, SpCT<R> Ws, SpCT<R> grads, R *FA);
template <class R> __global__ void k_face(DevMesh d, Params P, int32_t ns, int32_t s0, SpCT<R> Ws, SpCT<R> grads, SpCT<R> ghosts, R *FA, const R *uplus,
                       const R *Btot, const R *Bind, const R *Phi, const R *PhiAdd, R *Ef);
template <class R> __global__ void k_cell_update(DevMesh d, Params P, int32_t ns, int32_t s0, const R *FA, SpCT<R> rhs_in, SpCT<R> Wall,
                                                 const R *Btot, const R *J, const R *dt_dev, double dt_val, int32_t sources, int32_t update,
                                                 SpWT<R> Us, SpWT<R> rhs_out);
template <class R> __global__ void k_cell_update_many(DevMesh d, Params P, int32_t ns, const R *FA, SpCT<R> Wall, const R *Btot, const R *J,
                                                     const R *dt_dev, SpWT<R> Us);
__global__ void k_cell_vector(DevMesh d, const double *flux, double *b);
__global__ void k_local_dt(DevMesh d, Params P, SpC Ws, const double *Btot, double *dt_part);
template <class R> __global__ void k_edge_emf(DevMesh d, Params P, const R *Ef, const R *u_cell, const R *phi_ind, double cuct, int32_t vmax,
                           const R *E_in, R *E_out, int32_t apply_bc);
template <class R> __global__ void k_edge_emf_heavy(DevMesh d, Params P, const R *Ef, const R *u_cell, const R *phi_ind, double cuct,
                                 const int32_t *heavy, const R *E_in, R *E_out, int32_t apply_bc);
template <class R> __global__ void k_faraday(DevMesh d, const R *E, const double *dte, const R *dt_dev, double dt_val, R *B);
__global__ void k_div_b(DevMesh d, const double *B, const double *extra, double *part_abs, int32_t *part_cell, double *part_rel);
template <class R> __global__ void k_shell_energy(DevMesh d, const R *phi, const R *badd, double r0, double r1, R *part);
} // namespace mhdg
