#pragma once
// Device-side data of one context and the arithmetic helpers shared by the kernels.
//
// Incidence on the device is int32 CSR with the sign folded into the entry: v >= 0 is +1 on entity v,
// v < 0 is -1 on entity ~v (4 bytes per entry instead of 4 + 1).  Per-entity metric arrays are AoS
// (3 doubles per vector) as in the reference, which keeps a gathered vector in one cache line.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dual.h"

namespace mhdg
{
__device__ __forceinline__ int32_t pid(int32_t v) { return v >= 0 ? v : ~v; }
__device__ __forceinline__ double psg(int32_t v) { return v >= 0 ? 1.0 : -1.0; }
// std::min / std::max semantics (second operand wins only on strict comparison)
__device__ __forceinline__ double mn(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double mx(double a, double b) { return (a < b) ? b : a; }

// Face kinds and edge modes (mhd::FaceKind / BoundaryInfo::edge_mode)
enum : int8_t { kInterior = 0, kWall = 1, kInflow = 2, kOutflow = 3, kSkip = 4 };
constexpr int32_t kNoStencil = INT32_MIN;

template <class R> struct PrimT { R rho, u, v, w, p; };
using Prim = PrimT<double>;

struct DevMesh
{
    int32_t nc, nf, ne;
    int32_t cdeg, fdeg;             // faces per cell / edges per face when every row has that many (tets: 4, 3), else 0
    const int32_t *c2f_ptr, *c2f;   // cell -> faces, stored order
    const int32_t *cfs;             // cell -> faces sorted by face id (same ptr as c2f), signed
    const int32_t *f2e_ptr, *f2e;   // face -> edges, loop order
    const int32_t *e2f_ptr, *e2f;   // edge -> faces ascending (sign of the f2e entry)
    const int32_t *fL, *fR;         // cell seeing the face outward (+1) / inward (-1); -1 if none
    const int32_t *c2n;             // per c2f entry: the cell across that face (-1 unless both sides are cells)
    const int32_t *ec_ptr, *ec_cell;
    const int8_t *mat, *kind, *emode;
    const double *farea, *fcen, *famag, *ccen, *cvol, *evec, *emid, *elen, *ecw, *efw, *star2;
    // Stencil (reference Reconstruction.h:34-41), aligned with c2f entries
    const int32_t *snb;             // neighbour cell, -(f+1) virtual, kNoStencil for a Skip face
    const double *sw, *sd, *sinv;   // weight, offset (3), symmetric inverse (6 per cell: 0 1 2 4 5 8)
    const int8_t *svalid;
    // ImplicitInduction constants (reference ImplicitInduction.h:73-91)
    const double *recon;            // 3 per c2f entry
    const double *m_e;              // eta_e / star1_e
    const double *d2;               // sum of star2 over the faces of e
    const double *taper;            // per cell
    const double *atm;              // exp(-(|x_c| - r0) / H) per cell for the ion production, null without it
};

constexpr int kMaxSpecies = 10;

struct Params
{
    double gamma, inver_MA2, M_ref, M_H, M_Na, ne_floor, hall_coef;
    double cfl, rel_limit, rho_floor, p_floor, rho_floor_dt, cuct;
    int32_t second_order, lorentz, hall, resistive;
    double fs[kMaxSpecies][5];
    double E_inf[3];
    // tangent seeds of the free-stream states and of the imposed inflow EMF (used by the R = Dual kernels)
    double dfs[kMaxSpecies][5];
    double dE_inf[3];
    // ion species: count, molar masses (M_s[0] = M_H, M_s[1] = M_Na for the two-ion Mercury fluid)
    int32_t ns;
    double M_s[kMaxSpecies];
    // exponential-atmosphere ion production q_s = rate_s exp(-(|x| - r0) / H), new ions at rest at temperature T
    double src_rate[kMaxSpecies];
    double src_H, src_r0, src_T;
    double induce_coeff; // gyration split (0: off)
};

// Per-species array pointers passed by value (kernel arguments)
template <class R> struct SpCT { const R *p[kMaxSpecies]; };
template <class R> struct SpWT { R *p[kMaxSpecies]; };
using SpC = SpCT<double>;
using SpW = SpWT<double>;

__device__ __forceinline__ V3 L3(const double *p, long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// Common/physics/Euler.h:13-49
template <class R> __device__ __forceinline__ PrimT<R> to_prim(const R *U, double gamma, double rho_floor, double p_floor)
{
    PrimT<R> W;
    W.rho = mx(U[0], rho_floor);
    W.u = U[1] / W.rho; W.v = U[2] / W.rho; W.w = U[3] / W.rho;
    const R ke = 0.5 * W.rho * (W.u * W.u + W.v * W.v + W.w * W.w);
    W.p = mx((gamma - 1.0) * (U[4] - ke), p_floor);
    return W;
}
template <class R> __device__ __forceinline__ void to_cons(const PrimT<R> &W, double gamma, R *U)
{
    U[0] = W.rho; U[1] = W.rho * W.u; U[2] = W.rho * W.v; U[3] = W.rho * W.w;
    U[4] = W.p / (gamma - 1.0) + 0.5 * W.rho * (W.u * W.u + W.v * W.v + W.w * W.w);
}

// n_s = max(rho_s, 0) / M_s * M_ref (Common/physics/Scales.cpp:52-53 for any ion)
template <class R> __device__ __forceinline__ R ion_density(const Params &P, int s, const R &rho) { return mx(rho, 0.0) / P.M_s[s] * P.M_ref; }
// n_e = sum of the ion densities in species order (Scales.cpp:50-62: n_e = n_H + n_Na for the two Mercury ions,
// the same operations for any count of ions); n_e,eff = sqrt(n_e^2 + floor^2) is formed by the callers
template <class R> __device__ __forceinline__ R electron_density(const Params &P, const R *const *W, long c)
{
    R ne = ion_density(P, 0, W[0][5 * c]);
    for (int s = 1; s < P.ns; ++s) ne = ne + ion_density(P, s, W[s][5 * c]);
    return ne;
}

// Global thread index of a one-entity-per-thread launch.  (An XCD-contiguous block order, which keeps the
// blocks of one XCD on one index range, was measured on the gather kernels: 4 % faster for the FP32 edge
// kernel, which uses it, and no gain or a loss for the others, where the 256 MB MALL already serves the data
// the eight L2s share.)
__device__ __forceinline__ int32_t gid() { return static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); }
} // namespace mhdg
