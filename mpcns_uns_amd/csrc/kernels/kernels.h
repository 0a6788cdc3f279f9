#pragma once
// Kernel declarations shared between the kernel files and the context code.
#include "dev.h"

namespace mhdg
{
constexpr int32_t kHeavyMax = 1024; // largest edge valence handled by k_edge_emf_heavy (LDS arrays)
constexpr int32_t kLightMax = 32;   // edges up to this valence use one thread each

// explicit.hip
__global__ void k_prepare(DevMesh d, Params P, const double *U0, const double *U1, const double *Phi, const double *Badd, double *W0,
                          double *W1, double *Bind, double *Btot, double *uplus, double *alpha, double *dt_part);
__global__ void k_dt_final(const double *part, int32_t n, double dt_max, double *out);
__global__ void k_grad_current(DevMesh d, Params P, int32_t ns, int32_t s0, SpC Ws, SpC ghosts, SpW grads, int32_t limit,
                               const double *Bind, double *J);
__global__ void k_face(DevMesh d, Params P, int32_t ns, int32_t s0, SpC Ws, SpC grads, SpC ghosts, double *FA, const double *uplus,
                       const double *Btot, const double *Bind, const double *Phi, const double *PhiAdd, double *Ef);
__global__ void k_cell_update(DevMesh d, Params P, int32_t ns, int32_t s0, const double *FA, SpC rhs_in, const double *W0,
                              const double *W1, const double *Btot, const double *J, const double *dt_dev, double dt_val,
                              int32_t sources, int32_t update, SpW Us, SpW rhs_out);
__global__ void k_cell_vector(DevMesh d, const double *flux, double *b);
__global__ void k_local_dt(DevMesh d, Params P, const double *W0, const double *W1, const double *Btot, double *dt_part);
__global__ void k_edge_emf(DevMesh d, Params P, const double *Ef, const double *u_cell, const double *phi_ind, double cuct, int32_t vmax,
                           const double *E_in, double *E_out, int32_t apply_bc);
__global__ void k_edge_emf_heavy(DevMesh d, Params P, const double *Ef, const double *u_cell, const double *phi_ind, double cuct,
                                 const int32_t *heavy, const double *E_in, double *E_out, int32_t apply_bc);
__global__ void k_faraday(DevMesh d, const double *E, const double *dte, const double *dt_dev, double dt_val, double *B);
__global__ void k_div_b(DevMesh d, const double *B, const double *extra, double *part_abs, int32_t *part_cell, double *part_rel);
} // namespace mhdg
