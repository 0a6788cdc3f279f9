#pragma once
#include "kernels.h"

namespace mhdg
{
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024; // fixed reduction grid (determinism)

struct RedScratch
{
    double *partial;   // kRedBlocks
    unsigned *ticket;  // one counter, zero between launches
};

__global__ void k_finish(const double *partial, int32_t n, double *out, int op);
template <class T> __global__ void k_dot(const T *a, const T *b, int32_t n, RedScratch rs, double *out, int op);
template <class T> __global__ void k_mgs(T *w, const T *Vprev, const T *Vi, int32_t n, double *h, int32_t i, int32_t norm, RedScratch rs);
template <class T> __global__ void k_scale_div(const T *w, const double *h, int32_t n, T *out);
__global__ void k_div_scalar(const double *r, const double *beta, int32_t n, double *out);
__global__ void k_d1(DevMesh d, const double *v, const double *invdiag, const int8_t *dir, const double *dte, double *phi);
__global__ void k_bcell(DevMesh d, const double *phi, double *b);
__global__ void k_hall_g(DevMesh d, const double *b, const double *alpha, const double *B0, double *g);
constexpr int kCellLanes64 = 4; // lanes per cell of the FP64 cell kernels of apply_F
template <int G> __global__ void k_bcell_g(DevMesh d, const double *phi, double *b);
template <int G> __global__ void k_hall_g_g(DevMesh d, const double *b, const double *alpha, const double *B0, double *g);
__global__ void k_edge_apply(DevMesh d, int32_t resistive, int32_t hall, const double *phi, const double *g, const double *ide,
                             const int8_t *imask, const double *v, const double *invdiag, const int8_t *dir, double dt, int32_t mode,
                             int32_t mask, const double *bvec, double *out);
__global__ void k_edge_apply_heavy(DevMesh d, int32_t resistive, int32_t hall, const int32_t *heavy, const double *phi, const double *g,
                                   const double *ide, const int8_t *imask, const double *v, const double *invdiag, const int8_t *dir,
                                   double dt, int32_t mode, int32_t mask, const double *bvec, double *out);
// b = E_expl + F(Phi^n) and r0 = b - A x0 in one pass over the tables (the same bits as the separate products)
template <int G> __global__ void k_bcell2_g(DevMesh d, const double *phi1, const double *phi2, double *b1, double *b2);
template <int G>
__global__ void k_hall_g2_g(DevMesh d, const double *b1, const double *b2, const double *alpha, const double *B0, double *g1, double *g2);
__global__ void k_edge_rhs_residual(DevMesh d, int32_t resistive, int32_t hall, const double *phi1, const double *g1, const double *phi2,
                                    const double *g2, const double *E_expl, const double *x, const int8_t *dir, double dt, double *bvec,
                                    double *r);
__global__ void k_edge_rhs_residual_heavy(DevMesh d, int32_t resistive, int32_t hall, const int32_t *heavy, const double *phi1,
                                          const double *g1, const double *phi2, const double *g2, const double *E_expl, const double *x,
                                          const int8_t *dir, double dt, double *bvec, double *r);
__global__ void k_gmres_update(const double *const *V, const double *y, int32_t j, const double *invdiag, const int8_t *dir, int32_t n,
                               double *delta);
__global__ void k_solve_init(const double *E_expl, const double *corr, const int8_t *dir, const double *m_e, const double *d2, double dt,
                             const double *dte, int32_t n, double *x0, double *invdiag);
__global__ void k_add(const double *a, const double *b, int32_t n, double *out);
// tangent.hip
__global__ void k_field_of_flux(const double *Badd, const double *b, int64_t n, double *B0);
__global__ void k_hall_g_tangent(DevMesh d, const double *b, const double *db, const double *alpha, const double *Badd, double *g);
__global__ void k_newton_residual(const double *E, const double *E_expl, const double *G, const int8_t *dir, int32_t n, double *R);
__global__ void k_scaled(const double *x, double s, int32_t n, double *out);
__global__ void k_seed(const double *v, const double *dv, int64_t n, Dual *out);
__global__ void k_tangent_of(const Dual *x, int64_t n, double *d);
__global__ void k_solve_final(const double *x0, const double *delta, const double *E_expl, const int8_t *dir, int32_t n, double *E,
                              double *corr);
} // namespace mhdg
