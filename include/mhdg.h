/*
 * mhdg — MI355X (gfx950) Hall-MHD step on the signed cell -> face -> edge -> node complex.
 *
 * C-ABI of the device implementation of MPCNS_Uns's `mhd_common` hot path.  Plain pointers and sizes
 * only.  Every function returns 0 on success and a non-zero status otherwise; the message of the last
 * failure is available through mhdg_last_error().  Arrays follow the reference's layout: entity-numbered,
 * AoS `id*ncomp + comp` (reference Common/state/FieldSet.h:2), fluxes in the canonical face orientation,
 * EMFs as line integrals along the stored edge orientation.
 *
 * One context holds one mesh (the whole complex of one GPU: the reference's [owned | ghost] numbering
 * with no ghosts), its device-resident state and one HIP stream.  Calls on one context are not
 * thread-safe; contexts are independent.
 */
#ifndef MHDG_H
#define MHDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mhdg_ctx mhdg_ctx;

/* Signed CSR incidence table (reference include/uns/complex/Incidence.h:10-28): row r lists
 * idx[ptr[r] .. ptr[r+1]) with signs sgn (+1/-1). */
typedef struct mhdg_incidence {
    const int64_t *ptr;
    const int32_t *idx;
    const int8_t *sgn;
} mhdg_incidence;

/* Mesh, metric and boundary arrays (reference CellComplex.h:21-53, Geometry.h:25-43,
 * BoundaryInfo.h:21-29).  f2c and e2f must be the column-sorted transposes of c2f and f2e. */
typedef struct mhdg_mesh_desc {
    int32_t n_edges, n_faces, n_cells;
    mhdg_incidence f2e, c2f, e2f, f2c;
    const int32_t *cell_material;   /* 0 fluid, 1 solid (uns::Material) */
    const int8_t *face_kind;        /* 0 Interior, 1 Wall, 2 Inflow, 3 Outflow, 4 Skip (mhd::FaceKind) */
    const int8_t *edge_mode;        /* 0 free, 1 E = 0, 2 imposed free-stream EMF */
    const double *face_area;        /* 3F */
    const double *face_center;      /* 3F */
    const double *face_area_mag;    /* F  */
    const double *cell_center;      /* 3C */
    const double *cell_volume;      /* C  */
    const double *edge_vec;         /* 3E */
    const double *edge_mid;         /* 3E */
    const double *edge_len;         /* E  */
    const int64_t *ec_ptr;          /* E+1: edge -> incident cells, ascending */
    const int32_t *ec_cell;
    const double *ec_weight;        /* w_{e,c} */
    const double *ef_weight;        /* w_{e,f}, aligned with e2f */
    const double *star1;            /* E */
    const double *star2;            /* F */
    /* Edge -> node incidence (CellComplex::e2n; sign -1 the start node, +1 the end node) and the node count.
     * Optional (NULL ptr / 0): only the multigrid preconditioner (precision 2) reads them, to agglomerate nodes. */
    mhdg_incidence e2n;
    int32_t n_nodes;
} mhdg_mesh_desc;

/* Run configuration: the fields of mhd::Config / mhd::Scales the path reads
 * (reference Common/config/Config.h:19-96, Common/physics/Scales.h:13-41). */
typedef struct mhdg_config {
    double gamma, inver_MA2, M_ref, M_H, M_Na, ne_floor, hall_coef, inver_Rem;
    double cfl, recon_rel_limit, rho_floor, p_floor, rho_floor_dt;
    int32_t second_order, lorentz_force, hall, resistive;
    double eta_r_inner, eta_r_outer, eta_width, hall_r_min, hall_r_max;
    double ind_rtol;
    int32_t ind_max_it, ind_restart;
    double cuct;
    double fs[2][5];                /* free-stream primitive state (rho, u, v, w, p) per species */
    double u_inf[3], B_imf[3];
    /* Dirichlet edges of the implicit solve: 0 = imposed inflow edges only (mercury::Solver,
     * MercuryUns/driver/Solver.cpp:261-262); 1 = inflow and wall edges (mars::MarsSolver,
     * MarsUns/driver/MarsSolver.cpp:730), for a conducting wall that is the domain boundary. */
    int32_t wall_dirichlet;
    /* cfg.implicit_ideal (Config.h:52): the ideal EMF's part linear in the induced flux joins the implicit
     * operator once mhdg_ind_set_ideal_state has supplied u (ImplicitInduction.cpp:150-157). */
    int32_t implicit_ideal;
    /* Ion species of the fluid: 0 or 2 = the two Mercury ions (M_H, M_Na, fs); 3..10 = ion_mass[s] and the
     * free-stream primitive states ion_fs[s] (a ten-ion fluid in the manner of mars::MarsSolver). */
    int32_t n_species;
    double ion_mass[10];
    double ion_fs[10][5];
    /* Exponential-atmosphere ion production, added to the step's sources (no counterpart in the Mercury
     * step): q_s = src_rate[s] exp(-(|x| - src_r0) / src_scale_height) per unit volume and time, ions
     * created at rest with temperature src_T (energy q_s src_T / (gamma - 1)). */
    double src_rate[10];
    double src_scale_height, src_r0, src_T;
    /* Scales::induce_coeff (Scales.h:23; 0 in Scales::simple): with the Lorentz force on, the step applies the
     * exact gyration of the two ions' relative velocity about B as two half steps around the fluid update
     * (mercury::gyrate, MercuryUns/physics/Gyration.cpp:17-51, Solver.cpp:237-241, 272).  Two-ion fluids. */
    double induce_coeff;
    /* Agglomeration multigrid preconditioner of the mixed-precision solve (precision 2), named after the
     * reference's agglomeration-multigrid parameters (Config.h:56-63): mg_levels levels including the fine one;
     * mg_agglom_passes heavy-edge matching passes per level over the node graph (aggregates of ~2^passes
     * nodes); mg_pre / mg_post damped-Jacobi sweeps before / after the coarse correction, mg_coarse sweeps on the
     * coarsest level; mg_relax the Jacobi damping; mg_fine_hall 1: the fine-level smoother applies the full
     * operator (Hall and resistive), 0: its resistive part only.  Zero fields take the defaults 4, 3, 2, 2, 4,
     * 0.7 and 1 (mhdg_config_defaults sets them). */
    int32_t mg_levels, mg_agglom_passes, mg_pre, mg_post, mg_coarse;
    double mg_relax;
    int32_t mg_fine_hall;
} mhdg_config;

/* Fills `cfg` with the reference defaults (Config.h:19-96 initialisers, Scales::simple(5/3, 1, 1)). */
void mhdg_config_defaults(mhdg_config *cfg);

/* ---------------------------------------------------------------------------------------- context */
/* Uploads mesh + geometry and builds the derived tables of the reference constructors:
 * Stencil (Reconstruction.cpp:26-78), ImplicitInduction (ImplicitInduction.cpp:10-103), active faces
 * (Faraday.cpp:7-13).  Replaces the construction of mercury::Solver's members (Solver.cpp:36-44). */
int mhdg_create(const mhdg_mesh_desc *mesh, const mhdg_config *cfg, mhdg_ctx **out);
void mhdg_destroy(mhdg_ctx *ctx);
const char *mhdg_last_error(mhdg_ctx *ctx); /* ctx may be NULL: last creation error */
int mhdg_device_info(char *buf, int buflen);

/* ------------------------------------------------------------- per-call API (host arrays in/out) */
/* Each replaces the reference function named; inputs are copied to the device, outputs back. */

/* uns::reconstruct_cell_vector (src/uns/geometry/Reconstruct.cpp:8-46): flux F -> b 3C. */
int mhdg_cell_vector(mhdg_ctx *ctx, const double *flux, double *b);
/* mhd::cell_current (Common/induction/Current.cpp:9-31): B 3C -> J 3C. */
int mhdg_cell_current(mhdg_ctx *ctx, const double *B, double *J);
/* mhd::Stencil::limited_gradients (Common/fluid/Reconstruction.cpp:80-142): W 5C -> grad 15C.
 * `ghost` (nullable) is a per-face table of 5 primitives used for boundary faces instead of
 * ghost_state(); it replaces the GhostFn callback (Reconstruction.h:17-21). */
int mhdg_limited_gradients(mhdg_ctx *ctx, const double *W, int species, const double fs[5], int limit,
                           const double *ghost, double *grad);
/* mhd::Stencil::set_rel_limit (Common/fluid/Reconstruction.h:27): theta of the relative limiter. */
int mhdg_set_recon_rel_limit(mhdg_ctx *ctx, double theta);
/* mhd::fluid_flux_rhs (Common/fluid/FluxRHS.cpp:30-74): W 5C, grad 15C -> rhs 5C. */
int mhdg_flux_rhs(mhdg_ctx *ctx, int species, const double fs[5], const double *W, const double *grad,
                  const double *ghost, double *rhs);
/* mhd::add_ideal_emf (Common/induction/IdealEMF.cpp:25-129): adds into E (E entries). */
int mhdg_ideal_emf(mhdg_ctx *ctx, const double *u_cell, const double *B_cell_tot, const double *B_cell_ind,
                   const double *phi_tot, const double *phi_ind, double cuct, double *E);
/* mhd::faraday_update (Common/induction/Faraday.cpp:15-24): Phi -= dt d1 E on active faces. */
int mhdg_faraday(mhdg_ctx *ctx, const double *E, double *Phi, double dt);
/* mhd::faraday_update_edge_dt (Faraday.cpp:27-36): Phi -= d1(dt_e E). */
int mhdg_faraday_edge_dt(mhdg_ctx *ctx, const double *E, const double *dt_edge, double *Phi);
/* mhd::div_b (Common/diagnostics/DivB.cpp:8-24): max|d2 B|, max relative, worst cell. */
int mhdg_div_b(mhdg_ctx *ctx, const double *B, const double *extra, double *max_abs, double *max_rel,
               int64_t *worst_cell);
/* No counterpart in the reference: the allocation-free form of mhdg_div_b, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  B (n_faces doubles),
 * extra (n_faces doubles or NULL) and out (3 doubles: max|d2 B|, max relative, worst cell id or -1) are DEVICE
 * pointers; the work is queued on the context's stream and out is valid after it (mhdg_sync).  The reduction's
 * partials go to a fixed device array of the library (1.25 MiB per GPU, kept until the process ends); no device
 * memory is allocated.  At most 16,777,216 cells (larger meshes are refused with an error).  Calls that overlap on
 * one GPU share that array and race.  Same numbers as mhdg_div_b. */
int mhdg_div_b_noalloc(mhdg_ctx *ctx, const double *B, const double *extra, double *out);
/* mercury::cell_time_steps / local_time_step (MercuryUns/fluid/MercurySources.cpp:65-116). */
int mhdg_local_dt(mhdg_ctx *ctx, const double *W_H, const double *W_Na, const double *B_tot, double *dt);
/* mercury::add_sources, Lorentz part (MercurySources.cpp:20-63). */
int mhdg_add_sources(mhdg_ctx *ctx, int species, const double *W_H, const double *W_Na, const double *B_tot,
                     const double *J, double dt, double *rhs);

/* ImplicitInduction (Common/induction/ImplicitInduction.h:27-94). */
int mhdg_ind_set_frozen(mhdg_ctx *ctx, const double *B0_cell, const double *ne_eff_cell);
int mhdg_ind_apply_F(mhdg_ctx *ctx, const double *phi, double *out);
/* ImplicitInduction::active / has_ideal / has_hall (ImplicitInduction.h:31-33). */
int mhdg_ind_flags(mhdg_ctx *ctx, int32_t *active, int32_t *has_ideal, int32_t *has_hall);
/* ImplicitInduction::set_ideal_state (ImplicitInduction.cpp:113-117): u 3C, mask E (non-zero: the edge
 * carries the ideal EMF).  The arrays are copied at the call (the reference keeps the u pointer): call
 * again after u changes. */
int mhdg_ind_set_ideal_state(mhdg_ctx *ctx, const double *u_cell, const int8_t *mask);
/* ImplicitInduction::ideal_linear (ImplicitInduction.cpp:119-125): apply_F with only the ideal part. */
int mhdg_ind_ideal_linear(mhdg_ctx *ctx, const double *phi, double *out);
/* ImplicitInduction::solve (ImplicitInduction.cpp:299-390).  dt_edge (nullable, E values) replaces the
 * scalar dt: system (I + F d1 diag(dt_e)) E = b.  precision: 0 = FP64 GMRES(m) with Jacobi and modified
 * Gram-Schmidt (the reference iteration); 1 = mixed precision (FP32 operator and Krylov basis, classical
 * Gram-Schmidt with FP64 accumulation, inside FP64 iterative refinement to the same relative residual;
 * not available with the implicit ideal part); 2 = the mixed solve with flexible GMRES preconditioned by one FP32
 * agglomeration-multigrid V-cycle per step (mhdg_mg_apply; needs e2n in the mesh descriptor, scalar dt only); 3 = the
 * FP64 iteration of precision 0 with classical Gram-Schmidt (one multi-dot and one multi-axpy pass per step, as the
 * mixed solve orthogonalises) in place of modified Gram-Schmidt.  The warm-start correction is kept in the context
 * between calls, as the reference keeps corr_. */
int mhdg_ind_solve(mhdg_ctx *ctx, const double *Phi_n, const double *E_expl, const int8_t *dirichlet,
                   double dt, const double *dt_edge, int precision, double *E, int32_t *iterations,
                   double *rel_res);

/* Multigrid preconditioner of precision 2 (no reference counterpart: the reference preconditions with Jacobi or a
 * block-Jacobi ILU(0), ImplicitInduction.cpp:209-317).  The hierarchy (include/mhdg_mesh.h, mhdg_mg_build) is built on
 * the first use for a Dirichlet set and kept while later calls bring the same set.  mhdg_mg_apply: one V-cycle z = M v
 * (FP32 inside, FP64 in and out) with the frozen state of mhdg_ind_set_frozen and step dt; mhdg_mg_info: the number of
 * levels and the unknowns of each (up to max_levels entries of n_per_level). */
int mhdg_mg_apply(mhdg_ctx *ctx, const int8_t *dirichlet, double dt, const double *v, double *z);
int mhdg_mg_info(mhdg_ctx *ctx, const int8_t *dirichlet, int32_t max_levels, int32_t *levels, int32_t *n_per_level);

/* --------------------------------------------- nonlinear Hall induction: tangent-linear and Newton-Krylov */
/* No counterpart in the reference (it freezes the Hall field at B_tot^n, ImplicitInduction.cpp:105-111).
 * G(phi) is apply_F with the Hall field B0 = B_add + R phi of the iterate, R the recon map of
 * ImplicitInduction.cpp:150-160 (quadratic in phi).  mhdg_ind_jvp gives G'(phi_lin) dphi (hand-written
 * tangent kernel); mhdg_ind_newton solves E - E_expl - G(Phi_n - dt d1 E) = 0 on the non-Dirichlet edges by
 * Newton's method, each step by FP64 GMRES (Jacobi, MGS) on J = I + dt G'(Phi_k) d1 with relative forcing
 * `eta`, until |R| <= rtol |E_expl + G(Phi_n)|.  E0 (nullable, E values) is the initial iterate (default
 * E_expl).  The frozen alpha comes from mhdg_ind_set_frozen; the background B_add from mhdg_set_background. */
int mhdg_ind_apply_G(mhdg_ctx *ctx, const double *phi, double *out);
int mhdg_ind_jvp(mhdg_ctx *ctx, const double *phi_lin, const double *dphi, double *out);
int mhdg_ind_newton(mhdg_ctx *ctx, const double *Phi_n, const double *E_expl, const int8_t *dirichlet, double dt,
                    const double *E0, double rtol, int max_newton, double eta, double *E, int32_t *newton_its,
                    int32_t *gmres_its, double *rel_res);

/* ------------------------------------------------- tangent-linear of the explicit step (forward mode) */
/* No counterpart in the reference.  The explicit-stage kernels are templates on their real type; with a dual
 * number (value, derivative) they propagate one tangent direction exactly through the explicit stage of
 * step_first_order (prepare, CFL step, limited gradients, Rusanov fluxes, sources, forward Euler + floors,
 * UCT EMF with the boundary EMF, CT), branches taken on the values.
 * mhdg_shell_energy: J = sum over fluid cells with r0 <= |x_c| <= r1 of |b_c + BaddCell_c|^2 V_c, b the
 * reconstructed cell field of the resident Phi.
 * mhdg_explicit_tangent: from the resident state, seeds dU_H, dU_Na (5C), dPhi (F), dPhiAdd (F), dBadd (3C),
 * dfs (2 x 5: free-stream states), dE_inf (3: imposed inflow EMF), every one nullable = 0, runs n_steps
 * explicit steps on (value, tangent) pairs and returns J and dJ of the final state and, when non-null, the
 * tangents of the final U_H, U_Na, Phi.  The resident state is not changed; edge valence at most 512. */
int mhdg_shell_energy(mhdg_ctx *ctx, double r0, double r1, double *J);
int mhdg_explicit_tangent(mhdg_ctx *ctx, int n_steps, double dt_max, const double *dU_H, const double *dU_Na,
                          const double *dPhi, const double *dPhiAdd, const double *dBadd, const double *dfs,
                          const double *dE_inf, double r0, double r1, double *J, double *dJ, double *dU_H_out,
                          double *dU_Na_out, double *dPhi_out);

/* Reverse mode (adjoint) of mhdg_explicit_tangent: J, the shell magnetic energy after n_steps explicit steps from
 * the resident state (left unchanged), and its gradient with respect to the initial U_H, U_Na (5C each), Phi (F),
 * the background PhiAdd (F) and BaddCell (3C), the free-stream states (2x5, species-major) and the imposed inflow
 * EMF E_inf (3).  Every output may be NULL.  The forward sweep keeps each step's state (n_steps x (10C + F)
 * doubles of device memory) and the reverse sweep recomputes the step's limited gradients, fluxes and EMFs from
 * it.  Two-ion contexts.  No single reference function: the adjoint of step_first_order's explicit stage
 * (MercuryUns/driver/Solver.cpp:228-277), the counterpart of mhdg_ind_jvp on the explicit side. */
int mhdg_step_vjp(mhdg_ctx *ctx, int n_steps, double dt_max, double r0, double r1, double *J, double *gU_H, double *gU_Na,
                  double *gPhi, double *gPhiAdd, double *gBadd, double *gfs, double *gE_inf);

/* --------------------------------------------------------------- device-resident step (fused path) */
/* State: U_H, U_Na (5C each, AoS), induced flux Phi (F); background PhiAdd (F) and BaddCell (3C). */
int mhdg_set_state(mhdg_ctx *ctx, const double *U_H, const double *U_Na, const double *Phi);
/* Conservative state (5C) and primitive state of the last prepare (5C) of ion species s < n_species. */
int mhdg_set_species_state(mhdg_ctx *ctx, int species, const double *U);
int mhdg_get_species_state(mhdg_ctx *ctx, int species, double *U);
int mhdg_get_species_prim(mhdg_ctx *ctx, int species, double *W);
int mhdg_get_state(mhdg_ctx *ctx, double *U_H, double *U_Na, double *Phi);
int mhdg_set_background(mhdg_ctx *ctx, const double *PhiAdd, const double *BaddCell);
/* Reference scales and particle masses of MarsSolver's point-implicit sources (MarsParams,
 * MarsUns/driver/MarsSolver.h:42-59; MarsSolver.cpp:40-62): SI scales, rho_ref = Mref n / NA, p_ref = rho_ref U^2,
 * time_ref = L / U; ion_mass[s] the particle mass of ion species s [kg] (MARS::ParticleMass); n_neutral reactive
 * neutrals (<= 10, MARS::NumReactiveNeutral) with particle masses neutral_mass[n] [kg], read per cell from a
 * density array [cm^-3] of neutral_stride values per cell (MARS::NumNeutral = 12 in the reference). */
typedef struct mhdg_mars_params {
    double kB, q, L, U, B, n;
    double rho_ref, p_ref, time_ref;
    double rho_floor, p_floor;
    double ion_mass[10];
    int32_t n_neutral, neutral_stride;
    double neutral_mass[10];
} mhdg_mars_params;
/* MARS::MarsSolver::implicit_sources_ (MarsUns/driver/MarsSolver.cpp:422-527) on the resident state of every ion
 * species, with B_tot of the last prepare (mhdg_prepare or the step's own): Crank-Nicolson gyration of the coupled
 * ions about B, backward-Euler ion-neutral drag, and backward-Euler thermal relaxation with frictional heating, per
 * fluid cell.  neutral_cm3: neutral densities, neutral_stride per cell (host array). */
int mhdg_mars_implicit_sources(mhdg_ctx *ctx, const mhdg_mars_params *prm, const double *neutral_cm3, double dt);
/* Solver::prepare (MercuryUns/driver/Solver.cpp:94-125) on the resident state: primitives, B_ind, B_tot, u+ and
 * J of the current state for mhdg_get_derived (what Solver::write_vtk outputs, Solver.cpp:677-696). */
int mhdg_prepare(mhdg_ctx *ctx);
/* Device-side copy of the resident state (U of every species, Phi) and its restoration, which also drops the
 * implicit solve's warm start (as at the first step).  For repeated runs of one step from a fixed state; no
 * counterpart in the reference. */
int mhdg_state_save(mhdg_ctx *ctx);
int mhdg_state_restore(mhdg_ctx *ctx);
/* mercury::Solver::step_first_order (MercuryUns/driver/Solver.cpp:228-277) on the resident state:
 * prepare, CFL step (min with dt_max), limited gradients, Rusanov RHS, Lorentz sources, forward Euler
 * + floors, ideal UCT EMF + boundary EMF, [implicit Hall/resistive solve], CT.  `precision` as in
 * mhdg_ind_solve.  Outputs the dt used, GMRES iterations and final relative residual. */
int mhdg_step(mhdg_ctx *ctx, double dt_max, int precision, double *dt_used, int32_t *iterations,
              double *rel_res);
/* The same step with the implicit solve switched off (ideal MHD: E = E_expl), and the same step that
 * requires the implicit Hall/resistive solve to be active (error otherwise).  Fused device-resident
 * forms of the explicit and the Hall configurations of step_first_order. */
int mhdg_step_explicit(mhdg_ctx *ctx, double dt_max, double *dt_used);
int mhdg_step_hall(mhdg_ctx *ctx, double dt_max, int precision, double *dt_used, int32_t *iterations,
                   double *rel_res);
/* Derived cell fields of the last prepared state (Solver.h:97-101): which = 0 W_H, 1 W_Na (5C),
 * 2 B_ind, 3 B_tot, 4 J, 5 u_plus (3C), 6 E of the last step (E). */
int mhdg_get_derived(mhdg_ctx *ctx, int which, double *out);
/* Streaming calibration (no reference counterpart): kind 0 reads, 1 copies, 2 forms a = b + s c over arrays of `bytes`
 * bytes each with one float4 per thread over the full grid; the median of `repeats` timed launches (HIP events, after
 * one warm-up) in microseconds and the bytes one launch moves.  Allocates and frees its own buffers. */
int mhdg_stream_calibration(int kind, int64_t bytes, int repeats, double *median_us, double *bytes_moved);
/* Waits for the context's stream. */
int mhdg_sync(mhdg_ctx *ctx);
/* Per-kernel timing with HIP events recorded on the context's stream around every kernel the step and
 * the solves launch: enable = 1 on, 0 off, -1 clear the totals.  The report lists one line per kernel
 * name: "<name> <launches> <milliseconds>". */
int mhdg_profile(mhdg_ctx *ctx, int enable);
int mhdg_profile_report(mhdg_ctx *ctx, char *buf, int buflen);
/* Restricts the timing to the kernels of one name (NULL or "": every kernel), so that a timed region carries
 * events around the kernel of interest only. */
int mhdg_profile_only(mhdg_ctx *ctx, const char *name);

#ifdef __cplusplus
}
#endif
#endif
