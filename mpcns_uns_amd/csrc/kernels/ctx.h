#pragma once
// Device context: mesh tables, resident state, work buffers and the host drivers of the step and of the
// implicit induction solves.
#include <map>
#include <string>
#include <vector>

#include "../../../include/mhdg.h"
#include "induction.h"
#include "mixed.h"

namespace mhdg
{
void noalloc_setup(int device); // noalloc.hip: the device address of the fixed scratch array, once per device
}

struct mhdg_ctx
{
    hipStream_t stream = nullptr;
    int device = 0; // the device current at creation (its stream's)
    mhdg_config cfg{};
    mhdg::Params P{};
    mhdg::DevMesh d{};
    int32_t nc = 0, nf = 0, ne = 0, n_heavy = 0;
    size_t c2f_nnz = 0, ec_nnz = 0, e2f_nnz = 0;
    int32_t cdeg = 0, fdeg = 0; // uniform faces per cell / edges per face (0: mixed shapes)
    void *heavy = nullptr;
    std::vector<double> taper_h;
    int8_t *dir_mercury = nullptr;
    // resident state and derived fields
    double *U[mhdg::kMaxSpecies]{}, *Phi = nullptr, *PhiAdd = nullptr, *Badd = nullptr;
    double *W[mhdg::kMaxSpecies]{}, *grad[mhdg::kMaxSpecies]{}, *Bind = nullptr, *Btot = nullptr, *uplus = nullptr, *J = nullptr, *alpha = nullptr, *B0 = nullptr;
    double *FA = nullptr, *Ef = nullptr, *E_expl = nullptr, *E_last = nullptr, *dt_part = nullptr, *dt_dev = nullptr;
    // implicit induction work
    double *phi_w = nullptr, *bcell = nullptr, *gcell = nullptr, *corr = nullptr;
    bool have_corr = false;
    double *x0 = nullptr, *bvec = nullptr, *rvec = nullptr, *wvec = nullptr, *invdiag = nullptr, *delta = nullptr, *tmp = nullptr;
    double *scal = nullptr, *hcol = nullptr, *ydev = nullptr;
    // ImplicitInduction flags (ImplicitInduction.cpp:11) and the parts apply_op evaluates
    bool op_res = false, op_hall = false, ideal_cfg = false, skip_implicit = false;
    double *u_ideal = nullptr, *ide = nullptr, *Ef_i = nullptr; // implicit ideal part (set_ideal_state)
    int8_t *imask = nullptr;
    const double *dte_cur = nullptr;                              // per-edge pseudo-time step of the current solve
    // Newton-Krylov on the nonlinear Hall induction (tangent.hip)
    double *lin_b = nullptr;           // cell field at the linearisation point: apply_op then applies the tangent operator
    double rtol_override = 0;          // GMRES tolerance of a Newton step (0: cfg.ind_rtol)
    // tangent-linear of the explicit step (Dual state, tangent.hip)
    struct TangentWork { mhdg::Dual *U[2]{}, *Phi = nullptr, *PhiAdd = nullptr, *Badd = nullptr, *W[2]{}, *grad[2]{}, *Bind = nullptr,
                         *Btot = nullptr, *uplus = nullptr, *J = nullptr, *alpha = nullptr, *FA = nullptr, *Ef = nullptr, *E = nullptr,
                         *dt_part = nullptr, *dt_dev = nullptr, *part = nullptr; } tw;
    double *energy_part = nullptr;
    double *snapshot = nullptr; // mhdg_state_save: U of every species and Phi
    // reverse mode of the explicit step (adjoint.hip): gather maps, recomputed intermediates, adjoints, slots,
    // and the state of every step of the forward sweep
    struct AdjointWork { int32_t *fsl = nullptr, *f2e_pos = nullptr, *cec_ptr = nullptr, *cec = nullptr, *argc = nullptr;
                         double *U[2]{}, *W[2]{}, *G[2]{}, *Ub[2]{}, *Wb[2]{}, *Gb[2]{}, *sF[2]{}, *sG[2]{}, *sFA[2]{};
                         double *Phi = nullptr, *Bind = nullptr, *Btot = nullptr, *up = nullptr, *J = nullptr, *alpha = nullptr, *FA = nullptr,
                                *Ef = nullptr, *E = nullptr, *dt_part = nullptr, *dt_dev = nullptr;
                         double *Phib = nullptr, *PhiAddb = nullptr, *Baddb = nullptr, *Eb = nullptr, *Efb = nullptr, *FAb = nullptr, *upb = nullptr,
                                *Btb = nullptr, *Bib = nullptr, *Jb = nullptr, *sE = nullptr, *sPhi = nullptr, *eF = nullptr, *eC = nullptr,
                                *part = nullptr;
                         mhdg::Dual *wd = nullptr;
                         std::vector<double *> ck; } aw;
    struct NewtonWork { double *phik = nullptr, *G = nullptr, *R = nullptr, *rhs = nullptr, *delta = nullptr, *E = nullptr,
                        *lin_b = nullptr, *B0 = nullptr, *zero_phi = nullptr, *corr_save = nullptr; } nt;
    std::vector<double *> V;
    double **Vptrs = nullptr;
    int restart = 2;
    mhdg::RedScratch rs{};
    mhdg::MixedSolver *mixed = nullptr;
    // multigrid preconditioner (multigrid.hip): the host arrays the hierarchy is built from (kept when the mesh
    // descriptor carries e2n), the device hierarchy, and the Dirichlet set it was built for
    struct MgHost { int32_t n_nodes = 0; std::vector<int32_t> tail, head, f2e_ptr, f2e; std::vector<double> star1, star2, m_e; } mgh;
    std::vector<int8_t> dir_mercury_h, mg_dir_h;
    mhdg::MgDevice *mg = nullptr;
    void mg_build(const int8_t *dir_host);
    void mg_set_solve(const int8_t *dir_dev, float dt); // Dirichlet set (device) and step of the solve the cycles serve
    void mg_prepare(double dt);
    void mg_cycle(int l, const float *b, float *x);
    void op32(int mode, int hall_on, const float *pv, const float *v, const float *b, float omega, float *out);
    // pinned host memory: the step's time step at offset 0, small read-backs staged at kPinStageOffset
    static constexpr size_t kPinBytes = 4096, kPinStageOffset = 2048, kPinStageBytes = 2048;
    void *host_pin = nullptr;
    std::vector<void *> owned;
    std::map<int, std::pair<void *, size_t>> staging;
    std::string err;
    // per-kernel timing with HIP events on the context's stream (enabled by mhdg_profile)
    bool prof_on = false;
    struct ProfRec { const char *name; hipEvent_t a, b; };
    std::vector<ProfRec> prof_pending;
    std::vector<hipEvent_t> prof_pool; // events reused across launches (creating two per launch cost ~5% of a step)
    size_t prof_used = 0;
    std::string prof_only; // when set, only kernels of this name are timed
    bool prof_armed = false; // prof_begin recorded the start of the current launch
    std::map<std::string, std::pair<long, double>> prof_acc; // name -> (launches, milliseconds)
    void prof_begin(const char *name);
    void prof_end();
    void prof_flush();

    ~mhdg_ctx();
    void build(const mhdg_mesh_desc &m, const mhdg_config &c);
    void *alloc(size_t bytes);
    template <class T> T *upload(const std::vector<T> &v);
    template <class T> T *upload(const T *v, size_t n);
    double *stage(int slot, size_t n_doubles);
    double *stage_in(int slot, const void *host, size_t n_doubles);
    void d2h(void *host, const void *dev, size_t bytes);
    void check_launch(const char *what);
    void launch_edge_emf(const double *Ef, const double *u, const double *phi_ind, double cuct, const double *Ein, double *Eout, int bc);
    double dot(const double *a, const double *b, int op, int slot);
    void apply_op(const double *phi, const double *v, const double *inv, const int8_t *dir, double dt, int mode, int mask, const double *b,
                  double *out);
    // bvec = E_expl + F(Phi_n) and rvec = bvec - A x (no Jacobi scaling), the two products in one pass over the
    // tables where they allow it (no implicit ideal part, no tangent operator); the same bits as two apply_op calls
    void rhs_residual(const double *Phi_n, const double *E_expl, const double *x, const int8_t *dir, double dt);
    double *bcell2 = nullptr, *gcell2 = nullptr; // the second vector's cell fields of rhs_residual
    int solve64(const double *Phi_n, const double *E_expl, const int8_t *dir, double dt, double *E, double *rel_res);
    // the same system and iteration with classical Gram-Schmidt (cgs64.hip, precision 3)
    int solve64_cgs(const double *Phi_n, const double *E_expl, const int8_t *dir, double dt, double *E, double *rel_res);
    double *cgs_part = nullptr, *cgs_h = nullptr;
    void mixed_build();
    // use_mg: the inner solve is flexible GMRES preconditioned by one V-cycle of the multigrid hierarchy built for dir_host
    // (the host copy of dir); otherwise GMRES with the Jacobi preconditioner
    int solve_mixed(const double *Phi_n, const double *E_expl, const int8_t *dir, double dt, double *E, double *rel_res,
                    bool use_mg = false, const int8_t *dir_host = nullptr);
    double step(double dt_max, int precision, int *iters, double *res);
    void newton_alloc();
    void tangent_alloc();
    void explicit_tangent(int n_steps, double dt_max, const mhdg::Params &Pt);
    template <class R> void shell_energy(const R *phi, const R *badd, double r0, double r1, R *out);
    void vjp_alloc(int n_steps);
    double vjp_stage(double dt_max, bool update);
    double step_vjp(int n_steps, double dt_max, double r0, double r1, double *fs_b, double *einf_b);
    void apply_G(const double *phi, double *out);
    void jvp(const double *phi_lin, const double *dphi, double *out);
    int newton(const double *Phi_n, const double *E_expl, const int8_t *dir, double dt, const double *E0, double rtol, int max_newton,
               double eta, double *E, int *gmres_total, double *rel_res);
};
