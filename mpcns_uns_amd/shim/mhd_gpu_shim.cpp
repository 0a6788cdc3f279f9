// Drop-in replacement of MPCNS_Uns's mhd_common hot-path sources for a one-GPU (one MPI rank) build of the
// MercuryUns / MarsUns drivers.  Compile this file against the reference's headers and link it, with
// libmhdg.so, in place of
//
//   Common/fluid/Reconstruction.cpp   (mhd::ghost_state, mhd::Stencil)
//   Common/fluid/FluxRHS.cpp          (mhd::fluid_flux_rhs)
//   Common/induction/Current.cpp      (mhd::cell_current)
//   Common/induction/Faraday.cpp      (mhd::active_faces, faraday_update, faraday_update_edge_dt)
//   Common/induction/IdealEMF.cpp     (mhd::add_ideal_emf)
//   Common/induction/ImplicitInduction.cpp
//   Common/diagnostics/DivB.cpp       (mhd::div_b)
//
// Every function keeps the reference signature, argument meaning and error behaviour (a failing device call
// throws std::runtime_error with the library's message).  The device context of a mesh is created at the
// first call that needs it and is keyed by the mesh's CellComplex together with its entity counts and table
// addresses (a different mesh at a reused address gets a new context); it is rebuilt when a later call brings
// the boundary info or the configuration it was built without (setup only).  Config::ind_ilu is refused.  Arrays stay host-side
// std::vectors owned by the drivers: each call copies its inputs to the device and its outputs back.
//
// Scope: one rank (the Exchanger is the identity, every entity is owned); the fluid calls serve any species
// index below 10 with the caller's free-stream state (the ion slots of MarsSolver's ten-ion fluid).  A GhostFn is
// evaluated on the host once per boundary face and species into the ghost table of the device call, with the inside
// state the reference would hand it (the cell value for the gradients, the reconstructed face state for the fluxes),
// so any GhostFn gives the reference's result, MarsSolver's (MarsSolver.cpp:124-128) included.
// The implicit solve runs in FP64 (the reference iteration) unless MHDG_SOLVE_PRECISION=1 selects the
// mixed-precision solve (2: mixed precision with the multigrid preconditioner; 3: FP64 with classical Gram-Schmidt).
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "diagnostics/DivB.h"
#include "fluid/FluxRHS.h"
#include "fluid/Reconstruction.h"
#include "induction/Current.h"
#include "induction/Faraday.h"
#include "induction/IdealEMF.h"
#include "induction/ImplicitInduction.h"
#include "mhdg.h"

namespace
{
using uns::Id;

void ok(mhdg_ctx *ctx, int status, const char *what)
{
    if (status != 0) throw std::runtime_error(std::string(what) + ": " + mhdg_last_error(ctx));
}

const double *d3(const std::vector<uns::Vec3> &v) { return v.empty() ? nullptr : &v[0].x; }

mhdg_incidence inc(const uns::Incidence &t) { return {t.ptr.data(), t.idx.data(), t.sgn.data()}; }

// Config -> the C-ABI configuration (the fields of mhd::Config / mhd::Scales the path reads)
mhdg_config to_abi(const mhd::Config &k)
{
    mhdg_config c;
    mhdg_config_defaults(&c);
    c.gamma = k.sc.gamma; c.inver_MA2 = k.sc.inver_MA2; c.M_ref = k.sc.M_ref; c.M_H = k.sc.M_H; c.M_Na = k.sc.M_Na;
    c.ne_floor = k.sc.ne_floor; c.hall_coef = k.sc.hall_coef; c.inver_Rem = k.sc.inver_Rem;
    c.cfl = k.cfl; c.recon_rel_limit = k.recon_rel_limit; c.rho_floor = k.rho_floor; c.p_floor = k.p_floor;
    c.rho_floor_dt = k.rho_floor_dt;
    c.second_order = k.second_order; c.lorentz_force = k.lorentz_force; c.hall = k.hall; c.resistive = k.resistive;
    c.eta_r_inner = k.eta_r_inner; c.eta_r_outer = k.eta_r_outer; c.eta_width = k.eta_width;
    c.hall_r_min = k.hall_r_min; c.hall_r_max = k.hall_r_max;
    c.ind_rtol = k.ind_rtol; c.ind_max_it = k.ind_max_it; c.ind_restart = k.ind_restart;
    c.cuct = k.ideal.cuct;
    const uns::Vec3 uNa = k.u_Na();
    const double fsH[5] = {k.fsH.rho, k.u_inf.x, k.u_inf.y, k.u_inf.z, k.fsH.p};
    const double fsNa[5] = {k.fsNa.rho, uNa.x, uNa.y, uNa.z, k.fsNa.p};
    for (int m = 0; m < 5; ++m) { c.fs[0][m] = fsH[m]; c.fs[1][m] = fsNa[m]; }
    c.u_inf[0] = k.u_inf.x; c.u_inf[1] = k.u_inf.y; c.u_inf[2] = k.u_inf.z;
    c.B_imf[0] = k.B_imf.x; c.B_imf[1] = k.B_imf.y; c.B_imf[2] = k.B_imf.z;
    c.wall_dirichlet = 0; // the shim's solve takes the caller's Dirichlet array
    c.implicit_ideal = k.implicit_ideal;
    return c;
}

// What identifies the mesh a context was built from: the entity counts and the addresses of the arrays the
// device tables were copied from.  A complex destroyed and another created at the same address (drivers that
// new and delete their LocalMesh) differs in at least one of them unless it is the same allocation reused with
// the same sizes and buffers, in which case the arrays it describes are again the ones uploaded.
struct MeshKey
{
    int64_t n_cells = -1, n_faces = -1, n_edges = -1;
    const void *c2f = nullptr, *f2e = nullptr, *area = nullptr, *ec = nullptr;
    bool operator==(const MeshKey &o) const
    {
        return n_cells == o.n_cells && n_faces == o.n_faces && n_edges == o.n_edges && c2f == o.c2f && f2e == o.f2e && area == o.area &&
               ec == o.ec;
    }
};
MeshKey key_of(const uns::CellComplex &cx, const uns::Geometry *g)
{
    MeshKey k;
    k.n_cells = cx.n_cells(); k.n_faces = cx.n_faces(); k.n_edges = cx.n_edges();
    k.c2f = cx.c2f.idx.data(); k.f2e = cx.f2e.idx.data();
    if (g) { k.area = g->face_area.data(); k.ec = g->ec_cell.data(); }
    return k;
}

// One device context per mesh, with what it was built from
struct MeshRec
{
    const uns::LocalMesh *lm = nullptr;
    const uns::Geometry *geo = nullptr;
    const mhd::BoundaryInfo *bc = nullptr;
    std::unique_ptr<mhd::Config> cfg;
    mhdg_ctx *ctx = nullptr;
    MeshKey key; // of the mesh the context was built from
    bool built_with_bc = false, built_with_cfg = false;
    std::vector<char> no_kind, no_mode; // placeholders until the boundary info is known
    ~MeshRec() { if (ctx) mhdg_destroy(ctx); }
};
std::map<const uns::CellComplex *, std::unique_ptr<MeshRec>> g_mesh;

MeshRec &record(const uns::LocalMesh &lm, const uns::Geometry &geo)
{
    auto &r = g_mesh[&lm.mesh.cx];
    // a record left by a different mesh at this address is dropped with its context, bc and configuration
    if (!r || (r->ctx && !(r->key == key_of(lm.mesh.cx, &geo)))) r = std::make_unique<MeshRec>();
    r->lm = &lm;
    r->geo = &geo;
    return *r;
}

mhdg_ctx *context(MeshRec &r)
{
    const bool want_rebuild = r.ctx && ((r.bc && !r.built_with_bc) || (r.cfg && !r.built_with_cfg));
    if (r.ctx && !want_rebuild) return r.ctx;
    if (r.ctx) { mhdg_destroy(r.ctx); r.ctx = nullptr; }
    const uns::CellComplex &cx = r.lm->mesh.cx;
    const uns::Geometry &g = *r.geo;
    mhdg_mesh_desc d{};
    d.n_edges = cx.n_edges(); d.n_faces = cx.n_faces(); d.n_cells = cx.n_cells();
    d.f2e = inc(cx.f2e); d.c2f = inc(cx.c2f); d.e2f = inc(cx.e2f); d.f2c = inc(cx.f2c);
    d.cell_material = cx.cell_material.data();
    if (r.bc)
    {
        d.face_kind = reinterpret_cast<const int8_t *>(r.bc->face_kind.data());
        d.edge_mode = reinterpret_cast<const int8_t *>(r.bc->edge_mode.data());
    }
    else
    {
        r.no_kind.assign(cx.n_faces(), 0);
        r.no_mode.assign(cx.n_edges(), 0);
        d.face_kind = reinterpret_cast<const int8_t *>(r.no_kind.data());
        d.edge_mode = reinterpret_cast<const int8_t *>(r.no_mode.data());
    }
    d.face_area = d3(g.face_area); d.face_center = d3(g.face_center); d.face_area_mag = g.face_area_mag.data();
    d.cell_center = d3(g.cell_center); d.cell_volume = g.cell_volume.data();
    d.edge_vec = d3(g.edge_vec); d.edge_mid = d3(g.edge_mid); d.edge_len = g.edge_len.data();
    d.ec_ptr = g.ec_ptr.data(); d.ec_cell = g.ec_cell.data(); d.ec_weight = g.ec_weight.data(); d.ef_weight = g.ef_weight.data();
    d.star1 = g.star1.data(); d.star2 = g.star2.data();
    d.e2n = inc(cx.e2n); d.n_nodes = cx.n_nodes();
    mhdg_config c;
    if (r.cfg) c = to_abi(*r.cfg);
    else mhdg_config_defaults(&c);
    mhdg_ctx *ctx = nullptr;
    if (mhdg_create(&d, &c, &ctx) != 0) throw std::runtime_error(std::string("mhdg_create: ") + mhdg_last_error(nullptr));
    r.ctx = ctx;
    r.key = key_of(cx, r.geo);
    r.built_with_bc = r.bc != nullptr;
    r.built_with_cfg = r.cfg != nullptr;
    return ctx;
}

mhdg_ctx *context(const uns::CellComplex &cx)
{
    auto it = g_mesh.find(&cx);
    if (it == g_mesh.end() || !it->second->lm)
        throw std::runtime_error("mhd_gpu_shim: no mesh geometry registered for this complex (call a function taking LocalMesh first)");
    MeshRec &r = *it->second;
    // the calls that take only the complex cannot see the geometry: the counts and tables must still be the ones
    // the context was built from
    if (r.ctx)
    {
        MeshKey k = key_of(cx, nullptr), built = r.key;
        built.area = built.ec = nullptr;
        if (!(k == built))
            throw std::runtime_error("mhd_gpu_shim: this complex is not the mesh its context was built from (call a function taking LocalMesh first)");
    }
    return context(r);
}

// GhostFn -> per-face ghost table (5 primitives per face; zero rows for faces without a ghost).  The GhostFn sees
// the inside state the reference hands it: the cell value in the gradient call (grad == nullptr,
// Reconstruction.cpp:95-105), the linear reconstruction at the face centre with its positivity fallback in the flux
// call (FluxRHS.cpp:16-27, 57-65), restated here in the same evaluation order.
std::vector<double> ghost_table(const uns::LocalMesh &lm, const uns::Geometry &geo, const mhd::BoundaryInfo &bc, const double *W,
                                const double *grad, int species, const mhd::GhostFn &fn)
{
    const uns::CellComplex &cx = lm.mesh.cx;
    std::vector<double> t(5 * static_cast<size_t>(cx.n_faces()), 0.0);
    for (Id f = 0; f < cx.n_faces(); ++f)
    {
        const mhd::FaceKind k = bc.face_kind[f];
        if (k == mhd::FaceKind::Interior || k == mhd::FaceKind::Skip || cx.f2c.count(f) == 0) continue;
        // the fluid cell of the face and its outward normal
        Id c = cx.f2c.idx[cx.f2c.begin(f)];
        int sg = cx.f2c.sgn[cx.f2c.begin(f)];
        if (cx.f2c.count(f) == 2 && cx.cell_material[c] != 0) { c = cx.f2c.idx[cx.f2c.begin(f) + 1]; sg = cx.f2c.sgn[cx.f2c.begin(f) + 1]; }
        const double s = sg / geo.face_area_mag[f];
        const uns::Vec3 n(s * geo.face_area[f].x, s * geo.face_area[f].y, s * geo.face_area[f].z);
        const double *w = W + 5 * static_cast<size_t>(c);
        mhd::Prim in{w[0], w[1], w[2], w[3], w[4]};
        if (grad)
        {
            const double *gr = grad + 15 * static_cast<size_t>(c);
            const uns::Vec3 dx = geo.face_center[f] - geo.cell_center[c];
            mhd::Prim r;
            r.rho = w[0] + gr[0] * dx.x + gr[1] * dx.y + gr[2] * dx.z;
            r.u = w[1] + gr[3] * dx.x + gr[4] * dx.y + gr[5] * dx.z;
            r.v = w[2] + gr[6] * dx.x + gr[7] * dx.y + gr[8] * dx.z;
            r.w = w[3] + gr[9] * dx.x + gr[10] * dx.y + gr[11] * dx.z;
            r.p = w[4] + gr[12] * dx.x + gr[13] * dx.y + gr[14] * dx.z;
            if (r.rho > 0 && r.p > 0) in = r;
        }
        const mhd::Prim g = fn(f, species, in, n);
        double *q = &t[5 * static_cast<size_t>(f)];
        q[0] = g.rho; q[1] = g.u; q[2] = g.v; q[3] = g.w; q[4] = g.p;
    }
    return t;
}

int solve_precision()
{
    const char *p = std::getenv("MHDG_SOLVE_PRECISION");
    const int v = p ? std::atoi(p) : 0;
    return (v >= 1 && v <= 3) ? v : 0;
}

// ImplicitInduction instances -> a context of their own (each instance keeps its own warm-start correction,
// as the reference keeps corr_ per instance)
std::map<const mhd::ImplicitInduction *, std::unique_ptr<MeshRec>> g_ind;
MeshRec &ind_rec(const mhd::ImplicitInduction *self)
{
    auto it = g_ind.find(self);
    if (it == g_ind.end()) throw std::runtime_error("mhd_gpu_shim: unknown ImplicitInduction");
    return *it->second;
}
} // namespace

namespace mhd
{
// Reconstruction.cpp:9-24
Prim ghost_state(FaceKind kind, int species, const Prim &in, const uns::Vec3 &n, const Prim &fs)
{
    if (kind == FaceKind::Wall)
    {
        if (species == 1) return Prim{1e-12, 0, 0, 0, 1e-12};
        const double un = in.u * n.x + in.v * n.y + in.w * n.z;
        Prim g = in;
        g.u -= 2 * un * n.x; g.v -= 2 * un * n.y; g.w -= 2 * un * n.z;
        return g;
    }
    if (kind == FaceKind::Inflow) return fs;
    return in;
}

// Reconstruction.cpp:26-78: the stencil lives on the device (built with the context)
Stencil::Stencil(const uns::LocalMesh &lm, const uns::Geometry &geo, const BoundaryInfo &bc)
{
    MeshRec &r = record(lm, geo);
    r.bc = &bc;
}

// Reconstruction.cpp:80-142
void Stencil::limited_gradients(const uns::LocalMesh &lm, const uns::Geometry &geo, const BoundaryInfo &bc, const double *W, int species,
                                const Prim &fs, bool limit, double *grad, const GhostFn *ghost) const
{
    MeshRec &r = record(lm, geo);
    r.bc = &bc;
    mhdg_ctx *ctx = context(r);
    ok(ctx, mhdg_set_recon_rel_limit(ctx, rel_limit_), "Stencil::set_rel_limit");
    const double f5[5] = {fs.rho, fs.u, fs.v, fs.w, fs.p};
    std::vector<double> gt;
    if (ghost && *ghost) gt = ghost_table(lm, geo, bc, W, nullptr, species, *ghost);
    ok(ctx, mhdg_limited_gradients(ctx, W, species, f5, limit ? 1 : 0, gt.empty() ? nullptr : gt.data(), grad), "limited_gradients");
}

// FluxRHS.cpp:30-74
void fluid_flux_rhs(const uns::LocalMesh &lm, const uns::Geometry &geo, const BoundaryInfo &bc, const Config &cfg, int species,
                    const Prim &fs, const double *W, const double *grad, double *rhs, const GhostFn *ghost)
{
    MeshRec &r = record(lm, geo);
    r.bc = &bc;
    if (!r.cfg) r.cfg = std::make_unique<Config>(cfg);
    mhdg_ctx *ctx = context(r);
    const double f5[5] = {fs.rho, fs.u, fs.v, fs.w, fs.p};
    std::vector<double> gt;
    if (ghost && *ghost) gt = ghost_table(lm, geo, bc, W, grad, species, *ghost);
    ok(ctx, mhdg_flux_rhs(ctx, species, f5, W, grad, gt.empty() ? nullptr : gt.data(), rhs), "fluid_flux_rhs");
}

// Current.cpp:9-31
void cell_current(const uns::LocalMesh &lm, const uns::Geometry &geo, const double *B, double *J)
{
    mhdg_ctx *ctx = context(record(lm, geo));
    ok(ctx, mhdg_cell_current(ctx, B, J), "cell_current");
}

// IdealEMF.cpp:25-129 (adds into E)
void add_ideal_emf(const uns::LocalMesh &lm, const uns::Geometry &geo, const IdealInputs &in, const IdealParams &prm, double *E)
{
    mhdg_ctx *ctx = context(record(lm, geo));
    ok(ctx, mhdg_ideal_emf(ctx, in.u_cell, in.B_cell_tot, in.B_cell_ind, in.phi_tot, in.phi_ind, prm.cuct, E), "add_ideal_emf");
}

// Faraday.cpp:7-36: on one rank every cell is owned, so every face with a cell is active
std::vector<char> active_faces(const uns::CellComplex &cx)
{
    std::vector<char> a(cx.n_faces(), 0);
    for (Id f = 0; f < cx.n_faces(); ++f) a[f] = cx.f2c.count(f) > 0;
    return a;
}
// The device updates every face; the faces the caller's mask leaves inactive keep their values, as in the reference
// (the drivers pass active_faces, under which no face with a cell is inactive on one rank).
namespace
{
template <class F> void masked_update(const uns::CellComplex &cx, const std::vector<char> &active, double *B, F &&update)
{
    if (active.size() != static_cast<size_t>(cx.n_faces())) throw std::runtime_error("faraday_update: active mask of the wrong size");
    std::vector<std::pair<Id, double>> keep;
    for (Id f = 0; f < cx.n_faces(); ++f)
        if (!active[f]) keep.emplace_back(f, B[f]);
    update();
    for (const auto &[f, v] : keep) B[f] = v;
}
} // namespace
void faraday_update(const uns::CellComplex &cx, const std::vector<char> &active, const double *E, double *B, double dt)
{
    mhdg_ctx *ctx = context(cx);
    masked_update(cx, active, B, [&] { ok(ctx, mhdg_faraday(ctx, E, B, dt), "faraday_update"); });
}
void faraday_update_edge_dt(const uns::CellComplex &cx, const std::vector<char> &active, const double *E, const double *dt_edge, double *B)
{
    mhdg_ctx *ctx = context(cx);
    masked_update(cx, active, B, [&] { ok(ctx, mhdg_faraday_edge_dt(ctx, E, dt_edge, B), "faraday_update_edge_dt"); });
}

// DivB.cpp:8-24
DivBReport div_b(const uns::CellComplex &cx, const double *B, const double *scale_extra)
{
    mhdg_ctx *ctx = context(cx);
    DivBReport rep;
    int64_t worst = -1;
    ok(ctx, mhdg_div_b(ctx, B, scale_extra, &rep.max_abs, &rep.max_rel, &worst), "div_b");
    rep.worst_cell_gid = worst < 0 ? -1 : static_cast<long>(cx.ent[3].gid[worst]);
    return rep;
}

// ImplicitInduction.cpp:10-103: the tables are built on the device with the context
ImplicitInduction::ImplicitInduction(const uns::LocalMesh &lm, const uns::Geometry &geo, uns::comm::Exchanger &ex, const Config &cfg)
    : lm_(lm), geo_(geo), ex_(ex), resistive_(cfg.resistive && cfg.sc.inver_Rem > 0), hall_(cfg.hall && cfg.sc.hall_coef > 0),
      ideal_(cfg.implicit_ideal), ideal_prm_(cfg.ideal), rtol_(cfg.ind_rtol), max_it_(cfg.ind_max_it), restart_(std::max(2, cfg.ind_restart))
{
    // the block-Jacobi ILU(0) preconditioner (ImplicitInduction.cpp:209-295) is not on the device: refuse it rather than
    // solve with Jacobi and a different iteration history
    if (cfg.ind_ilu) throw std::runtime_error("mhd_gpu_shim: induction_ilu is not supported by the device solve (Jacobi only)");
    auto r = std::make_unique<MeshRec>();
    r->lm = &lm;
    r->geo = &geo;
    r->cfg = std::make_unique<Config>(cfg);
    // an instance constructed at the address of a destroyed one replaces its record (and releases its context);
    // the reference class declares no destructor for the shim to hook, so records of destroyed instances whose
    // address is not reused stay until the process ends
    g_ind[this] = std::move(r);
    MeshRec &m = record(lm, geo); // the mesh's free functions use the same configuration
    if (!m.cfg) m.cfg = std::make_unique<Config>(cfg);
}
// ImplicitInduction.cpp:105-111 (alpha from taper, hall_coef and n_e,eff on the device side)
void ImplicitInduction::set_frozen_state(const double *B0_cell, const double *ne_eff_cell)
{
    mhdg_ctx *ctx = context(ind_rec(this));
    ok(ctx, mhdg_ind_set_frozen(ctx, B0_cell, ne_eff_cell), "set_frozen_state");
}
// ImplicitInduction.cpp:113-117
void ImplicitInduction::set_ideal_state(const double *u_cell, const std::vector<char> &mask)
{
    mhdg_ctx *ctx = context(ind_rec(this));
    ok(ctx, mhdg_ind_set_ideal_state(ctx, u_cell, reinterpret_cast<const int8_t *>(mask.data())), "set_ideal_state");
}
// ImplicitInduction.cpp:119-125
void ImplicitInduction::ideal_linear(const double *phi, double *edge_out)
{
    mhdg_ctx *ctx = context(ind_rec(this));
    ok(ctx, mhdg_ind_ideal_linear(ctx, phi, edge_out), "ideal_linear");
}
// ImplicitInduction.cpp:127-183
void ImplicitInduction::apply_F(const double *phi, double *edge_out)
{
    mhdg_ctx *ctx = context(ind_rec(this));
    ok(ctx, mhdg_ind_apply_F(ctx, phi, edge_out), "apply_F");
}
// ImplicitInduction.cpp:299-390
int ImplicitInduction::solve(const double *Phi_n, const double *E_expl, const std::vector<char> &dir, double dt, double *E, double *rel_res,
                             const double *dt_edge)
{
    mhdg_ctx *ctx = context(ind_rec(this));
    int32_t it = 0;
    double rr = 0;
    ok(ctx, mhdg_ind_solve(ctx, Phi_n, E_expl, reinterpret_cast<const int8_t *>(dir.data()), dt, dt_edge, solve_precision(), E, &it, &rr),
       "ImplicitInduction::solve");
    if (rel_res) *rel_res = rr;
    return it;
}
} // namespace mhd
