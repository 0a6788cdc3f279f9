// ORACLE — test infrastructure only.  C-ABI harness around the compiled reference (MPCNS_Uns sources
// built in place from /root/reference by oracle/Makefile into oracle/_ref/).  It constructs the
// reference's own objects (ComplexBuilder, compute_geometry, BoundaryInfo, Stencil, ImplicitInduction,
// mercury::Solver) on a mesh given as plain arrays and calls them, one rank (MPI_COMM_SELF).
// Used only by tests/ to pin the oracle restatement and the host setup against the reference.
#include <mpi.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "boundary/BoundaryInfo.h"
#include "config/Config.h"
#include "diagnostics/DivB.h"
#include "driver/Solver.h"
#include "fluid/FluxRHS.h"
#include "fluid/MercurySources.h"
#include "fluid/Reconstruction.h"
#include "induction/Current.h"
#include "induction/EdgeEMF.h"
#include "induction/Faraday.h"
#include "induction/IdealEMF.h"
#include "induction/ImplicitInduction.h"
#include "initial/Initial.h"
#include "uns/comm/Exchanger.h"
#include "uns/complex/ComplexBuilder.h"
#include "uns/complex/Validate.h"
#include "uns/geometry/Geometry.h"
#include "uns/geometry/Reconstruct.h"
#include "uns/io/MeshIO.h"
#include "uns/io/VtkWriter.h"
#include "uns/partition/Partitioner.h"
#include "MarsUns/physics/MarsDipoles.h"

#include <cmath>
#include <iomanip>
#include <sstream>

#include "../include/mhdg.h"
#include "ref_config.h"

using namespace uns;
using namespace mhd;
using refh::build_mesh;
using refh::make_config;

struct ref_ctx
{
    LocalMesh lm;
    Geometry geo;
    Config cfg;
    std::unique_ptr<comm::Exchanger> ex;
    std::unique_ptr<BoundaryInfo> bc;
    std::unique_ptr<Stencil> st;
    std::unique_ptr<ImplicitInduction> ind;
    std::unique_ptr<mercury::Solver> solver;
    std::vector<char> active;
    std::vector<double> PhiAdd, BaddCell;
    std::vector<double> u_ideal; // set_ideal_state keeps the pointer: the harness owns the array
    std::string err;
};

namespace
{
Prim prim5(const double *p) { return Prim{p[0], p[1], p[2], p[3], p[4]}; }
template <class T> void put(const std::vector<T> &v, T *out) { std::memcpy(out, v.data(), v.size() * sizeof(T)); }
void put3(const std::vector<Vec3> &v, double *out)
{
    for (size_t i = 0; i < v.size(); ++i) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; }
}
} // namespace

extern "C" {
// cells: cell c has loops [cell_ptr[c], cell_ptr[c+1]) ; loop l has nodes loop_nodes[loop_ptr[l] .. loop_ptr[l+1])
// tags: tag t has nodes tag_nodes[tag_ptr[t] .. tag_ptr[t+1]) and id tag_id[t]; tag names "inner" = 1, "outer" = 2, box faces 1..6
int ref_create(int32_t n_nodes, const double *xyz, int64_t n_cells, const int32_t *cell_ptr, const int32_t *loop_ptr,
               const int32_t *loop_nodes, const int32_t *material, int64_t n_tags, const int32_t *tag_ptr, const int32_t *tag_nodes,
               const int32_t *tag_id, int tagset, const mhdg_config *cfg, int dipole, const double *dm, const double *dc, double soft,
               ref_ctx **out)
{
    try
    {
        int init = 0;
        MPI_Initialized(&init);
        if (!init) MPI_Init(nullptr, nullptr);
        auto *r = new ref_ctx();
        r->lm.mesh = build_mesh(n_nodes, xyz, n_cells, cell_ptr, loop_ptr, loop_nodes, material, n_tags, tag_ptr, tag_nodes, tag_id, tagset);
        r->cfg = make_config(*cfg, dipole, dm, dc, soft);
        r->geo = compute_geometry(r->lm.mesh);
        r->ex = std::make_unique<comm::Exchanger>(MPI_COMM_SELF, r->lm);
        r->bc = std::make_unique<BoundaryInfo>(BoundaryInfo::build(r->lm, r->geo, r->cfg));
        r->st = std::make_unique<Stencil>(r->lm, r->geo, *r->bc);
        r->st->set_rel_limit(r->cfg.recon_rel_limit);
        r->ind = std::make_unique<ImplicitInduction>(r->lm, r->geo, *r->ex, r->cfg);
        r->active = active_faces(r->lm.mesh.cx);
        mercury::set_background(r->lm, r->geo, r->cfg, r->PhiAdd, r->BaddCell);
        *out = r;
        return 0;
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "ref_create: %s\n", e.what());
        return 1;
    }
}
void ref_destroy(ref_ctx *r) { delete r; }

// counts: nodes, edges, faces, cells, nnz(e2n), nnz(f2e), nnz(c2f), nnz(ec)
void ref_counts(ref_ctx *r, int64_t *n)
{
    const CellComplex &cx = r->lm.mesh.cx;
    n[0] = cx.n_nodes(); n[1] = cx.n_edges(); n[2] = cx.n_faces(); n[3] = cx.n_cells();
    n[4] = cx.e2n.nnz(); n[5] = cx.f2e.nnz(); n[6] = cx.c2f.nnz(); n[7] = static_cast<int64_t>(r->geo.ec_cell.size());
}
// which: 0 e2n 1 f2e 2 c2f 3 e2f 4 f2c
void ref_incidence(ref_ctx *r, int which, int64_t *ptr, int32_t *idx, int8_t *sgn)
{
    const CellComplex &cx = r->lm.mesh.cx;
    const Incidence *t[5] = {&cx.e2n, &cx.f2e, &cx.c2f, &cx.e2f, &cx.f2c};
    put(t[which]->ptr, ptr); put(t[which]->idx, idx); put(t[which]->sgn, sgn);
}
void ref_face_tags(ref_ctx *r, int32_t *tags) { put(r->lm.mesh.cx.face_tag, tags); }
// topology report: euler, boundary faces, d1d0 violations, d2d1 violations, ok
void ref_validate(ref_ctx *r, int64_t *out)
{
    const TopologyReport t = validate_topology(r->lm.mesh.cx);
    out[0] = t.euler; out[1] = t.boundary_faces; out[2] = t.d1d0_violations; out[3] = t.d2d1_violations; out[4] = t.ok;
}
// geometry field by name
int ref_geometry(ref_ctx *r, const char *name, void *out)
{
    const Geometry &g = r->geo;
    const std::string n(name);
    if (n == "edge_vec") put3(g.edge_vec, (double *)out);
    else if (n == "edge_mid") put3(g.edge_mid, (double *)out);
    else if (n == "edge_len") put(g.edge_len, (double *)out);
    else if (n == "face_area") put3(g.face_area, (double *)out);
    else if (n == "face_center") put3(g.face_center, (double *)out);
    else if (n == "face_area_mag") put(g.face_area_mag, (double *)out);
    else if (n == "cell_center") put3(g.cell_center, (double *)out);
    else if (n == "cell_volume") put(g.cell_volume, (double *)out);
    else if (n == "edge_dual_area") put3(g.edge_dual_area, (double *)out);
    else if (n == "ec_ptr") put(g.ec_ptr, (int64_t *)out);
    else if (n == "ec_cell") put(g.ec_cell, (int32_t *)out);
    else if (n == "ec_weight") put(g.ec_weight, (double *)out);
    else if (n == "ef_weight") put(g.ef_weight, (double *)out);
    else if (n == "star1") put(g.star1, (double *)out);
    else if (n == "star2") put(g.star2, (double *)out);
    else if (n == "star3") put(g.star3, (double *)out);
    else return 1;
    return 0;
}
void ref_boundary(ref_ctx *r, int8_t *face_kind, int8_t *edge_mode)
{
    for (size_t f = 0; f < r->bc->face_kind.size(); ++f) face_kind[f] = static_cast<int8_t>(r->bc->face_kind[f]);
    for (size_t e = 0; e < r->bc->edge_mode.size(); ++e) edge_mode[e] = static_cast<int8_t>(r->bc->edge_mode[e]);
}
void ref_background(ref_ctx *r, double *PhiAdd, double *BaddCell) { put(r->PhiAdd, PhiAdd); put(r->BaddCell, BaddCell); }

void ref_cell_vector(ref_ctx *r, const double *flux, double *b)
{
    const CellComplex &cx = r->lm.mesh.cx;
    put3(reconstruct_cell_vector(cx, r->geo, flux, cx.n_cells()), b);
}
void ref_cell_current(ref_ctx *r, const double *B, double *J) { cell_current(r->lm, r->geo, B, J); }
void ref_limited_gradients(ref_ctx *r, const double *W, int species, const double *fs, int limit, double *grad)
{
    r->st->limited_gradients(r->lm, r->geo, *r->bc, W, species, prim5(fs), limit != 0, grad);
}
void ref_flux_rhs(ref_ctx *r, int species, const double *fs, const double *W, const double *grad, double *rhs)
{
    fluid_flux_rhs(r->lm, r->geo, *r->bc, r->cfg, species, prim5(fs), W, grad, rhs);
}
// The same two calls with a GhostFn (Reconstruction.h:17-21) that returns a per-face table of 5 primitives: the
// form of MarsSolver's cached wall state (MarsSolver.cpp:124-128, 148-167).  The callback also records the
// largest |inside| it was handed, so a test can see that the call went through the GhostFn branch.
static GhostFn table_ghost(const double *table, double *seen)
{
    return [table, seen](Id f, int, const Prim &in, const Vec3 &) {
        if (seen) *seen = std::max(*seen, std::abs(in.rho) + std::abs(in.p));
        const double *g = table + 5 * static_cast<std::size_t>(f);
        return Prim{g[0], g[1], g[2], g[3], g[4]};
    };
}
void ref_limited_gradients_ghost(ref_ctx *r, const double *W, int species, const double *fs, int limit, const double *table, double *grad,
                                 double *seen)
{
    const GhostFn g = table_ghost(table, seen);
    r->st->limited_gradients(r->lm, r->geo, *r->bc, W, species, prim5(fs), limit != 0, grad, &g);
}
void ref_flux_rhs_ghost(ref_ctx *r, int species, const double *fs, const double *W, const double *grad, const double *table, double *rhs,
                        double *seen)
{
    const GhostFn g = table_ghost(table, seen);
    fluid_flux_rhs(r->lm, r->geo, *r->bc, r->cfg, species, prim5(fs), W, grad, rhs, &g);
}
// The same two calls with a GhostFn of MarsSolver's form (MarsSolver.cpp:124-128): the cached per-face state on wall
// faces, ghost_state(kind, species, inside, n, fs) on every other boundary face, so the ghost of an outflow face is the
// inside state the reference hands the GhostFn (the cell value for the gradients, the reconstructed face state for the
// fluxes).
static GhostFn wall_table_ghost(const ref_ctx *r, const double *table, const Prim &fs)
{
    return [r, table, fs](Id f, int species, const Prim &in, const Vec3 &n) {
        const FaceKind k = r->bc->face_kind[f];
        if (k == FaceKind::Wall)
        {
            const double *g = table + 5 * static_cast<std::size_t>(f);
            return Prim{g[0], g[1], g[2], g[3], g[4]};
        }
        return ghost_state(k, species, in, n, fs);
    };
}
void ref_limited_gradients_wall_ghost(ref_ctx *r, const double *W, int species, const double *fs, int limit, const double *table,
                                      double *grad)
{
    const GhostFn g = wall_table_ghost(r, table, prim5(fs));
    r->st->limited_gradients(r->lm, r->geo, *r->bc, W, species, prim5(fs), limit != 0, grad, &g);
}
void ref_flux_rhs_wall_ghost(ref_ctx *r, int species, const double *fs, const double *W, const double *grad, const double *table,
                             double *rhs)
{
    const GhostFn g = wall_table_ghost(r, table, prim5(fs));
    fluid_flux_rhs(r->lm, r->geo, *r->bc, r->cfg, species, prim5(fs), W, grad, rhs, &g);
}
void ref_ideal_emf(ref_ctx *r, const double *u, const double *Bt, const double *Bi, const double *pt, const double *pi, double cuct, double *E)
{
    IdealInputs in;
    in.u_cell = u; in.B_cell_tot = Bt; in.B_cell_ind = Bi; in.phi_tot = pt; in.phi_ind = pi;
    IdealParams p;
    p.cuct = cuct;
    add_ideal_emf(r->lm, r->geo, in, p, E);
}
void ref_faraday(ref_ctx *r, const double *E, double *Phi, double dt) { faraday_update(r->lm.mesh.cx, r->active, E, Phi, dt); }
// faraday_update with the caller's face mask (nonzero = active) in place of active_faces
void ref_faraday_masked(ref_ctx *r, const double *E, double *Phi, double dt, const int8_t *mask)
{
    const std::vector<char> a(mask, mask + r->lm.mesh.cx.n_faces());
    faraday_update(r->lm.mesh.cx, a, E, Phi, dt);
}
void ref_faraday_edge_dt(ref_ctx *r, const double *E, const double *dt_edge, double *Phi)
{
    faraday_update_edge_dt(r->lm.mesh.cx, r->active, E, dt_edge, Phi);
}
void ref_div_b(ref_ctx *r, const double *B, double *max_abs, double *max_rel, int64_t *worst)
{
    const DivBReport d = div_b(r->lm.mesh.cx, B);
    *max_abs = d.max_abs; *max_rel = d.max_rel; *worst = d.worst_cell_gid;
}
double ref_local_dt(ref_ctx *r, const double *WH, const double *WNa, const double *B) { return mercury::local_time_step(r->lm, r->geo, r->cfg, WH, WNa, B); }
void ref_add_sources(ref_ctx *r, int s, const double *WH, const double *WNa, const double *up, const double *B, const double *J, double dt, double *rhs)
{
    mercury::SourceFields sf;
    mercury::add_sources(r->lm, r->cfg, s, WH, WNa, up, B, J, sf, dt, rhs);
}
void ref_ind_set_frozen(ref_ctx *r, const double *B0, const double *ne) { r->ind->set_frozen_state(B0, ne); }
void ref_ind_apply_F(ref_ctx *r, const double *phi, double *out) { r->ind->apply_F(phi, out); }
int ref_ind_solve(ref_ctx *r, const double *Phi_n, const double *E_expl, const int8_t *dir, double dt, const double *dte, double *E,
                  double *rel_res)
{
    std::vector<char> d(dir, dir + r->lm.mesh.cx.n_edges());
    return r->ind->solve(Phi_n, E_expl, d, dt, E, rel_res, dte);
}
void ref_ind_set_ideal_state(ref_ctx *r, const double *u, const int8_t *mask)
{
    r->u_ideal.assign(u, u + 3 * size_t(r->lm.mesh.cx.n_cells()));
    std::vector<char> m(mask, mask + r->lm.mesh.cx.n_edges());
    r->ind->set_ideal_state(r->u_ideal.data(), m);
}
void ref_ind_ideal_linear(ref_ctx *r, const double *phi, double *out) { r->ind->ideal_linear(phi, out); }

// output directory of the solver's VTK files (before ref_solver_init) and Solver::write_vtk
void ref_set_output_dir(ref_ctx *r, const char *dir) { r->cfg.output_dir = dir; }
int ref_write_vtk(ref_ctx *r)
{
    try { r->solver->write_vtk(); return 0; }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_write_vtk: %s\n", e.what()); return 1; }
}
// ---- full step through mercury::Solver (MercuryUns/driver/Solver.cpp:228-277)
int ref_solver_init(ref_ctx *r)
{
    try
    {
        r->solver = std::make_unique<mercury::Solver>(MPI_COMM_SELF, r->lm, r->cfg);
        r->solver->initialize();
        return 0;
    }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_solver_init: %s\n", e.what()); return 1; }
}
void ref_set_state(ref_ctx *r, const double *UH, const double *UNa, const double *Phi)
{
    auto &s = r->solver->state();
    std::memcpy(s.U[0].data(), UH, s.U[0].size() * 8);
    std::memcpy(s.U[1].data(), UNa, s.U[1].size() * 8);
    std::memcpy(s.Phi.data(), Phi, s.Phi.size() * 8);
}
void ref_get_state(ref_ctx *r, double *UH, double *UNa, double *Phi)
{
    const auto &s = r->solver->state();
    put(s.U[0], UH); put(s.U[1], UNa); put(s.Phi, Phi);
}
int ref_step(ref_ctx *r, double dt_max, double *dt)
{
    try { *dt = r->solver->step_first_order(dt_max); return 0; }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_step: %s\n", e.what()); return 1; }
}
// 0 W_H 1 W_Na 2 B_ind 3 B_tot 4 J
void ref_get_derived(ref_ctx *r, int which, double *out)
{
    switch (which)
    {
    case 0: put(r->solver->prim(0), out); break;
    case 1: put(r->solver->prim(1), out); break;
    case 2: put(r->solver->B_induced_cell(), out); break;
    case 3: put(r->solver->B_total_cell(), out); break;
    case 4: put(r->solver->current_cell(), out); break;
    }
}
double ref_last_induction_residual(ref_ctx *r) { return r->solver->last_induction_residual(); }

// uns::io (src/uns/io/MeshIO.cpp): the reference's complex and field files
int ref_write_mesh(ref_ctx *r, const char *path)
{
    try { uns::io::write_local_mesh(path, r->lm); return 0; }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_write_mesh: %s\n", e.what()); return 1; }
}
// reads a complex file (the reader verifies the stored hash); counts of nodes, edges, faces, cells and the hash
int ref_read_mesh(const char *path, int64_t counts[4], uint64_t *hash)
{
    try
    {
        const LocalMesh lm = uns::io::read_local_mesh(path);
        for (int d = 0; d < 4; ++d) counts[d] = lm.mesh.cx.count(static_cast<Dim>(d));
        *hash = topology_hash(lm.mesh.cx);
        return 0;
    }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_read_mesh: %s\n", e.what()); return 1; }
}
uint64_t ref_topology_hash(ref_ctx *r) { return topology_hash(r->lm.mesh.cx); }
// uns::io VTK writers (src/uns/io/VtkWriter.cpp): arrays as (name, ncomp, values of `counts` doubles)
static std::vector<uns::io::VtkField> vtk_fields(int n, const char *const *names, const int32_t *ncomp, const double *const *values,
                                                 const int64_t *counts)
{
    std::vector<uns::io::VtkField> f(size_t(n > 0 ? n : 0));
    for (int i = 0; i < n; ++i)
    {
        f[size_t(i)].name = names[i];
        f[size_t(i)].ncomp = ncomp[i];
        if (values && values[i]) f[size_t(i)].values.assign(values[i], values[i] + counts[i]);
    }
    return f;
}
int ref_write_vtu(ref_ctx *r, const char *path, int nc, const char *const *cn, const int32_t *cc, const double *const *cv, const int64_t *ccount,
                  int np, const char *const *pn, const int32_t *pc, const double *const *pv, const int64_t *pcount)
{
    try
    {
        uns::io::write_vtu(path, r->lm.mesh, r->lm.mesh.cx.count(Dim::Cell), vtk_fields(nc, cn, cc, cv, ccount), vtk_fields(np, pn, pc, pv, pcount));
        return 0;
    }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_write_vtu: %s\n", e.what()); return 1; }
}
int ref_write_pvtu(const char *path, int n, const char *const *pieces, int nc, const char *const *cn, const int32_t *cc, int np,
                   const char *const *pn, const int32_t *pc)
{
    try
    {
        uns::io::write_pvtu(path, std::vector<std::string>(pieces, pieces + n), vtk_fields(nc, cn, cc, nullptr, nullptr),
                            vtk_fields(np, pn, pc, nullptr, nullptr));
        return 0;
    }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_write_pvtu: %s\n", e.what()); return 1; }
}
int ref_write_pvd(const char *path, int n, const double *times, const char *const *files)
{
    try
    {
        std::vector<std::pair<double, std::string>> e;
        for (int i = 0; i < n; ++i) e.emplace_back(times[i], files[i]);
        uns::io::write_pvd(path, e);
        return 0;
    }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_write_pvd: %s\n", e.what()); return 1; }
}
int ref_write_fields(const char *path, uint64_t hash, int n, const char *const *names, const int32_t *dims, const int32_t *ncomps,
                     const double *const *values, const int64_t *counts)
{
    try
    {
        std::vector<uns::io::FieldRecord> f(n);
        for (int i = 0; i < n; ++i)
        {
            f[i].name = names[i];
            f[i].dim = static_cast<Dim>(dims[i]);
            f[i].ncomp = ncomps[i];
            f[i].values.assign(values[i], values[i] + counts[i]);
        }
        uns::io::write_fields(path, hash, f);
        return 0;
    }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_write_fields: %s\n", e.what()); return 1; }
}
// the values of field `name` (count in/out: capacity / size), its dim and ncomp, and the file's hash
int ref_read_field(const char *path, const char *name, double *out, int64_t *count, int32_t *dim, int32_t *ncomp, uint64_t *hash)
{
    try
    {
        for (const auto &f : uns::io::read_fields(path, hash))
            if (f.name == name)
            {
                if (static_cast<int64_t>(f.values.size()) > *count) return 2;
                std::memcpy(out, f.values.data(), f.values.size() * 8);
                *count = static_cast<int64_t>(f.values.size());
                *dim = static_cast<int32_t>(f.dim);
                *ncomp = f.ncomp;
                return 0;
            }
        return 3;
    }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_read_field: %s\n", e.what()); return 1; }
}

// uns::partition_cells with the Hilbert method (src/uns/partition/Partitioner.cpp:37-108, 188-211) on n unit-weight
// cells at the given centres, cut into n parts without refinement: the part of each cell is its position in the
// Hilbert order, so order[part[c]] = c.
int ref_hilbert_order(int32_t n, const double *centres, int32_t *order)
{
    try
    {
        CellGraph g;
        g.xadj.assign(size_t(n) + 1, 0);
        g.vw.assign(size_t(n), 1.0);
        g.centre.resize(size_t(n));
        for (int32_t i = 0; i < n; ++i) g.centre[size_t(i)] = {centres[3 * size_t(i)], centres[3 * size_t(i) + 1], centres[3 * size_t(i) + 2]};
        PartOptions opt;
        opt.method = PartMethod::Hilbert;
        opt.refine_passes = 0;
        const std::vector<int> part = partition_cells(g, n, opt);
        std::vector<char> hit(size_t(n), 0);
        for (int32_t c = 0; c < n; ++c)
        {
            if (part[size_t(c)] < 0 || part[size_t(c)] >= n || hit[size_t(part[size_t(c)])]) return 2;
            hit[size_t(part[size_t(c)])] = 1;
            order[part[size_t(c)]] = c;
        }
        return 0;
    }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_hilbert_order: %s\n", e.what()); return 1; }
}

// MARS::DipoleModel (MarsUns/physics/MarsDipoles.cpp): the reference's crustal-dipole edge integrals, the model read
// from its own text format with unit reference scales (L = B = 1, mu0 = 4 pi: no rescaling)
int ref_dipole_edge_integrals(int n, const double *pos, const double *mom, const double *core, double cutoff, int64_t ne, const double *xa,
                              const double *xb, double *out)
{
    try
    {
        std::ostringstream s;
        s << std::setprecision(17) << "MARS_DIPOLES_V1 " << n << "\n";
        for (int i = 0; i < n; ++i)
            s << pos[3 * i] << " " << pos[3 * i + 1] << " " << pos[3 * i + 2] << " " << mom[3 * i] << " " << mom[3 * i + 1] << " "
              << mom[3 * i + 2] << " " << core[i] << "\n";
        MARS::DipoleModel model;
        std::istringstream in(s.str());
        model.Read(in, 1.0, 1.0, 4.0 * std::acos(-1.0), "harness");
        model.SetCutoff(cutoff);
        for (int64_t e = 0; e < ne; ++e)
            out[e] = model.EdgeIntegral({xa[3 * e], xa[3 * e + 1], xa[3 * e + 2]}, {xb[3 * e], xb[3 * e + 1], xb[3 * e + 2]});
        return 0;
    }
    catch (const std::exception &e) { std::fprintf(stderr, "ref_dipole_edge_integrals: %s\n", e.what()); return 1; }
}
}
