// ORACLE — test infrastructure only.  Helpers shared by the reference harness (ref_harness.cpp) and the MPI
// baseline program (ref_mpi_bench.cpp): the reference's Config from the C-ABI mhdg_config, and the reference's
// global Mesh built through its own ComplexBuilder from cells given as outward node loops.
#pragma once
#include <cstdint>
#include <vector>

#include "config/Config.h"
#include "uns/complex/ComplexBuilder.h"

#include "../include/mhdg.h"

namespace refh
{
using namespace uns;
using namespace mhd;

inline Config make_config(const mhdg_config &c, int dipole, const double *m, const double *cc, double soft)
{
    Config k;
    k.sc = Scales::simple(c.gamma, c.inver_MA2, c.M_Na / c.M_H);
    k.sc.M_ref = c.M_ref; k.sc.M_H = c.M_H; k.sc.M_Na = c.M_Na;
    k.sc.ne_floor = c.ne_floor;
    k.sc.hall_coef = c.hall_coef;
    k.sc.induce_coeff = c.induce_coeff;
    k.sc.inver_Rem = c.inver_Rem;
    k.cfl = c.cfl;
    k.second_order = c.second_order != 0;
    k.recon_rel_limit = c.recon_rel_limit;
    k.rho_floor = c.rho_floor; k.p_floor = c.p_floor; k.rho_floor_dt = c.rho_floor_dt;
    k.lorentz_force = c.lorentz_force != 0;
    k.time_order = 1;
    k.hall = c.hall != 0;
    k.resistive = c.resistive != 0;
    k.eta_r_inner = c.eta_r_inner; k.eta_r_outer = c.eta_r_outer; k.eta_width = c.eta_width;
    k.hall_r_min = c.hall_r_min; k.hall_r_max = c.hall_r_max;
    k.ind_rtol = c.ind_rtol; k.ind_max_it = c.ind_max_it; k.ind_restart = c.ind_restart;
    k.ind_ilu = false;
    k.ideal.cuct = c.cuct;
    k.implicit_ideal = c.implicit_ideal != 0;
    k.fsH = {c.fs[0][0], c.fs[0][4]};
    k.fsNa = {c.fs[1][0], c.fs[1][4]};
    k.u_inf = {c.u_inf[0], c.u_inf[1], c.u_inf[2]};
    k.na_seed_at_rest = !(c.fs[1][1] == c.u_inf[0] && c.fs[1][2] == c.u_inf[1] && c.fs[1][3] == c.u_inf[2]);
    k.B_imf = {c.B_imf[0], c.B_imf[1], c.B_imf[2]};
    k.dipole = dipole != 0;
    if (m) k.dipole_m = {m[0], m[1], m[2]};
    if (cc) k.dipole_c = {cc[0], cc[1], cc[2]};
    k.dipole_soft = soft;
    k.write_vtk = false; k.write_fields = false; k.write_profile = false; k.log_every = 0;
    k.bc_default = "outflow";
    k.bc_by_tag["inner"] = "wall";
    k.bc_by_tag["outer"] = "farfield";
    return k;
}

// cells: cell c has loops [cell_ptr[c], cell_ptr[c+1]); loop l has nodes loop_nodes[loop_ptr[l] .. loop_ptr[l+1]).
// tags: tag t has nodes tag_nodes[tag_ptr[t] .. tag_ptr[t+1]) and id tag_id[t]; tagset 0: "inner" = 1, "outer" = 2
// (the shell), else the six faces of a box.
inline Mesh build_mesh(int32_t n_nodes, const double *xyz, int64_t n_cells, const int32_t *cell_ptr, const int32_t *loop_ptr,
                       const int32_t *loop_nodes, const int32_t *material, int64_t n_tags, const int32_t *tag_ptr, const int32_t *tag_nodes,
                       const int32_t *tag_id, int tagset)
{
    Mesh m;
    ComplexBuilder b(n_nodes);
    std::vector<std::vector<Id>> loops;
    for (int64_t c = 0; c < n_cells; ++c)
    {
        loops.clear();
        for (int32_t l = cell_ptr[c]; l < cell_ptr[c + 1]; ++l) loops.emplace_back(loop_nodes + loop_ptr[l], loop_nodes + loop_ptr[l + 1]);
        b.add_cell(loops, material ? material[c] : kFluid);
    }
    for (int64_t t = 0; t < n_tags; ++t) b.tag_face(std::vector<Id>(tag_nodes + tag_ptr[t], tag_nodes + tag_ptr[t + 1]), tag_id[t]);
    m.cx = b.finish();
    m.x.resize(size_t(n_nodes));
    for (int32_t i = 0; i < n_nodes; ++i) m.x[size_t(i)] = {xyz[3 * size_t(i)], xyz[3 * size_t(i) + 1], xyz[3 * size_t(i) + 2]};
    if (tagset == 0) m.tag_names = {"", "inner", "outer"};
    else m.tag_names = {"", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
    return m;
}
} // namespace refh
