#pragma once
// Device tables derived once per mesh on the host: packed signed incidence, face -> (L, R) cells,
// face-sorted cell rows for the owner-gather, the reconstruction stencil and the implicit-induction
// constants.  These are the data the reference builds in its constructors (Stencil, ImplicitInduction).
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/mhdg.h"

namespace mhdg
{
struct Tables
{
    int32_t nc = 0, nf = 0, ne = 0;
    std::vector<int32_t> c2f_ptr, c2f, cfs, f2e_ptr, f2e, e2f_ptr, e2f, fL, fR, ec_ptr, ec_cell;
    std::vector<int8_t> mat8;
    std::vector<int32_t> snb;
    std::vector<double> sw, sd, sinv;
    std::vector<int8_t> svalid;
    std::vector<double> recon, m_e, d2, taper;
    std::vector<int32_t> heavy; // edges whose valence exceeds the one-thread limit
    int32_t max_valence = 0;
};

// Throws std::runtime_error on inconsistent input.
// Checks every count, CSR pointer and index of a descriptor (throws std::runtime_error naming the defect).
void check_desc(const mhdg_mesh_desc &d);
Tables build_tables(const mhdg_mesh_desc &d, const mhdg_config &cfg, int32_t light_max);
} // namespace mhdg
