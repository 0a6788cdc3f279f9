// ORACLE — test infrastructure only.  CPU baseline: the compiled reference's Mercury step on several MPI ranks of
// the host, built from the reference sources in place (oracle/Makefile, target ref-mpi) and never part of the product.
//
//   ref_mpi_bench parts DIR NRANKS      one process: reads DIR/input.bin (mesh, configuration, seeded state; written by
//                                       oracle/cpu_baseline.py), builds the global mesh with the reference's
//                                       ComplexBuilder and geometry, partitions it as the reference's preprocessing
//                                       does (Preprocess/Pipeline.cpp:133-156: cell graph, Hilbert partition with its
//                                       default refinement, Distributor) and writes one UNSBIN part file per rank
//                                       (uns::io::write_local_mesh) and that rank's state in its local numbering.
//   mpiexec -n NRANKS ref_mpi_bench run DIR STEPS
//                                       every rank reads its part and state, builds mercury::Solver on MPI_COMM_WORLD
//                                       (MercuryUns/main.cpp), and times STEPS first steps from the seeded state, the
//                                       solver rebuilt before each so every implicit solve starts cold; the step's
//                                       wall time is the solver's own wall_s_ (Solver.cpp:231, 275), maximum over
//                                       ranks.  Rank 0 prints one JSON line.
#include <mpi.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "driver/Solver.h"
#include "uns/geometry/Geometry.h"
#include "uns/io/MeshIO.h"
#include "uns/parallel/Distributor.h"
#include "uns/partition/CellGraph.h"
#include "uns/partition/Partitioner.h"

#include "ref_config.h"

using namespace uns;

namespace
{
struct Reader
{
    std::ifstream in;
    explicit Reader(const std::string &path) : in(path, std::ios::binary)
    {
        if (!in) throw std::runtime_error("cannot open " + path);
    }
    template <class T> void get(T *p, size_t n)
    {
        if (!in.read(reinterpret_cast<char *>(p), std::streamsize(n * sizeof(T)))) throw std::runtime_error("short read");
    }
    template <class T> T one()
    {
        T v;
        get(&v, 1);
        return v;
    }
    template <class T> std::vector<T> vec(size_t n)
    {
        std::vector<T> v(n);
        get(v.data(), n);
        return v;
    }
};

// configuration block: sizeof(mhdg_config) bytes, then dipole flag, moment, centre and softening
struct RunConfig
{
    mhdg_config cfg{};
    int32_t dipole = 0;
    double dm[3] = {0, 0, 0}, dc[3] = {0, 0, 0}, soft = 0;
    void read(Reader &r)
    {
        const int64_t size = r.one<int64_t>();
        if (size != int64_t(sizeof(mhdg_config))) throw std::runtime_error("mhdg_config size mismatch");
        r.get(&cfg, 1);
        dipole = r.one<int32_t>();
        r.get(dm, 3);
        r.get(dc, 3);
        soft = r.one<double>();
    }
    void write(std::ofstream &o) const
    {
        const int64_t size = sizeof(mhdg_config);
        o.write(reinterpret_cast<const char *>(&size), 8);
        o.write(reinterpret_cast<const char *>(&cfg), sizeof cfg);
        o.write(reinterpret_cast<const char *>(&dipole), 4);
        o.write(reinterpret_cast<const char *>(dm), 24);
        o.write(reinterpret_cast<const char *>(dc), 24);
        o.write(reinterpret_cast<const char *>(&soft), 8);
    }
};

std::string part_file(const std::string &dir, const char *stem, int r)
{
    char b[64];
    std::snprintf(b, sizeof b, "/%s_%05d.unsb", stem, r);
    return dir + b;
}

// input.bin: "MHDGBIN1", int64 n_nodes, n_cells, n_faces, n_tags; xyz (3 n_nodes doubles); tets (4 n_cells int32,
// positively oriented); tagged faces (3 n_tags int32 node ids, then n_tags int32 tag ids); the configuration
// block; U_H, U_Na (5 n_cells doubles each) and Phi (n_faces doubles) in the numbering of the builder.
int parts(const std::string &dir, int nranks)
{
    Reader r(dir + "/input.bin");
    char magic[8];
    r.get(magic, 8);
    if (std::memcmp(magic, "MHDGBIN1", 8) != 0) throw std::runtime_error("input.bin: bad magic");
    const int64_t nn = r.one<int64_t>(), nc = r.one<int64_t>(), nf = r.one<int64_t>(), nt = r.one<int64_t>();
    const auto xyz = r.vec<double>(size_t(3 * nn));
    const auto tets = r.vec<int32_t>(size_t(4 * nc));
    const auto tag_nodes = r.vec<int32_t>(size_t(3 * nt));
    const auto tag_id = r.vec<int32_t>(size_t(nt));
    RunConfig rc;
    rc.read(r);
    const auto UH = r.vec<double>(size_t(5 * nc)), UNa = r.vec<double>(size_t(5 * nc));
    const auto Phi = r.vec<double>(size_t(nf));
    // outward loops of a positive tet (a c b), (a b d), (a d c), (b c d): the loops the shell generator's cells have
    std::vector<int32_t> cell_ptr(size_t(nc) + 1), loop_ptr(size_t(4 * nc) + 1), loop_nodes(size_t(12 * nc));
    static const int L[12] = {0, 2, 1, 0, 1, 3, 0, 3, 2, 1, 2, 3};
    for (int64_t c = 0; c <= nc; ++c) cell_ptr[size_t(c)] = int32_t(4 * c);
    for (int64_t l = 0; l <= 4 * nc; ++l) loop_ptr[size_t(l)] = int32_t(3 * l);
    for (int64_t c = 0; c < nc; ++c)
        for (int k = 0; k < 12; ++k) loop_nodes[size_t(12 * c + k)] = tets[size_t(4 * c + L[k])];
    std::vector<int32_t> tag_ptr(size_t(nt) + 1);
    for (int64_t t = 0; t <= nt; ++t) tag_ptr[size_t(t)] = int32_t(3 * t);
    const Mesh m = refh::build_mesh(int32_t(nn), xyz.data(), nc, cell_ptr.data(), loop_ptr.data(), loop_nodes.data(), nullptr, nt,
                                    tag_ptr.data(), tag_nodes.data(), tag_id.data(), 0);
    if (m.cx.n_faces() != nf) throw std::runtime_error("input.bin: face count differs from the built complex");
    const Geometry geo = compute_geometry(m);
    const CellGraph cg = build_cell_graph(m, geo, CostModel());
    const PartOptions po; // Hilbert, 8 refinement passes, 3 % tolerance: the preprocessing defaults
    const std::vector<int> part = partition_cells(cg, nranks, po);
    Distributor dist(m, part, nranks);
    for (int k = 0; k < nranks; ++k)
    {
        const LocalMesh lm = dist.build(k);
        io::write_local_mesh(part_file(dir, "part", k), lm);
        const CellComplex &cx = lm.mesh.cx;
        std::vector<io::FieldRecord> fr(3);
        fr[0] = {"U_H", Dim::Cell, 5, {}};
        fr[1] = {"U_Na", Dim::Cell, 5, {}};
        fr[2] = {"Phi", Dim::Face, 1, {}};
        for (Id c = 0; c < cx.n_cells(); ++c)
        {
            const size_t g = size_t(cx.ent[3].gid[size_t(c)]);
            for (int q = 0; q < 5; ++q)
            {
                fr[0].values.push_back(UH[5 * g + size_t(q)]);
                fr[1].values.push_back(UNa[5 * g + size_t(q)]);
            }
        }
        for (Id f = 0; f < cx.n_faces(); ++f)
            fr[2].values.push_back(double(cx.ent[2].alias_sign[size_t(f)]) * Phi[size_t(cx.ent[2].gid[size_t(f)])]);
        io::write_fields(part_file(dir, "state", k), topology_hash(cx), fr);
    }
    std::ofstream o(dir + "/run.bin", std::ios::binary);
    rc.write(o);
    std::printf("{\"parts\": %d, \"cells\": %lld}\n", nranks, static_cast<long long>(nc));
    return 0;
}

int run(const std::string &dir, int steps)
{
    int rank = 0, size = 1;
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    const LocalMesh lm = io::read_local_mesh(part_file(dir, "part", rank));
    if (lm.nranks != size) throw std::runtime_error("part files were written for another rank count");
    RunConfig rc;
    {
        Reader r(dir + "/run.bin");
        rc.read(r);
    }
    const mhd::Config cfg = refh::make_config(rc.cfg, rc.dipole, rc.dm, rc.dc, rc.soft);
    std::uint64_t hash = 0;
    const auto fr = io::read_fields(part_file(dir, "state", rank), &hash);
    if (hash != topology_hash(lm.mesh.cx) || fr.size() != 3) throw std::runtime_error("state file does not belong to the part file");
    long long owned = lm.n_owned(Dim::Cell), cells = 0;
    MPI_Allreduce(&owned, &cells, 1, MPI_LONG_LONG, MPI_SUM, MPI_COMM_WORLD);
    std::vector<double> wall, outer, res;
    std::vector<long> iters;
    for (int s = 0; s < steps; ++s)
    {
        mercury::Solver solver(MPI_COMM_WORLD, lm, cfg);
        solver.initialize();
        auto &st = solver.state();
        if (st.U[0].size() != fr[0].values.size() || st.Phi.size() != fr[2].values.size()) throw std::runtime_error("state size mismatch");
        st.U[0] = fr[0].values;
        st.U[1] = fr[1].values;
        st.Phi = fr[2].values;
        solver.reset_profile();
        MPI_Barrier(MPI_COMM_WORLD);
        const double t0 = MPI_Wtime();
        solver.step_first_order(1e9);
        MPI_Barrier(MPI_COMM_WORLD);
        const double t1 = MPI_Wtime() - t0;
        const auto p = solver.profile();
        double w = p.wall, wmax = 0, r = solver.last_induction_residual(), rmax = 0;
        MPI_Reduce(&w, &wmax, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
        MPI_Reduce(&r, &rmax, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
        wall.push_back(wmax);
        outer.push_back(t1);
        res.push_back(rmax);
        iters.push_back(p.cg_iterations);
    }
    if (rank == 0)
    {
        std::string s = "{\"ranks\": " + std::to_string(size) + ", \"cells\": " + std::to_string(cells) + ", \"step_wall_s\": [";
        for (size_t i = 0; i < wall.size(); ++i) s += (i ? ", " : "") + std::to_string(wall[i]);
        s += "], \"barrier_to_barrier_s\": [";
        for (size_t i = 0; i < outer.size(); ++i) s += (i ? ", " : "") + std::to_string(outer[i]);
        s += "], \"gmres_iterations\": [";
        for (size_t i = 0; i < iters.size(); ++i) s += (i ? ", " : "") + std::to_string(iters[i]);
        char b[64];
        std::snprintf(b, sizeof b, "], \"max_rel_residual\": %.3e}", *std::max_element(res.begin(), res.end()));
        s += b;
        std::printf("%s\n", s.c_str());
        std::fflush(stdout);
    }
    return 0;
}
} // namespace

int main(int argc, char **argv)
{
    const std::string mode = argc > 1 ? argv[1] : "";
    if (argc < 4 || (mode != "parts" && mode != "run"))
    {
        std::fprintf(stderr, "usage: ref_mpi_bench parts DIR NRANKS | run DIR STEPS\n");
        return 2;
    }
    MPI_Init(&argc, &argv);
    int rc = 0;
    try
    {
        rc = mode == "parts" ? parts(argv[2], std::atoi(argv[3])) : run(argv[2], std::max(1, std::atoi(argv[3])));
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "ref_mpi_bench %s: %s\n", mode.c_str(), e.what());
        std::fflush(stderr);
        if (mode == "run") MPI_Abort(MPI_COMM_WORLD, 1); // the other ranks may wait in a collective
        rc = 1;
    }
    MPI_Finalize();
    return rc;
}
