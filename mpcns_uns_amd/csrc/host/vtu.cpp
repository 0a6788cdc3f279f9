// VTK XML writers (reference src/uns/io/VtkWriter.cpp:42-116): the element order, attribute text and the
// appended-data layout (UInt64 byte count before each payload, offsets counting those 8 bytes) are the
// file format and are kept as the reference writes them.
#include "vtu.h"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace mhdg::vtk
{
namespace
{
// payloads of the <AppendedData> block; each DataArray tag records the running offset of its payload
struct Payloads
{
    std::vector<std::pair<const void *, uint64_t>> parts;
    uint64_t offset = 0;
    std::string tag(const char *type, const std::string &name, int ncomp, const void *data, uint64_t bytes)
    {
        std::ostringstream s;
        s << "<DataArray type=\"" << type << "\" Name=\"" << name << "\" NumberOfComponents=\"" << ncomp
          << "\" format=\"appended\" offset=\"" << offset << "\"/>";
        parts.emplace_back(data, bytes);
        offset += 8 + bytes;
        return s.str();
    }
};

// sorted distinct nodes of the cell's face loops, and the loops oriented outward (reversed when the
// c2f sign is negative), as CellComplex::cell_nodes / face_loop_outward give them
void cell_topology(const Complex &cx, int32_t c, std::vector<int64_t> &nodes, std::vector<int64_t> &faces)
{
    int32_t loop[64];
    int n = 0;
    nodes.clear();
    const int64_t k0 = cx.c2f.ptr[c], k1 = cx.c2f.ptr[c + 1];
    faces.push_back(k1 - k0);
    for (int64_t k = k0; k < k1; ++k)
    {
        const int32_t f = cx.c2f.idx[k];
        if (cx.f2e.ptr[f + 1] - cx.f2e.ptr[f] > 64) throw std::runtime_error("vtu: face with more than 64 edges");
        cx.face_loop(f, loop, n);
        nodes.insert(nodes.end(), loop, loop + n);
        if (cx.c2f.sgn[k] < 0) std::reverse(loop, loop + n);
        faces.push_back(n);
        faces.insert(faces.end(), loop, loop + n);
    }
    std::sort(nodes.begin(), nodes.end());
    nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
}
} // namespace

void write_vtu(const std::string &path, const Complex &cx, const std::vector<double> &xyz, int32_t n_cells,
               const std::vector<Array> &cell_data, const std::vector<Array> &point_data)
{
    if (n_cells < 0 || n_cells > cx.nc) throw std::runtime_error("vtu: cell count out of range");
    const size_t n_nodes = xyz.size() / 3;
    std::vector<int64_t> conn, offsets, faces, face_offsets, cn;
    std::vector<uint8_t> types(size_t(n_cells), 42);
    for (int32_t c = 0; c < n_cells; ++c)
    {
        cell_topology(cx, c, cn, faces);
        conn.insert(conn.end(), cn.begin(), cn.end());
        offsets.push_back(int64_t(conn.size()));
        face_offsets.push_back(int64_t(faces.size()));
    }
    Payloads p;
    std::ostringstream h;
    h << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
      << "<UnstructuredGrid>\n<Piece NumberOfPoints=\"" << n_nodes << "\" NumberOfCells=\"" << n_cells << "\">\n";
    h << "<Points>" << p.tag("Float64", "Points", 3, xyz.data(), 8 * xyz.size()) << "</Points>\n<Cells>\n";
    h << p.tag("Int64", "connectivity", 1, conn.data(), 8 * conn.size()) << "\n";
    h << p.tag("Int64", "offsets", 1, offsets.data(), 8 * offsets.size()) << "\n";
    h << p.tag("UInt8", "types", 1, types.data(), types.size()) << "\n";
    h << p.tag("Int64", "faces", 1, faces.data(), 8 * faces.size()) << "\n";
    h << p.tag("Int64", "faceoffsets", 1, face_offsets.data(), 8 * face_offsets.size()) << "\n</Cells>\n";
    auto section = [&](const char *name, const std::vector<Array> &arrays, size_t count) {
        h << "<" << name << ">\n";
        for (const Array &a : arrays) h << p.tag("Float64", a.name, a.ncomp, a.values, 8 * count * size_t(a.ncomp)) << "\n";
        h << "</" << name << ">\n";
    };
    section("CellData", cell_data, size_t(n_cells));
    section("PointData", point_data, n_nodes);
    h << "</Piece>\n</UnstructuredGrid>\n";
    std::ofstream os(path, std::ios::binary);
    if (!os) throw std::runtime_error("cannot write " + path);
    os << h.str() << "<AppendedData encoding=\"raw\">\n_";
    for (const auto &part : p.parts)
    {
        os.write(reinterpret_cast<const char *>(&part.second), 8);
        if (part.second) os.write(static_cast<const char *>(part.first), std::streamsize(part.second));
    }
    os << "\n</AppendedData>\n</VTKFile>\n";
    if (!os) throw std::runtime_error("write failed: " + path);
}

void write_pvtu(const std::string &path, const std::vector<std::string> &pieces, const std::vector<Array> &cell_data,
                const std::vector<Array> &point_data)
{
    std::ofstream os(path);
    if (!os) throw std::runtime_error("cannot write " + path);
    os << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
       << "<PUnstructuredGrid GhostLevel=\"0\">\n<PPoints><PDataArray type=\"Float64\" NumberOfComponents=\"3\"/></PPoints>\n";
    for (const auto &sec : {std::make_pair("PCellData", &cell_data), std::make_pair("PPointData", &point_data)})
    {
        os << "<" << sec.first << ">\n";
        for (const Array &a : *sec.second)
            os << "<PDataArray type=\"Float64\" Name=\"" << a.name << "\" NumberOfComponents=\"" << a.ncomp << "\"/>\n";
        os << "</" << sec.first << ">\n";
    }
    for (const std::string &piece : pieces) os << "<Piece Source=\"" << piece << "\"/>\n";
    os << "</PUnstructuredGrid>\n</VTKFile>\n";
}

void write_pvd(const std::string &path, const std::vector<std::pair<double, std::string>> &entries)
{
    std::ofstream os(path);
    if (!os) throw std::runtime_error("cannot write " + path);
    os << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n<Collection>\n";
    for (const auto &e : entries) os << "<DataSet timestep=\"" << e.first << "\" part=\"0\" file=\"" << e.second << "\"/>\n";
    os << "</Collection>\n</VTKFile>\n";
}
} // namespace mhdg::vtk
