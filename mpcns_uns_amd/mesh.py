"""Host-side mesh: complex, geometry, boundary classification, background field and seeded states.

Wraps include/mhdg_mesh.h.  All arrays are copied into numpy so a Mesh outlives nothing native except the
handle used to build descriptors.
"""
import ctypes as C

import numpy as np

from ._lib import MeshDesc, Config, lib, ptr, f64

BC_TYPES = {"wall": 0, "inflow": 1, "outflow": 2, "farfield": 3}


class Mesh:
    """A cell complex with its geometry (one handle of the native library)."""

    def __init__(self, handle):
        self._h = handle
        cnt = np.zeros(8, np.int64)
        lib().mhdg_mesh_counts(self._h, ptr(cnt))
        self.n_nodes, self.n_edges, self.n_faces, self.n_cells = (int(x) for x in cnt[:4])
        self._desc = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().mhdg_mesh_free(h)
            self._h = None

    # ---------------------------------------------------------------- construction
    @classmethod
    def shell(cls, n_lat=24, n_lon=48, n_r=9, r_in=1.0, r_out=4.0, ratio=1.05, jitter=0.05, seed=20261202):
        """Seeded tetrahedral lat-lon shell (tags 1 = inner, 2 = outer)."""
        h = C.c_void_p()
        if lib().mhdg_mesh_shell(n_lat, n_lon, n_r, r_in, r_out, ratio, jitter, seed, C.byref(h)):
            raise RuntimeError(lib().mhdg_mesh_error().decode())
        return cls(h)

    def hilbert_order(self):
        """Cells in the order of the Hilbert key of their centres (mhdg_mesh_hilbert_order): order[k] = old cell id."""
        order = np.zeros(self.n_cells, np.int32)
        if lib().mhdg_mesh_hilbert_order(self._h, ptr(order)):
            raise RuntimeError(lib().mhdg_mesh_error().decode())
        return order

    def renumbered(self, order):
        """The same mesh with its cells in `order` and faces and edges numbered by first appearance
        (mhdg_mesh_renumber_cells), and the maps back: (mesh, cell_map, face_map, edge_map), x_map[new id] = old id.
        The boundary classification is not carried over: call boundary() on the new mesh."""
        order = np.ascontiguousarray(order, dtype=np.int32)
        if order.size != self.n_cells:
            raise ValueError("order needs one entry per cell")
        h = C.c_void_p()
        if lib().mhdg_mesh_renumber_cells(self._h, ptr(order), C.byref(h)):
            raise RuntimeError(lib().mhdg_mesh_error().decode())
        m = Mesh(h)
        # a cell keeps its loops, a face its loop start: rows correspond entry by entry
        face_map = np.empty(self.n_faces, np.int32)
        old_c2f = self.array("c2f.idx").reshape(self.n_cells, -1)
        face_map[m.array("c2f.idx").reshape(m.n_cells, -1)] = old_c2f[order]
        edge_map = np.empty(self.n_edges, np.int32)
        old_f2e = self.array("f2e.idx").reshape(self.n_faces, -1)
        edge_map[m.array("f2e.idx").reshape(m.n_faces, -1)] = old_f2e[face_map]
        return m, order.copy(), face_map, edge_map

    @classmethod
    def from_cells(cls, xyz, cell_loops, material=None, tags=()):
        """cell_loops: per cell a list of outward node loops; tags: iterable of (nodes, tag_id)."""
        xyz = f64(xyz).reshape(-1, 3)
        cell_ptr, loop_ptr, nodes = [0], [0], []
        for loops in cell_loops:
            for lp in loops:
                nodes.extend(int(v) for v in lp)
                loop_ptr.append(len(nodes))
            cell_ptr.append(len(loop_ptr) - 1)
        tag_ptr, tag_nodes, tag_id = [0], [], []
        for nds, t in tags:
            tag_nodes.extend(int(v) for v in nds)
            tag_ptr.append(len(tag_nodes))
            tag_id.append(int(t))
        a = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.int32))
        cp, lp, nd, tp, tn, ti = a(cell_ptr), a(loop_ptr), a(nodes), a(tag_ptr), a(tag_nodes), a(tag_id)
        mat = None if material is None else a(material)
        h = C.c_void_p()
        if lib().mhdg_mesh_from_cells(len(xyz), ptr(xyz), len(cell_loops), ptr(cp), ptr(lp), ptr(nd), ptr(mat),
                                      len(ti), ptr(tp), ptr(tn), ptr(ti), C.byref(h)):
            raise RuntimeError(lib().mhdg_mesh_error().decode())
        m = cls(h)
        m._loops = (cp, lp, nd, tp, tn, ti, mat)
        return m

    @classmethod
    def read_unsbin(cls, path):
        """Complex file of the reference (uns::io::read_local_mesh) written for one rank: stored numbering and
        orientation kept, geometry recomputed, topology hash verified."""
        h = C.c_void_p()
        if lib().mhdg_mesh_read_unsbin(str(path).encode(), C.byref(h)):
            raise RuntimeError(lib().mhdg_mesh_error().decode())
        return cls(h)

    @classmethod
    def read_unsbin_parts(cls, paths):
        """The part files of an n-rank partition (the reference's preprocessing output) merged into the global
        complex, in its global numbering; hashes and ghost copies checked, geometry recomputed."""
        enc = [str(p).encode() for p in paths]
        arr = (C.c_char_p * len(enc))(*enc)
        h = C.c_void_p()
        if lib().mhdg_mesh_read_unsbin_parts(arr, len(enc), C.byref(h)):
            raise RuntimeError(lib().mhdg_mesh_error().decode())
        return cls(h)

    def write_unsbin(self, path):
        """Complex file readable by the reference (uns::io::write_local_mesh layout, one rank)."""
        if lib().mhdg_mesh_write_unsbin(self._h, str(path).encode()):
            raise RuntimeError(lib().mhdg_mesh_error().decode())

    def write_vtu(self, path, cell_data=(), point_data=()):
        """VTK unstructured-grid file of every cell as a polyhedron (uns::io::write_vtu layout).
        cell_data / point_data: [(name, ncomp, values)] with ncomp values per cell / node."""
        keep = []
        cells, points = _vtk_arrays(cell_data, self.n_cells, keep), _vtk_arrays(point_data, self.n_nodes, keep)
        if lib().mhdg_mesh_write_vtu(self._h, str(path).encode(), len(cell_data), cells, len(point_data), points):
            raise RuntimeError(lib().mhdg_mesh_error().decode())

    @property
    def topology_hash(self):
        h = C.c_uint64()
        lib().mhdg_mesh_topology_hash(self._h, C.byref(h))
        return h.value

    @property
    def tag_names(self):
        buf = C.create_string_buffer(1 << 16)
        lib().mhdg_mesh_tag_names(self._h, buf, len(buf))
        return buf.value.decode().split("\n")[:-1]

    @tag_names.setter
    def tag_names(self, names):
        lib().mhdg_mesh_set_tag_names(self._h, "".join(n + "\n" for n in names).encode())

    # ---------------------------------------------------------------- arrays
    def array(self, name):
        p, n, es = C.c_void_p(), C.c_int64(), C.c_int32()
        if lib().mhdg_mesh_array(self._h, name.encode(), C.byref(p), C.byref(n), C.byref(es)):
            raise KeyError(name)
        if n.value == 0:
            return np.zeros(0)
        if name.endswith(".ptr") or name == "ec_ptr":
            dt = np.int64
        elif name.endswith(".sgn") or name in ("face_kind", "edge_mode"):
            dt = np.int8
        elif name.endswith(".idx") or name in ("tets", "face_tag", "material", "ec_cell", "ec_misoriented"):
            dt = np.int32
        else:
            dt = np.float64
        assert np.dtype(dt).itemsize == es.value, name
        buf = (C.c_char * (n.value * es.value)).from_address(p.value)
        return np.frombuffer(buf, dtype=dt).copy()

    def boundary(self, tag_types=None, default="outflow", u_inf=(-1.0, 0.0, 0.0)):
        """Face kinds and edge modes from face tags (mhd::BoundaryInfo::build); default: inner wall, outer farfield."""
        if tag_types is None:
            tag_types = {1: "wall", 2: "farfield"}
        n = max(tag_types) + 1 if tag_types else 1
        tt = np.full(n, -1, np.int32)
        for k, v in tag_types.items():
            tt[k] = BC_TYPES[v]
        u = f64(u_inf)
        if lib().mhdg_mesh_boundary(self._h, ptr(tt), n, BC_TYPES[default], ptr(u)):
            raise RuntimeError(lib().mhdg_mesh_error().decode())
        self._desc = None
        return self

    def desc(self):
        """mhdg_mesh_desc pointing into this mesh (valid while the Mesh lives)."""
        if self._desc is None:
            d = MeshDesc()
            if lib().mhdg_mesh_desc_fill(self._h, C.byref(d)):
                raise RuntimeError(lib().mhdg_mesh_error().decode())
            self._desc = d
        return self._desc

    def mg_hierarchy(self, cfg, dirichlet=None):
        """Levels of the multigrid hierarchy the device builds for this mesh and Dirichlet set (mhdg_mg_build): a list of
        dicts with n, n_nodes, mass, the end nodes tail/head of every unknown and, from level 1 on, agg (the aggregate of
        every finer node), map (per finer unknown: signed index, v < 0 = -1 on ~v, INT32_MIN none) and the CSR of K_l
        (k_ptr, k_col, k_val)."""
        d = None if dirichlet is None else np.ascontiguousarray(np.asarray(dirichlet) != 0, dtype=np.int8)
        if d is not None and d.size != self.n_edges:
            raise ValueError("dirichlet needs one entry per edge")
        h = C.c_void_p()
        if lib().mhdg_mg_build(C.byref(self.desc()), C.byref(cfg), ptr(d), C.byref(h)):
            raise RuntimeError(lib().mhdg_mesh_error().decode())
        try:
            out = []
            for lvl in range(lib().mhdg_mg_levels(h)):
                n, nn, nnz = C.c_int32(), C.c_int32(), C.c_int64()
                lib().mhdg_mg_level_size(h, lvl, C.byref(n), C.byref(nn), C.byref(nnz))
                L = dict(n=n.value, n_nodes=nn.value, mass=np.zeros(n.value), tail=np.zeros(n.value, np.int32),
                         head=np.zeros(n.value, np.int32))
                if lvl > 0:
                    L["agg"] = np.zeros(out[-1]["n_nodes"], np.int32)
                lib().mhdg_mg_level_nodes(h, lvl, ptr(L["tail"]), ptr(L["head"]), ptr(L.get("agg")))
                if lvl == 0:
                    lib().mhdg_mg_level_arrays(h, lvl, None, ptr(L["mass"]), None, None, None)
                else:
                    L.update(map=np.zeros(out[-1]["n"], np.int32), k_ptr=np.zeros(n.value + 1, np.int32),
                             k_col=np.zeros(nnz.value, np.int32), k_val=np.zeros(nnz.value))
                    lib().mhdg_mg_level_arrays(h, lvl, ptr(L["map"]), ptr(L["mass"]), ptr(L["k_ptr"]), ptr(L["k_col"]), ptr(L["k_val"]))
                out.append(L)
            return out
        finally:
            lib().mhdg_mg_free(h)

    def background_dipoles(self, B_imf, positions, moments, cores=None, cutoff=0.0):
        """Background of the IMF plus a crustal field of point dipoles (MarsSolver::build_background_ recipe)."""
        pos, mom = f64(positions).reshape(-1, 3), f64(moments).reshape(-1, 3)
        core = f64(np.zeros(len(pos)) if cores is None else cores)
        PhiAdd, Badd = np.zeros(self.n_faces), np.zeros(3 * self.n_cells)
        if lib().mhdg_mesh_background_dipoles(self._h, ptr(f64(B_imf)), len(pos), ptr(pos), ptr(mom), ptr(core), cutoff, ptr(PhiAdd),
                                              ptr(Badd)):
            raise RuntimeError(lib().mhdg_mesh_error().decode())
        return PhiAdd, Badd

    def background(self, B_imf=(0, 0, 0), dipole_m=None, dipole_c=(0, 0, 0), soft=1e-3):
        PhiAdd = np.zeros(self.n_faces)
        Badd = np.zeros(3 * self.n_cells)
        B, m, c = f64(B_imf), f64(dipole_m if dipole_m is not None else (0, 0, 0)), f64(dipole_c)
        lib().mhdg_mesh_background(self._h, ptr(B), int(dipole_m is not None), ptr(m), ptr(c), soft, ptr(PhiAdd), ptr(Badd))
        return PhiAdd, Badd

    def cell_vector(self, flux):
        out = np.zeros(3 * self.n_cells)
        lib().mhdg_mesh_cell_vector(self._h, ptr(f64(flux)), ptr(out))
        return out

    def seeded_state(self, cfg, noise, seed, Badd):
        UH, UNa = np.zeros(5 * self.n_cells), np.zeros(5 * self.n_cells)
        Phi = np.zeros(self.n_faces)
        lib().mhdg_mesh_seeded_state(self._h, C.byref(cfg), noise, seed, ptr(f64(Badd)), ptr(UH), ptr(UNa), ptr(Phi))
        return UH, UNa, Phi

    def seeded_species_state(self, cfg, fs_list, noise, seed, Badd):
        """Conservative states of len(fs_list) species (free-stream primitives with seeded relative noise, as
        seeded_state, two species per draw with seeds seed, seed+1, ...) and the seeded induced flux of the first draw."""
        states, Phi = [], None
        for k in range(0, len(fs_list), 2):
            pair = list(fs_list[k:k + 2])
            if len(pair) == 1:
                pair.append(pair[0])
            c = Config.from_buffer_copy(cfg)
            for s in range(2):
                for m in range(5):
                    c.fs[s][m] = float(pair[s][m])
            a, b, phi = self.seeded_state(c, noise, seed + k // 2, Badd)
            states += [a, b]
            if Phi is None:
                Phi = phi
        return states[:len(fs_list)], Phi


def default_config(**kw):
    """mhdg_config with the reference defaults, then keyword overrides.

    Array fields take sequences: fs (2x5), ion_fs (up to 10x5), ion_mass and src_rate (up to 10), u_inf and
    B_imf (3); missing trailing entries stay zero."""
    c = Config()
    lib().mhdg_config_defaults(C.byref(c))
    for k, v in kw.items():
        if k in ("fs", "ion_fs"):
            arr = getattr(c, k)
            for s, row in enumerate(v):
                for m in range(5):
                    arr[s][m] = float(row[m])
        elif k in ("u_inf", "B_imf", "ion_mass", "src_rate"):
            arr = getattr(c, k)
            for i, x in enumerate(v):
                arr[i] = float(x)
        else:
            setattr(c, k, v)
    return c


# --------------------------------------------------------------------------- VTK output
class _VtkArray(C.Structure):
    _fields_ = [("name", C.c_char_p), ("ncomp", C.c_int32), ("values", C.c_void_p)]


def _vtk_arrays(arrays, count, keep):
    out = (_VtkArray * max(len(arrays), 1))()
    for i, (name, ncomp, values) in enumerate(arrays):
        v = None if values is None else f64(values)
        if v is not None and count is not None and v.size != count * ncomp:
            raise ValueError(f"{name}: {v.size} values, expected {count} x {ncomp}")
        keep.append(v)
        out[i] = _VtkArray(name.encode(), int(ncomp), None if v is None else v.ctypes.data)
    return out


def write_pvtu(path, pieces, cell_data=(), point_data=()):
    """Parallel master file over piece files (uns::io::write_pvtu); arrays as [(name, ncomp)]."""
    keep = []
    cells = _vtk_arrays([(n, c, None) for n, c in cell_data], None, keep)
    points = _vtk_arrays([(n, c, None) for n, c in point_data], None, keep)
    names = (C.c_char_p * max(len(pieces), 1))(*[str(p).encode() for p in pieces])
    if lib().mhdg_write_pvtu(str(path).encode(), len(pieces), names, len(cell_data), cells, len(point_data), points):
        raise RuntimeError(lib().mhdg_mesh_error().decode())


def write_pvd(path, entries):
    """Time-series collection (uns::io::write_pvd): entries = [(time, file)]."""
    times = np.array([float(t) for t, _ in entries] or [0.0])
    files = (C.c_char_p * max(len(entries), 1))(*[str(f).encode() for _, f in entries])
    if lib().mhdg_write_pvd(str(path).encode(), len(entries), times.ctypes.data, files):
        raise RuntimeError(lib().mhdg_mesh_error().decode())


# --------------------------------------------------------------------------- field files
DIMS = {"node": 0, "edge": 1, "face": 2, "cell": 3}


class _Field(C.Structure):
    _fields_ = [("name", C.c_char_p), ("dim", C.c_int32), ("ncomp", C.c_int32), ("count", C.c_int64), ("values", C.c_void_p)]


def write_fields(path, fields, topology_hash=0):
    """UNSBIN field file (uns::io::write_fields): fields = [(name, dim, ncomp, values)], dim 0 node .. 3 cell."""
    keep, recs = [], (_Field * len(fields))()
    for i, (name, dim, ncomp, values) in enumerate(fields):
        v = f64(values)
        keep.append(v)
        recs[i] = _Field(name.encode(), DIMS.get(dim, dim), int(ncomp), v.size, v.ctypes.data)
    if lib().mhdg_fields_write(str(path).encode(), C.c_uint64(topology_hash), recs, len(fields)):
        raise RuntimeError(lib().mhdg_mesh_error().decode())


def read_fields(path):
    """(topology hash, {name: (dim, ncomp, values)}) of a UNSBIN field file (uns::io::read_fields)."""
    h, n, hs = C.c_void_p(), C.c_int32(), C.c_uint64()
    if lib().mhdg_fields_read(str(path).encode(), C.byref(h), C.byref(hs), C.byref(n)):
        raise RuntimeError(lib().mhdg_mesh_error().decode())
    out = {}
    try:
        for i in range(n.value):
            f = _Field()
            lib().mhdg_fields_get(h, i, C.byref(f))
            vals = np.ctypeslib.as_array(C.cast(f.values, C.POINTER(C.c_double)), (f.count,)).copy() if f.count else np.zeros(0)
            out[f.name.decode()] = (f.dim, f.ncomp, vals)
    finally:
        lib().mhdg_fields_free(h)
    return hs.value, out


def dipole_edge_integrals(positions, moments, cores, xa, xb, cutoff=0.0):
    """Edge integrals of the crustal-dipole vector potential (MarsDipoles.cpp:63-107) for edges xa -> xb."""
    pos, mom, core = f64(positions).reshape(-1, 3), f64(moments).reshape(-1, 3), f64(cores)
    xa, xb = f64(xa).reshape(-1, 3), f64(xb).reshape(-1, 3)
    out = np.zeros(len(xa))
    if lib().mhdg_dipole_edge_integrals(len(pos), ptr(pos), ptr(mom), ptr(core), cutoff, len(xa), ptr(xa), ptr(xb), ptr(out)):
        raise RuntimeError(lib().mhdg_mesh_error().decode())
    return out
