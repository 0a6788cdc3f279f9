"""VTK XML output (the reference's src/uns/io/VtkWriter.cpp layout): polyhedral unstructured-grid files, the
parallel master file and the time-series collection.

Against the compiled reference (oracle/_ref): for the same mesh and arrays both writers produce the same bytes,
on the tetrahedral shell and on a hexahedral box.  Without the reference: the file's own structure (appended
offsets, cell connectivity and outward face loops, array payloads).
"""
import ctypes as C
import re

import numpy as np
import pytest

from meshes import box_mesh


def _arrays(m, rng):
    cell = [("U_H", 5, rng.normal(size=5 * m.n_cells)), ("B", 3, rng.normal(size=3 * m.n_cells))]
    point = [("phi", 1, rng.normal(size=m.n_nodes))]
    return cell, point


def _ref_write(ref, path, cell, point):
    from oracle.oracle import Reference

    L = Reference.lib()
    L.ref_write_vtu.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                C.c_void_p, C.c_void_p, C.c_void_p]
    keep = []

    def pack(arrs):
        names = (C.c_char_p * max(len(arrs), 1))(*[a[0].encode() for a in arrs])
        nc = np.array([a[1] for a in arrs] or [0], dtype=np.int32)
        vals = [np.ascontiguousarray(a[2], dtype=np.float64) for a in arrs]
        keep.extend(vals)
        ptrs = (C.c_void_p * max(len(arrs), 1))(*[v.ctypes.data for v in vals])
        cnt = np.array([v.size for v in vals] or [0], dtype=np.int64)
        keep.extend([nc, cnt])
        return names, nc.ctypes.data, ptrs, cnt.ctypes.data

    cn, cc, cv, ck = pack(cell)
    pn, pc, pv, pk = pack(point)
    assert L.ref_write_vtu(ref._h, str(path).encode(), len(cell), cn, cc, cv, ck, len(point), pn, pc, pv, pk) == 0


def _ref():
    from oracle.oracle import have_ref

    if not have_ref():
        pytest.skip("reference library oracle/_ref not built")


def test_vtu_matches_the_reference_shell(tmp_path):
    _ref()
    from mpcns_uns_amd import workloads
    from oracle.oracle import reference_for_mesh

    w = workloads.build(net=(6, 12, 2), hall=False)
    m = w["mesh"]
    ref = reference_for_mesh(m, w["cfg"])
    cell, point = _arrays(m, np.random.default_rng(3))
    m.write_vtu(tmp_path / "ours.vtu", cell, point)
    _ref_write(ref, tmp_path / "ref.vtu", cell, point)
    assert (tmp_path / "ours.vtu").read_bytes() == (tmp_path / "ref.vtu").read_bytes()


def test_vtu_matches_the_reference_hex_box(tmp_path):
    _ref()
    from mpcns_uns_amd import default_config
    from oracle.oracle import Reference

    m = box_mesh(3, 2, 2, hi=(2.0, 1.5, 1.0))
    xyz, cells, tags = m._cells
    ref = Reference(xyz, cells, default_config(), tags=tags, tagset=1)
    cell, point = _arrays(m, np.random.default_rng(4))
    m.write_vtu(tmp_path / "ours.vtu", cell, [])
    _ref_write(ref, tmp_path / "ref.vtu", cell, [])
    assert (tmp_path / "ours.vtu").read_bytes() == (tmp_path / "ref.vtu").read_bytes()


def test_master_and_collection_match_the_reference(tmp_path):
    _ref()
    from mpcns_uns_amd import write_pvd, write_pvtu
    from oracle.oracle import Reference

    L = Reference.lib()
    pieces = ["step_0_r0.vtu", "step_0_r1.vtu"]
    write_pvtu(tmp_path / "ours.pvtu", pieces, [("U_H", 5), ("B", 3)], [("phi", 1)])
    names = lambda xs: (C.c_char_p * len(xs))(*[x.encode() for x in xs])
    L.ref_write_pvtu.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    cc, pc = np.array([5, 3], np.int32), np.array([1], np.int32)
    assert L.ref_write_pvtu(str(tmp_path / "ref.pvtu").encode(), 2, names(pieces), 2, names(["U_H", "B"]), cc.ctypes.data, 1,
                            names(["phi"]), pc.ctypes.data) == 0
    assert (tmp_path / "ours.pvtu").read_bytes() == (tmp_path / "ref.pvtu").read_bytes()
    entries = [(0.0, "s0.pvtu"), (0.125, "s1.pvtu"), (1.0 / 3.0, "s2.pvtu"), (12345.678, "s3.pvtu")]
    write_pvd(tmp_path / "ours.pvd", entries)
    t = np.array([e[0] for e in entries])
    L.ref_write_pvd.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p]
    assert L.ref_write_pvd(str(tmp_path / "ref.pvd").encode(), len(entries), t.ctypes.data, names([e[1] for e in entries])) == 0
    assert (tmp_path / "ours.pvd").read_bytes() == (tmp_path / "ref.pvd").read_bytes()


def test_vtu_structure(tmp_path):
    """The appended payloads sit at the offsets the tags give; the polyhedron faces of every cell list its face loops
    (outward) and its connectivity is the sorted set of those nodes; the cell arrays are stored as given."""
    from mpcns_uns_amd import Mesh

    m = Mesh.shell(4, 8, 2)
    cell, point = _arrays(m, np.random.default_rng(5))
    p = tmp_path / "s.vtu"
    m.write_vtu(p, cell, point)
    raw = p.read_bytes()
    head, _, tail = raw.partition(b'<AppendedData encoding="raw">\n_')
    tags = re.findall(rb'type="(\w+)" Name="(\w+)" NumberOfComponents="(\d+)" format="appended" offset="(\d+)"', head)
    payload = {}
    for typ, name, nc, off in tags:
        off = int(off)
        n = int.from_bytes(tail[off:off + 8], "little")
        dt = {b"Float64": np.float64, b"Int64": np.int64, b"UInt8": np.uint8}[typ]
        payload[name.decode()] = np.frombuffer(tail[off + 8:off + 8 + n], dtype=dt)
    assert np.array_equal(payload["Points"], m.array("xyz"))
    assert np.array_equal(payload["U_H"], cell[0][2]) and np.array_equal(payload["phi"], point[0][2])
    assert np.all(payload["types"] == 42) and payload["offsets"][-1] == payload["connectivity"].size
    tets = m.array("tets").reshape(-1, 4)
    conn, offs = payload["connectivity"], np.concatenate([[0], payload["offsets"]])
    faces, foffs = payload["faces"], np.concatenate([[0], payload["faceoffsets"]])
    xyz = m.array("xyz").reshape(-1, 3)
    cc = m.array("cell_center").reshape(-1, 3)
    for c in range(m.n_cells):
        assert sorted(conn[offs[c]:offs[c + 1]]) == sorted(tets[c])
        f = faces[foffs[c]:foffs[c + 1]]
        assert f[0] == 4
        k = 1
        for _ in range(4):
            n = f[k]
            loop = f[k + 1:k + 1 + n]
            k += 1 + n
            a, b, d = xyz[loop[0]], xyz[loop[1]], xyz[loop[2]]
            normal = np.cross(b - a, d - a)
            assert normal @ (a - cc[c]) > 0  # outward


def test_solution_files_match_the_reference_solver(tmp_path):
    """mercury::Solver::write_vtk after 2 steps against SolutionWriter fed with the oracle's prepared fields (the oracle
    reproduces the solver bit for bit): the same .vtu, .pvtu and .pvd bytes."""
    _ref()
    from mpcns_uns_amd import workloads
    from mpcns_uns_amd.output import SolutionWriter, solution_fields
    from oracle.oracle import Oracle, reference_for_mesh

    w = workloads.build(net=(6, 12, 2), hall=False)
    m = w["mesh"]
    ref = reference_for_mesh(m, w["cfg"], dipole_m=workloads.DIPOLE_M)
    ref.set_output_dir(tmp_path / "ref")
    ref.solver_init()
    ref.set_state(*w["state"])
    orc = Oracle(m, w["cfg"])
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    out = SolutionWriter(tmp_path / "ours")
    t = 0.0
    for step in (1, 2):
        dt = ref.step()
        assert dt == orc.step()[0]
        t += dt  # Solver::time_ accumulates the same steps in the same order
        ref.write_vtk(tmp_path / "ref")
        orc.prepare()
        out.write(m, solution_fields(orc, m), step, t)
    for name in ("sol_0000001_r00000.vtu", "sol_0000002_r00000.vtu", "sol_0000002.pvtu", "solution.pvd"):
        assert (tmp_path / "ours" / name).read_bytes() == (tmp_path / "ref" / name).read_bytes(), name
