"""UNSBIN mesh and field files (the reference's src/uns/io/{BinContainer,MeshIO}.cpp format).

Round trips need no reference.  Against the compiled reference (oracle/_ref): the two writers produce byte-identical
complex and field files for the same mesh and data, each side reads the other's files, and the topology hashes agree.
"""
import ctypes as C
import os

import numpy as np
import pytest

from meshes import box_mesh

ARRAYS = ("e2n.ptr", "e2n.idx", "e2n.sgn", "f2e.ptr", "f2e.idx", "f2e.sgn", "c2f.ptr", "c2f.idx", "c2f.sgn", "e2f.idx",
          "e2f.sgn", "f2c.idx", "f2c.sgn", "face_tag", "material", "face_area", "cell_volume", "ec_weight", "ef_weight", "star1",
          "star2")


def _ref():
    from oracle.oracle import have_ref

    if not have_ref():
        pytest.skip("reference library oracle/_ref not built")


def test_mesh_round_trip(tmp_path):
    from mpcns_uns_amd import Mesh

    m = Mesh.shell(8, 16, 3)
    p = tmp_path / "shell.unsb"
    m.write_unsbin(p)
    r = Mesh.read_unsbin(p)
    assert (r.n_nodes, r.n_edges, r.n_faces, r.n_cells) == (m.n_nodes, m.n_edges, m.n_faces, m.n_cells)
    for a in ARRAYS:
        assert np.array_equal(r.array(a), m.array(a)), a
    assert r.topology_hash == m.topology_hash and r.tag_names == ["", "inner", "outer"]
    # the box (hexahedra) too, with its own tag names
    b = box_mesh(3, 2, 2)
    b.tag_names = ["", "xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
    b.write_unsbin(tmp_path / "box.unsb")
    rb = Mesh.read_unsbin(tmp_path / "box.unsb")
    for a in ARRAYS:
        assert np.array_equal(rb.array(a), b.array(a)), a
    assert rb.tag_names == b.tag_names


def test_bad_files_are_refused(tmp_path):
    from mpcns_uns_amd import Mesh

    m = Mesh.shell(6, 12, 2)
    p = tmp_path / "m.unsb"
    m.write_unsbin(p)
    raw = bytearray(p.read_bytes())
    (tmp_path / "trunc.unsb").write_bytes(raw[: len(raw) // 2])
    with pytest.raises(RuntimeError, match="truncated"):
        Mesh.read_unsbin(tmp_path / "trunc.unsb")
    bad = bytearray(raw)
    bad[0] = ord("X")
    (tmp_path / "magic.unsb").write_bytes(bad)
    with pytest.raises(RuntimeError, match="not an UNSBIN"):
        Mesh.read_unsbin(tmp_path / "magic.unsb")
    # meta record comes first: u32 len(4) "meta" u8 type u8[3] u32 cols u64 rows, then the values
    off = 16 + 4 + 4 + 4 + 4 + 8
    meta = np.frombuffer(bytes(raw[off:off + 96]), np.int64).copy()
    two = bytearray(raw)
    meta2 = meta.copy()
    meta2[2] = 2
    two[off:off + 96] = meta2.tobytes()
    (tmp_path / "part.unsb").write_bytes(two)
    with pytest.raises(RuntimeError, match="partition"):
        Mesh.read_unsbin(tmp_path / "part.unsb")
    h = bytearray(raw)
    meta3 = meta.copy()
    meta3[11] ^= 1
    h[off:off + 96] = meta3.tobytes()
    (tmp_path / "hash.unsb").write_bytes(h)
    with pytest.raises(RuntimeError, match="hash"):
        Mesh.read_unsbin(tmp_path / "hash.unsb")


def test_fields_round_trip(tmp_path):
    from mpcns_uns_amd import Mesh, read_fields, write_fields

    m = Mesh.shell(6, 12, 2)
    rng = np.random.default_rng(1)
    f = [("U_H", "cell", 5, rng.normal(size=5 * m.n_cells)), ("B", "face", 1, rng.normal(size=m.n_faces)), ("_time", "node", 1, [0.25])]
    write_fields(tmp_path / "f.unsb", f, m.topology_hash)
    h, got = read_fields(tmp_path / "f.unsb")
    assert h == m.topology_hash and list(got) == ["U_H", "B", "_time"]
    for name, dim, nc, v in f:
        assert got[name][0] == {"cell": 3, "face": 2, "node": 0}[dim] and got[name][1] == nc
        assert np.array_equal(got[name][2], np.asarray(v, dtype=float))


def test_complex_files_match_the_reference(tmp_path):
    _ref()
    from mpcns_uns_amd import workloads
    from mpcns_uns_amd import Mesh
    from oracle.oracle import Reference, reference_for_mesh

    w = workloads.build(net=(8, 16, 3), hall=False)
    m = w["mesh"]
    ref = reference_for_mesh(m, w["cfg"])
    L = Reference.lib()
    pr, po = tmp_path / "ref.unsb", tmp_path / "ours.unsb"
    assert L.ref_write_mesh(ref._h, str(pr).encode()) == 0
    m.write_unsbin(po)
    assert pr.read_bytes() == po.read_bytes()  # byte-identical complex files
    assert L.ref_topology_hash(ref._h) == m.topology_hash
    r = Mesh.read_unsbin(pr)
    for a in ARRAYS:
        assert np.array_equal(r.array(a), m.array(a)), a
    counts, h = np.zeros(4, np.int64), C.c_uint64()
    assert L.ref_read_mesh(str(po).encode(), counts.ctypes.data, C.byref(h)) == 0
    assert tuple(counts) == (m.n_nodes, m.n_edges, m.n_faces, m.n_cells) and h.value == m.topology_hash


def test_field_files_match_the_reference(tmp_path):
    _ref()
    from mpcns_uns_amd import Mesh, read_fields, write_fields
    from oracle.oracle import Reference

    L = Reference.lib()
    m = Mesh.shell(6, 12, 2)
    rng = np.random.default_rng(2)
    recs = [("U_H", 3, 5, rng.normal(size=5 * m.n_cells)), ("B", 2, 1, rng.normal(size=m.n_faces)), ("_time", 0, 1, np.array([1.5]))]
    po, pr = tmp_path / "ours.unsb", tmp_path / "ref.unsb"
    write_fields(po, recs, m.topology_hash)
    names = (C.c_char_p * 3)(*[r[0].encode() for r in recs])
    dims = np.array([r[1] for r in recs], np.int32)
    ncs = np.array([r[2] for r in recs], np.int32)
    vals = [np.ascontiguousarray(r[3], dtype=float) for r in recs]
    vp = (C.c_void_p * 3)(*[v.ctypes.data for v in vals])
    cnt = np.array([v.size for v in vals], np.int64)
    assert L.ref_write_fields(str(pr).encode(), C.c_uint64(m.topology_hash), 3, names, dims.ctypes.data, ncs.ctypes.data, vp,
                              cnt.ctypes.data) == 0
    assert pr.read_bytes() == po.read_bytes()
    h, got = read_fields(pr)
    assert h == m.topology_hash
    for name, dim, nc, v in recs:
        out = np.zeros(v.size)
        n, d, c, hh = C.c_int64(v.size), C.c_int32(), C.c_int32(), C.c_uint64()
        assert L.ref_read_field(str(po).encode(), name.encode(), out.ctypes.data, C.byref(n), C.byref(d), C.byref(c), C.byref(hh)) == 0
        assert np.array_equal(out, v) and (d.value, c.value, hh.value) == (dim, nc, m.topology_hash)
        assert np.array_equal(got[name][2], v)
