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


def _records(raw):
    """{name: (dtype, cols, rows, payload offset)} of an UNSBIN file's records."""
    n, off, out = int(np.frombuffer(raw[8:16], np.uint64)[0]), 16, {}
    size = {0: 1, 1: 4, 2: 8, 3: 4, 4: 8}
    for _ in range(n):
        ln = int(np.frombuffer(raw[off:off + 4], np.uint32)[0])
        name = bytes(raw[off + 4:off + 4 + ln]).decode()
        off += 4 + ln
        dt, cols, rows = raw[off], int(np.frombuffer(raw[off + 4:off + 8], np.uint32)[0]), int(np.frombuffer(raw[off + 8:off + 16], np.uint64)[0])
        off += 16
        out[name] = (dt, cols, rows, off)
        nbytes = rows * cols * size[dt]
        off += nbytes + (8 - nbytes % 8) % 8
    return out


def _reference_parts(tmp_path, nranks):
    """Part files of the shell workload written by the reference's own preprocessing steps (build_mesh,
    partition_cells, Distributor::build, write_local_mesh) through oracle/_ref/ref_mpi_bench."""
    import subprocess

    from mpcns_uns_amd import workloads
    from oracle.cpu_baseline import MPI_BENCH, write_mpi_input

    if not os.path.exists(MPI_BENCH):
        pytest.skip("oracle/_ref/ref_mpi_bench not built")
    w = workloads.build(net=(8, 16, 3), hall=True)
    d = tmp_path / f"np{nranks}"
    d.mkdir()
    write_mpi_input(str(d / "input.bin"), w, workloads.DIPOLE_M)
    subprocess.run([MPI_BENCH, "parts", str(d), str(nranks)], check=True, capture_output=True, timeout=300)
    return w["mesh"], [d / f"part_{k:05d}.unsb" for k in range(nranks)]


@pytest.mark.parametrize("nranks", [1, 3, 4])
def test_partition_merges_to_the_global_complex(tmp_path, nranks):
    """The part files of an n-rank partition, merged through their global ids, give the complex the reference
    partitioned: every table, sign, tag, coordinate and the recomputed geometry bit for bit."""
    from mpcns_uns_amd import Mesh

    m, paths = _reference_parts(tmp_path, nranks)
    for order in (paths, paths[::-1]):
        g = Mesh.read_unsbin_parts(order)
        assert (g.n_nodes, g.n_edges, g.n_faces, g.n_cells) == (m.n_nodes, m.n_edges, m.n_faces, m.n_cells)
        for a in ARRAYS + ("xyz",):
            assert np.array_equal(g.array(a), m.array(a)), a
        assert g.topology_hash == m.topology_hash and g.tag_names == ["", "inner", "outer"]
    if nranks == 1:
        assert Mesh.read_unsbin(paths[0]).topology_hash == m.topology_hash
    else:
        with pytest.raises(RuntimeError, match="partition"):
            Mesh.read_unsbin(paths[0])


def test_partition_refusals(tmp_path):
    from mpcns_uns_amd import Mesh

    _, paths = _reference_parts(tmp_path, 3)
    with pytest.raises(RuntimeError, match="3-rank partition, 2 files given"):
        Mesh.read_unsbin_parts(paths[:2])
    with pytest.raises(RuntimeError, match="once each"):
        Mesh.read_unsbin_parts([paths[0], paths[0], paths[2]])
    with pytest.raises(RuntimeError, match="no part files"):
        Mesh.read_unsbin_parts([])
    raw = bytearray(paths[1].read_bytes())
    rec = _records(raw)
    meta = np.frombuffer(bytes(raw[rec["meta"][3]:rec["meta"][3] + 96]), np.int64)
    n_nodes, owned_nodes = int(meta[3]), int(meta[7])
    assert n_nodes > owned_nodes  # the part has ghost nodes
    bad = tmp_path / "bad"
    bad.mkdir()
    # a ghost copy of a node moved: the owner's coordinates no longer match
    x = bytearray(raw)
    o = rec["x"][3] + 8 * 3 * (n_nodes - 1)
    x[o:o + 8] = (np.frombuffer(bytes(raw[o:o + 8]), np.float64) + 1e-9).tobytes()
    (bad / "x.unsb").write_bytes(x)
    with pytest.raises(RuntimeError, match="node coordinates"):
        Mesh.read_unsbin_parts([paths[0], bad / "x.unsb", paths[2]])
    # an owned node given the global id of another part's owned node: two owners
    g0 = _records(bytearray(paths[0].read_bytes()))["ent0.gid"][3]
    other = int(np.frombuffer(paths[0].read_bytes()[g0:g0 + 8], np.int64)[0])
    dup = bytearray(raw)
    o = rec["ent0.gid"][3]
    dup[o:o + 8] = np.int64(other).tobytes()
    (bad / "dup.unsb").write_bytes(dup)
    with pytest.raises(RuntimeError, match="owned once each"):
        Mesh.read_unsbin_parts([paths[0], bad / "dup.unsb", paths[2]])
    # a row of a part edited: the stored topology hash catches it
    e = bytearray(raw)
    o = rec["e2n.sgn"][3]
    e[o] = (e[o] ^ 0xFE) & 0xFF
    (bad / "sgn.unsb").write_bytes(e)
    with pytest.raises(RuntimeError, match="hash"):
        Mesh.read_unsbin_parts([paths[0], bad / "sgn.unsb", paths[2]])
