"""Host mesh layer invariants and the C-ABI library surface (no GPU needed).

Integer invariants of the complex are exact (d1 d0 = 0, d2 d1 = 0, Euler characteristic 2 of the shell,
every interior face shared by two cells with opposite signs); the metric ones are at round-off (closure,
sum of the dual weights, positive volumes and Hodge stars).
"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _csr(m, name):
    return m.array(name + ".ptr"), m.array(name + ".idx"), m.array(name + ".sgn")


def _dense_products_zero(m, upper, lower, n_lower_cols):
    p, i, s = _csr(m, upper)
    q, j, t = _csr(m, lower)
    import scipy.sparse as sp

    U = sp.csr_matrix((s.astype(np.int64), i, p))
    L = sp.csr_matrix((t.astype(np.int64), j, q), shape=(U.shape[1], n_lower_cols))
    P = (U @ L).tocsr()
    P.eliminate_zeros()
    return P.nnz == 0


@pytest.fixture(scope="module")
def shell():
    from mpcns_uns_amd import workloads

    return workloads.build(net=(10, 20, 4), hall=False)["mesh"]


def test_integer_invariants(shell):
    m = shell
    assert _dense_products_zero(m, "f2e", "e2n", m.n_nodes)
    assert _dense_products_zero(m, "c2f", "f2e", m.n_edges)
    assert m.n_nodes - m.n_edges + m.n_faces - m.n_cells == 2
    fp, _, fs = _csr(m, "f2c")
    cnt = np.diff(fp)
    assert set(np.unique(cnt)) == {1, 2}
    two = np.nonzero(cnt == 2)[0]
    assert np.all(fs[fp[two]] + fs[fp[two] + 1] == 0)
    tags = m.array("face_tag")
    assert set(np.unique(tags[cnt == 1])) == {1, 2} and np.all(tags[cnt == 2] == 0)


def test_metric_invariants(shell):
    m = shell
    c2f, c2s = m.array("c2f.idx").reshape(-1, 4), m.array("c2f.sgn").reshape(-1, 4)
    A = m.array("face_area").reshape(-1, 3)
    closure = np.abs((c2s[..., None] * A[c2f]).sum(1)).max() / np.abs(A).max()
    assert closure < 1e-12
    ecp, ecw = m.array("ec_ptr"), m.array("ec_weight")
    s = np.add.reduceat(ecw, ecp[:-1])
    assert np.abs(s - 1).max() < 1e-12
    assert m.array("cell_volume").min() > 0 and m.array("star1").min() > 0 and m.array("star2").min() > 0
    assert m.array("ec_misoriented").max() == 0
    # total volume of the planar-faced shell: below the sphere volume, within the polyhedral error
    vol = m.array("cell_volume").sum()
    assert 0.9 * (4 / 3) * np.pi * (4 ** 3 - 1) < vol < (4 / 3) * np.pi * (4 ** 3 - 1)


def test_sorted_builder_equals_incremental_builder():
    """The shell generator's sort-based complex builder numbers faces and edges exactly as the incremental one."""
    from mpcns_uns_amd import Mesh

    a = Mesh.shell(6, 12, 3)
    tets = a.array("tets").reshape(-1, 4)
    loops = [[(t0, t2, t1), (t0, t1, t3), (t0, t3, t2), (t1, t2, t3)] for t0, t1, t2, t3 in tets]
    ftag, f2e = a.array("face_tag"), a.array("f2e.idx").reshape(-1, 3)
    e2n = a.array("e2n.idx").reshape(-1, 2)
    tags = [(sorted(set(e2n[f2e[f]].ravel())), ftag[f]) for f in np.nonzero(ftag)[0]]
    b = Mesh.from_cells(a.array("xyz"), loops, tags=tags)
    for tb in ("e2n", "f2e", "c2f", "e2f", "f2c"):
        for part in ("ptr", "idx", "sgn"):
            assert np.array_equal(a.array(f"{tb}.{part}"), b.array(f"{tb}.{part}"))
    assert np.array_equal(a.array("face_tag"), b.array("face_tag"))
    for g in ("face_area", "cell_center", "ec_weight", "ef_weight", "star1", "star2"):
        assert np.array_equal(a.array(g), b.array(g))


def test_library_exports_every_declared_symbol():
    from mpcns_uns_amd._lib import LIB_PATH

    lib = ctypes.CDLL(LIB_PATH)
    names = set()
    for h in ("mhdg.h", "mhdg_mesh.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"^(?:int|void|const char \*)\s*(mhdg_[a-z0-9_]+)\(", text, re.MULTILINE))
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_create_reports_errors_without_device():
    """mhdg_create returns a non-zero status and a message when it cannot run (no device here, or bad input)."""
    import ctypes as C

    from mpcns_uns_amd import default_config
    from mpcns_uns_amd._lib import MeshDesc, lib
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("a device is present: creation succeeds")
    d, cfg, h = MeshDesc(), default_config(), C.c_void_p()
    assert lib().mhdg_create(C.byref(d), C.byref(cfg), C.byref(h)) != 0
    assert lib().mhdg_last_error(None)


def test_shim_compiles_against_reference_headers():
    """The C++ drop-in shim keeps the reference's mhd_common signatures: it compiles against the reference's own
    headers (where the reference sources are present)."""
    import shutil
    import subprocess

    ref = "/root/reference"
    if not os.path.isdir(os.path.join(ref, "Common")):
        pytest.skip("reference sources not present")
    if not shutil.which("g++"):
        pytest.skip("no g++")
    cmd = ["g++", "-std=c++20", "-fsyntax-only", f"-I{ref}/include", f"-I{ref}/Common", "-I/opt/conda/include",
           "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "mpcns_uns_amd", "shim", "mhd_gpu_shim.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.parametrize("defect, message", [
    ("f2e_index", "f2e.idx out of range"),
    ("c2f_sign", "c2f.sgn not"),
    ("e2f_ptr", "e2f.ptr not monotonic"),
    ("ec_cell", "ec_cell out of range"),
    ("null_area", "missing"),
    ("transpose", "transposes"),
])
def test_create_rejects_bad_descriptors(defect, message):
    """Every index a kernel follows is checked on the host before any device call: a corrupt descriptor is an
    error with a message, never an out-of-bounds access on the GPU (checked here without a GPU)."""
    import ctypes as C

    from mpcns_uns_amd import workloads
    from mpcns_uns_amd._lib import MeshDesc, lib

    w = workloads.build(net=(6, 12, 2), hall=False)
    m, cfg = w["mesh"], w["cfg"]
    d = MeshDesc.from_buffer_copy(m.desc())
    keep = []

    def patched(arr, dtype, ctype):
        a = np.array(arr, dtype=dtype)
        keep.append(a)
        return a, a.ctypes.data_as(C.POINTER(ctype))

    if defect == "f2e_index":
        a, p = patched(m.array("f2e.idx"), np.int32, C.c_int32)
        a[7] = m.n_edges
        d.f2e.idx = p
    elif defect == "c2f_sign":
        a, p = patched(m.array("c2f.sgn"), np.int8, C.c_int8)
        a[3] = 0
        d.c2f.sgn = p
    elif defect == "e2f_ptr":
        a, p = patched(m.array("e2f.ptr"), np.int64, C.c_int64)
        a[5], a[6] = a[6], a[5] - 1
        d.e2f.ptr = p
    elif defect == "ec_cell":
        a, p = patched(m.array("ec_cell"), np.int32, C.c_int32)
        a[-1] = -1
        d.ec_cell = p
    elif defect == "null_area":
        d.face_area = None
    elif defect == "transpose":
        a, p = patched(m.array("e2f.ptr"), np.int64, C.c_int64)
        a[-1] = a[-2]  # last edge loses its faces: e2f no longer matches f2e
        d.e2f.ptr = p
    h = C.c_void_p()
    assert lib().mhdg_create(C.byref(d), C.byref(cfg), C.byref(h)) != 0
    assert message in lib().mhdg_last_error(None).decode()


def test_converted_inputs_outlive_the_call():
    """ptr() of a converted temporary keeps the converted array alive (the pointer holds a reference), so a float32 or
    list input reaches C intact; Device checks every input's size before its pointer crosses the C-ABI."""
    import ctypes as C

    from mpcns_uns_amd import Device
    from mpcns_uns_amd._lib import f64, ptr

    q = ptr(f64([1.0, 2.0, 3.0, 4.0, 5.0]))
    junk = [np.full(5, 9.0) for _ in range(50)]  # reuses freed buffers of the same size, if any
    assert C.cast(q, C.POINTER(C.c_double))[0:5] == [1.0, 2.0, 3.0, 4.0, 5.0]
    del junk
    d = Device.__new__(Device)  # size checks only: no device context
    d.nc, d.nf, d.ne = 4, 6, 7
    assert d._in(np.arange(12, dtype=np.float32), "c3").dtype == np.float64
    assert d._in(list(range(7)), "e").tolist() == list(range(7))
    assert d._in(np.arange(24.0).reshape(6, 4)[:, 0], "f").flags["C_CONTIGUOUS"]
    assert d._in([True, False] * 3 + [True], "e", np.int8).tolist() == [1, 0, 1, 0, 1, 0, 1]
    with pytest.raises(ValueError):
        d._in(np.zeros(11), "c3")
    m = __import__("mpcns_uns_amd").Mesh.shell(6, 12, 2, seed=3)
    flux = np.random.default_rng(0).normal(size=m.n_faces)
    assert np.array_equal(m.cell_vector(flux.astype(np.float32)), m.cell_vector(flux.astype(np.float32).astype(np.float64)))
    assert np.array_equal(m.cell_vector(list(flux)), m.cell_vector(flux))
