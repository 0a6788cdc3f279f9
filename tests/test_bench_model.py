"""bench.py's byte models, checked on the CPU: the whole-step compulsory-byte model against its terms summed by hand
for the 8M mixed first step (40 iterations in cycles of 30 and 10), and the per-launch algorithmic bytes of the FP32
edge kernel against its arrays counted on a small shell.
"""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_step_model_mixed_first_step(bench):
    # explicit 2 ions: 80*2 + 437; 40 FP32 products at 190; FP64 products: b, r0 and one residual per cycle (2);
    # Arnoldi steps j = 0..29 and 0..9 at (2j + 5) 9.33 / 2; Jacobi 28 per iteration
    arnoldi = (sum(2 * j + 5 for j in range(30)) + sum(2 * j + 5 for j in range(10))) * 9.33 / 2
    expect = 597 + 40 * 190 + 4 * 329 + arnoldi + 40 * 28
    assert bench.step_compulsory_bytes(40.0, 30, 2, True) == pytest.approx(expect, rel=1e-12)
    assert 15900 < expect < 16200


def test_step_model_fp64_and_vcycle(bench):
    fp64 = bench.step_compulsory_bytes(18.0, 30, 2, False)
    expect = 597 + 18 * 329 + 3 * 329 + sum(2 * j + 5 for j in range(18)) * 9.33 + 18 * 28
    assert fp64 == pytest.approx(expect, rel=1e-12)
    # a V-cycle per iteration adds its 1.3 kB; an exact multiple of the restart gives whole cycles only
    assert bench.step_compulsory_bytes(30.0, 30, 2, True, True) - bench.step_compulsory_bytes(30.0, 30, 2, True) == pytest.approx(30 * 1300.0)
    assert bench.step_compulsory_bytes(60.0, 30, 10, True) == pytest.approx(
        80 * 10 + 437 + 60 * 190 + 4 * 329 + 2 * sum(2 * j + 5 for j in range(30)) * 9.33 / 2 + 60 * 28)


def test_edge_kernel_bytes_count_its_arrays(bench):
    from mpcns_uns_amd.mesh import Mesh

    m = Mesh.shell(12, 24, 3)
    C, F, E = m.n_cells, m.n_faces, m.n_edges
    # the row lengths from the mesh's own tables (the model takes three edges per face)
    nf2e, nec = len(m.array("e2f.idx")), len(m.array("ec_cell"))
    assert nf2e == 3 * F
    # dir 1, two row-pointer arrays, (index, weight) pairs of both rows, phi, edge vector, g, m_e, pv, v, out
    expect = E + 2 * 4 * (E + 1) + 8 * nf2e + 8 * nec + 4 * F + 12 * E + 12 * C + 4 * E * 4
    assert bench.kernel_bytes(m, "edge_apply_f32") == expect
