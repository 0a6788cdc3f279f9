"""Crustal field of point dipoles (MarsUns/physics/MarsDipoles.cpp:63-107, MarsSolver.cpp:169-200).

Edge integrals bit for bit against the compiled reference's MARS::DipoleModel (random dipoles, cores, cutoff); the
background flux is an exact d1 of edge values (div B at round-off); far from a single dipole the reconstructed cell
field is the analytic dipole field.
"""
import ctypes as C

import numpy as np
import pytest


def _dipoles(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    pos = 0.98 * d / np.linalg.norm(d, axis=1)[:, None]
    mom = rng.normal(size=(n, 3)) * 1e-3
    core = rng.uniform(0, 0.02, n)
    return pos, mom, core


@pytest.mark.parametrize("cutoff", [0.0, 0.5])
def test_edge_integrals_match_the_reference(cutoff):
    from oracle.oracle import Reference, have_ref

    if not have_ref():
        pytest.skip("reference library oracle/_ref not built")
    from mpcns_uns_amd.mesh import dipole_edge_integrals

    pos, mom, core = _dipoles(100, 3)
    rng = np.random.default_rng(4)
    xa = rng.normal(size=(500, 3))
    xa *= (1.0 + rng.uniform(0, 2, 500) / np.linalg.norm(xa, axis=1))[:, None]
    xb = xa + rng.normal(scale=0.05, size=(500, 3))
    xb[0] = xa[0]  # a zero-length edge
    ours = dipole_edge_integrals(pos, mom, core, xa, xb, cutoff)
    ref = np.zeros(500)
    L = Reference.lib()
    f = lambda a: np.ascontiguousarray(a, dtype=float)
    assert L.ref_dipole_edge_integrals(100, f(pos).ctypes.data, f(mom).ctypes.data, f(core).ctypes.data, cutoff, 500,
                                       f(xa).ctypes.data, f(xb).ctypes.data, ref.ctypes.data) == 0
    assert np.array_equal(ours, ref) and ours[0] == 0.0 and np.abs(ours).max() > 0


def test_background_is_divergence_free_and_converges_to_the_dipole_field():
    from mpcns_uns_amd import Mesh

    pos, mom = np.array([[0.0, 0.0, 0.3]]), np.array([[0.02, -0.01, 0.05]])
    errs = []
    for net in [(16, 32, 4), (32, 64, 8)]:
        m = Mesh.shell(*net)
        PhiAdd, Badd = m.background_dipoles((0.0, 0.0, -0.05), pos, mom, np.zeros(1))
        c2f, c2s = m.array("c2f.idx").reshape(-1, 4), m.array("c2f.sgn").reshape(-1, 4)
        assert np.abs((c2s * PhiAdd[c2f]).sum(1)).max() < 1e-13 * np.abs(PhiAdd).max()  # d2 d1 = 0
        cc = m.array("cell_center").reshape(-1, 3)
        r = cc - pos[0]
        d = np.linalg.norm(r, axis=1)
        rh = r / d[:, None]
        dip = (3 * rh * (rh @ mom[0])[:, None] - mom[0]) / d[:, None] ** 3
        sel = d > 2.5
        got = Badd.reshape(-1, 3)[sel] - np.array([0.0, 0.0, -0.05])  # the IMF part is exact
        errs.append(np.median(np.linalg.norm(got - dip[sel], axis=1) / np.linalg.norm(dip[sel], axis=1)))
    # first-order convergence of the reconstructed cell field of the dipole flux
    assert errs[1] < 0.12 and errs[0] / errs[1] > 1.6
