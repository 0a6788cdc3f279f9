"""Why the Hall workloads hold the wall's EMF at zero in the implicit solve, and why the benchmark times first steps.

With Mercury's Dirichlet rule (only the imposed inflow edges, MercuryUns/driver/Solver.cpp:261-262) the compiled
reference's own step_first_order grows the induced field next to the south pole at the inner wall of a small shell;
the oracle reproduces that trajectory bit for bit, so the growth is the reference algorithm's, not a port error.  With
MarsSolver's rule (the conducting wall's edges Dirichlet as well, MarsUns/driver/MarsSolver.cpp:730; the
ImplicitInduction::solve called with that set is pinned bitwise in test_oracle_reference.py) the same run stays flat.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["have_ref"]).have_ref(),
                                reason="reference library oracle/_ref not built")

NET, STEPS = (24, 48, 6), 10


def _run(x, derived):
    peak = []
    for _ in range(STEPS):
        x.step()
        peak.append(float(np.abs(derived()).max()))
    return np.array(peak)


def test_mercury_rule_grows_in_the_reference_and_the_oracle():
    from mpcns_uns_amd import workloads
    from oracle.oracle import Oracle, reference_for_mesh

    w = workloads.build(net=NET, hall=True, wall_dirichlet=0)
    ref = reference_for_mesh(w["mesh"], w["cfg"], dipole_m=workloads.DIPOLE_M)
    ref.solver_init()
    ref.set_state(*w["state"])
    orc = Oracle(w["mesh"], w["cfg"])
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    peak_r, peak_o = [], []
    for _ in range(STEPS):
        assert ref.step() == orc.step()[0]
        for a, b in zip(ref.get_state(), orc.get_state()):
            assert np.array_equal(a, b)
        peak_r.append(float(np.abs(ref.derived("B_ind")).max()))
        peak_o.append(float(np.abs(orc.derived("B_ind")).max()))
    assert peak_r == peak_o
    growth = peak_r[-1] / min(peak_r)
    print(f"Mercury rule: max |B_ind| {min(peak_r):.4f} -> {peak_r[-1]:.4f} over {STEPS} steps (x{growth:.2f})")
    assert growth > 1.3
    # where: the inner wall next to the south pole
    cc = w["mesh"].array("cell_center").reshape(-1, 3)
    i = int(np.argmax(np.abs(orc.derived("B_ind")).reshape(-1, 3).max(axis=1)))
    r = np.linalg.norm(cc[i])
    assert r < 1.2 and np.degrees(np.arccos(cc[i, 2] / r)) > 160.0


def test_mars_rule_stays_flat():
    from mpcns_uns_amd import workloads
    from oracle.oracle import Oracle

    w = workloads.build(net=NET, hall=True, wall_dirichlet=1)
    orc = Oracle(w["mesh"], w["cfg"])
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    peak = _run(orc, lambda: orc.derived("B_ind"))
    print(f"Mars rule: max |B_ind| {peak.min():.4f} -> {peak[-1]:.4f} over {STEPS} steps")
    assert peak.max() / peak.min() < 1.05
