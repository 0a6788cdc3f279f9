"""Device against the oracle at the workloads' own sizes where the oracle finishes in seconds: the explicit
two-ion step on shell_60k (59,616 tets) bit for bit, and the FP64 Hall step on shell_2m_hall (1.95M tets) to
1e-12 per field with div B at round-off, over consecutive steps.
"""
import numpy as np
import pytest

from conftest import maxrel

pytestmark = pytest.mark.gpu


def _pair(name):
    from mpcns_uns_amd import Device, workloads
    from oracle.oracle import Oracle

    print(f"\n[config size] building {name}", flush=True)
    w = workloads.build(name)
    dev, orc = Device(w["mesh"], w["cfg"]), Oracle(w["mesh"], w["cfg"])
    for x in (dev, orc):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    return w, dev, orc


def test_shell_60k_explicit_steps_bitwise():
    w, dev, orc = _pair("shell_60k")
    assert w["mesh"].n_cells == 59616
    for k in range(3):
        assert dev.step() == orc.step()
        for a, b in zip(dev.get_state(), orc.get_state()):
            assert np.array_equal(a, b)
        for n in ("W_H", "W_Na", "B_ind", "B_tot", "J", "u_plus", "E"):
            assert np.array_equal(dev.derived(n), orc.derived(n)), n
        print(f"[config size] shell_60k step {k + 1} bitwise", flush=True)


def test_shell_2m_hall_fp64_steps():
    w, dev, orc = _pair("shell_2m_hall")
    assert w["mesh"].n_cells == 1950720
    for k in range(3):
        dt_d, it_d, r_d = dev.step(precision=0)
        dt_o, it_o, r_o = orc.step()
        sd, so = dev.get_state(), orc.get_state()
        rel = [maxrel(a, b) for a, b in zip(sd, so)]
        _, div, _ = dev.div_b(sd[2], w["background"][0])
        print(f"[config size] shell_2m_hall step {k + 1}: dt {dt_d:.6e}/{dt_o:.6e}, GMRES {it_d}/{it_o}, residual "
              f"{r_d:.2e}/{r_o:.2e}, max rel (U_H, U_Na, Phi) " + ", ".join(f"{x:.1e}" for x in rel) + f", div B {div:.1e}",
              flush=True)
        assert abs(dt_d - dt_o) <= 1e-12 * dt_o
        assert r_d <= 1e-8 and r_o <= 1e-8 and abs(it_d - it_o) <= 1
        assert max(rel) < 1e-12
        assert div < 1e-12
