"""Tangent-linear Hall kernel and Newton-Krylov on the nonlinear implicit induction (no reference counterpart).

G(phi) is the induction operator with the Hall field B_add + R phi of its own argument, quadratic in phi, so a
central difference of G is exact up to rounding and pins the hand-written tangent kernel; at phi = 0 with B_add as
the frozen field the tangent is the semi-implicit operator itself (bit for bit).  The Newton solution is checked by an
independent evaluation of the nonlinear residual with the face flux computed on the host.
"""
import numpy as np
import pytest

from conftest import maxrel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(small_hall):
    from mpcns_uns_amd import Device
    from oracle.oracle import Oracle

    w = small_hall
    orc = Oracle(w["mesh"], w["cfg"])
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    orc.step()
    WH, WNa = orc.derived("W_H"), orc.derived("W_Na")
    ne_eff = np.sqrt((WH[0::5] + WNa[0::5] / 23.0) ** 2 + 1e-4)
    dev = Device(w["mesh"], w["cfg"])
    dev.set_background(*w["background"])
    dev.ind_set_frozen(w["background"][1], ne_eff)  # frozen field = B_add: the semi-implicit operator at phi = 0
    return dev, w


def d1(m, E):
    f2e, f2s = m.array("f2e.idx").reshape(-1, 3), m.array("f2e.sgn").reshape(-1, 3)
    return (f2s * E[f2e]).sum(1)


def test_tangent_against_exact_central_difference(setup):
    dev, w = setup
    m = w["mesh"]
    rng = np.random.default_rng(21)
    phi = w["state"][2] + d1(m, rng.normal(size=m.n_edges)) * 1e-2
    v = d1(m, rng.normal(size=m.n_edges))
    h = 1e-3 * np.linalg.norm(phi) / np.linalg.norm(v)
    fd = (dev.ind_apply_G(phi + h * v) - dev.ind_apply_G(phi - h * v)) / (2 * h)
    jv = dev.ind_jvp(phi, v)
    assert np.abs(jv).max() > 0
    assert maxrel(jv, fd) < 1e-8
    u = d1(m, rng.normal(size=m.n_edges))
    assert maxrel(dev.ind_jvp(phi, 2.0 * v + u), 2.0 * jv + dev.ind_jvp(phi, u)) < 1e-12


def test_tangent_at_zero_is_the_semi_implicit_operator(setup):
    dev, w = setup
    m = w["mesh"]
    v = d1(m, np.random.default_rng(22).normal(size=m.n_edges))
    assert np.array_equal(dev.ind_jvp(np.zeros(m.n_faces), v), dev.ind_apply_F(v))


def test_newton_krylov(setup):
    dev, w = setup
    m = w["mesh"]
    rng = np.random.default_rng(23)
    Phi = w["state"][2]
    E_expl = rng.normal(size=m.n_edges) * 1e-2
    dirichlet = (m.array("edge_mode") != 0).astype(np.int8)
    dt = 5e-3
    res = [dev.ind_newton(Phi, E_expl, dirichlet, dt, rtol=1e-12, max_newton=k, eta=1e-10)[3] for k in range(4)]
    assert res[1] < 1e-2 * res[0] and res[2] < 1e-2 * res[1]  # superlinear: each step squares the error
    E, k, gm, rr = dev.ind_newton(Phi, E_expl, dirichlet, dt, rtol=1e-10)
    assert rr < 1e-10 and k <= 6 and gm > 0
    R = E - E_expl - dev.ind_apply_G(Phi - dt * d1(m, E))
    R[dirichlet != 0] = 0.0
    b = E_expl + dev.ind_apply_G(Phi)
    assert np.linalg.norm(R) / np.linalg.norm(b) < 1e-9
    assert np.array_equal(E[dirichlet != 0], E_expl[dirichlet != 0])
