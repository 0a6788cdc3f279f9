"""Tangent-linear (forward mode) of the explicit step: dual-number instantiation of the explicit kernels.

Objective: the magnetic energy J = sum |B|^2 V over the shell 1.2 <= r <= 1.5 after 5 explicit steps, as a function of
the upstream IMF (3 components: background flux and field, imposed inflow EMF) and the upstream H+ density (inflow
states and the initial density at fixed velocity and pressure).  Checked against central finite differences with step
1e-6 (each evaluation a fresh device run of the double kernels), for linearity in the seed and for the value.  (The
torch.autograd.Function over the step, whose backward is the adjoint, is in tests/test_gpu_adjoint.py.)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R0, R1, STEPS = 1.2, 1.5, 5


@pytest.fixture(scope="module")
def problem():
    from mpcns_uns_amd import workloads

    return workloads.build(net=(12, 48, 6), hall=False)


def _inputs(w, p):
    """Background, state and configuration at p = (B_imf x, y, z, rho_H upstream)."""
    from mpcns_uns_amd import workloads

    cfg = type(w["cfg"]).from_buffer_copy(w["cfg"])
    cfg.B_imf[0], cfg.B_imf[1], cfg.B_imf[2] = p[0], p[1], p[2]
    rho0 = w["cfg"].fs[0][0]
    cfg.fs[0][0] = p[3]
    m = w["mesh"]
    PhiAdd, Badd = m.background(tuple(p[:3]), workloads.DIPOLE_M, (0.0, 0.0, 0.0), 1e-3)
    UH, UNa, Phi = (x.copy() for x in w["state"])
    U = UH.reshape(-1, 5)
    ke = 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / U[:, 0]
    s = p[3] / rho0
    U[:, 4] = (U[:, 4] - ke) + s * ke
    U[:, :4] *= s
    return cfg, (PhiAdd, Badd), (UH, UNa, Phi)


def energy(w, p):
    from mpcns_uns_amd import Device

    cfg, bg, st = _inputs(w, p)
    dev = Device(w["mesh"], cfg)
    dev.set_background(*bg)
    dev.set_state(*st)
    for _ in range(STEPS):
        dev.step()
    return dev.shell_energy(R0, R1)


def tangent(w, p, k):
    """(J, dJ/dp_k) by one forward-mode run."""
    from mpcns_uns_amd import Device

    cfg, bg, st = _inputs(w, p)
    m = w["mesh"]
    dev = Device(m, cfg)
    dev.set_background(*bg)
    dev.set_state(*st)
    kw = {}
    if k < 3:
        dPhiAdd = np.ascontiguousarray(m.array("face_area").reshape(-1, 3)[:, k])
        e = np.zeros(3)
        e[k] = 1.0
        kw = dict(dPhiAdd=dPhiAdd, dBadd=m.cell_vector(dPhiAdd), dE_inf=-np.cross(np.array(cfg.u_inf[:]), e))
    else:
        U = st[0].reshape(-1, 5)
        ke = 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / U[:, 0]
        dU = np.column_stack([U[:, 0], U[:, 1], U[:, 2], U[:, 3], ke]) / p[3]
        dfs = np.zeros(10)
        dfs[0] = 1.0
        kw = dict(dU_H=dU.ravel(), dfs=dfs)
    return dev.explicit_tangent(STEPS, R0, R1, **kw)


def p0(w):
    return np.array([*w["cfg"].B_imf[:], w["cfg"].fs[0][0]])


def test_tangent_value_is_the_double_run(problem):
    p = p0(problem)
    J, _ = tangent(problem, p, 0)
    assert J > 0 and J == energy(problem, p)


@pytest.mark.parametrize("k", [0, 1, 2, 3])
def test_tangent_against_central_differences(problem, k):
    p = p0(problem)
    h = 1e-6
    e = np.zeros(4)
    e[k] = h
    fd = (energy(problem, p + e) - energy(problem, p - e)) / (2 * h)
    _, dJ = tangent(problem, p, k)
    assert abs(dJ) > 0
    assert abs(dJ - fd) <= 1e-5 * abs(dJ) + 1e-9 * energy(problem, p)


def test_tangent_is_linear_in_the_seed(problem):
    from mpcns_uns_amd import Device

    w = problem
    m = w["mesh"]
    dev = Device(m, w["cfg"])
    dev.set_background(*w["background"])
    dev.set_state(*w["state"])
    rng = np.random.default_rng(31)
    a, b = rng.normal(size=5 * m.n_cells) * 1e-3, rng.normal(size=m.n_faces) * 1e-3
    _, d1 = dev.explicit_tangent(STEPS, R0, R1, dU_H=a)
    _, d2 = dev.explicit_tangent(STEPS, R0, R1, dPhi=b)
    _, d12 = dev.explicit_tangent(STEPS, R0, R1, dU_H=2 * a, dPhi=-3 * b)
    assert abs(d12 - (2 * d1 - 3 * d2)) <= 1e-10 * (abs(2 * d1) + abs(3 * d2))
