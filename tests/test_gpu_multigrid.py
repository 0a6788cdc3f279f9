"""The multigrid preconditioner on the device: one V-cycle against the dense float64 restatement of its stage
matrices (oracle/multigrid.py), and the flexible mixed-precision solve it preconditions (precision 2) against the
FP64 solve and the oracle.
"""
import numpy as np
import pytest

from conftest import maxrel

pytestmark = pytest.mark.gpu


def _frozen(m, seed=31):
    rng = np.random.default_rng(seed)
    B0 = np.tile([0.1, -0.2, 0.3], m.n_cells) + 0.05 * rng.normal(size=3 * m.n_cells)
    ne_eff = 1.0 + 0.1 * rng.uniform(size=m.n_cells)
    return B0, ne_eff


@pytest.mark.parametrize("cells", ["shell", "hex_pyramid"])
@pytest.mark.parametrize("fine_hall", [1, 0])
def test_vcycle_matches_stage_matrices(fine_hall, cells):
    """On a small tetrahedral shell and on a box of hexahedra and pyramids (cells of five and six faces)."""
    from meshes import mixed_workload
    from mpcns_uns_amd import Device, workloads
    from oracle.multigrid import operator_F, vcycle
    from oracle.oracle import Oracle

    if cells == "shell":
        w = workloads.build(net=(6, 12, 3), hall=True, mg_fine_hall=fine_hall)
    else:
        w = mixed_workload(True, nx=8, ny=6, nz=6, mg_fine_hall=fine_hall)
    m, cfg = w["mesh"], w["cfg"]
    B0, ne_eff = _frozen(m)
    dirichlet = m.array("edge_mode") != 0
    dt = 2e-3
    orc = Oracle(m, cfg)
    orc.ind_set_frozen(B0, ne_eff)
    res = type(cfg).from_buffer_copy(cfg)
    res.hall = 0
    K = operator_F(m, Oracle(m, res).ind_apply_F)
    F = operator_F(m, orc.ind_apply_F) if fine_hall else K
    A0 = np.eye(m.n_edges) + dt * F
    invdiag = 1.0 / (1.0 + dt * np.diag(K))  # 1 / (1 + dt m_e d2_e): the resistive diagonal
    levels = m.mg_hierarchy(cfg, dirichlet)
    dev = Device(m, cfg)
    dev.ind_set_frozen(B0, ne_eff)
    assert dev.mg_info(dirichlet) == [lv["n"] for lv in levels]
    rng = np.random.default_rng(5)
    for _ in range(2):
        v = rng.normal(size=m.n_edges)
        z = dev.mg_apply(v, dirichlet, dt)
        ref = vcycle(A0, invdiag, levels, dt, v, dirichlet, omega=cfg.mg_relax, pre=cfg.mg_pre, post=cfg.mg_post,
                     coarse=cfg.mg_coarse)
        assert np.all(z[dirichlet] == 0)
        assert maxrel(z, ref) < 2e-5, maxrel(z, ref)
    # linear in v (up to FP32 rounding)
    v1, v2 = rng.normal(size=m.n_edges), rng.normal(size=m.n_edges)
    assert maxrel(dev.mg_apply(v1 + 2 * v2, dirichlet, dt), dev.mg_apply(v1, dirichlet, dt) + 2 * dev.mg_apply(v2, dirichlet, dt)) < 1e-5


def test_mg_solve_matches_fp64_and_oracle(small_hall):
    from mpcns_uns_amd import Device
    from oracle.oracle import Oracle

    w = small_hall
    m = w["mesh"]
    B0, ne_eff = _frozen(m)
    dirichlet = (m.array("edge_mode") != 0).astype(np.int8)
    rng = np.random.default_rng(8)
    E_expl = rng.normal(size=m.n_edges) * 1e-3
    out = {}
    for prec in (0, 1, 2):
        dev = Device(m, w["cfg"])
        dev.ind_set_frozen(B0, ne_eff)
        out[prec] = dev.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3, precision=prec)
    orc = Oracle(m, w["cfg"])
    orc.ind_set_frozen(B0, ne_eff)
    Eo, ito, ro = orc.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3)
    for prec in (0, 1, 2):
        E, it, rr = out[prec]
        print(f"precision {prec}: {it} iterations, residual {rr:.2e}, vs oracle {maxrel(E, Eo):.1e}")
        assert rr <= 1e-8 and it > 0
        assert maxrel(E, Eo) < 1e-6
    # the fused step with the multigrid solve against the FP64 step
    dev = Device(m, w["cfg"])
    for x in (dev,):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    ref = Device(m, w["cfg"])
    ref.set_background(*w["background"])
    ref.set_state(*w["state"])
    for k in range(3):
        dt2, it2, r2 = dev.step(precision=2)
        dt0, it0, r0 = ref.step(precision=0)
        assert r2 <= 1e-8 and r0 <= 1e-8
        # the first step's dt comes from the same state; later ones from states that differ within the solver tolerance
        assert dt2 == dt0 if k == 0 else abs(dt2 - dt0) <= 1e-8 * dt0
    # three steps, each solve stopped at a 1e-8 relative residual: the differences add up over the steps
    for a, b in zip(dev.get_state(), ref.get_state()):
        assert maxrel(a, b) < 3e-8
    _, rel, _ = dev.div_b(dev.get_state()[2], w["background"][0])
    assert rel < 1e-12


def test_mg_refuses_what_it_cannot_serve(small_hall):
    from mpcns_uns_amd import Device

    w = small_hall
    m = w["mesh"]
    dev = Device(m, w["cfg"])
    dev.ind_set_frozen(*_frozen(m))
    d = (m.array("edge_mode") != 0).astype(np.int8)
    with pytest.raises(RuntimeError, match="per-edge"):
        dev.ind_solve(w["state"][2], np.zeros(m.n_edges), d, 2e-3, precision=2, dt_edge=np.full(m.n_edges, 2e-3))
    with pytest.raises(RuntimeError, match="precision"):
        dev.ind_solve(w["state"][2], np.zeros(m.n_edges), d, 2e-3, precision=4)
