"""Branches the seeded states never take, forced from the host, on the device against the oracle (which matches the
compiled reference on the same inputs: tests/test_oracle_branches.py): floors, the face-state fallback, the invalid
time step, GMRES truncated at ind_max_it, and a GhostFn served through the drop-in shim.
"""
import numpy as np
import pytest

from conftest import maxrel
from meshes import fallback_faces, floored_state, rough_gradients

pytestmark = pytest.mark.gpu


def _pair(w, state=None):
    from mpcns_uns_amd import Device
    from oracle.oracle import Oracle

    dev, orc = Device(w["mesh"], w["cfg"]), Oracle(w["mesh"], w["cfg"])
    for x in (dev, orc):
        x.set_background(*w["background"])
        x.set_state(*(state or w["state"]))
    return dev, orc


def test_floors_forced_bitwise(small_explicit):
    w = small_explicit
    state, rho_cells, p_cells = floored_state(w)
    dev, orc = _pair(w, state)
    cfg = w["cfg"]
    for k in range(2):
        assert dev.step() == orc.step()
        for a, b in zip(dev.get_state(), orc.get_state()):
            assert np.array_equal(a, b)
        if k == 0:
            # the first step took the density branch: stagnant floor state in the seeded cells
            Na = orc.get_state()[1].reshape(-1, 5)
            assert np.all(Na[rho_cells] == [cfg.rho_floor, 0, 0, 0, cfg.p_floor / (cfg.gamma - 1.0)])
            # and the pressure branch: H+ re-clamped to the pressure floor
            H = orc.get_state()[0].reshape(-1, 5)
            ke = 0.5 * (H[p_cells, 1] ** 2 + H[p_cells, 2] ** 2 + H[p_cells, 3] ** 2) / H[p_cells, 0]
            assert np.allclose((cfg.gamma - 1.0) * (H[p_cells, 4] - ke), cfg.p_floor, rtol=1e-6, atol=1e-15)


@pytest.mark.parametrize("species", [0, 1])
def test_face_state_fallback_bitwise(small_explicit, species):
    w = small_explicit
    dev, orc = _pair(w)
    orc.prepare()
    W = orc.derived("W_H" if species == 0 else "W_Na")
    fs = np.array(w["cfg"].fs[species][:])
    g = rough_gradients(w, scale=50.0 if species == 0 else 0.05)
    assert fallback_faces(w["mesh"], W, g) > 100
    assert np.array_equal(dev.flux_rhs(species, fs, W, g), orc.flux_rhs(species, fs, W, g))


@pytest.mark.parametrize("gyration", [False, True])
def test_invalid_time_step_leaves_state(gyration):
    """dt <= 0 raises "invalid time step" before the state changes (also with the gyration split, whose first half
    rotates U); a NaN cap is no cap, as in the reference."""
    from mpcns_uns_amd import workloads

    w = workloads.build(net=(12, 48, 3), hall=True, **({"induce_coeff": 40.0} if gyration else {}))
    dev, orc = _pair(w)
    for dt_max in (-1.0, 0.0):
        with pytest.raises(RuntimeError, match="invalid time step"):
            dev.step(dt_max)
        for a, b in zip(dev.get_state(), w["state"]):
            assert np.array_equal(a, b)
    dd, _, rd = dev.step(float("nan"))
    do, _, ro = orc.step(float("nan"))
    assert abs(dd - do) <= 1e-14 * do and rd < 1e-8
    with pytest.raises(RuntimeError, match="precision"):
        dev.step(precision=7)


def test_gmres_truncated_at_max_it():
    from mpcns_uns_amd import Device, workloads
    from oracle.oracle import Oracle

    w = workloads.build(net=(12, 48, 3), hall=True, ind_max_it=3)
    m = w["mesh"]
    rng = np.random.default_rng(31)
    B0 = np.tile([0.1, -0.2, 0.3], m.n_cells) + 0.05 * rng.normal(size=3 * m.n_cells)
    ne_eff = 1.0 + 0.1 * rng.uniform(size=m.n_cells)
    dev, mix, orc = Device(m, w["cfg"]), Device(m, w["cfg"]), Oracle(m, w["cfg"])
    for x in (dev, mix, orc):
        x.ind_set_frozen(B0, ne_eff)
    dirichlet = (m.array("edge_mode") != 0).astype(np.int8)
    E_expl = rng.normal(size=m.n_edges) * 1e-3
    Ed, itd, rd = dev.ind_solve(w["state"][2], E_expl, dirichlet, 5e-2, precision=0)
    Eo, ito, ro = orc.ind_solve(w["state"][2], E_expl, dirichlet, 5e-2)
    assert itd == ito == 3
    assert rd > w["cfg"].ind_rtol and abs(rd - ro) <= 1e-10 * ro
    assert maxrel(Ed, Eo) < 1e-12
    Em, itm, rm = mix.ind_solve(w["state"][2], E_expl, dirichlet, 5e-2, precision=1)
    assert itm == 3 and rm > w["cfg"].ind_rtol


@pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["x"]).have_ref_on_device(),
                    reason="oracle/_ref/libmpcns_ref_gpu.so not built")
@pytest.mark.parametrize("species", [0, 1])
def test_shim_ghostfn_table(small_explicit, species):
    """The reference's Stencil::limited_gradients / fluid_flux_rhs called with a GhostFn (a per-face table, the form
    MarsSolver uses) through the drop-in shim on the device: equal to the oracle's ghost-table path."""
    from mpcns_uns_amd import workloads
    from oracle.oracle import reference_for_mesh

    w = small_explicit
    dev, orc = _pair(w)
    orc.prepare()
    W = orc.derived("W_H" if species == 0 else "W_Na")
    fs = np.array(w["cfg"].fs[species][:])
    table = np.abs(np.random.default_rng(3 + species).normal(1.0, 0.2, 5 * w["mesh"].n_faces))
    hyb = reference_for_mesh(w["mesh"], w["cfg"], dipole_m=workloads.DIPOLE_M, on_device=True)
    for limit in (True, False):
        gh, seen = hyb.limited_gradients(W, species, fs, limit, ghost=table)
        assert seen > 0
        go = orc.limited_gradients(W, species, fs, limit, table)
        assert np.array_equal(gh, go)
        rh, _ = hyb.flux_rhs(species, fs, W, go, ghost=table)
        assert np.array_equal(rh, orc.flux_rhs(species, fs, W, go, table))


@pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["x"]).have_ref_on_device(),
                    reason="oracle/_ref/libmpcns_ref_gpu.so not built")
@pytest.mark.parametrize("species", [0, 1])
def test_shim_ghostfn_of_marssolver_form(small_explicit, species):
    """A GhostFn of MarsSolver's form (MarsSolver.cpp:124-128: a cached state on wall faces, ghost_state of the inside
    state on the others) through the shim on the device, against the compiled reference on the CPU, bit for bit.  The
    reference hands the GhostFn the cell value in the gradients and the reconstructed face state in the fluxes
    (FluxRHS.cpp:57-65), so the ghost of an outflow face differs between the two calls; a table built from the cell
    values alone gives other fluxes."""
    from mpcns_uns_amd import workloads
    from oracle.oracle import reference_for_mesh

    w = small_explicit
    m = w["mesh"]
    _, orc = _pair(w)
    orc.prepare()
    W = orc.derived("W_H" if species == 0 else "W_Na")
    fs = np.array(w["cfg"].fs[species][:])
    wall = np.abs(np.random.default_rng(5 + species).normal(1.0, 0.2, 5 * m.n_faces))
    kind = m.array("face_kind")
    assert np.count_nonzero(kind == 3) > 0  # outflow faces, whose ghost is the inside state
    ref = reference_for_mesh(m, w["cfg"], dipole_m=workloads.DIPOLE_M)
    hyb = reference_for_mesh(m, w["cfg"], dipole_m=workloads.DIPOLE_M, on_device=True)
    for limit in (True, False):
        gr = ref.limited_gradients_wall_ghost(W, species, fs, wall, limit)
        assert np.array_equal(hyb.limited_gradients_wall_ghost(W, species, fs, wall, limit), gr)
        rr = ref.flux_rhs_wall_ghost(species, fs, W, gr, wall)
        assert np.array_equal(hyb.flux_rhs_wall_ghost(species, fs, W, gr, wall), rr)
        # the same GhostFn evaluated with the cell value as the inside state of every boundary face
        ptr, idx = m.array("f2c.ptr"), m.array("f2c.idx")
        first = np.zeros(5 * m.n_faces)
        for f in np.nonzero((kind != 0) & (kind != 4))[0]:
            c = idx[ptr[f]]
            first[5 * f:5 * f + 5] = wall[5 * f:5 * f + 5] if kind[f] == 1 else (fs if kind[f] == 2 else W[5 * c:5 * c + 5])
        assert not np.array_equal(ref.flux_rhs(species, fs, W, gr, ghost=first)[0], rr)


@pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["x"]).have_ref_on_device(),
                    reason="oracle/_ref/libmpcns_ref_gpu.so not built")
def test_shim_faraday_keeps_inactive_faces(small_explicit):
    """faraday_update (Faraday.cpp:15-25) with a caller's mask other than active_faces: through the shim on the device
    the inactive faces keep their values and the active ones are updated, bit for bit as the compiled reference."""
    from mpcns_uns_amd import workloads
    from oracle.oracle import reference_for_mesh

    w = small_explicit
    m = w["mesh"]
    rng = np.random.default_rng(11)
    E, Phi = rng.normal(size=m.n_edges), rng.normal(size=m.n_faces)
    mask = (rng.random(m.n_faces) < 0.5).astype(np.int8)
    ref = reference_for_mesh(m, w["cfg"], dipole_m=workloads.DIPOLE_M)
    hyb = reference_for_mesh(m, w["cfg"], dipole_m=workloads.DIPOLE_M, on_device=True)
    r, h = ref.faraday_masked(E, Phi, 0.3, mask), hyb.faraday_masked(E, Phi, 0.3, mask)
    assert np.array_equal(h, r)
    assert np.array_equal(r[mask == 0], Phi[mask == 0]) and not np.array_equal(r[mask == 1], Phi[mask == 1])
