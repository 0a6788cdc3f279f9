"""Branches of the reference path that seeded noisy states never take, forced from the host and compared between
the oracle and the compiled reference (oracle/_ref) bit for bit; plus the reference's GhostFn path and
faraday_update_edge_dt, each against the reference and against the invariant it must keep.
"""
import numpy as np
import pytest

from meshes import d2_of, fallback_faces, floored_state, rough_gradients

pytestmark = pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["have_ref"]).have_ref(),
                                reason="reference library oracle/_ref not built")


def _shell(hall=False, **kw):
    from mpcns_uns_amd import workloads
    from oracle.oracle import Oracle, reference_for_mesh

    w = workloads.build(net=(12, 48, 3), hall=hall, **kw)
    ref = reference_for_mesh(w["mesh"], w["cfg"], dipole_m=workloads.DIPOLE_M)
    return w, ref, Oracle(w["mesh"], w["cfg"])


@pytest.fixture(scope="module")
def explicit():
    return _shell()


def _ghost_table(w, seed=3):
    rng = np.random.default_rng(seed)
    return np.abs(rng.normal(1.0, 0.2, 5 * w["mesh"].n_faces))


@pytest.mark.parametrize("species", [0, 1])
@pytest.mark.parametrize("limit", [True, False])
def test_ghost_table_is_the_reference_ghostfn(explicit, species, limit):
    """Stencil::limited_gradients and fluid_flux_rhs with a GhostFn (Reconstruction.h:17-21) that returns a per-face
    table, against the oracle's ghost-table branch."""
    w, ref, orc = explicit
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    orc.prepare()
    W = orc.derived("W_H" if species == 0 else "W_Na")
    fs = np.array(w["cfg"].fs[species][:])
    table = _ghost_table(w, 3 + species)
    gr, seen_g = ref.limited_gradients(W, species, fs, limit, ghost=table)
    assert seen_g > 0  # the GhostFn was called (boundary cells have virtual neighbours)
    go = orc.limited_gradients(W, species, fs, limit, table)
    assert np.array_equal(go, gr)
    assert not np.array_equal(go, orc.limited_gradients(W, species, fs, limit))  # the table changed the result
    rr, seen_f = ref.flux_rhs(species, fs, W, gr, ghost=table)
    assert seen_f > 0
    ro = orc.flux_rhs(species, fs, W, gr, table)
    assert np.array_equal(ro, rr)
    assert not np.array_equal(ro, orc.flux_rhs(species, fs, W, gr))


def test_faraday_edge_dt_reference_and_d2_invariant(explicit):
    """faraday_update_edge_dt (Faraday.cpp:27-36): bitwise with the reference; d2 B is unchanged for any per-edge
    weights (d2 d1 = 0, the comment at Faraday.cpp:26), checked with integer incidence.  A wrong edge sign or a weight
    applied per face instead of per edge breaks the invariant."""
    w, ref, orc = explicit
    m = w["mesh"]
    rng = np.random.default_rng(9)
    E = rng.normal(size=m.n_edges)
    dte = rng.uniform(0.1, 2.0, m.n_edges)
    Phi = w["state"][2]
    out = orc.faraday_edge_dt(E, dte, Phi)
    assert np.array_equal(out, ref.faraday_edge_dt(E, dte, Phi))
    assert np.abs(out - Phi).max() > 1e-3  # the update moved the flux
    d0, d1 = d2_of(m, Phi), d2_of(m, out)
    scale = np.abs(out).max()
    assert np.abs(d1 - d0).max() <= 1e-14 * scale
    # the same update with the edge signs of the face loops dropped does not keep d2 B
    f2e_ptr, f2e = m.array("f2e.ptr"), m.array("f2e.idx")
    wrong = Phi - np.add.reduceat((dte * E)[f2e], f2e_ptr[:-1])
    assert np.abs(d2_of(m, wrong) - d0).max() > 1e-6
    # scalar dt is the special case dt_e = dt
    assert np.array_equal(orc.faraday_edge_dt(E, np.full(m.n_edges, 0.3), Phi), orc.faraday(0.3 * E, Phi, 1.0))


def test_floors_forced(explicit):
    """enforce_floors (Solver.cpp:171-191): a step from a state with negative Na+ densities and negative H+ internal
    energies takes both branches; oracle and reference agree bit for bit."""
    w, ref, orc = explicit
    state, rho_cells, p_cells = floored_state(w)
    ref.solver_init()
    ref.set_state(*state)
    orc.set_background(*w["background"])
    orc.set_state(*state)
    assert ref.step() == orc.step()[0]
    so, sr = orc.get_state(), ref.get_state()
    for a, b in zip(so, sr):
        assert np.array_equal(a, b)
    cfg = w["cfg"]
    floor = np.array([cfg.rho_floor, 0.0, 0.0, 0.0, cfg.p_floor / (cfg.gamma - 1.0)])
    Na = so[1].reshape(-1, 5)
    assert np.all(Na[rho_cells] == floor)  # density branch: stagnant floor state
    H = so[0].reshape(-1, 5)
    ke = 0.5 * (H[p_cells, 1] ** 2 + H[p_cells, 2] ** 2 + H[p_cells, 3] ** 2) / H[p_cells, 0]
    p = (cfg.gamma - 1.0) * (H[p_cells, 4] - ke)
    assert np.allclose(p, cfg.p_floor, rtol=1e-6, atol=1e-15)  # pressure branch: re-clamped to the floor


@pytest.mark.parametrize("species", [0, 1])
def test_face_state_fallback_forced(explicit, species):
    """fluid_flux_rhs with gradients that make many linear face states non-positive (the fallback of
    FluxRHS.cpp:26): bitwise with the reference, and the fallback changes the result."""
    w, ref, orc = explicit
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    orc.prepare()
    W = orc.derived("W_H" if species == 0 else "W_Na")
    fs = np.array(w["cfg"].fs[species][:])
    g = rough_gradients(w, scale=50.0 if species == 0 else 0.05)
    assert fallback_faces(w["mesh"], W, g) > 100
    out = orc.flux_rhs(species, fs, W, g)
    assert np.array_equal(out, ref.flux_rhs(species, fs, W, g))
    assert np.all(np.isfinite(out))


def test_invalid_time_step_raises_and_keeps_state(explicit):
    """step_first_order throws "invalid time step" for dt <= 0 or a non-finite dt (Solver.cpp:236) before it
    changes the state; the oracle does the same."""
    w, ref, orc = explicit
    ref.solver_init()
    ref.set_state(*w["state"])
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    for dt_max in (-1.0, 0.0):
        with pytest.raises(RuntimeError):
            ref.step(dt_max)
        with pytest.raises(RuntimeError, match="invalid time step"):
            orc.step(dt_max)
    for a, b, c in zip(ref.get_state(), orc.get_state(), w["state"]):
        assert np.array_equal(a, c) and np.array_equal(b, c)
    # a NaN cap is no cap: std::min(gdt, NaN) keeps gdt (Solver.cpp:235), in both
    assert ref.step(float("nan")) == orc.step(float("nan"))[0] > 0
    for a, b in zip(ref.get_state(), orc.get_state()):
        assert np.array_equal(a, b)


def test_gmres_truncated_at_max_it():
    """ImplicitInduction::solve stops at ind_max_it (ImplicitInduction.cpp:329) and reports rel_res > rtol; the
    oracle stops at the same iterate bit for bit."""
    w, ref, orc = _shell(hall=True, ind_max_it=3)
    m = w["mesh"]
    rng = np.random.default_rng(31)
    B0 = np.tile([0.1, -0.2, 0.3], m.n_cells) + 0.05 * rng.normal(size=3 * m.n_cells)
    ne_eff = 1.0 + 0.1 * rng.uniform(size=m.n_cells)
    for x in (orc, ref):
        x.ind_set_frozen(B0, ne_eff)
    dirichlet = (m.array("edge_mode") != 0).astype(np.int8)
    E_expl = rng.normal(size=m.n_edges) * 1e-3
    Eo, ito, ro = orc.ind_solve(w["state"][2], E_expl, dirichlet, 5e-2)
    Er, itr, rr = ref.ind_solve(w["state"][2], E_expl, dirichlet, 5e-2)
    assert ito == itr == 3 and ro == rr and np.array_equal(Eo, Er)
    assert rr > w["cfg"].ind_rtol
