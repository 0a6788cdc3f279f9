"""Parity of the HIP path (C-ABI) with the CPU oracle on seeded shell meshes.

The oracle reproduces the reference bit for bit (tests/test_oracle_reference.py); the device kernels mirror
the same evaluation order without contraction, so every explicit operation is compared for exact equality.
The GMRES solve uses parallel (fixed-tree) dot products, so its result is compared to a tolerance.
"""
import numpy as np
import pytest

from conftest import maxrel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(small_hall):
    from mpcns_uns_amd import Device
    from oracle.oracle import Oracle

    w = small_hall
    dev, orc = Device(w["mesh"], w["cfg"]), Oracle(w["mesh"], w["cfg"])
    for x in (dev, orc):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    orc.step()  # derived fields of the initial state
    fields = {n: orc.derived(n) for n in ("W_H", "W_Na", "B_ind", "B_tot", "J", "u_plus")}
    return dev, orc, w, fields


def test_cell_vector_and_current(pair):
    dev, orc, w, f = pair
    Phi = w["state"][2]
    assert np.array_equal(dev.cell_vector(Phi), orc.cell_vector(Phi))
    assert np.array_equal(dev.cell_current(f["B_ind"]), orc.cell_current(f["B_ind"]))


@pytest.mark.parametrize("species", [0, 1])
@pytest.mark.parametrize("limit", [True, False])
def test_limited_gradients(pair, species, limit):
    dev, orc, w, f = pair
    W = f["W_H"] if species == 0 else f["W_Na"]
    fs = np.array(w["cfg"].fs[species][:])
    assert np.array_equal(dev.limited_gradients(W, species, fs, limit), orc.limited_gradients(W, species, fs, limit))


@pytest.mark.parametrize("species", [0, 1])
def test_flux_rhs(pair, species):
    dev, orc, w, f = pair
    W = f["W_H"] if species == 0 else f["W_Na"]
    fs = np.array(w["cfg"].fs[species][:])
    g = orc.limited_gradients(W, species, fs, True)
    assert np.array_equal(dev.flux_rhs(species, fs, W, g), orc.flux_rhs(species, fs, W, g))


def test_flux_rhs_ghost_table(pair):
    dev, orc, w, f = pair
    rng = np.random.default_rng(3)
    ghost = np.abs(rng.normal(1.0, 0.1, 5 * dev.nf))
    fs = np.array(w["cfg"].fs[0][:])
    g = orc.limited_gradients(f["W_H"], 0, fs, True, ghost)
    assert np.array_equal(dev.limited_gradients(f["W_H"], 0, fs, True, ghost), g)
    assert np.array_equal(dev.flux_rhs(0, fs, f["W_H"], g, ghost), orc.flux_rhs(0, fs, f["W_H"], g, ghost))


def test_ideal_emf_heavy_edges(pair):
    """Random u and B: every edge class (one thread per edge, pole axes with valence n_lon by a block)."""
    dev, orc, w, f = pair
    rng = np.random.default_rng(11)
    u = rng.uniform(-1, 1, 3 * dev.nc)
    Bt, Bi = rng.normal(size=3 * dev.nc), rng.normal(size=3 * dev.nc)
    pt, pi = rng.normal(size=dev.nf), rng.normal(size=dev.nf)
    E0 = rng.normal(size=dev.ne)
    assert np.array_equal(dev.ideal_emf(u, Bt, Bi, pt, pi, 0.5, E0), orc.ideal_emf(u, Bt, Bi, pt, pi, 0.5, E0))
    valence = np.diff(w["mesh"].array("e2f.ptr"))
    assert valence.max() == 48 > 32  # pole axes of the 12 x 48 net go through the block-per-edge kernel


def test_faraday_divb_dt_sources(pair):
    dev, orc, w, f = pair
    rng = np.random.default_rng(5)
    E = rng.normal(size=dev.ne)
    Phi = w["state"][2]
    assert np.array_equal(dev.faraday(E, Phi, 0.37), orc.faraday(E, Phi, 0.37))
    dte = rng.uniform(0.1, 1, dev.ne)
    assert np.array_equal(dev.faraday_edge_dt(E, dte, Phi), orc.faraday_edge_dt(E, dte, Phi))
    B = dev.faraday(E, Phi, 0.37)
    assert dev.div_b(B) == orc.div_b(B)
    PA = w["background"][0]
    assert dev.div_b(B, PA) == orc.div_b(B, PA)
    assert dev.local_dt(f["W_H"], f["W_Na"], f["B_tot"]) == orc.local_dt(f["W_H"], f["W_Na"], f["B_tot"])
    rhs = rng.normal(size=5 * dev.nc)
    for s in (0, 1):
        assert np.array_equal(dev.add_sources(s, f["W_H"], f["W_Na"], f["B_tot"], f["J"], 1e-3, rhs),
                              orc.add_sources(s, f["W_H"], f["W_Na"], f["B_tot"], f["J"], 1e-3, rhs))


def test_apply_F_and_solve(pair):
    dev, orc, w, f = pair
    rng = np.random.default_rng(7)
    ne_eff = np.sqrt((f["W_H"][0::5] + f["W_Na"][0::5] / 23.0) ** 2 + 1e-4)
    for x in (dev, orc):
        x.ind_set_frozen(f["B_tot"], ne_eff)
    phi = rng.normal(size=dev.nf)
    assert np.array_equal(dev.ind_apply_F(phi), orc.ind_apply_F(phi))
    E_expl = rng.normal(size=dev.ne) * 1e-3
    dirichlet = (w["mesh"].array("edge_mode") == 2).astype(np.int8)
    Ed, itd, rd = dev.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3)
    Eo, ito, ro = orc.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3)
    assert rd < 1e-8 and ro < 1e-8
    assert abs(itd - ito) <= 1
    assert maxrel(Ed, Eo) < 1e-7


def _fresh(w):
    from mpcns_uns_amd import Device
    from oracle.oracle import Oracle

    dev, orc = Device(w["mesh"], w["cfg"]), Oracle(w["mesh"], w["cfg"])
    for x in (dev, orc):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    return dev, orc


def test_step_explicit_bitwise(small_explicit):
    dev, orc = _fresh(small_explicit)
    for _ in range(3):
        dd = dev.step()
        do = orc.step()
        assert dd == do
        for a, b in zip(dev.get_state(), orc.get_state()):
            assert np.array_equal(a, b)
        for n in ("W_H", "W_Na", "B_ind", "B_tot", "J", "u_plus", "E"):
            assert np.array_equal(dev.derived(n), orc.derived(n)), n


def test_step_hall(small_hall):
    dev, orc = _fresh(small_hall)
    for _ in range(3):
        dt_d, it_d, r_d = dev.step()
        dt_o, it_o, r_o = orc.step()
        assert r_d < 1e-8 and abs(it_d - it_o) <= 1
        sd, so = dev.get_state(), orc.get_state()
        for a, b in zip(sd, so):
            assert maxrel(a, b) < 1e-12
        _, rel, _ = dev.div_b(sd[2], small_hall["background"][0])
        assert rel < 1e-12


def test_mixed_solve_matches_fp64(pair):
    """Mixed precision (FP32 inner GMRES, FP64 refinement) reaches the same FP64 relative residual."""
    dev, orc, w, f = pair
    rng = np.random.default_rng(8)
    ne_eff = np.sqrt((f["W_H"][0::5] + f["W_Na"][0::5] / 23.0) ** 2 + 1e-4)
    dev.ind_set_frozen(f["B_tot"], ne_eff)
    E_expl = rng.normal(size=dev.ne) * 1e-3
    dirichlet = (w["mesh"].array("edge_mode") != 0).astype(np.int8)
    from mpcns_uns_amd import Device

    d64 = Device(w["mesh"], w["cfg"])
    d64.ind_set_frozen(f["B_tot"], ne_eff)
    E64, it64, r64 = d64.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3, precision=0)
    E32, it32, r32 = dev.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3, precision=1)
    assert r64 < 1e-8 and r32 < 1e-8
    # both solutions are within the solver tolerance of the exact one
    assert maxrel(E32, E64) < 1e-6


def test_16_bit_and_32_bit_ids_agree(pair, monkeypatch):
    """The FP32 operator's 16-bit ids (the edge kernel's compact row slices: offsets from a slice's base id in 12-byte
    pairs; the cell kernels' face and neighbour offsets) and its 32-bit ids (MHDG_WIDE_IDS=1) give the same mixed solve,
    and the same multigrid-preconditioned solve (whose V-cycle runs the same kernels), bit for bit."""
    from mpcns_uns_amd import Device

    _, _, w, f = pair
    rng = np.random.default_rng(9)
    ne_eff = np.sqrt((f["W_H"][0::5] + f["W_Na"][0::5] / 23.0) ** 2 + 1e-4)
    dirichlet = (w["mesh"].array("edge_mode") != 0).astype(np.int8)
    E_expl = rng.normal(size=w["mesh"].n_edges) * 1e-3
    runs = []
    for wide in ("0", "1"):
        monkeypatch.setenv("MHDG_WIDE_IDS", wide)
        d = Device(w["mesh"], w["cfg"])
        d.ind_set_frozen(f["B_tot"], ne_eff)
        runs.append([d.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3, precision=p) for p in (1, 2)])
    for (Ea, ia, ra), (Eb, ib, rb) in zip(*runs):
        assert ia == ib and ra == rb and ra < 1e-8
        assert np.array_equal(Ea, Eb)


def test_step_mixed_vs_fp64(small_hall):
    from mpcns_uns_amd import Device

    w = small_hall
    a, b = Device(w["mesh"], w["cfg"]), Device(w["mesh"], w["cfg"])
    for x in (a, b):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    for _ in range(3):
        dta, _, ra = a.step(precision=0)
        dtb, _, rb = b.step(precision=1)
        assert ra < 1e-8 and rb < 1e-8
    for x, y in zip(a.get_state(), b.get_state()):
        assert maxrel(y, x) < 1e-8
    _, rel, _ = b.div_b(b.get_state()[2], w["background"][0])
    assert rel < 1e-12


def test_step_fp64_classical_gram_schmidt(small_hall):
    """precision 3 (the FP64 iteration with classical Gram-Schmidt) against the oracle's reference iteration (modified
    Gram-Schmidt): the same Krylov spaces, so the same iteration counts and states to rounding."""
    from mpcns_uns_amd import Device
    from oracle.oracle import Oracle

    w = small_hall
    dev, orc = Device(w["mesh"], w["cfg"]), Oracle(w["mesh"], w["cfg"])
    for x in (dev, orc):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    for _ in range(3):
        dt_d, it_d, r_d = dev.step(precision=3)
        dt_o, it_o, r_o = orc.step()
        assert abs(dt_d - dt_o) <= 1e-12 * dt_o and abs(it_d - it_o) <= 1 and r_d <= 1e-8
        for a, b in zip(dev.get_state(), orc.get_state()):
            assert maxrel(a, b) < 1e-10
    _, rel, _ = dev.div_b(dev.get_state()[2], w["background"][0])
    assert rel < 1e-12


@pytest.mark.parametrize("cells", ["hex", "hex_pyramid"])
def test_box_steps(cells):
    """The device on a jittered hexahedral box (six faces per cell, four edges per face: fixed row lengths other than
    the tetrahedron's) and on the same box with every other hexahedron split into six pyramids (rows of varying length,
    the kernels' row-pointer code; 31-33 GMRES iterations per step, across a restart) against the oracle, which matches
    the compiled reference bit for bit on both: explicit steps bitwise, the FP64 Hall step to 1e-12, the mixed and
    FP64-CGS steps within 1e-6 (every solve stops at a 1e-8 relative residual; with hall_coef 50 the system's
    conditioning puts the state difference of two such solves at a few 1e-8)."""
    from mpcns_uns_amd import Device
    from meshes import hex_workload, mixed_workload
    from oracle.oracle import Oracle

    make = hex_workload if cells == "hex" else mixed_workload
    we = make(False, nx=8, ny=6, nz=6)
    dev, orc = Device(we["mesh"], we["cfg"]), Oracle(we["mesh"], we["cfg"])
    for x in (dev, orc):
        x.set_background(*we["background"])
        x.set_state(*we["state"])
    for _ in range(3):
        assert dev.step() == orc.step()
        for a, b in zip(dev.get_state(), orc.get_state()):
            assert np.array_equal(a, b)
    w = make(True, nx=8, ny=6, nz=6, hall_coef=50.0)
    orc = Oracle(w["mesh"], w["cfg"])
    devs = {p: Device(w["mesh"], w["cfg"]) for p in (0, 1, 3)}
    for x in (orc, *devs.values()):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    for k in range(3):
        dt_o, it_o, r_o = orc.step()
        so = orc.get_state()
        for p, dev in devs.items():
            dt_d, it_d, r_d = dev.step(precision=p)
            print(f"[{cells} box] step {k + 1} precision {p}: GMRES {it_d}/{it_o}, residual {r_d:.2e}", flush=True)
            # the states of the mixed and CGS paths differ at the solver tolerance, and the next CFL step with them
            tol = 1e-12 if p == 0 else 1e-6
            assert abs(dt_d - dt_o) <= tol * dt_o and r_d <= 1e-8 and abs(it_d - it_o) <= (2 if p == 1 else 1)
            for a, b in zip(dev.get_state(), so):
                assert maxrel(a, b) < tol
            _, rel, _ = dev.div_b(dev.get_state()[2], w["background"][0])
            assert rel < 1e-12


def test_implicit_ideal_and_edge_dt():
    """set_ideal_state / ideal_linear (bitwise: the same face and edge EMF kernels as the explicit stage) and the
    per-edge pseudo-time solve, FP64 and mixed."""
    from mpcns_uns_amd import Device, workloads
    from oracle.oracle import Oracle

    w = workloads.build(net=(12, 48, 3), hall=True, implicit_ideal=1)
    m = w["mesh"]
    dev, orc = Device(m, w["cfg"]), Oracle(m, w["cfg"])
    assert dev.ind_flags() == (True, True, True)
    rng = np.random.default_rng(12)
    B0 = np.tile([0.1, -0.2, 0.3], m.n_cells) + 0.05 * rng.normal(size=3 * m.n_cells)
    ne_eff = 1.0 + 0.1 * rng.uniform(size=m.n_cells)
    u = rng.uniform(-1, 1, 3 * m.n_cells)
    mask = (m.array("edge_mode") == 0).astype(np.int8)
    phi = rng.normal(size=m.n_faces)
    for x in (dev, orc):
        x.ind_set_frozen(B0, ne_eff)
        x.ind_set_ideal_state(u, mask)
    L = orc.ind_ideal_linear(phi)
    assert np.abs(L).max() > 0
    assert np.array_equal(dev.ind_ideal_linear(phi), L)
    assert np.array_equal(dev.ind_apply_F(phi), orc.ind_apply_F(phi))
    dirichlet = (m.array("edge_mode") != 0).astype(np.int8)
    E_expl = rng.normal(size=m.n_edges) * 1e-3
    dte = 2e-3 * rng.uniform(0.5, 2.0, m.n_edges)
    Ed, itd, rd = dev.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3, dt_edge=dte)
    Eo, ito, ro = orc.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3, dt_edge=dte)
    assert rd < 1e-8 and ro < 1e-8 and abs(itd - ito) <= 1
    assert maxrel(Ed, Eo) < 1e-7
    with pytest.raises(RuntimeError, match="ideal"):
        dev.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3, precision=1)
    # without the ideal part the mixed solve supports dt_edge too
    w2 = workloads.build(net=(12, 48, 3), hall=True)
    a, b = Device(w2["mesh"], w2["cfg"]), Device(w2["mesh"], w2["cfg"])
    for x in (a, b):
        x.ind_set_frozen(B0, ne_eff)
    E64, _, r64 = a.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3, precision=0, dt_edge=dte)
    E32, _, r32 = b.ind_solve(w["state"][2], E_expl, dirichlet, 2e-3, precision=1, dt_edge=dte)
    assert r64 < 1e-8 and r32 < 1e-8 and maxrel(E32, E64) < 1e-6


def test_fused_step_entries(small_hall):
    """mhdg_step_explicit on a Hall configuration is the ideal step of the same configuration without the implicit
    part (bitwise against the oracle with Hall and resistivity off); mhdg_step_hall is mhdg_step."""
    from mpcns_uns_amd import Device
    from oracle.oracle import Oracle

    w = small_hall
    cfg_off = type(w["cfg"]).from_buffer_copy(w["cfg"])
    cfg_off.hall, cfg_off.resistive = 0, 0
    dev = Device(w["mesh"], w["cfg"])
    orc = Oracle(w["mesh"], cfg_off)
    for x in (dev, orc):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    assert dev.step_explicit() == orc.step()[0]
    for a, b in zip(dev.get_state(), orc.get_state()):
        assert np.array_equal(a, b)
    ref = Device(w["mesh"], w["cfg"])
    for x in (dev, ref):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    assert dev.step_hall(precision=0) == ref.step(precision=0)
    for a, b in zip(dev.get_state(), ref.get_state()):
        assert np.array_equal(a, b)
    off = Device(w["mesh"], cfg_off)
    with pytest.raises(RuntimeError, match="step_hall"):
        off.step_hall()


def test_restart_from_field_file(small_explicit, tmp_path):
    """A state written as a UNSBIN restart file (U_H, U_Na, B) and loaded on the device steps exactly as the arrays."""
    from mpcns_uns_amd import write_fields

    w = small_explicit
    UH, UNa, Phi = w["state"]
    p = tmp_path / "fields_final.unsb"
    write_fields(p, [("U_H", "cell", 5, UH), ("U_Na", "cell", 5, UNa), ("B", "face", 1, Phi), ("_time", "node", 1, [0.0])])
    dev, orc = _fresh(w)
    dev.set_state_from_fields(p)
    assert dev.step() == orc.step()
    for a, b in zip(dev.get_state(), orc.get_state()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("species", [2, 5, 9])
def test_fluid_calls_for_ten_ion_species(pair, species):
    """The per-call fluid entries serve any species index below 10 with the caller's free-stream state (the ion
    slots of a ten-ion run): bitwise against the oracle."""
    dev, orc, w, f = pair
    rng = np.random.default_rng(40 + species)
    W = f["W_H"].copy().reshape(-1, 5)
    W[:, 0] *= rng.uniform(0.5, 1.5, W.shape[0])
    W = W.ravel()
    fs = np.array([1e-4 * species, -1.0, 0.0, 0.0, 1e-5])
    for limit in (True, False):
        g = orc.limited_gradients(W, species, fs, limit)
        assert np.array_equal(dev.limited_gradients(W, species, fs, limit), g)
        assert np.array_equal(dev.flux_rhs(species, fs, W, g), orc.flux_rhs(species, fs, W, g))


def _ten_ion_pair(hall, net=(12, 24, 3)):
    from mpcns_uns_amd import Device, workloads
    from oracle.oracle import Oracle

    w = workloads.build("shell_4m_mars", net=net, hall=hall)
    dev, orc = Device(w["mesh"], w["cfg"]), Oracle(w["mesh"], w["cfg"])
    workloads.load(dev, w)
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    for s, U in enumerate(w["species_state"][2:], start=2):
        orc.set_species_state(s, U)
    return dev, orc, w


def test_ten_ion_step_explicit_bitwise():
    """Fused step of the ten-ion fluid (crustal dipoles, ion production, no Hall): bitwise for every species."""
    dev, orc, w = _ten_ion_pair(False)
    for _ in range(3):
        assert dev.step() == orc.step()
    for s in range(10):
        assert np.array_equal(dev.get_species_state(s), orc.get_species_state(s)), s
        assert np.array_equal(dev.get_species_prim(s), orc.get_species_prim(s)), s
    assert np.array_equal(dev.get_state()[2], orc.get_state()[2])
    for n in ("B_ind", "B_tot", "J", "u_plus", "E"):
        assert np.array_equal(dev.derived(n), orc.derived(n)), n


def test_ten_ion_step_hall_fp64_and_mixed():
    from mpcns_uns_amd import Device, workloads

    dev, orc, w = _ten_ion_pair(True)
    mix = Device(w["mesh"], w["cfg"])
    workloads.load(mix, w)
    for _ in range(3):
        dt_d, it_d, r_d = dev.step(precision=0)
        dt_o, it_o, r_o = orc.step()
        dt_m, it_m, r_m = mix.step(precision=1)
        assert r_d < 1e-8 and r_m < 1e-8 and abs(it_d - it_o) <= 1
    for s in range(10):
        assert maxrel(dev.get_species_state(s), orc.get_species_state(s)) < 1e-12, s
        assert maxrel(mix.get_species_state(s), dev.get_species_state(s)) < 1e-8, s
    Phi = dev.get_state()[2]
    assert maxrel(Phi, orc.get_state()[2]) < 1e-12
    assert maxrel(mix.get_state()[2], Phi) < 1e-8
    _, rel, _ = dev.div_b(Phi, w["background"][0])
    assert rel < 1e-12


def test_species_entries_reject_out_of_range():
    dev, _, _ = _ten_ion_pair(False, net=(6, 12, 2))
    with pytest.raises(RuntimeError):
        dev.get_species_state(10)
    with pytest.raises(RuntimeError):
        dev.set_species_state(-1, np.zeros(5 * dev.nc))
    with pytest.raises(RuntimeError):  # the two-ion per-call entries refuse a ten-ion context
        dev.local_dt(np.zeros(5 * dev.nc), np.zeros(5 * dev.nc), np.zeros(3 * dev.nc))


def test_prepare_and_solution_files(small_explicit, tmp_path):
    """Solver::prepare on the resident state and the Solver::write_vtk file set written from the device: the same bytes
    as from the oracle's fields (which match the reference solver's own files, tests/test_vtu.py)."""
    from mpcns_uns_amd.output import SolutionWriter, solution_fields

    dev, orc = _fresh(small_explicit)
    m = small_explicit["mesh"]
    a, b = SolutionWriter(tmp_path / "dev"), SolutionWriter(tmp_path / "orc")
    t = 0.0
    for step in (1, 2):
        dt = dev.step()[0]
        assert dt == orc.step()[0]
        t += dt
        dev.prepare()
        orc.prepare()
        for n in ("W_H", "W_Na", "B_ind", "B_tot", "J", "u_plus"):
            assert np.array_equal(dev.derived(n), orc.derived(n)), n
        a.write_device(dev, step, t)
        b.write(m, solution_fields(orc, m), step, t)
    for name in ("sol_0000002_r00000.vtu", "sol_0000002.pvtu", "solution.pvd"):
        assert (tmp_path / "dev" / name).read_bytes() == (tmp_path / "orc" / name).read_bytes(), name


@pytest.mark.parametrize("hall", [False, True])
def test_gyration_split_step(hall):
    """The fused step with the gyration split (induce_coeff > 0) against the oracle, which matches the reference bit for
    bit (tests/test_oracle_reference.py).  The rotation uses the device's cos and sin, so agreement is to round-off."""
    from mpcns_uns_amd import Device, workloads
    from oracle.oracle import Oracle

    w = workloads.build(net=(12, 48, 3), hall=hall, induce_coeff=40.0)
    dev, orc = Device(w["mesh"], w["cfg"]), Oracle(w["mesh"], w["cfg"])
    for x in (dev, orc):
        x.set_background(*w["background"])
        x.set_state(*w["state"])
    for _ in range(3):
        dd, _, rd = dev.step()
        do, _, ro = orc.step()
        assert abs(dd - do) <= 1e-14 * do
    for a, b in zip(dev.get_state(), orc.get_state()):
        assert maxrel(a, b) < 1e-12
    with pytest.raises(RuntimeError):
        dev.step_vjp(1, 1.2, 1.5)


@pytest.mark.parametrize("hall", [False, True])
def test_renumbered_mesh_step(hall):
    """The step on the Hilbert-renumbered mesh: bitwise (explicit) / 1e-12 (Hall) against the oracle on the same
    numbering, and within 1e-12 of the original numbering's step mapped back."""
    from mpcns_uns_amd import Device, workloads
    from oracle.oracle import Oracle

    w = workloads.build(net=(12, 48, 3), hall=hall)
    r = workloads.renumber(w)
    cm, fm, em = r["maps"]
    dev, orc, org = Device(r["mesh"], r["cfg"]), Oracle(r["mesh"], r["cfg"]), Device(w["mesh"], w["cfg"])
    for x, y in ((dev, r), (orc, r), (org, w)):
        x.set_background(*y["background"])
        x.set_state(*y["state"])
    for _ in range(3):
        dd, _, rd = dev.step()
        do, _, ro = orc.step()
        dg, _, rg = org.step()
        assert abs(dd - dg) <= 1e-14 * dg
        sd, so, sg = dev.get_state(), orc.get_state(), org.get_state()
        for a, b in zip(sd, so):
            assert np.array_equal(a, b) if not hall else maxrel(a, b) < 1e-12
        assert maxrel(sd[0].reshape(-1, 5), sg[0].reshape(-1, 5)[cm]) < 1e-12
        assert maxrel(sd[1].reshape(-1, 5), sg[1].reshape(-1, 5)[cm]) < 1e-12
        assert maxrel(sd[2], sg[2][fm]) < 1e-12
