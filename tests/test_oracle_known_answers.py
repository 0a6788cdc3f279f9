"""Known-answer and invariant tests of the oracle that need no reference build.

* ideal UCT EMF on a regular hexahedral box against the legacy structured Mercury formula, the transcription
  the reference's own test holds (tests/test_ideal_emf.cpp:112-220; tolerance 1e-13 there and here);
* uniform B: the transverse upwind part vanishes on planar meshes; uniform u and B: E = -(u x B).dl
  (tests/test_ideal_emf.cpp:223-281);
* free-stream preservation of the finite-volume RHS (tests/test_solver.cpp:68-98);
* least-squares gradients exact for linear fields; cell vector exact for uniform fields; J = 0 for uniform B;
* constrained transport keeps d2 B at round-off for any E;
* the GMRES solve against a dense direct solve of the same operator assembled column by column.
"""
import numpy as np
import pytest

from meshes import box_mesh


def interior_edges(m):
    return np.diff(m.array("e2f.ptr")) == np.diff(m.array("ec_ptr"))


def oracle_on(m, **cfg_kw):
    from mpcns_uns_amd import default_config
    from oracle.oracle import Oracle

    return Oracle(m, default_config(**cfg_kw))


def test_ideal_emf_legacy_regression():
    n, h = 6, 0.5
    m = box_mesh(n, n, n, hi=(n * h, n * h, n * h))
    orc = oracle_on(m)
    rng = np.random.default_rng(5)
    f2e, f2s = m.array("f2e.idx").reshape(-1, 4), m.array("f2e.sgn").reshape(-1, 4)
    A = rng.uniform(-1, 1, m.n_edges)
    Phi = (f2s * A[f2e]).sum(1)
    u = rng.uniform(-1, 1, 3 * m.n_cells)
    Bc = orc.cell_vector(Phi)
    E = orc.ideal_emf(u, Bc, Bc, Phi, Phi, 0.5)
    cc = m.array("cell_center").reshape(-1, 3)
    uc = np.zeros((n, n, n, 3))
    ijk = np.floor(cc / h).astype(int)
    uc[ijk[:, 0], ijk[:, 1], ijk[:, 2]] = u.reshape(-1, 3) / h
    Fl = [np.zeros((n + 1, n, n)), np.zeros((n, n + 1, n)), np.zeros((n, n, n + 1))]
    fc, fa = m.array("face_center").reshape(-1, 3), m.array("face_area").reshape(-1, 3)
    for f in range(m.n_faces):
        ax = int(np.argmax(np.abs(fa[f])))
        i, j, k = np.floor(fc[f] / h + 1e-9).astype(int)
        Fl[ax][i, j, k] = np.sign(fa[f, ax]) * Phi[f]
    ucon = lambda a, i, j, k: uc[i, j, k, a]
    bind = lambda a, i, j, k: 0.5 * (Fl[a][i, j, k] + Fl[a][i + (a == 0), j + (a == 1), k + (a == 2)])

    def G(beta, fax, L, R, bface):
        GL = ucon(fax, *L) * bind(beta, *L) - ucon(beta, *L) * bface
        GR = ucon(fax, *R) * bind(beta, *R) - ucon(beta, *R) * bface
        rad = max(abs(ucon(fax, *L)), abs(ucon(fax, *R)))
        return 0.5 * (GL + GR) - 0.5 * rad * (bind(beta, *R) - bind(beta, *L))

    def Eface(axis, i, j, k, comp):
        L, R, bf = (i - (axis == 0), j - (axis == 1), k - (axis == 2)), (i, j, k), Fl[axis][i, j, k]
        e = [0.0, 0.0, 0.0]
        if axis == 0: e[1], e[2] = G(2, 0, L, R, bf), -G(1, 0, L, R, bf)
        if axis == 1: e[0], e[2] = -G(2, 1, L, R, bf), G(0, 1, L, R, bf)
        if axis == 2: e[0], e[1] = G(1, 2, L, R, bf), -G(0, 2, L, R, bf)
        return e[comp]

    amax = lambda a, *cs: max(abs(ucon(a, *c)) for c in cs)
    C = 0.5

    def legacy(axis, i, j, k):
        fx, fy, fz = Fl
        if axis == 0:
            Ec = 0.25 * (Eface(1, i, j, k, 0) + Eface(1, i, j, k - 1, 0) + Eface(2, i, j, k, 0) + Eface(2, i, j - 1, k, 0))
            cs = ((i, j, k), (i, j - 1, k), (i, j, k - 1), (i, j - 1, k - 1))
            return Ec + C * (0.5 * amax(1, *cs) * (fz[i, j, k] - fz[i, j - 1, k]) - 0.5 * amax(2, *cs) * (fy[i, j, k] - fy[i, j, k - 1]))
        if axis == 1:
            Ec = 0.25 * (Eface(0, i, j, k, 1) + Eface(0, i, j, k - 1, 1) + Eface(2, i, j, k, 1) + Eface(2, i - 1, j, k, 1))
            cs = ((i, j, k), (i - 1, j, k), (i, j, k - 1), (i - 1, j, k - 1))
            return Ec + C * (0.5 * amax(2, *cs) * (fx[i, j, k] - fx[i, j, k - 1]) - 0.5 * amax(0, *cs) * (fz[i, j, k] - fz[i - 1, j, k]))
        Ec = 0.25 * (Eface(0, i, j, k, 2) + Eface(0, i, j - 1, k, 2) + Eface(1, i, j, k, 2) + Eface(1, i - 1, j, k, 2))
        cs = ((i, j, k), (i - 1, j, k), (i, j - 1, k), (i - 1, j - 1, k))
        return Ec + C * (0.5 * amax(0, *cs) * (fy[i, j, k] - fy[i - 1, j, k]) - 0.5 * amax(1, *cs) * (fx[i, j, k] - fx[i, j - 1, k]))

    ev, em = m.array("edge_vec").reshape(-1, 3), m.array("edge_mid").reshape(-1, 3)
    err, scale, cnt = 0.0, 0.0, 0
    for e in range(m.n_edges):
        ax = 0 if abs(ev[e, 0]) > 0.1 else (1 if abs(ev[e, 1]) > 0.1 else 2)
        i, j, k = np.floor(em[e] / h + 1e-9).astype(int)
        o = [x for a, x in enumerate((i, j, k)) if a != ax]
        if not all(1 <= x <= n - 1 for x in o):
            continue
        ref = legacy(ax, i, j, k)
        err, scale, cnt = max(err, abs(E[e] - ref)), max(scale, abs(ref)), cnt + 1
    assert cnt > 100
    assert err < 1e-13 * max(scale, 1.0)


@pytest.mark.parametrize("kind", ["box", "shell"])
def test_uniform_fields(kind):
    from mpcns_uns_amd import workloads

    if kind == "box":
        m = box_mesh(5, 4, 4, hi=(5.0, 4.0, 4.0))
    else:
        m = workloads.build(net=(8, 16, 3), hall=False)["mesh"]
    orc = oracle_on(m)
    b0, u0 = np.array([0.4, -0.9, 0.6]), np.array([0.7, 0.2, -0.5])
    Phi = m.array("face_area").reshape(-1, 3) @ b0
    Bc = np.tile(b0, m.n_cells)
    # cell vector exact, J = 0 for a uniform field
    assert np.abs(orc.cell_vector(Phi) - Bc).max() < 1e-12
    assert np.abs(orc.cell_current(Bc)).max() < 1e-11
    inter = interior_edges(m)
    ur = np.random.default_rng(9).uniform(-1, 1, 3 * m.n_cells)
    Ea, Eb = orc.ideal_emf(ur, Bc, Bc, Phi, Phi, 0.5), orc.ideal_emf(ur, Bc, Bc, Phi, Phi, 0.0)
    assert np.abs(Ea - Eb)[inter].max() < 1e-12 * max(np.abs(Ea[inter]).max(), 1.0)
    Eu = orc.ideal_emf(np.tile(u0, m.n_cells), Bc, Bc, Phi, Phi, 0.5)
    exact = -(m.array("edge_vec").reshape(-1, 3) @ np.cross(u0, b0))
    assert np.abs(Eu - exact)[inter].max() < 1e-12 * max(np.abs(exact).max(), 1.0)


def test_free_stream_and_linear_gradients():
    m = box_mesh(5, 5, 4, hi=(1.0, 1.0, 0.8), default="inflow")
    fs = np.array([1.3, 0.4, -0.2, 0.1, 0.8])
    orc = oracle_on(m, fs=(tuple(fs), tuple(fs)), recon_rel_limit=0.25)
    W = np.tile(fs, m.n_cells)
    g = orc.limited_gradients(W, 0, fs, True)
    assert np.abs(g).max() < 1e-13
    rhs = orc.flux_rhs(0, fs, W, g)
    assert np.abs(rhs).max() < 1e-12
    # linear field, limiter off: exact gradient in cells whose stencil has no boundary (virtual) neighbour
    cc = m.array("cell_center").reshape(-1, 3)
    slope = np.array([[0.3, -0.2, 0.5], [0.1, 0.2, 0.3], [-0.4, 0.1, 0.0], [0.2, 0.2, -0.1], [0.05, -0.3, 0.2]])
    W = (fs[None, :] + cc @ slope.T).ravel()
    g = orc.limited_gradients(W, 0, fs, False).reshape(-1, 5, 3)
    inner = np.all(np.diff(m.array("f2c.ptr"))[m.array("c2f.idx").reshape(-1, 6)] == 2, axis=1)
    assert inner.sum() > 4
    assert np.abs(g[inner] - slope[None]).max() < 1e-12


def test_ct_keeps_div_b():
    from mpcns_uns_amd import workloads

    w = workloads.build(net=(8, 16, 3), hall=False)
    m = w["mesh"]
    orc = oracle_on(m)
    rng = np.random.default_rng(2)
    Phi = w["state"][2]
    for _ in range(5):
        Phi = orc.faraday(rng.normal(size=m.n_edges), Phi, 0.7)
    _, rel, _ = orc.div_b(Phi, w["background"][0])
    assert rel < 1e-13


def test_gmres_against_dense_solve():
    """ImplicitInduction::solve on a small box with Hall + resistivity equals a dense direct solve."""
    m = box_mesh(3, 3, 2, hi=(1.0, 1.0, 0.7))
    orc = oracle_on(m, hall=1, hall_coef=0.5, hall_r_min=-100.0, hall_r_max=-99.0, resistive=1, inver_Rem=0.05,
                    eta_r_inner=-100.0, eta_r_outer=100.0, eta_width=1.0, ind_rtol=1e-13, ind_max_it=500, ind_restart=30)
    rng = np.random.default_rng(3)
    B0 = np.tile([1.0, 0.3, -0.2], m.n_cells) + 0.1 * rng.normal(size=3 * m.n_cells)
    orc.ind_set_frozen(B0, np.ones(m.n_cells))
    f2e, f2s = m.array("f2e.idx").reshape(-1, 4), m.array("f2e.sgn").reshape(-1, 4)
    d1 = lambda x: (f2s * x[f2e]).sum(1)
    dt, ne = 0.05, m.n_edges
    A = np.zeros((ne, ne))
    for e in range(ne):
        x = np.zeros(ne)
        x[e] = 1.0
        A[:, e] = x + dt * orc.ind_apply_F(d1(x))
    dirichlet = np.zeros(ne, np.int8)
    dirichlet[rng.choice(ne, 10, replace=False)] = 1
    Phi_n = d1(rng.normal(size=ne))
    E_expl = rng.normal(size=ne)
    b = E_expl + orc.ind_apply_F(Phi_n)
    free = dirichlet == 0
    E_ref = E_expl.copy()
    rhs = b[free] - A[np.ix_(free, ~free)] @ E_expl[~free]
    E_ref[free] = np.linalg.solve(A[np.ix_(free, free)], rhs)
    E, it, res = orc.ind_solve(Phi_n, E_expl, dirichlet, dt)
    assert res < 1e-13
    assert np.abs(E - E_ref).max() < 1e-10 * np.abs(E_ref).max()
