"""The agglomeration hierarchy of the multigrid preconditioner (host side, no GPU), pinned by mathematics on a small
shell: every coarse operator is the Galerkin triple product of the level above, and the first one is P^T K P with K
the resistive curl-curl assembled column by column from the oracle's ImplicitInduction::apply_F; the masses are P^T P.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module", params=["shell", "hex_pyramid"])
def shell(request):
    """A small tetrahedral shell, or a box of hexahedra and pyramids (cells of five and six faces)."""
    from meshes import mixed_workload
    from mpcns_uns_amd import workloads
    from oracle.multigrid import operator_F
    from oracle.oracle import Oracle

    if request.param == "shell":
        w = workloads.build(net=(6, 12, 3), hall=True)
    else:
        w = mixed_workload(True, nx=8, ny=6, nz=6)
    m = w["mesh"]
    res = type(w["cfg"]).from_buffer_copy(w["cfg"])
    res.hall = 0  # the resistive part only: what the coarse levels carry
    orc = Oracle(m, res)
    K0 = operator_F(m, orc.ind_apply_F)
    dirichlet = m.array("edge_mode") != 0
    return w, K0, dirichlet


def test_galerkin_coarse_operators(shell):
    from oracle.multigrid import csr_dense, prolongation

    w, K0, dirichlet = shell
    levels = w["mesh"].mg_hierarchy(w["cfg"], dirichlet)
    assert len(levels) >= 3 and levels[0]["n"] == w["mesh"].n_edges
    free = (~dirichlet).astype(float)
    assert np.array_equal(levels[0]["mass"], free)
    Kf, Mf = K0 * free[:, None] * free[None, :], np.diag(free)
    for lvl in levels[1:]:
        P = prolongation(lvl, lvl["n"])
        # every finer unknown maps to at most one coarse unknown, with sign +-1
        assert np.all(np.abs(P).sum(axis=1) <= 1)
        Kc = csr_dense(lvl)
        ref = P.T @ Kf @ P
        assert np.abs(Kc - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.array_equal(np.diag(P.T @ Mf @ P), lvl["mass"])
        assert np.count_nonzero(np.diag(lvl["mass"])) == lvl["n"]  # no empty coarse unknown
        Kf, Mf = Kc, np.diag(lvl["mass"])


def test_aggregation_ratio_and_dirichlet(shell):
    w, K0, dirichlet = shell
    levels = w["mesh"].mg_hierarchy(w["cfg"], dirichlet)
    for a, b in zip(levels[:-1], levels[1:]):
        if b["n_nodes"] >= 8:
            assert 4.0 <= a["n_nodes"] / b["n_nodes"] <= 12.0  # three matching passes: aggregates of ~8 nodes
        assert b["n"] < a["n"]
    m1 = levels[1]["map"]
    assert np.all(m1[dirichlet] == np.iinfo(np.int32).min)  # Dirichlet edges are outside every coarse space


def test_prolongation_carries_gradients(shell):
    """P maps the gradient of any coarse nodal field to the gradient of the same field spread over the aggregates (on the
    free edges; 0 inside an aggregate): the orientation and sign of every map entry, independently of the Galerkin
    products.  A flipped sign or a wrong coarse index fails it."""
    from oracle.multigrid import prolongation

    w, _, dirichlet = shell
    m = w["mesh"]
    levels = m.mg_hierarchy(w["cfg"], dirichlet)
    e2n = m.array("e2n.idx").reshape(-1, 2)
    sg = m.array("e2n.sgn").reshape(-1, 2)
    tail0 = np.where(sg[:, 0] < 0, e2n[:, 0], e2n[:, 1])
    head0 = np.where(sg[:, 0] < 0, e2n[:, 1], e2n[:, 0])
    assert np.array_equal(levels[0]["tail"], tail0) and np.array_equal(levels[0]["head"], head0)
    rng = np.random.default_rng(2)
    free = ~dirichlet
    for f, c in zip(levels[:-1], levels[1:]):
        phi_c = rng.normal(size=c["n_nodes"])
        phi_f = phi_c[c["agg"]]
        g_f = phi_f[f["head"]] - phi_f[f["tail"]]
        g_c = phi_c[c["head"]] - phi_c[c["tail"]]
        mask = free if f is levels[0] else np.ones(f["n"], bool)
        Pg = prolongation(c, c["n"]) @ g_c
        assert np.abs((Pg - g_f)[mask]).max() <= 1e-12 * np.abs(g_f).max()
        free = np.ones(c["n"], bool)


def test_hierarchy_refuses_a_mesh_without_nodes(shell):
    w, _, _ = shell
    from mpcns_uns_amd._lib import MeshDesc

    d = MeshDesc.from_buffer_copy(w["mesh"].desc())
    d.n_nodes = 0
    import ctypes as C

    from mpcns_uns_amd._lib import lib

    h = C.c_void_p()
    assert lib().mhdg_mg_build(C.byref(d), C.byref(w["cfg"]), None, C.byref(h)) != 0
    assert b"e2n" in lib().mhdg_mesh_error()
