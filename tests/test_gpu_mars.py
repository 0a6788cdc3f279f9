"""MarsSolver's point-implicit sources (MarsUns/driver/MarsSolver.cpp:422-527) on the device against the oracle
(tests/test_oracle_mars_sources.py pins the oracle): the ten-ion shell with seeded velocities, a seeded neutral
atmosphere and the step's own field, both gyration and drag/relaxation branches, 1e-12 relative per species (the
complex divisions round differently on the device), and the explicit step that follows them.
"""
import numpy as np
import pytest

from conftest import maxrel

pytestmark = pytest.mark.gpu


def test_mars_implicit_sources_match_oracle():
    from mpcns_uns_amd import Device, workloads
    from oracle.oracle import Oracle
    from test_oracle_mars_sources import _stirred

    w = workloads.build("shell_4m_mars", net=(12, 24, 3))
    mesh, cfg = w["mesh"], w["cfg"]
    ns = cfg.n_species
    st = _stirred(w, seed=12)
    cold = np.arange(0, mesh.n_cells, 7)  # cells in the explicit relaxation branch (T < 50 K)
    for U in st[2:]:
        U = U.reshape(-1, 5)
        ke = 0.5 * (U[cold, 1] ** 2 + U[cold, 2] ** 2 + U[cold, 3] ** 2) / U[cold, 0]
        U[cold, 4] = ke + 1e-9
    prm = workloads.mars_params()
    nd = workloads.neutral_atmosphere(mesh)
    dev, orc = Device(mesh, cfg), Oracle(mesh, cfg)
    for x in (dev, orc):
        x.set_background(*w["background"])
        x.set_state(st[0], st[1], w["state"][2])
        for s in range(2, ns):
            x.set_species_state(s, st[s])
        x.prepare()
    for dt in (2e-3, 5e-3):
        dev.mars_implicit_sources(prm, nd, dt)
        orc.mars_implicit_sources(prm, nd, dt)
        for s in range(ns):
            a, b = dev.get_species_state(s), orc.get_species_state(s)
            assert maxrel(a, b) < 1e-12, (s, maxrel(a, b))
    # an explicit step from the result agrees as the ten-ion steps do
    dt_d, _, _ = dev.step()
    dt_o, _, _ = orc.step()
    assert abs(dt_d - dt_o) <= 1e-12 * dt_o
    for s in range(ns):
        assert maxrel(dev.get_species_state(s), orc.get_species_state(s)) < 1e-10
