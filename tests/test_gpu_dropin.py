"""The drop-in: the reference's own mercury::Solver (MercuryUns/driver/Solver.cpp) linked, through the C++ shim
mpcns_uns_amd/shim/mhd_gpu_shim.cpp, against the device library in place of the mhd_common hot-path sources.

The driver keeps its CPU parts (prepare's cell-vector reconstruction, sources, CFL step, EMF assembly) and calls the
device for the limited gradients, fluxes, cell current, ideal EMF, implicit solve and CT.  The hybrid step must equal
the oracle's (itself bit-identical to the reference): bit for bit without Hall, to 1e-12 with the FP64 Hall solve.
The linked library is built by `make -C oracle ref` where the reference sources are present.
"""
import numpy as np
import pytest

from conftest import maxrel

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not __import__("oracle.oracle", fromlist=["x"]).have_ref_on_device(),
                                 reason="oracle/_ref/libmpcns_ref_gpu.so not built")]


@pytest.mark.parametrize("hall", [False, True])
def test_reference_solver_on_device(hall):
    from mpcns_uns_amd import workloads
    from oracle.oracle import Oracle, reference_for_mesh

    w = workloads.build(net=(12, 48, 3), hall=hall, wall_dirichlet=0)
    hyb = reference_for_mesh(w["mesh"], w["cfg"], dipole_m=workloads.DIPOLE_M, on_device=True)
    hyb.solver_init()
    hyb.set_state(*w["state"])
    orc = Oracle(w["mesh"], w["cfg"])
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    for _ in range(3):
        dt_h = hyb.step()
        dt_o = orc.step()[0]
        sh, so = hyb.get_state(), orc.get_state()
        if hall:
            assert abs(dt_h - dt_o) <= 1e-12 * dt_o
            for a, b in zip(sh, so):
                assert maxrel(a, b) < 1e-12
        else:
            assert dt_h == dt_o
            for a, b in zip(sh, so):
                assert np.array_equal(a, b)


@pytest.mark.parametrize("hall", [False, True])
def test_reference_solver_on_device_hexahedron_pyramid_box(hall):
    """The same drop-in on a box of hexahedra and pyramids (rows of varying length; every boundary face outflow)."""
    from meshes import mixed_workload
    from mpcns_uns_amd import workloads
    from oracle.oracle import Oracle, ReferenceOnDevice

    w = mixed_workload(hall, nx=8, ny=6, nz=6, wall_dirichlet=0, hall_coef=50.0)
    hyb = ReferenceOnDevice(w["xyz"], w["cells"], w["cfg"], tags=w["tags"], tagset=1, dipole_m=workloads.DIPOLE_M,
                            dipole_c=(0.0, 0.0, 0.0), soft=1e-3)
    hyb.solver_init()
    hyb.set_state(*w["state"])
    orc = Oracle(w["mesh"], w["cfg"])
    orc.set_background(*w["background"])
    orc.set_state(*w["state"])
    for _ in range(3):
        dt_h = hyb.step()
        dt_o = orc.step()[0]
        sh, so = hyb.get_state(), orc.get_state()
        if hall:
            assert abs(dt_h - dt_o) <= 1e-12 * dt_o
            for a, b in zip(sh, so):
                assert maxrel(a, b) < 1e-12
        else:
            assert dt_h == dt_o
            for a, b in zip(sh, so):
                assert np.array_equal(a, b)
