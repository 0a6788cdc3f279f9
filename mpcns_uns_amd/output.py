"""Solution output for visual checks: the cell fields and file set of mercury::Solver::write_vtk
(MercuryUns/driver/Solver.cpp:677-718) written by the VTK writers of mhdg_mesh.h.

Per call: `sol_<step:07d>_r00000.vtu` (one piece: one GPU), the master `sol_<step:07d>.pvtu` and the
collection `solution.pvd` listing every output so far with its time.
"""
import os

import numpy as np

from .mesh import write_pvd, write_pvtu


def solution_fields(source, mesh):
    """(name, ncomp, values) of Solver::write_vtk from an object with derived(name) after a prepare."""
    W = [np.asarray(source.derived(n)).reshape(-1, 5) for n in ("W_H", "W_Na")]
    return [("rho_H", 1, W[0][:, 0]), ("rho_Na", 1, W[1][:, 0]), ("p_H", 1, W[0][:, 4]), ("p_Na", 1, W[1][:, 4]),
            ("u_H", 3, W[0][:, 1:4].ravel()), ("u_Na", 3, W[1][:, 1:4].ravel()), ("B_induced", 3, source.derived("B_ind")),
            ("B_total", 3, source.derived("B_tot")), ("J", 3, source.derived("J")), ("u_plus", 3, source.derived("u_plus")),
            ("material", 1, mesh.array("material").astype(np.float64))]


class SolutionWriter:
    """The output files of one run in `out_dir`."""

    def __init__(self, out_dir):
        self.out_dir = str(out_dir)
        self.entries = []

    def write(self, mesh, fields, step, time):
        os.makedirs(self.out_dir, exist_ok=True)
        base = "sol_%07d" % step
        piece = base + "_r00000.vtu"
        mesh.write_vtu(os.path.join(self.out_dir, piece), fields)
        write_pvtu(os.path.join(self.out_dir, base + ".pvtu"), [piece], [(n, c) for n, c, _ in fields])
        self.entries.append((time, base + ".pvtu"))
        write_pvd(os.path.join(self.out_dir, "solution.pvd"), self.entries)

    def write_device(self, dev, step, time):
        """Solver::write_vtk of the resident state of a Device (prepare on the device, then the derived fields)."""
        dev.prepare()
        self.write(dev.mesh, solution_fields(dev, dev.mesh), step, time)
