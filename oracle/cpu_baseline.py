"""ORACLE — test infrastructure only: CPU baseline of bench.py (run as a subprocess).

Times the Mercury step with the Hall solve on the CPU, on the benchmark's own workload or a stated sample of it
(same generator, physics and solver settings), and prints one JSON line.  Kinds:

  mpi        the compiled reference (oracle/_ref/ref_mpi_bench, the reference's sources with its Release flags) on
             NRANKS MPI ranks of the host: the mesh partitioned by the reference's own preprocessing steps and
             mercury::Solver on MPI_COMM_WORLD, as MercuryUns runs (MercuryUns/main.cpp); the step's time is the
             solver's own wall clock (Solver.cpp:231, 275), maximum over ranks.  Every timed step is the first step
             from the seeded state with a cold implicit solve, the device benchmark's protocol.
  reference  the compiled reference in this process on one rank (oracle/_ref/libmpcns_ref.so), same protocol.
  port       the oracle restatement (OpenMP over the host cores), same protocol.

The reference's Mercury solver keeps only the inflow edges Dirichlet in the implicit solve (Solver.cpp:262), whatever
the workload's wall rule; the GMRES iterations it took are reported beside its time.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[0] = os.path.dirname(HERE)  # repository root first (this directory holds oracle.py)
MPI_BENCH = os.path.join(HERE, "_ref", "ref_mpi_bench")


def find_mpiexec():
    for p in (shutil.which("mpiexec.hydra"), "/opt/conda/bin/mpiexec.hydra", shutil.which("mpiexec"), "/opt/conda/bin/mpiexec"):
        if p and os.path.exists(p):
            return p
    return None


def tagged_faces(mesh):
    """Node triples (ascending) and tag ids of the tagged boundary faces (triangles)."""
    ftag = mesh.array("face_tag")
    f = np.nonzero(ftag)[0]
    ptr, idx = mesh.array("f2e.ptr"), mesh.array("f2e.idx")
    e2n = mesh.array("e2n.idx").reshape(-1, 2)
    if np.any(ptr[f + 1] - ptr[f] != 3):
        raise ValueError("tagged faces must be triangles")
    edges = idx[ptr[f][:, None] + np.arange(3)[None, :]]
    nodes = np.sort(e2n[edges].reshape(-1, 6), axis=1)[:, ::2]  # each node of a triangle appears in two of its edges
    return np.ascontiguousarray(nodes, dtype=np.int32), np.ascontiguousarray(ftag[f], dtype=np.int32)


def write_mpi_input(path, w, dipole_m):
    """input.bin of ref_mpi_bench (its header comment gives the layout)."""
    import ctypes as C

    mesh, cfg = w["mesh"], w["cfg"]
    UH, UNa, Phi = w["state"]
    xyz = np.ascontiguousarray(mesh.array("xyz"), dtype=np.float64)
    tets = np.ascontiguousarray(mesh.array("tets"), dtype=np.int32)
    tn, tid = tagged_faces(mesh)
    with open(path, "wb") as f:
        f.write(b"MHDGBIN1")
        f.write(np.array([len(xyz.reshape(-1, 3)), mesh.n_cells, mesh.n_faces, len(tid)], np.int64).tobytes())
        for a in (xyz, tets, tn, tid):
            f.write(a.tobytes())
        f.write(np.array([C.sizeof(cfg)], np.int64).tobytes())
        f.write(bytes(cfg))
        f.write(np.array([1], np.int32).tobytes())
        f.write(np.array([*dipole_m, 0.0, 0.0, 0.0, 1e-3], np.float64).tobytes())
        for a in (UH, UNa, Phi):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())


def run_mpi(w, ranks, steps, dipole_m, workdir=None):
    mpiexec = find_mpiexec()
    if not (os.path.exists(MPI_BENCH) and mpiexec):
        raise RuntimeError("ref_mpi_bench or mpiexec not present")
    d = tempfile.mkdtemp(prefix="mhdg_mpi_", dir=workdir)
    try:
        t = time.perf_counter()
        write_mpi_input(os.path.join(d, "input.bin"), w, dipole_m)
        subprocess.run([MPI_BENCH, "parts", d, str(ranks)], check=True, capture_output=True, text=True, timeout=900)
        prep = time.perf_counter() - t
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([mpiexec, "-n", str(ranks), MPI_BENCH, "run", d, str(steps)], capture_output=True, text=True,
                           timeout=1800, env=env)
        if r.returncode != 0:
            raise RuntimeError(f"mpiexec exit {r.returncode}: {(r.stderr or r.stdout)[-600:]}")
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1]
        out = json.loads(line)
        out["prep_s"] = round(prep, 1)
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="shell_8m_mixed")
    ap.add_argument("--net", type=int, nargs=3, default=None, help="a smaller sample net (n_lat n_lon n_r) of the workload")
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--kind", default="auto", choices=["auto", "mpi", "reference", "port"])
    ap.add_argument("--ranks", type=int, default=0, help="MPI ranks (default: the usable host cores, at most 16)")
    a = ap.parse_args()
    from mpcns_uns_amd import workloads

    w = workloads.build(a.workload, net=tuple(a.net) if a.net else None)
    mesh, cfg = w["mesh"], w["cfg"]
    net = w["spec"]["net"]
    sample = (f"{a.steps} first steps (seeded state, cold implicit solve) of {a.workload} on the "
              f"{net[0]}x{net[1]}x{net[2]} shell, {mesh.n_cells} cells")
    kind = a.kind
    if kind == "auto":
        from oracle.oracle import have_ref

        kind = "mpi" if (os.path.exists(MPI_BENCH) and find_mpiexec()) else ("reference" if have_ref() else "port")
    if kind == "mpi":
        ranks = a.ranks or min(16, len(os.sched_getaffinity(0)))
        r = run_mpi(w, ranks, a.steps, workloads.DIPOLE_M)
        step_s = float(np.median(r["step_wall_s"]))
        print(json.dumps({"value": round(mesh.n_cells / step_s, 1), "unit": "cell-steps/s", "cores": ranks, "kind": "reference",
                          "runner": f"mpiexec -n {ranks} (MPICH), one rank per core", "sample": sample,
                          "step_wall_s": r["step_wall_s"], "gmres_iterations": r["gmres_iterations"],
                          "max_rel_residual": r["max_rel_residual"], "dirichlet": "inflow edges (Mercury, Solver.cpp:262)",
                          "prep_s": r["prep_s"]}))
        return
    from oracle.oracle import Oracle, reference_for_mesh

    times, iters = [], []
    for _ in range(a.steps):
        if kind == "reference":
            x = reference_for_mesh(mesh, cfg, dipole_m=workloads.DIPOLE_M)
            x.solver_init()
        else:
            x = Oracle(mesh, cfg)
            x.set_background(*w["background"])
        x.set_state(*w["state"])
        t = time.perf_counter()
        out = x.step()
        times.append(time.perf_counter() - t)
        if isinstance(out, tuple) and len(out) > 1:
            iters.append(int(out[1]))
    cores = 1 if kind == "reference" else int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    step_s = float(np.median(times))
    print(json.dumps({"value": round(mesh.n_cells / step_s, 1), "unit": "cell-steps/s", "cores": cores,
                      "kind": "reference" if kind == "reference" else "port", "sample": sample,
                      "step_wall_s": [round(t, 3) for t in times], "gmres_iterations": iters or None}))


if __name__ == "__main__":
    main()
