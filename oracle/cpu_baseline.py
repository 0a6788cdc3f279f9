"""ORACLE — test infrastructure only: CPU baseline of bench.py (run as a subprocess).

Times the Mercury step with the Hall solve on a bounded sample of the benchmark workload (same generator
and physics, smaller net): the compiled reference (oracle/_ref, one MPI rank) when it is present, else the
oracle restatement (OpenMP over the host cores).  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[0] = os.path.dirname(HERE)  # repository root first (this directory holds oracle.py)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--net", type=int, nargs=3, default=(48, 96, 10))
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--kind", default="auto", choices=["auto", "reference", "port"])
    a = ap.parse_args()
    from mpcns_uns_amd import workloads
    from oracle.oracle import Oracle, have_ref, reference_for_mesh

    w = workloads.build(net=tuple(a.net), hall=True)
    mesh, cfg = w["mesh"], w["cfg"]
    kind = "reference" if (a.kind == "reference" or (a.kind == "auto" and have_ref())) else "port"
    if kind == "reference":
        x = reference_for_mesh(mesh, cfg, dipole_m=workloads.DIPOLE_M)
        x.solver_init()
        x.set_state(*w["state"])
        x.step()  # warm-up step (first solve starts cold, as in the device measurement's warmup)
        t = time.perf_counter()
        for _ in range(a.steps):
            x.step()
        el = time.perf_counter() - t
        cores = 1
    else:
        x = Oracle(mesh, cfg)
        x.set_background(*w["background"])
        x.set_state(*w["state"])
        x.step()
        t = time.perf_counter()
        for _ in range(a.steps):
            x.step()
        el = time.perf_counter() - t
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    print(json.dumps({"value": round(mesh.n_cells * a.steps / el, 1), "unit": "cell-steps/s", "cores": cores, "kind": kind,
                      "sample": f"{a.steps} Hall steps (after 1 warm-up) on the {a.net[0]}x{a.net[1]}x{a.net[2]} shell, "
                                f"{mesh.n_cells} cells, same physics and solver settings", "seconds": round(el, 2)}))


if __name__ == "__main__":
    main()
