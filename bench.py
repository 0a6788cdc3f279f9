"""Benchmark of the Hall-MHD step on one MI355X: cell-steps/s of the Mercury step (explicit stage + implicit
Hall/resistive GMRES solve to 1e-8 + constrained transport) on the seeded 7.83M-cell tetrahedral shell.

Prints one JSON line.  `value` = cells x timed steps / wall seconds of the timed region (state resident in
HBM).  Every timed step is the workload's first step: the seeded state is restored on the device (a
device-to-device copy, inside the timed region) and the implicit solve starts cold, because the run itself
diverges at the poles of the lat-lon net within a few steps (the reference's discrete Hall operator is not
energy-neutral on the polar wedge cells); `trajectory` reports the consecutive steps of that run as well.  `roofline` is the dominant kernel's algorithmic HBM bytes per launch over its average launch time,
measured with HIP events on the context stream during the timed region.  `cpu_baseline` times the CPU
code (the compiled reference where present, else the oracle restatement) on a bounded sample of the same
workload in a subprocess.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM = 8.0e12  # MI355X HBM3E, bytes/s

# HBM bytes per launch moved by a kernel on a workload, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs scaled by
# what the same counters read for the library's streaming kernels over 2 GiB arrays (x2.00 and x1.00); the command
# and the per-kernel table are in profiles/hall_step_shell_8m_hbm_counters.md.  Counter runs cannot share a process
# with this measurement.
PMC_TRAFFIC = {("shell_8m_mixed", "lat", "edge_apply_f32"): 1239373251}


def step_compulsory_bytes(its, restart, ns, mixed, vcycle=False):
    """Compulsory HBM bytes per cell-step of the Hall step (every input read once, every output written once; tet
    ratios F = 2C, E = 7C/6): the explicit stage 80 ns + 437 B; a matrix-free apply_A 329 B in FP64, 190 B in FP32;
    Arnoldi step j (2j + 5) 9.33 B in FP64, half in FP32; the Jacobi product 28 B per iteration. The solve takes `its`
    Krylov iterations in cycles of at most `restart` and three kinds of FP64 product besides: b, the first residual
    and one residual per cycle. With the multigrid preconditioner each iteration adds one V-cycle, about 1.3 kB (six
    FP32 operator products and the transfers over a hierarchy of ratio 8)."""
    its = int(round(its))
    cycles = [restart] * (its // restart) + ([its % restart] if its % restart else [])
    arnoldi = sum((2 * j + 5) * 9.33 for n in cycles for j in range(n)) * (0.5 if mixed else 1.0)
    krylov = its * (190.0 if mixed else 329.0)
    return 80.0 * ns + 437.0 + krylov + (2 + len(cycles)) * 329.0 + arnoldi + 28.0 * its + (1300.0 * its if vcycle else 0.0)


def kernel_bytes(mesh, name, precision_bytes=8, ns=2):
    """Compulsory HBM bytes of one launch of a kernel: every array it reads or writes counted once (ns ion species)."""
    C, F, E = mesh.n_cells, mesh.n_faces, mesh.n_edges
    nc2f, nf2e, nec = 4 * C, 3 * F, int(mesh.array("ec_ptr")[-1])
    b = precision_bytes
    table = {
        # x (E) read, phi (F) write, f2e ptr+idx
        "d1": E * b + F * b + 4 * (F + 1) + 4 * nf2e,
        # phi (F), recon (3 per c2f entry), b (3C) write, c2f ptr+idx
        "bcell": F * b + 3 * nc2f * 8 + 3 * C * b + 4 * (C + 1) + 4 * nc2f,
        # b (3C), alpha (C), B0 (3C), area (3F), f2c (2F), volume, g (3C) write, c2f
        "hall_g": 3 * C * b + C * 8 + 3 * C * 8 + 3 * F * 8 + 8 * F + C * 8 + 3 * C * b + 4 * (C + 1) + 4 * nc2f,
        # phi, star2, e2f, g (3C), ec lists + weights, edge vec, m_e, v, invdiag, dir, out
        "edge_apply": F * b + F * 8 + 4 * (E + 1) + 4 * nf2e + 3 * C * b + 4 * (E + 1) + 12 * nec + 24 * E + 8 * E
                      + 2 * E * b + 8 * E + E + E * b,
        # one fused MGS step: w read + write, two basis vectors
        "mgs": 4 * E * b,
        "prepare": ns * 5 * C * 8 * 2 + F * 8 + 3 * C * 8 * 5 + C * 8 + 4 * F * 8 + 4 * nc2f + 4 * (C + 1) + C * 8 + C,
        "grad_current": ns * (5 * C * 8 + 15 * C * 8) + 3 * C * 8 * 2 + 6 * C * 8 + nc2f * (4 + 8 + 24) + 3 * F * 8 * 2 + 3 * C * 8
                        + 8 * F + C * 8 + C + 4 * (C + 1),
        "face": ns * (5 * C * 8 + 15 * C * 8) + 5 * ns * F * 8 + 3 * F * 8 + 3 * C * 8 * 3 + F * 8 * 2 + 3 * F * 8 * 2 + F * 8 + 8 * F
                + 3 * C * 8 + F,
        "cell_update": 5 * ns * F * 8 + ns * 5 * C * 8 * 3 + 3 * C * 8 * 2 + C * 8 + C + 4 * nc2f + 4 * (C + 1),
        "edge_emf": 3 * F * 8 + 3 * C * 8 + F * 8 + 4 * (E + 1) + 4 * nf2e + 8 * nf2e + 4 * (E + 1) + 4 * nec + 3 * F * 8 * 2 + F * 8
                    + 3 * E * 8 * 2 + E * 8 + E * 8 + E,
        "faraday": E * 8 + F * 8 * 2 + 4 * (F + 1) + 4 * nf2e + 8 * F,
        # FP32 operator of the mixed solve (int32 incidence, FP32 values)
        "d1_f32": 4 * (F + 1) + 4 * nf2e + 4 * E + 4 * E + E + 4 * F,
        "bcell_f32": 4 * (C + 1) + 4 * nc2f + 12 * nc2f + 4 * F + 12 * C,
        # alpha B0 (3C), b (3C), c2f ptr + idx, f2c, area, volume, g (3C) write
        "hall_g_f32": 12 * C + 4 * (C + 1) + 4 * nc2f + 8 * F + 12 * F + 4 * C + 12 * C + 12 * C,
        # dir, the two rows in compact CSR form (row offsets + (index, weight) entries: e2f with signed star2, edge-cell
        # with weights), phi, edge vec, g, m_e, pv, v, out.  The kernel streams the rows in 64-edge slices of 12-byte
        # pairs with 16-bit offsets, padded to each slice's longest row, plus a valence flag per edge (1.135 GB at
        # 8M against these 1.269 GB); the count stays the CSR size with 32-bit ids.
        "edge_apply_f32": E + 4 * (E + 1) + 8 * nf2e + 4 * F + 12 * E + 4 * (E + 1) + 8 * nec + 12 * C + 4 * E * 4,
        # the same gather with the multigrid outputs: A x (x, out), a Jacobi sweep or a residual (x, b, invdiag, out)
        "mg_apply_f32": E + 4 * (E + 1) + 8 * nf2e + 4 * F + 12 * E + 4 * (E + 1) + 8 * nec + 12 * C + 4 * E * 3,
        "mg_smooth_f32": E + 4 * (E + 1) + 8 * nf2e + 4 * F + 12 * E + 4 * (E + 1) + 8 * nec + 12 * C + 4 * E * 5,
        "mg_resid_f32": E + 4 * (E + 1) + 8 * nf2e + 4 * F + 12 * E + 4 * (E + 1) + 8 * nec + 12 * C + 4 * E * 4,
    }
    return table.get(name)


KERNEL_CLASSES = {
    "explicit": ("prepare", "dt_final", "grad_current", "face", "face_emf", "cell_update", "edge_emf", "edge_emf_heavy", "faraday",
                 "gyrate"),
    "operator": ("d1", "bcell", "hall_g", "edge_apply", "edge_apply_heavy", "d1_f32", "bcell_f32", "hall_g_f32", "edge_apply_f32",
                 "mg_apply_f32"),
    "orthogonalisation": ("mgs", "finish", "dot", "multidot", "multidot_final", "multiaxpy", "scale", "multidot64", "multidot64_final",
                          "multiaxpy64", "norm64"),
    "multigrid": ("mg_first", "mg_smooth_f32", "mg_resid_f32", "mg_sweep", "mg_resid", "mg_restrict", "mg_prolong", "mg_dinv"),
}


def split_by_class(breakdown, steps):
    """Milliseconds per step of each kernel class (explicit stage, Krylov operator, orthogonalisation, multigrid,
    the rest) from a per-kernel breakdown over `steps` steps."""
    out = {k: 0.0 for k in (*KERNEL_CLASSES, "other")}
    for name, (_, ms) in (breakdown or {}).items():
        cls = next((k for k, names in KERNEL_CLASSES.items() if name in names), "other")
        out[cls] += ms / max(steps, 1)
    return {k: round(v, 3) for k, v in out.items()}


def stream_calibration(gib=2.0):
    """The device's streaming rates (GB/s) from the library's read, copy and triad kernels over `gib` GiB arrays."""
    import ctypes as C

    from mpcns_uns_amd._lib import lib

    out = {}
    for kind, name in enumerate(("read", "copy", "triad")):
        us, moved = C.c_double(), C.c_double()
        if lib().mhdg_stream_calibration(kind, int(gib * 2 ** 30), 5, C.byref(us), C.byref(moved)):
            return None
        out[name + "_GBps"] = round(moved.value / (us.value * 1e-6) / 1e9, 1)
    out["array_GiB"] = gib
    return out


def run_steps(dev, steps, precision, dt_max, restore=True):
    its, res, dts = [], [], []
    for _ in range(steps):
        if restore:
            dev.state_restore()
        dt, it, rr = dev.step(dt_max, precision)
        its.append(it)
        res.append(rr)
        dts.append(dt)
    return its, res, dts


def cpu_baseline(args):
    """The CPU leg, in a subprocess: by default the compiled reference under MPI on the host cores, on the benchmark's
    own workload (oracle/cpu_baseline.py)."""
    cmd = [sys.executable, "-u", os.path.join(HERE, "oracle", "cpu_baseline.py"), "--workload", args.workload,
           "--steps", str(args.cpu_steps)]
    if args.cpu_net:
        cmd += ["--net", *map(str, args.cpu_net)]
    elif args.net:
        cmd += ["--net", *map(str, args.net)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
        return json.loads(line[-1]) if line else {"error": (r.stderr or r.stdout)[-400:]}
    except Exception as e:  # the baseline is reported, never required
        return {"error": str(e)}


def same_step_as_reference(w, precision, steps=2):
    """The device step the CPU reference times: the first step from the seeded state under the reference Mercury
    solver's Dirichlet rule (inflow edges only, Solver.cpp:262), so the two solves run the same iteration."""
    from mpcns_uns_amd import Device, workloads

    cfg = type(w["cfg"]).from_buffer_copy(w["cfg"])
    cfg.wall_dirichlet = 0
    dev = Device(w["mesh"], cfg)
    try:
        workloads.load(dev, w)
        dev.state_save()
        run_steps(dev, 1, precision, 1e9)
        dev.sync()
        t = time.perf_counter()
        its, res, _ = run_steps(dev, steps, precision, 1e9)
        dev.sync()
        el = (time.perf_counter() - t) / steps
    finally:
        dev.close()
    return dict(ms_per_step=round(el * 1e3, 3), value=round(w["mesh"].n_cells / el, 1), gmres_iterations_per_step=float(np.mean(its)),
                max_rel_residual=float(max(res)), precision={0: "fp64", 1: "mixed", 2: "mg"}[precision],
                dirichlet="inflow edges (Mercury)")


def differentiation(args):
    """Gradient of J = sum |B|^2 V over 1.2 <= r <= 1.5 after `steps` explicit steps with respect to the 3 IMF
    components and the upstream density: 4 tangent sweeps against one adjoint sweep (all inputs at once)."""
    from mpcns_uns_amd import Device, workloads

    w = workloads.build(args.workload, net=tuple(args.net) if args.net else None, hall=False)
    mesh = w["mesh"]
    dev = Device(mesh, w["cfg"])
    workloads.load(dev, w)
    r0, r1, n = 1.2, 1.5, args.steps
    dPhiAdd = np.ascontiguousarray(mesh.array("face_area").reshape(-1, 3)[:, 0])
    dBadd = mesh.cell_vector(dPhiAdd)
    dev.step_vjp(n, r0, r1)  # warm-up: allocation and gather maps
    dev.explicit_tangent(n, r0, r1, dPhiAdd=dPhiAdd, dBadd=dBadd)
    dev.sync()
    t = time.perf_counter()
    J, g = dev.step_vjp(n, r0, r1)
    dev.sync()
    t_adj = time.perf_counter() - t
    t = time.perf_counter()
    _, dJ = dev.explicit_tangent(n, r0, r1, dPhiAdd=dPhiAdd, dBadd=dBadd)
    dev.sync()
    t_tan = time.perf_counter() - t
    agree = abs(g["PhiAdd"] @ dPhiAdd + g["Badd"] @ dBadd - dJ) / abs(dJ)
    print(json.dumps({"metric": "shell-energy gradient after n explicit steps, wall s", "workload": args.workload,
                      "cells": mesh.n_cells, "steps": n, "adjoint_s": round(t_adj, 4), "tangent_sweep_s": round(t_tan, 4),
                      "tangent_4_params_s": round(4 * t_tan, 4), "J": J,
                      "adjoint_vs_tangent_rel": agree}))


PRECISIONS = {"fp64": 0, "mixed": 1, "mg": 2, "fp64-cgs": 3}
PRECONDITIONER = {0: "jacobi", 1: "jacobi", 2: "multigrid", 3: "jacobi"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="shell_8m_mixed")
    ap.add_argument("--net", type=int, nargs=3, default=None, help="override the lat-lon net (n_lat n_lon n_r)")
    ap.add_argument("--precision", default="mixed", choices=sorted(PRECISIONS),
                    help="fp64: the reference iteration; mixed: FP32 GMRES with Jacobi in FP64 refinement; mg: the same "
                         "with the agglomeration-multigrid V-cycle as preconditioner")
    ap.add_argument("--order", default="lat", choices=["lat", "hilbert"],
                    help="cell numbering: the generator's lat-row order or the Hilbert order of the cell centres")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 step and the other preconditioner")
    ap.add_argument("--no-fp64-reference", action="store_true", help="skip the plain FP64 step measurement")
    ap.add_argument("--cpu-baseline", dest="cpu", action="store_true", default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu", action="store_false")
    ap.add_argument("--cpu-net", type=int, nargs=3, default=None,
                    help="a smaller sample net for the CPU leg (default: the workload's own net)")
    ap.add_argument("--cpu-steps", type=int, default=1, help="first steps the CPU leg times")
    ap.add_argument("--no-stream", dest="stream", action="store_false", default=True,
                    help="skip the streaming calibration (read, copy, triad kernels)")
    ap.add_argument("--stream-only", action="store_true",
                    help="only run the library's streaming calibration kernels (read, copy, triad over 2 GiB arrays)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only run a device-to-device copy of GIB GiB five times (scale of the HBM byte counters)")
    ap.add_argument("--differentiation", action="store_true",
                    help="only time the gradient of the shell energy after --steps explicit steps (forward mode: one "
                         "tangent sweep per parameter; reverse mode: mhdg_step_vjp) on the workload's mesh, Hall off")
    args = ap.parse_args()
    if args.differentiation:
        differentiation(args)
        return
    if args.stream_only:
        print(json.dumps(stream_calibration()))
        return
    if args.copy_calibration > 0:
        import torch

        n = int(args.copy_calibration * 2 ** 30) // 4
        x = torch.ones(n, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        ev[0].record()
        for i in range(5):
            y.copy_(x)
            ev[i + 1].record()
        torch.cuda.synchronize()
        us = sorted(ev[i].elapsed_time(ev[i + 1]) * 1e3 for i in range(5))
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "repeats": 5, "median_us": round(us[2], 1),
                          "copy_GBps": round(8 * n / (us[2] * 1e-6) / 1e9, 1)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark measures one GPU; the step would shard by mesh partition with a halo "
                                   "exchange per step, which is not part of this benchmark", "n_gpus": args.gpus}))
        sys.exit(2)

    from mpcns_uns_amd import Device, workloads

    t0 = time.time()
    w = workloads.build(args.workload, net=tuple(args.net) if args.net else None, order=None if args.order == "lat" else "hilbert")
    mesh, cfg = w["mesh"], w["cfg"]
    setup_s = time.time() - t0
    print(f"# workload {args.workload} ({args.order} order) cells={mesh.n_cells} faces={mesh.n_faces} edges={mesh.n_edges} "
          f"setup={setup_s:.1f}s", file=sys.stderr, flush=True)
    dev = Device(mesh, cfg)
    ns = max(int(cfg.n_species), 2)
    precision = PRECISIONS[args.precision]
    dt_max = 1e9

    def measure(prec, profile, restore=True):
        """Warmup steps with every kernel timed (the per-kernel breakdown), then the timed steps with HIP events
        around the dominant kernel only (its launch durations for the roofline).  Returns the state after the last
        step as well (with restore: the first step's result)."""
        workloads.load(dev, w)
        dev.state_save()
        # the first warmup step carries one-time work (solver set-up, hierarchy build, first touch): not in the breakdown
        first = 1 if args.warmup > 1 else 0
        run_steps(dev, first, prec, dt_max, restore)
        if profile:
            dev.profile_only(None)
            dev.profile(None)
            dev.profile(True)
        run_steps(dev, args.warmup - first, prec, dt_max, restore)
        dev.sync()
        breakdown, dom = None, None
        if profile:
            dev.profile(False)
            breakdown = dev.profile_report()
            dev.profile(None)
            if breakdown:
                dom = max(breakdown, key=lambda k: breakdown[k][1] if kernel_bytes(mesh, k, ns=ns) else -1)
                dev.profile_only(dom)
            dev.profile(True)  # without profiled warmup steps: every kernel in the timed region
        t = time.perf_counter()
        its, res, dts = run_steps(dev, args.steps, prec, dt_max, restore)
        dev.sync()
        el = time.perf_counter() - t
        timed = None
        if profile:
            dev.profile(False)
            timed = dev.profile_report()
            dev.profile_only(None)
            if not breakdown:
                breakdown = timed
                dom = max(timed, key=lambda k: timed[k][1] if kernel_bytes(mesh, k, ns=ns) else -1)
        return dict(el=el, its=its, res=res, prof=(breakdown, dom, timed), dts=dts, state=dev.get_state())

    nbd = args.warmup - 1 if args.warmup > 1 else args.warmup  # steps in a breakdown

    def summary(m):
        return dict(value=round(mesh.n_cells * args.steps / m["el"], 1), ms_per_step=round(m["el"] / args.steps * 1e3, 3),
                    gmres_iterations_per_step=float(np.mean(m["its"])), max_rel_residual=float(max(m["res"])),
                    ms_per_step_by_class=split_by_class(m["prof"][0], nbd))

    # main measurement; the timed region carries events around the dominant kernel's launches only
    main_m = measure(precision, True)
    el, its, res, prof = main_m["el"], main_m["its"], main_m["res"], main_m["prof"]
    value = mesh.n_cells * args.steps / el
    print(f"# {args.precision}: {el / args.steps * 1e3:.2f} ms/step, GMRES it/step {np.mean(its):.1f}, max res {max(res):.2e}",
          file=sys.stderr, flush=True)
    fp64, fp64_cgs, field_err, precond = None, None, None, {}
    dirichlet = mesh.array("edge_mode") != 0 if cfg.wall_dirichlet else mesh.array("edge_mode") == 2
    if precision != 0 and not (args.no_compare or args.no_fp64_reference):
        m64 = measure(0, True)
        fp64 = summary(m64)
        print(f"# fp64: {fp64['ms_per_step']:.2f} ms/step, GMRES it/step {fp64['gmres_iterations_per_step']:.1f}", file=sys.stderr, flush=True)
        # the FP64 iteration with the mixed solve's orthogonalisation (classical Gram-Schmidt): splits the mixed gain
        # into its precision part (fp64_cgs / mixed) and its orthogonalisation part (fp64 / fp64_cgs)
        fp64_cgs = summary(measure(3, True))
        print(f"# fp64 CGS: {fp64_cgs['ms_per_step']:.2f} ms/step, GMRES it/step {fp64_cgs['gmres_iterations_per_step']:.1f}",
              file=sys.stderr, flush=True)
        # per-field difference of the first step's result against the all-FP64 step (both solves to 1e-8)
        field_err = {}
        for name, a, b in zip(("U_H", "U_Na", "Phi"), main_m["state"], m64["state"]):
            field_err[name] = float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))
    if precision != 0:
        precond[PRECONDITIONER[precision]] = summary(main_m)
        if not args.no_compare:
            other = 2 if precision == 1 else 1
            mo = measure(other, True)
            precond[PRECONDITIONER[other]] = summary(mo)
            print(f"# {PRECONDITIONER[other]}: {precond[PRECONDITIONER[other]]['ms_per_step']:.2f} ms/step, GMRES it/step "
                  f"{np.mean(mo['its']):.1f}", file=sys.stderr, flush=True)
        if "multigrid" in precond:
            precond["multigrid"]["unknowns_per_level"] = dev.mg_info(dirichlet)
            precond["multigrid"]["cycle"] = dict(levels=int(cfg.mg_levels), agglomeration_passes=int(cfg.mg_agglom_passes),
                                                 pre=int(cfg.mg_pre), post=int(cfg.mg_post), coarse=int(cfg.mg_coarse),
                                                 relax=float(cfg.mg_relax), fine_smoother="full operator" if cfg.mg_fine_hall
                                                 else "resistive part")

    # roofline of the dominant kernel (largest total time in the warmup breakdown), timed in the timed region
    roof = None
    kernels = {}
    breakdown, dom, timed = prof
    stream = stream_calibration() if args.stream else None
    if breakdown and dom in (timed or {}):
        total_ms = sum(v[1] for v in breakdown.values())
        for k, (n, ms) in sorted(breakdown.items(), key=lambda kv: -kv[1][1]):
            kernels[k] = dict(launches=n, ms_total=round(ms, 4), share=round(ms / total_ms, 4))
        n, ms = timed[dom]
        byt = kernel_bytes(mesh, dom, ns=ns)
        avg_s = ms / n / 1e3
        achieved = byt / avg_s
        traffic = PMC_TRAFFIC.get((args.workload, args.order, dom)) if not args.net else None
        roof = dict(kernel=dom, bound="hbm", achieved=round(achieved / 1e9, 1), peak=PEAK_HBM / 1e9, unit="GB/s",
                    frac=round(achieved / PEAK_HBM, 4), traffic=traffic, bytes_per_launch=int(byt),
                    avg_launch_us=round(avg_s * 1e6, 2), launches=n)
        if stream:
            roof["stream_copy_GBps"] = stream["copy_GBps"]
            roof["frac_of_stream_copy"] = round(achieved / 1e9 / stream["copy_GBps"], 4)

    # the whole step against the compulsory-byte model of the path (no intermediates, no index overhead)
    step_comp = None
    if w["spec"]["hall"]:
        bpc = step_compulsory_bytes(np.mean(its), int(cfg.ind_restart), ns, precision in (1, 2), precision == 2)
        step_comp = dict(bytes_per_cell_step=round(bpc, 1), achieved_GBps=round(bpc * value / 1e9, 1),
                         frac=round(bpc * value / PEAK_HBM, 4))

    # the run's own consecutive steps after the warmup (no restore): dt collapses as the polar field grows
    mt = measure(precision, False, restore=False)
    dtt = mt["dts"]
    trajectory = dict(ms_per_step=round(mt["el"] / args.steps * 1e3, 3), gmres_iterations_per_step=float(np.mean(mt["its"])),
                      max_rel_residual=float(max(mt["res"])), dt_first=dtt[0], dt_last=dtt[-1])
    print(f"# trajectory: {trajectory['ms_per_step']:.2f} ms/step, GMRES it/step {trajectory['gmres_iterations_per_step']:.1f}, "
          f"dt {dtt[0]:.2e} -> {dtt[-1]:.2e}", file=sys.stderr, flush=True)
    cpu = None
    if args.cpu:
        # the device on the step the CPU reference runs, then the CPU reference itself
        same = same_step_as_reference(w, precision) if w["spec"]["hall"] else None
        cpu = cpu_baseline(args)
        if same and "value" in cpu and not args.cpu_net:
            cpu["gpu_same_step"] = same
            cpu["gpu_over_cpu_same_step"] = round(same["value"] / cpu["value"], 1)
    out = {
        "metric": "cell-steps/s, mixed-prec vs FP64 GMRES+MG Hall step at 1e-8 rel. residual",
        "value": round(value, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision in (0, 3) else "f64+f32",
        "data": "synthetic (seeded tetrahedral shell, seeded noisy state)",
        "config": {"workload": args.workload, "cells": mesh.n_cells, "faces": mesh.n_faces, "edges": mesh.n_edges,
                   "net": list(w["spec"]["net"]), "cell_order": args.order, "precision": args.precision,
                   "preconditioner": PRECONDITIONER[precision], "hall": bool(w["spec"]["hall"]), "ion_species": ns,
                   "wall_dirichlet": int(cfg.wall_dirichlet), "gmres_restart": int(cfg.ind_restart), "rtol": float(cfg.ind_rtol),
                   "gmres_iterations_per_step": float(np.mean(its)), "max_rel_residual": float(max(res)),
                   "setup_s": round(setup_s, 1), "field_error_vs_fp64": field_err},
        "protocol": "every timed step is the first step from the seeded state (device-side restore, cold solve)",
        "ms_per_step_by_class": split_by_class(breakdown, nbd),
        "trajectory": trajectory,
        "fp64_step": fp64,
        "vs_fp64": round(value / fp64["value"], 3) if fp64 else None,
        "fp64_cgs_step": fp64_cgs,
        "vs_fp64_cgs": round(value / fp64_cgs["value"], 3) if fp64_cgs else None,
        "field_error_vs_fp64": field_err,
        "preconditioners": precond or None,
        "roofline": roof,
        "step_compulsory": step_comp,
        "stream": stream,
        "kernels": kernels,  # per-kernel totals over the warmup steps
        "kernels_over_steps": nbd,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
