"""Benchmark of the LTE time step on one MI355X.

Workload (default): the ab3Explicit step of GeodesicODIS on the 163842-cell
icosahedral-geodesic grid (reference GRID_LVL 8, "level 7" in cells
10*4^(l)+2 numbering), Enceladus ocean (R = 252.1 km, g = 0.113 m/s^2,
Omega = 5.307e-5 rad/s, h = 38 km, linear drag 1e-8 s^-1, eccentricity tide
e = 0.0047), rest start, FP64 plain method, dt = P/18300 = 6.46962 s.  The
time step is the reference's quantised dt (P / N_s, N_s a multiple of 100)
for a 6.5 s target: the reference's scheme (AB3 momentum, then AB3
continuity on the updated velocity) is stable for omega_max dt < 0.546 and
omega_max = 0.078 s^-1 on this grid, so 8 s steps diverge.  A "step" is one
pass of the loop body over all cells and faces.

Prints one JSON line: cell-steps/s of the whole job, the roofline of the
dominant kernel (the face kernel, or the fused step kernel with --fused;
HIP-event timed on every 16th step inside the timed region) and a
CPU baseline (the FP64 C oracle, OpenMP, on a bounded sample of the same
workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def face_bytes(E, N, nfields, p):
    """Algorithmic bytes of one face-kernel launch: every operand and state
    array touched once (p = bytes per value, int32 indices)."""
    per_face = 10 * 4 + 10 * p + 2 * 4 + 2 * p + p + p + nfields * p  # operands
    per_face += p + 2 * p + p + p  # v read, 2 history reads, history write, v write
    return E * per_face + N * p     # + eta read once


def cell_bytes(E, N, p):
    per_cell = 6 * 4 + 6 * p + p + p + 2 * p + p  # operands, eta r/w, history r2/w1
    return N * per_cell + E * p                   # + new v read once


def fused_bytes(E, N, nfields, p):
    """Algorithmic bytes of one fused step-kernel launch: the face kernel's
    set plus the continuity operands and state of the cells (eta^{n-1} read,
    eta^n written, two history reads and one write) and the owner byte per
    face; v is read once for both updates."""
    per_cell = 6 * 4 + 6 * p + p + 2 * p + p
    return face_bytes(E, N, nfields, p) + N * per_cell + E * 1


def measured_traffic(kname, tname, n_cells):
    """HBM bytes per launch of the kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, copy-calibrated), when one exists for
    this kernel, type and grid; else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rows = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    for r in rows:
        if r.get("kernel") == "%s<%s>" % (kname, tname) and r.get("cells") == n_cells:
            return r.get("hbm_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--level", type=int, default=8, help="reference grid level (8 = 163842 cells)")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU work of the baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fused", action="store_true",
                    help="one fused step kernel per step instead of the face + cell kernel pair")
    ap.add_argument("--graphs", action="store_true", help="replay 64-step chunks as HIP graphs")
    ap.add_argument("--timing-stride", type=int, default=16,
                    help="dispatch-stamped HIP events on the kernels of every T-th step (0: none, "
                         "roofline achieved then null)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (20 times), to calibrate "
                         "the FETCH_SIZE / WRITE_SIZE counters, and exit")
    args = ap.parse_args()
    if args.copy_calibration_mib:
        import torch
        n = args.copy_calibration_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(20):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 8, "launches": 20}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; multi-GPU runs are not part of it "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import geodesicodis_amd as G

    t0 = time.time()
    grid = G.Grid.generate(args.level, 64)
    p, steps_per_orbit = G.make_params(time_step=6.5)
    mesh = G.Mesh(grid, p.radius)
    prec = G.FP64 if args.precision == "fp64" else G.FP32
    pbytes = 8 if prec == G.FP64 else 4
    ctx = G.Context(mesh, p, precision=prec)
    ctx.use_fused(args.fused)
    ctx.use_graphs(args.graphs)
    print("bench: grid %d cells / %d faces ready in %.1fs" % (mesh.N, mesh.E, time.time() - t0),
          file=sys.stderr, flush=True)

    ctx.step(args.warmup, sync=True)
    ctx.enable_timing(args.timing_stride)  # dispatch-stamped HIP events on the kernels
    ctx.sync()
    t1 = time.perf_counter()
    ctx.step(args.steps, sync=True)
    ctx.sync()
    elapsed = time.perf_counter() - t1
    tm = ctx.timing()
    ctx.enable_timing(0)
    diss = float(ctx.dissipation()[0])
    ms_per_step = 1e3 * elapsed / args.steps
    value = mesh.N * args.steps / elapsed

    nfields = 2
    cb = cell_bytes(mesh.E, mesh.N, pbytes)
    if not args.fused:
        fb = face_bytes(mesh.E, mesh.N, nfields, pbytes)
        kname, step_bytes = "k_face", fb + cb
    else:
        fb = fused_bytes(mesh.E, mesh.N, nfields, pbytes)
        kname, step_bytes = "k_fused", fb
    achieved = fb / (tm["face_ms"] * 1e-3) / 1e9 if tm["face_ms"] > 0 else None
    tname = "double" if prec == G.FP64 else "float"
    traffic = measured_traffic(kname, tname, mesh.N)
    out = {
        "metric": "cell-steps/s at level 7 (163842 cells)",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == G.FP64 else "f32",
        "data": "synthetic (generated SCVT grid, rest start)",
        "config": {
            "workload": "GeodesicODIS LTE AB3 step, plain method, single instance, "
                        "Enceladus free-surface ocean, eccentricity tide, no self-gravity",
            "cells": mesh.N, "faces": mesh.E, "instances": 1,
            "grid": "icosahedral bisection + 64 Lloyd sweeps, reference level %d" % args.level,
            "dt_s": p.dt, "steps_per_orbit": steps_per_orbit,
            "launch": "fused step kernel, k_cell once per 4096-step chunk" if args.fused else
                      "face kernel + cell kernel per step",
            "graphs": bool(args.graphs),
            "kernel_ms": {"momentum": tm["face_ms"], "cell": tm["cell_ms"], "step_events": tm["step_ms"]},
            "bytes_per_step_alg": step_bytes,
            "dissipation_W_m2": diss,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": traffic,
            "kernel": "%s<%s>" % (kname, tname),
            "alg_bytes_per_launch": fb,
        },
    }
    ctx.close()

    if not args.no_cpu_baseline:
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        om = O.Mesh(grid.arrays(), p.radius)
        sim = O.Sim(om, O.make_params(g=p.g, h=p.h, alpha=p.alpha, omega=p.omega, radius=p.radius,
                                      ecc=p.ecc, dt=p.dt, tide=p.tide))
        sim.step(5)
        n, t_cpu = 0, 0.0
        batch = 10
        while t_cpu < args.cpu_seconds and n < 5000:
            ts = time.perf_counter()
            sim.step(batch)
            t_cpu += time.perf_counter() - ts
            n += batch
        out["cpu_baseline"] = {
            "value": mesh.N * n / t_cpu,
            "unit": "cell-steps/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", threads)),
            "kind": "port",
            "sample": "FP64 C oracle (OpenMP rows), same %d-cell grid and parameters, %d steps "
                      "after 5 warm-up steps" % (mesh.N, n),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
