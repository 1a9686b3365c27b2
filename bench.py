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
dominant kernel (the one-kernel step, or the longer of the face and cell
kernels; HIP-event timed on every 16th step inside the timed region) and a
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


def face_bytes(E, N, nfields, p, compact):
    """Algorithmic bytes of one face-kernel launch: every operand and state
    array touched once (p = bytes per value, int32 indices).
    Per face, both forms: 2 gradient indices and coefficients, Coriolis,
    area; v read, 2 history reads, history write, v write; eta once per cell.
    Tabulated: 10 TRiSK indices and coefficients and nfields forcing
    gradients per face.
    Compact: d_e and the 2 TRiSK partial sums per face, nfields forcing
    fields per cell."""
    common = E * (2 * 4 + 2 * p + p + p + p + 2 * p + p + p) + N * p
    if compact:
        return common + E * (p + 2 * p) + N * nfields * p
    return common + E * (10 * 4 + 10 * p + nfields * p)


def cell_bytes(E, N, p, compact):
    """Algorithmic bytes of one cell-kernel launch.
    Tabulated: 6 face indices and divergence coefficients, eta read and
    write, 2 history reads and 1 write per cell; the new v once per face.
    Compact: face cycle (6 indices), 6 kite fractions, cycle info, area, the
    same eta and history traffic per cell; l_e and the new v read and the 2
    TRiSK partial sums written once per face."""
    if compact:
        return N * (6 * 4 + 6 * p + 4 + p + 2 * p + 3 * p) + E * (p + p + 2 * p)
    return N * (6 * 4 + 6 * p + 2 * p + 3 * p) + E * p


def step_kernel_bytes(E, N, nfields, p):
    """Algorithmic bytes of one launch of the one-kernel step: the compact
    face update (plus l_e per face for the patch store), then per cell the
    patch-local face cycle (6 uint16), 6 kite fractions, cycle info, area,
    eta written and the history (2 reads, 1 write), and per face the 2 TRiSK
    partial sums written (eta^n is counted once, in the face update)."""
    return face_bytes(E, N, nfields, p, True) + E * p + N * (6 * 2 + 6 * p + 4 + p + p + 3 * p) + E * 2 * p


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
    ap.add_argument("--tabulated", action="store_true",
                    help="tabulated TRiSK operator (10 indices and coefficients per face) instead of the "
                         "cell-local form")
    ap.add_argument("--face-cell", action="store_true",
                    help="a face and a cell kernel per step instead of the one-kernel step")
    ap.add_argument("--graphs", action="store_true", help="replay 64-step chunks as HIP graphs")
    ap.add_argument("--patch-cells", type=int, default=0,
                    help="cells per patch of the one-kernel step (64, 128, 256; 0: library default)")
    ap.add_argument("--step-lanes", type=int, default=0,
                    help="lanes per cell of the one-kernel step (2 or 4; 0: library default)")
    ap.add_argument("--step-blocks", type=int, default=0,
                    help="workgroups of the one-kernel step (0: library default, one per patch)")
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
    ctx.use_compact(not args.tabulated)
    one = not args.tabulated and not args.face_cell
    ctx.use_one_kernel(one)
    if one and args.patch_cells:
        ctx.set_option(G.OPT_PATCH_CELLS, args.patch_cells)
    if one and args.step_lanes:
        ctx.set_option(G.OPT_STEP_LANES, args.step_lanes)
    if one and args.step_blocks:
        ctx.set_option(G.OPT_STEP_BLOCKS, args.step_blocks)
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
    compact = not args.tabulated
    fb = face_bytes(mesh.E, mesh.N, nfields, pbytes, compact)
    cb = cell_bytes(mesh.E, mesh.N, pbytes, compact)
    tname = ("double" if prec == G.FP64 else "float") + (",compact" if compact else ",tabulated")
    if one:
        sb = step_kernel_bytes(mesh.E, mesh.N, nfields, pbytes)
        fb, cb = sb, 0
        kname, kbytes, kms = "k_step", sb, tm["face_ms"]
    elif tm["cell_ms"] > tm["face_ms"]:  # dominant kernel: the longer of the two per step
        kname, kbytes, kms = "k_cell", cb, tm["cell_ms"]
    else:
        kname, kbytes, kms = "k_face", fb, tm["face_ms"]
    achieved = kbytes / (kms * 1e-3) / 1e9 if kms > 0 else None
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
            "trisk": "cell-local (kite fractions)" if compact else "tabulated",
            "launch": ("one kernel per step (patches of <= %d cells, %d lanes per cell, %d workgroups)"
                       % (ctx.option(G.OPT_PATCH_CELLS), ctx.option(G.OPT_STEP_LANES),
                          ctx.option(G.OPT_STEP_BLOCKS))) if one
                      else "face + cell kernel per step",
            "graphs": bool(args.graphs),
            "kernel_ms": {"face": tm["face_ms"], "cell": tm["cell_ms"], "step_events": tm["step_ms"]},
            "alg_bytes": {"face": fb, "cell": cb, "step": fb + cb},
            "hbm_gbs_step_events": (fb + cb) / (tm["step_ms"] * 1e-3) / 1e9 if tm["step_ms"] > 0 else None,
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
            "alg_bytes_per_launch": kbytes,
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
