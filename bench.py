"""Benchmark of the LTE time step on one MI355X.

Default workload ("shell"): the ab3Explicit step of GeodesicODIS on the
163842-cell icosahedral-geodesic grid (reference GRID_LVL 8, "level 7" in
10*4^l+2 numbering).  An Enceladus ocean lies under a 20 km elastic ice
shell (LID_MEMBR: R = 252.1 km body, ocean top 232.1 km, h = 38 km,
linear drag 1e-8 s^-1, eccentricity tide e = 0.0047).  The pressure term
sees the self-gravity / shell coupling through the spherical-harmonic
projection up to degree 32 (1089 harmonics), weighted by the membrane
factors beta_l.  Rest start.  A "step" is one pass of the loop body over
all cells and faces, SH projection included.

Time step: the reference's quantised dt (P / N_s, N_s a multiple of 100) for
a 0.4 s target (N_s = 296000).  The shell stiffens the high degrees
(beta_32 = 1.7e4), so the degree-32 gravity wave has omega = 1.2 s^-1.  The
reference scheme (AB3 momentum, then AB3 continuity on the new velocity) is
stable for omega dt < 0.546.

`value` is the FP64 plain method (the drop-in).  Its SH basis is evaluated
on the fly by the Legendre recurrence in FP64 (`--sh-basis stored` reads a
stored FP64 basis instead).  The same line carries:
  - the FP32 form on the same workload (FP32 stencils and recurrence, or a
    stored split-BF16 basis with --sh-basis stored), with its error against
    FP64 after the same steps;
  - FP32 stencils with the SH passes as split-BF16 products on the matrix
    cores (one instance), with its error against FP64 and the matrix-core
    utilisation;
  - the free-surface ocean without the shell (workload "free": dt = P/18300,
    where the stencils alone are timed);
  - a batched FP32 sweep: 256 instances on the 40962-cell grid, seeded
    thickness and drag draws, against one plain FP64 instance;
  - an uneven FP32 batch: 512 instances over the 2562- to 163842-cell
    grids, FULL tide, quadratic drag, half under a shell with SH coupling
    (the batched split-BF16 products for shell groups of 4 or more).
`--workload converge` instead times the plain-method case (10242 cells) to
converged orbit-averaged dissipation, plain FP64 and FP32; `--workload
sweep` and `--workload uneven` run one batched extra alone (for profiling).

Prints one JSON line: cell-steps/s of the whole job; the roofline of the
dominant kernel (HIP events stamped on the kernel dispatches of every 64th
step of a long run, inside the timed region); and a CPU baseline (the FP64 C oracle with
OpenMP, with the same projection as dense products, on a bounded sample of
the same workload).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense bf16 matrix cores
SH_DEGREE = 32
SHELL_INPUT = """radius; 252.1e3; top radius;
k2; 0.0; k2;
h2; 0.0; h2;
love reduction factor; 1.0; set by the surface condition;
ANGULAR VELOCITY; 5.307e-5; mean motion;
surface gravity; 0.113; g;
eccentricity; 0.0047; e;
obliquity; 0.0; degrees;
ocean thickness; 38e3; h_o;
shell thickness; 20e3; h_s;
sh degree; %d; l_max;
friction coefficient; 1e-8; alpha;
friction type; LINEAR; friction;
potential; ECC; potential;
surface type; LID_MEMBR; shell;
advection; false; advection;
solver type; AB3; solver;
geodesic grid level; 8; level;
output time; 10; out;
initial conditions; NONE; ic;
simulation end time; 20; orbits;
time step; 0.4; dt target;
""" % SH_DEGREE


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


def sh_bytes(N, K, p, pb, chunks=8):
    """Algorithmic bytes of the SH kernels per step: analysis (basis once,
    eta once, chunk partials written), coefficients (the K x K inverse Gram
    matrix, partials, factors), synthesis (basis once, eta read, eta_eff
    written).  pb = bytes per basis entry (8 FP64, 4 split-BF16 word; 0 for
    the basis evaluated on the fly, which reads 4 doubles per cell)."""
    cell = 4 * 8 if pb == 0 else 0
    return {"sh_analysis": K * N * pb + N * (p + cell) + chunks * K * 8,
            "sh_coef": K * K * 8 + chunks * K * 8 + 2 * K * 8,
            "sh_synth": K * N * pb + N * (2 * p + cell) + K * 8}


def sh_alg_flops(N, K):
    """Algorithmic flops of one SH pass (analysis b = Y eta or synthesis
    Y^T c): one multiply-add per harmonic and cell, 2 K N."""
    return 2 * K * N


def sh_flops(N, L):
    """Flops one on-the-fly pass issues (analysis or synthesis) with the
    rescaled Legendre recurrence q_l = mu q_{l-1} + beta_l q_{l-2}: per cell
    and (l, m) the recurrence step (one multiplication, one multiply-add:
    3 flops) and the multiply-adds of the cos part and, for m > 0, the sin
    part (2 flops each); 1.77 times sh_alg_flops at degree 32."""
    nlm = (L + 1) * (L + 2) // 2
    return N * (3 * nlm + 2 * nlm + 2 * (nlm - (L + 1)))


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


def timed_run(ctx, steps, warmup, stride):
    """Warm-up, then `steps` timed steps between synchronisations; returns
    (seconds, kernel times)."""
    # the warm-up steps are sampled too (their times dropped), so that the
    # events' first use is not inside the timed steps
    ctx.enable_timing(stride)
    ctx.step(warmup, sync=True)
    ctx.timing()
    ctx.sync()
    t1 = time.perf_counter()
    ctx.step(steps, sync=True)
    ctx.sync()
    elapsed = time.perf_counter() - t1
    tm = ctx.timing()
    ctx.enable_timing(0)
    return elapsed, tm


def shell_setup(G):
    """Params of the shell workload through the reference's input.in rules
    (membraneNuBeta moves g and R to the ocean top) and its beta_l."""
    d = tempfile.mkdtemp(prefix="odis_bench_")
    with open(os.path.join(d, "input.in"), "w") as f:
        f.write(SHELL_INPUT)
    p, info = G.read_config(d)
    beta = G.surface_beta(G.LID_MEMBR, SH_DEGREE, 252.1e3, 0.113, 20e3, 38e3)
    return p, info["steps_per_orbit"], beta


def apply_launch_options(G, ctx, args):
    ctx.use_compact(not args.tabulated)
    ctx.use_one_kernel(not args.tabulated and not args.face_cell)
    ctx.use_graphs(args.graphs)
    if not args.tabulated and not args.face_cell:
        if args.patch_cells:
            ctx.set_option(G.OPT_PATCH_CELLS, args.patch_cells)
        if args.step_lanes:
            ctx.set_option(G.OPT_STEP_LANES, args.step_lanes)
        if args.step_blocks:
            ctx.set_option(G.OPT_STEP_BLOCKS, args.step_blocks)


def stencil_roofline(G, ctx, mesh, tm, p_bytes, one):
    """Kernel name, bytes and time of the dominant stencil kernel, and the
    algorithmic bytes of the stencil kernels per step."""
    nfields = 2
    compact = ctx.option(G.OPT_COMPACT) == 1
    if one and compact:
        sb = step_kernel_bytes(mesh.E, mesh.N, nfields, p_bytes)
        return "k_step", sb, tm["face_ms"], {"step": sb}
    fb = face_bytes(mesh.E, mesh.N, nfields, p_bytes, compact)
    cb = cell_bytes(mesh.E, mesh.N, p_bytes, compact)
    if tm["cell_ms"] > tm["face_ms"]:
        return "k_cell", cb, tm["cell_ms"], {"face": fb, "cell": cb}
    return "k_face", fb, tm["face_ms"], {"face": fb, "cell": cb}


_GRIDS = {}


def generated_grid(G, level):
    """The generated grid of a level (icosahedral bisection + 64 Lloyd
    sweeps), built once per run."""
    if level not in _GRIDS:
        _GRIDS[level] = G.Grid.generate(level, 64)
    return _GRIDS[level]


SWEEP_SEED = 20261203
# largest gravity-wave frequency at 40962 cells for a 38 km Enceladus ocean
# (twice the 0.0193 s^-1 measured at 10242 cells); it scales as sqrt(h)
OMEGA_MAX_40962_38KM = 0.0386


def sweep_params(G, n, seed=SWEEP_SEED):
    """The batched-sweep instances: thickness 10^U(3, 5) m and linear drag
    10^U(-11, -7) s^-1 drawn in that order per instance from
    numpy.random.default_rng(seed); eccentricity tide, free surface; each
    instance's dt is the quantised step below 0.35 / omega_max(h)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ps = []
    for _ in range(n):
        h = 10.0 ** rng.uniform(3.0, 5.0)
        alpha = 10.0 ** rng.uniform(-11.0, -7.0)
        w = OMEGA_MAX_40962_38KM * (h / 38e3) ** 0.5
        p, _ = G.make_params(h=h, alpha=alpha, time_step=0.35 / w)
        ps.append(p)
    return ps


def sweep_run(G, args):
    """Batched sweep: 256 instances on the 40962-cell grid in one FP32
    context (every kernel takes the instances along its grid's y axis),
    against one plain FP64 instance of the median thickness on the same
    grid."""
    grid = generated_grid(G, 7)
    ps = sweep_params(G, args.sweep_instances)
    mesh = G.Mesh(grid, ps[0].radius)
    n_steps, n_warm = max(20, args.steps // 10), 10
    tr = None
    ctx = G.Context(mesh, ps, precision=G.FP32)
    if args.il_order >= 0:
        ctx.set_option(G.OPT_IL_ORDER, args.il_order)
    elif not args.no_tune:
        from geodesicodis_amd import tune as T
        t0 = time.perf_counter()
        tr = T.tune(mesh, ps, precision=G.FP32, log=lambda m: print("bench: " + m, file=sys.stderr, flush=True))
        tr["tune_s"] = time.perf_counter() - t0
        T.apply(ctx, T.option_ids(tr["options"]))
    el, _ = timed_run(ctx, n_steps, n_warm, 0)
    ok = bool(all(abs(x) < float("inf") for x in ctx.dissipation()))
    ctx.close()
    hs = sorted(p.h for p in ps)
    p1 = [p for p in ps if p.h == hs[len(hs) // 2]][0]
    c1 = G.Context(mesh, p1, precision=G.FP64)
    el1, _ = timed_run(c1, args.steps, args.warmup, 0)
    c1.close()
    batched = len(ps) * mesh.N * n_steps / el
    single = mesh.N * args.steps / el1
    return {"instances": len(ps), "cells": mesh.N, "seed": SWEEP_SEED, "precision": "fp32 state, fp64 dissipation",
            "value": batched, "ms_per_step": 1e3 * el / n_steps, "steps": n_steps,
            "plain_fp64_single_instance": single, "batched_over_single": batched / single,
            "finite": ok,
            "autotune": ({x: tr[x] for x in ("options", "default", "ms_per_step", "default_ms_per_step", "speedup",
                                             "candidates", "cached", "tune_s", "timings_ms")} if tr else "off")}


def mfma_flops(n_cells, l_max, n_inst):
    """Dense bf16 flops one batched split-BF16 pass issues on the matrix
    cores (three products per term, harmonics padded to 32, cells to 32 or
    16, instances to 16), and the useful flops 2 K N n of the same product."""
    K = (l_max + 1) ** 2
    Kp = (K + 31) // 32 * 32
    it = (n_inst + 15) // 16 * 16
    ana = 3 * 2 * Kp * ((n_cells + 31) // 32 * 32) * it
    syn = 3 * 2 * Kp * ((n_cells + 15) // 16 * 16) * it
    return {"sh_analysis": ana, "sh_synth": syn, "useful": 2 * K * n_cells * n_inst}


def mfma_utilisation(tm, n_cells, l_max, n_inst):
    """Matrix-core utilisation of the two product passes against the dense
    bf16 peak (2.5 PFLOP/s)."""
    fl = mfma_flops(n_cells, l_max, n_inst)
    out = {}
    for k in ("sh_analysis", "sh_synth"):
        ms = tm[k + "_ms"]
        if ms > 0:
            out[k] = {"issued_tflops": fl[k] / (ms * 1e-3) / 1e12,
                      "useful_tflops": fl["useful"] / (ms * 1e-3) / 1e12,
                      "frac_of_dense_bf16_peak": fl[k] / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS}
    return out


UNEVEN_SEED = 20261204
UNEVEN_SH_DEGREE = 16


def uneven_groups(G, n, seed=UNEVEN_SEED):
    """The uneven-batch instances grouped by (level, shell): each instance
    draws (level, obliquity, quadratic-drag coefficient, surface) in that
    order from numpy.random.default_rng(seed).  Returns ({(level, lid):
    [Params]}, beta) with beta the degree-16 LID_MEMBR factors of the 20 km
    shell; dt is each group's quantised step below 0.35 / omega_max."""
    import numpy as np
    rng = np.random.default_rng(seed)
    draws = {}
    for _ in range(n):
        level = int(rng.choice([4, 5, 6, 7]))
        theta = rng.uniform(0.0, 1.0)
        cd = rng.uniform(1e-3, 5e-3)
        lid = bool(rng.uniform() < 0.5)
        draws.setdefault((level, lid), []).append((theta, cd))
    L = UNEVEN_SH_DEGREE
    beta = np.array(G.surface_beta(G.LID_MEMBR, L, 252.1e3, 0.113, 20e3, 38e3))
    ls = np.arange(L + 1)
    w_shell = float(np.max(np.sqrt(0.1166 * 38e3 * np.maximum(beta, 1.0) * ls * (ls + 1)) / 232.1e3))
    groups = {}
    for (level, lid), inst in draws.items():
        n_cells = 10 * 4 ** level + 2
        w = 0.0193 * (n_cells / 10242.0) ** 0.5
        if lid:
            w = max(w, w_shell)
        ps = []
        for theta, cd in inst:
            kw = dict(tide=G.FULL, obliquity=np.radians(theta), friction=G.QUADRATIC, alpha=cd, quad_drag=1,
                      time_step=0.35 / w)
            if lid:
                kw.update(radius=232.1e3, g=0.1166, surface=G.LID_MEMBR, shell_thickness=20e3, love_reduct=0.99)
            ps.append(G.make_params(**kw)[0])
        groups[(level, lid)] = ps
    return groups, beta


def uneven_run(G, args):
    """Uneven batch: args.uneven_instances instances, each drawing (level,
    obliquity, quadratic-drag coefficient, surface) in that order from
    numpy.random.default_rng(20261204): grid level from {4, 5, 6, 7} of the
    reference's numbering (2562 .. 163842 cells), FULL tide (e = 0.0047,
    obliquity U(0, 1) degrees), quadratic drag c_D = U(1e-3, 5e-3), and a
    free surface or a 20 km LID_MEMBR shell (SH coupling to degree 16) with
    probability 1/2.  Instances on one grid and surface share a context
    (the kernels take them along their grid's y axis); a shell group of at
    least 4 instances is coupled through the batched split-BF16 matrix
    products, a smaller one through the recurrence (args.uneven_sh overrides
    this); each instance's dt is
    the quantised step below 0.35 / omega_max, omega_max the larger of the
    grid-scale gravity-wave frequency (0.0193 s^-1 at 10242 cells, scaling
    as sqrt(N)) and, under the shell, the degree-16 shell-mode frequency
    sqrt(g h beta_l l (l + 1)) / R.  FP32 state.  Each context is timed
    alone (the per-group rows), then the 8 contexts' steps are enqueued
    together, one stream each, which gives `value`."""
    import numpy as np
    groups, beta = uneven_groups(G, args.uneven_instances)
    L = UNEVEN_SH_DEGREE
    total_cells, elapsed, n_steps, rows, ctxs = 0, 0.0, max(10, args.steps // 40), [], []
    for (level, lid), ps in sorted(groups.items()):
        grid = generated_grid(G, level + 1)
        n_cells = grid.n
        mesh = G.Mesh(grid, ps[0].radius)
        ctx = G.Context(mesh, ps, precision=G.FP32)
        sh = None
        if lid:
            sh = args.uneven_sh if args.uneven_sh != "auto" else ("mfma" if len(ps) >= 4 else "recur")
            ctx.set_sh(L, np.tile(beta, (len(ps), 1)), mode=G.SH_MFMA if sh == "mfma" else G.SH_RECUR)
        el, tm = timed_run(ctx, n_steps, 3, 4 if lid else 0)
        ok = bool(all(abs(x) < float("inf") for x in ctx.dissipation()))
        ctxs.append(ctx)
        elapsed += el
        total_cells += n_cells * len(ps)
        row = {"cells": n_cells, "surface": "LID_MEMBR" if lid else "FREE", "instances": len(ps),
               "ms_per_step": 1e3 * el / n_steps, "finite": ok}
        if sh:
            row["sh"] = {"mfma": "split-BF16 matrix products", "recur": "on-the-fly recurrence"}[sh]
            row["kernel_ms"] = {k[:-3]: tm[k] for k in ("step_ms", "sh_analysis_ms", "sh_coef_ms", "sh_synth_ms")}
            if sh == "mfma":
                row["mfma"] = mfma_utilisation(tm, n_cells, L, len(ps))
        rows.append(row)
    # all groups at once: every context's steps enqueued on its own stream
    for c in ctxs:
        c.sync()
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(n_steps, sync=False)
    for c in ctxs:
        c.sync()
    conc = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    return {"instances": args.uneven_instances, "seed": UNEVEN_SEED, "precision": "fp32 state, fp64 dissipation",
            "value": total_cells * n_steps / conc, "ms_per_step_all_groups": 1e3 * conc / n_steps,
            "launch": "the 8 contexts' steps enqueued together, one stream each",
            "value_groups_in_turn": total_cells * n_steps / elapsed,
            "ms_per_step_groups_in_turn": 1e3 * elapsed / n_steps, "steps": n_steps, "groups": rows}


def converge_run(G, args):
    """Time to converged orbit-averaged dissipation for one instance of the
    plain-method case (Enceladus free surface, 38 km, ECC, linear drag
    args.alpha) on the 10242-cell grid, plain FP64 and the FP32 form.
    D_k is the mean per-step dissipation over orbit k; the run stops when
    |D_k - D_{k-1}| / D_k < 1e-4 for 3 consecutive orbits (D* = D_k) or at
    args.max_orbits (capped).  Time-to-converged: wall time until
    |D_k - D*| / D* <= 1 % and it stays so for 3 orbits, D* of plain FP64."""
    import numpy as np
    grid = generated_grid(G, 6)
    # 4200 steps per orbit: the quantised 30 s target (4000) is unstable on this grid
    p, n_s = G.make_params(alpha=args.alpha, time_step=28.2)
    mesh = G.Mesh(grid, p.radius)
    res = {}
    for name, prec in (("fp64", G.FP64), ("fp32", G.FP32)):
        ctx = G.Context(mesh, p, precision=prec)
        D, t_end, stable = [], [], 0
        ctx.step(1, sync=True)  # state 0 -> 1 outside the clock
        t0 = time.perf_counter()
        first = ctx.get_state()["iter"]
        k = 0
        while k < args.max_orbits:
            ctx.step(n_s, sync=True)
            t_end.append(time.perf_counter() - t0)
            ser = ctx.dissipation_series(0, first + k * n_s, n_s)
            D.append(float(np.mean(ser)))
            if k > 0 and abs(D[-1] - D[-2]) / abs(D[-1]) < 1e-4:
                stable += 1
            else:
                stable = 0
            k += 1
            if k % 100 == 0:
                print("bench: converge %s orbit %d D %.6e (%.1f s)" % (name, k, D[-1], t_end[-1]),
                      file=sys.stderr, flush=True)
            if stable >= 3:
                break
        ctx.close()
        res[name] = {"orbits": k, "converged": stable >= 3, "D_star": D[-1], "D": D, "t": t_end,
                     "wall_s": t_end[-1], "steps_per_s": k * n_s / t_end[-1]}
    dstar = res["fp64"]["D_star"]

    def t_conv(r):
        D, t = r["D"], r["t"]
        for i in range(len(D) - 2):
            if all(abs(D[j] - dstar) <= 0.01 * abs(dstar) for j in range(i, len(D))):
                return t[i], i + 1
        return None, None

    for r in res.values():
        r["time_to_1pct_s"], r["orbits_to_1pct"] = t_conv(r)
    out = {"metric": "time-to-converged dissipation (+-1%) vs plain FP64", "value": res["fp32"]["time_to_1pct_s"],
           "unit": "s", "higher_is_better": False, "n_gpus": 1,
           "config": {"workload": "plain-method Enceladus free-surface ocean, 38 km, ECC, linear drag %.0e s^-1, "
                                  "one instance, rest start" % args.alpha,
                      "cells": mesh.N, "steps_per_orbit": n_s, "dt_s": p.dt, "max_orbits": args.max_orbits}}
    for name, r in res.items():
        out["config"][name] = {k: r[k] for k in ("orbits", "converged", "D_star", "wall_s", "steps_per_s",
                                                  "time_to_1pct_s", "orbits_to_1pct")}
    out["config"]["fp32_D_star_rel_err"] = abs(res["fp32"]["D_star"] - dstar) / abs(dstar)
    out["config"]["fp64_time_over_fp32_time"] = (res["fp64"]["time_to_1pct_s"] / res["fp32"]["time_to_1pct_s"]
                                                 if res["fp64"]["time_to_1pct_s"] and res["fp32"]["time_to_1pct_s"]
                                                 else None)
    return out


def shell_orbit_run(G, args):
    """The shell workload (163842 cells, 20 km LID_MEMBR shell, SH coupling
    to degree 32, ECC, rest start) stepped args.orbits whole orbits (296000
    steps each) in plain FP64, in the FP32 form (FP32 stencils and
    recurrence) and in FP32 + split-BF16 MFMA products; the two fast forms'
    relative L2 errors in eta and v and the relative error of the dissipation
    against FP64, after every orbit."""
    import numpy as np
    grid = generated_grid(G, args.level)
    p, steps_per_orbit, beta = shell_setup(G)
    mesh = G.Mesh(grid, p.radius)
    forms = (("fp64", G.FP64, G.SH_RECUR), ("fp32", G.FP32, G.SH_RECUR), ("fp32_mfma", G.FP32, G.SH_MFMA))
    wanted = {"fp64"} | set(args.orbit_forms.split(","))
    forms = tuple(f for f in forms if f[0] in wanted)
    ctxs = {}
    for name, prec, mode in forms:
        c = G.Context(mesh, p, precision=prec)
        c.set_sh(SH_DEGREE, beta, mode=mode)
        ctxs[name] = c
    n_orbits = max(1, int(round(args.orbits)))
    rows, walls = [], {k: 0.0 for k in ctxs}
    chunk = steps_per_orbit // 8
    keys = ("v", "dvdt", "eta", "detadt")
    k0 = 0
    if args.orbit_resume:  # continue runs saved after whole orbits
        for name, c in ctxs.items():
            st = {x: np.load(os.path.join(args.orbit_resume, "%s_%s.npy" % (name, x))) for x in keys}
            it = int(np.load(os.path.join(args.orbit_resume, "%s_iter.npy" % name)))
            c.set_state(0, st["v"], st["dvdt"], st["eta"], st["detadt"], it)
            k0 = it // steps_per_orbit
    for k in range(k0, k0 + n_orbits):
        for name, c in ctxs.items():
            t0 = time.perf_counter()
            for q in range(8):
                c.step(chunk)
                print("bench: shell-orbit %s orbit %d, %d / 8" % (name, k + 1, q + 1), file=sys.stderr, flush=True)
            walls[name] += time.perf_counter() - t0
        ref = ctxs["fp64"].get_state()
        d_ref = float(ctxs["fp64"].dissipation()[0])
        row = {"orbit": k + 1, "steps": (k + 1) * steps_per_orbit, "dissipation_fp64_W_m2": d_ref}
        for name in (f for f in ("fp32", "fp32_mfma") if f in ctxs):
            st = ctxs[name].get_state()
            d = float(ctxs[name].dissipation()[0])
            row[name] = {"eta": rel_l2(st["eta"], ref["eta"]), "v": rel_l2(st["v"], ref["v"]),
                         "dissipation": abs(d - d_ref) / abs(d_ref)}
        rows.append(row)
    if args.orbit_save:
        os.makedirs(args.orbit_save, exist_ok=True)
        for name, c in ctxs.items():
            st = c.get_state()
            for x in keys:
                np.save(os.path.join(args.orbit_save, "%s_%s.npy" % (name, x)), st[x])
            np.save(os.path.join(args.orbit_save, "%s_iter.npy" % name), np.array(st["iter"]))
    for c in ctxs.values():
        c.close()
    return {"workload": "shell, 163842 cells, degree 32, orbits %d-%d of %d steps, rest start" % (
                k0 + 1, k0 + n_orbits, steps_per_orbit),
            "errors_vs_fp64": rows, "wall_s": walls}


WARMSTART_MODEL = os.path.join(ROOT, "geodesicodis_amd", "data", "warmstart_c5.npz")


def converged_leg(G, args):
    """Second half of the metric: wall time to the converged dissipation band
    (orbit means within 1 % of D* for 3 orbits) of held-out sweep instances
    (seed 20261206: thickness 10^U(3,5) m, drag 10^U(-11,-7) s^-1, ECC or
    FULL tide) on the 40962-cell grid.  Batched FP32 from the learned warm
    start (the committed network, trained on the GPU from 2000 converged
    runs) against batched FP32 from rest, with retirement, and against the
    plain FP64 method run one instance after another (orbits from rest x
    steps per orbit x the measured FP64 step time; a lower bound where the
    rest start reached its cap).  D* is each instance's periodic state,
    solved on the 10242-cell grid."""
    from geodesicodis_amd import warmstart as W
    model = W.WarmStartModel.load(G, WARMSTART_MODEL)
    n = args.converged_instances if args.workload != "warmstart" else max(args.converged_instances, 256)
    rows = W.draw_instances(G, n, W.HELDOUT_SEED)
    t0 = time.perf_counter()
    ev = W.evaluate(G, model, rows, level=7, solve_level=6, cap_warm=30, cap_rest=args.converged_cap_rest,
                    workers=min(16, os.cpu_count() or 1),
                    log=lambda m: print("bench: " + m, file=sys.stderr, flush=True))
    ev["wall_s"] = time.perf_counter() - t0
    ev["seed"] = W.HELDOUT_SEED
    ev["model"] = os.path.relpath(WARMSTART_MODEL, ROOT)
    return ev


def rel_l2(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", choices=["shell", "free", "converge", "sweep", "uneven", "warmstart",
                                           "shell-orbit"],
                    default="shell",
                    help="shell: ocean under a 20 km ice shell with SH coupling to degree 32 (default); "
                         "free: free-surface ocean, stencils only; converge: time to converged dissipation "
                         "of the plain-method case, FP64 and FP32; sweep / uneven: the batched extras alone; "
                         "warmstart: time to the converged dissipation band of held-out sweep instances "
                         "from the learned warm start and from rest; shell-orbit: the shell workload's FP32 "
                         "and FP32 + split-BF16 MFMA forms against FP64 after whole orbits")
    ap.add_argument("--orbits", type=float, default=1.0,
                    help="orbits of the shell-orbit workload (fast forms against FP64 after whole orbits)")
    ap.add_argument("--orbit-save", default="",
                    help="shell-orbit: save each form's state after the last orbit into this directory")
    ap.add_argument("--orbit-resume", default="",
                    help="shell-orbit: continue from the states saved in this directory")
    ap.add_argument("--orbit-forms", default="fp32,fp32_mfma",
                    help="fast forms the shell-orbit workload steps beside FP64 (comma-separated: fp32, fp32_mfma)")
    ap.add_argument("--no-converged", action="store_true",
                    help="skip the time-to-converged leg (learned warm start vs rest, held-out instances)")
    ap.add_argument("--converged-instances", type=int, default=64,
                    help="held-out sweep instances of the time-to-converged leg")
    ap.add_argument("--converged-cap-rest", type=int, default=10,
                    help="orbit cap of the rest start in the time-to-converged leg")
    ap.add_argument("--alpha", type=float, default=1e-8, help="linear drag of the converge workload (s^-1)")
    ap.add_argument("--max-orbits", type=int, default=3000, help="orbit cap of the converge workload")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-sweep extra")
    ap.add_argument("--sweep-instances", type=int, default=256, help="instances of the batched-sweep extra")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batch extra")
    ap.add_argument("--uneven-instances", type=int, default=512, help="instances of the uneven-batch extra")
    ap.add_argument("--uneven-sh", choices=("auto", "recur", "mfma"), default="auto",
                    help="SH coupling of the uneven batch's shell groups: split-BF16 matrix products, the "
                         "on-the-fly recurrence, or the products for groups of at least 4 instances (auto)")
    ap.add_argument("--level", type=int, default=8, help="reference grid level (8 = 163842 cells)")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="precision of the headline value (fp32: FP32 stencils, split-BF16 SH basis)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU work of the baseline sample")
    ap.add_argument("--cpu-sizes", action="store_true",
                    help="also time the FP64 C oracle on the plain-method free-surface step at 10242, 40962 and "
                         "163842 cells, with all threads and with one (cpu_baseline_sizes)")
    ap.add_argument("--cpu-size-seconds", type=float, default=2.0,
                    help="approximate CPU work of each --cpu-sizes sample")
    ap.add_argument("--sh-basis", choices=["recur", "stored"], default="recur",
                    help="SH basis evaluated on the fly by the Legendre recurrence (default) or stored "
                         "(FP64, or split-BF16 for fp32)")
    ap.add_argument("--sh-cells", type=int, default=0,
                    help="cells per lane of the on-the-fly SH analysis (4, 8, 16; 0: library default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the headline measurement (no fast form / free-surface lines)")
    ap.add_argument("--tabulated", action="store_true",
                    help="tabulated TRiSK operator (10 indices and coefficients per face) instead of the "
                         "cell-local form")
    ap.add_argument("--face-cell", action="store_true",
                    help="a face and a cell kernel per step instead of the one-kernel step")
    ap.add_argument("--graphs", action="store_true", help="replay 64-step chunks as HIP graphs")
    ap.add_argument("--il-order", type=int, default=-1,
                    help="workgroup order of the sweep's interleaved step (ODIS_OPT_IL_ORDER; implies --no-tune "
                         "for the sweep)")
    ap.add_argument("--no-tune", action="store_true",
                    help="launch options from the flags / library defaults instead of the autotuner's choice "
                         "(the tuner runs when no launch flag is given)")
    ap.add_argument("--patch-cells", type=int, default=0,
                    help="cells per patch of the one-kernel step (64, 128, 256; 0: library default)")
    ap.add_argument("--step-lanes", type=int, default=0,
                    help="lanes per cell of the one-kernel step (2 or 4; 0: library default)")
    ap.add_argument("--step-blocks", type=int, default=0,
                    help="workgroups of the one-kernel step (0: library default, one per patch)")
    ap.add_argument("--timing-stride", type=int, default=-1,
                    help="dispatch-stamped HIP events on the kernels of every T-th step (0: none, "
                         "roofline achieved then null; default: 64 from 1024 steps on, else 16, or half "
                         "the steps for short runs, which samples two)")
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
    if args.timing_stride < 0:
        # every stamped step costs ~25 us (its kernels cannot overlap their
        # neighbours): short runs sample two steps, long ones every 64th (16
        # samples or more)
        args.timing_stride = 64 if args.steps >= 1024 else max(1, min(16, args.steps // 2))
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; multi-GPU runs are not part of it "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import geodesicodis_amd as G

    host = host_cpu_info()  # before any OpenMP runtime binds this thread to one place
    if args.workload == "converge":
        print(json.dumps(converge_run(G, args)), flush=True)
        return
    if args.workload == "warmstart":
        print(json.dumps(converged_leg(G, args)), flush=True)
        return
    if args.workload == "shell-orbit":
        print(json.dumps(shell_orbit_run(G, args)), flush=True)
        return
    if args.workload in ("sweep", "uneven"):
        r = sweep_run(G, args) if args.workload == "sweep" else uneven_run(G, args)
        print(json.dumps({"metric": "cell-steps/s, batched FP32 (%s)" % args.workload, "value": r["value"],
                          "unit": "cell-steps/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f32", "data": "synthetic (generated SCVT grids, rest start)",
                          "config": r}), flush=True)
        return

    t0 = time.time()
    grid = generated_grid(G, args.level)
    shell = args.workload == "shell"
    if shell:
        p, steps_per_orbit, beta = shell_setup(G)
    else:
        p, steps_per_orbit = G.make_params(time_step=6.5)
        beta = None
    mesh = G.Mesh(grid, p.radius)
    print("bench: %s workload, grid %d cells / %d faces ready in %.1fs" % (args.workload, mesh.N, mesh.E,
                                                                          time.time() - t0),
          file=sys.stderr, flush=True)
    K = (SH_DEGREE + 1) ** 2

    from geodesicodis_amd import tune as T
    explicit = (args.tabulated or args.face_cell or args.graphs or args.patch_cells or args.step_lanes or
                args.step_blocks)
    tuned = {}

    def tuned_options(prec):
        """The autotuner's launch options of the headline workload at prec."""
        if args.no_tune or explicit:
            return None
        if prec not in tuned:
            sh = None
            if shell:
                sh = (SH_DEGREE, beta, G.SH_RECUR if args.sh_basis == "recur" else
                      (G.SH_FP64 if prec == G.FP64 else G.SH_SPLIT_BF16))
            t0 = time.perf_counter()
            # a timed region that stamps kernels with events replays no graphs
            tuned[prec] = T.tune(mesh, [p], precision=prec, sh=sh, graphs=args.timing_stride == 0,
                                 log=lambda m: print("bench: " + m, file=sys.stderr, flush=True))
            tuned[prec]["tune_s"] = time.perf_counter() - t0
        return tuned[prec]

    def make_ctx(prec):
        ctx = G.Context(mesh, p, precision=prec)
        apply_launch_options(G, ctx, args)
        tr = tuned_options(prec)
        if tr:
            T.apply(ctx, T.option_ids(tr["options"]))
        if shell:
            if args.sh_basis == "recur":
                if args.sh_cells:
                    ctx.set_option(G.OPT_SH_CELLS, args.sh_cells)
                ctx.set_sh(SH_DEGREE, beta, mode=G.SH_RECUR)
            else:
                ctx.set_sh(SH_DEGREE, beta, mode=G.SH_FP64 if prec == G.FP64 else G.SH_SPLIT_BF16)
        return ctx

    prec = G.FP64 if args.precision == "fp64" else G.FP32
    p_bytes = 8 if prec == G.FP64 else 4
    ctx = make_ctx(prec)
    elapsed, tm = timed_run(ctx, args.steps, args.warmup, args.timing_stride)
    ms_per_step = 1e3 * elapsed / args.steps
    value = mesh.N * args.steps / elapsed
    diss = ctx.dissipation()
    state = ctx.get_state()
    print("bench: headline %.3e cell-steps/s (%.1f us/step)" % (value, 1e3 * ms_per_step),
          file=sys.stderr, flush=True)
    one = ctx.option(G.OPT_ONE_KERNEL) == 1
    tname = ("double" if prec == G.FP64 else "float") + (",compact" if ctx.option(G.OPT_COMPACT) else ",tabulated")
    kname, kbytes, kms, stencil_alg = stencil_roofline(G, ctx, mesh, tm, p_bytes, one)
    kernels = {"stencil": {"kernel": "%s<%s>" % (kname, tname), "ms": kms, "alg_bytes": kbytes}}
    alg = dict(stencil_alg)
    recur = shell and args.sh_basis == "recur"
    if shell:
        pb = 0 if recur else (8 if prec == G.FP64 else 4)
        compiled = SH_DEGREE in (8, 16, 32)
        # partials: one per workgroup of 4 waves x 64 lanes x 16 cells (pair
        # analysis, compiled degrees) or of 8 chunks of 64 x 8 cells
        cpl = 4 * 64 * (args.sh_cells or 16) if compiled else 8 * 64 * (args.sh_cells or 8)
        chunks = (mesh.N + cpl - 1) // cpl if recur else 8
        shb = sh_bytes(mesh.N, K, p_bytes, pb, chunks)
        alg.update(shb)
        sh_t = {"sh_analysis": tm["sh_analysis_ms"], "sh_coef": tm["sh_coef_ms"], "sh_synth": tm["sh_synth_ms"]}
        tn = "double" if prec == G.FP64 else "float"
        for k, v in sh_t.items():
            if recur and k != "sh_coef":
                kname_sh = ({"sh_analysis": "k_sh_analysis_pair", "sh_synth": "k_sh_synth_chain"}[k] if compiled
                            else "k_%s_rec" % k)
                kernels[k] = {"kernel": "%s<%s>" % (kname_sh, tn), "ms": v, "alg_bytes": shb[k],
                              "alg_flops": sh_alg_flops(mesh.N, K),
                              "issued_flops": sh_flops(mesh.N, SH_DEGREE)}
            else:
                kernels[k] = {"kernel": "k_%s<%s>" % (k, tn if recur else ("double" if prec == G.FP64 else
                                                                           "split_bf16")),
                              "ms": v, "alg_bytes": shb[k]}
    # dominant kernel: the longest per step; the on-the-fly SH passes are
    # bound by FP64 (FP32) vector arithmetic, the others by HBM bytes
    dom = max(kernels.values(), key=lambda r: r["ms"])
    issued = None
    if "alg_flops" in dom:
        # the arithmetic roof: these passes issue FP64 (FP32) vector FMAs and
        # no MFMA; the peak is MI355X's published FP64 (FP32) rate, which the
        # vector units share with the FP64 matrix cores.  achieved counts the
        # algorithmic 2 K N flops per pass; issued the recurrence's own.
        peak = 78.6 if prec == G.FP64 else 157.3
        achieved = dom["alg_flops"] / (dom["ms"] * 1e-3) / 1e12 if dom["ms"] > 0 else None
        bound, unit = ("fp64-valu" if prec == G.FP64 else "fp32-valu"), "TFLOP/s"
        compute = ("%s vector FMA (VALU), no MFMA; peak = the published %s rate"
                   % (("fp64", "FP64") if prec == G.FP64 else ("fp32", "FP32")))
        if dom["ms"] > 0:
            it = dom["issued_flops"] / (dom["ms"] * 1e-3) / 1e12
            issued = {"flops_per_launch": dom["issued_flops"], "tflops": it, "frac": it / peak}
    else:
        peak = HBM_PEAK_GBS
        achieved = dom["alg_bytes"] / (dom["ms"] * 1e-3) / 1e9 if dom["ms"] > 0 else None
        bound, unit, compute = "hbm", "GB/s", None
    dk, dt_ = dom["kernel"].split("<", 1)
    traffic = measured_traffic(dk, dt_[:-1], mesh.N)
    step_bytes = sum(alg.values())
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
            "workload": ("GeodesicODIS LTE AB3 step, plain method, single instance, Enceladus ocean under a 20 km "
                         "ice shell (LID_MEMBR), self-gravity / shell coupling by SH projection to degree %d "
                         "(%d harmonics), eccentricity tide" % (SH_DEGREE, K)) if shell else
                        ("GeodesicODIS LTE AB3 step, plain method, single instance, Enceladus free-surface ocean, "
                         "eccentricity tide, no self-gravity"),
            "cells": mesh.N, "faces": mesh.E, "instances": 1,
            "grid": "icosahedral bisection + 64 Lloyd sweeps, reference level %d" % args.level,
            "dt_s": p.dt, "steps_per_orbit": steps_per_orbit,
            "sh_basis": (("on the fly (Legendre recurrence, %s)" % ("fp64" if prec == G.FP64 else "fp32")) if recur
                         else ("stored fp64" if prec == G.FP64 else "stored split-bf16")) if shell else None,
            "trisk": "cell-local (kite fractions)" if ctx.option(G.OPT_COMPACT) else "tabulated",
            "launch": ("SH analysis, partial sums, coefficients, synthesis + " if shell else "") +
                      (("one stencil kernel per step (patches of <= %d cells, %d lanes per cell, %d workgroups)"
                        % (ctx.option(G.OPT_PATCH_CELLS), ctx.option(G.OPT_STEP_LANES),
                           ctx.option(G.OPT_STEP_BLOCKS))) if one else "face + cell kernel per step"),
            "graphs": bool(ctx.option(G.OPT_GRAPHS)),
            "autotune": ({("fp64" if k == G.FP64 else "fp32"): {
                x: v[x] for x in ("options", "default", "ms_per_step", "default_ms_per_step", "speedup",
                                  "candidates", "cached", "tune_s")} for k, v in tuned.items()}
                         if tuned else "off (launch flags given or --no-tune)"),
            "kernel_ms": {k: r["ms"] for k, r in kernels.items()},
            "step_events_ms": tm["step_ms"],
            "alg_bytes": alg,
            "alg_bytes_per_step": step_bytes,
            "hbm_gbs_per_step": step_bytes / (ms_per_step * 1e-3) / 1e9,
            "dissipation_W_m2": float(diss[0]),
        },
        "roofline": {
            "bound": bound,
            "achieved": achieved,
            "peak": peak,
            "unit": unit,
            "frac": achieved / peak if achieved else None,
            "traffic": traffic,
            "kernel": dom["kernel"],
            "kernel_ms": dom["ms"],
            "compute": compute,
            "alg_flops_per_launch": dom.get("alg_flops"),
            "issued": issued,
            "alg_bytes_per_launch": dom["alg_bytes"],
        },
    }
    ctx.close()

    if not args.no_extra:
        if shell and prec == G.FP64:
            # the fast form on the same workload, and its error against FP64
            fctx = make_ctx(G.FP32)
            fel, ftm = timed_run(fctx, args.steps, args.warmup, args.timing_stride)
            fs = fctx.get_state()
            fd = fctx.dissipation()
            out["config"]["fp32_form"] = {
                "sh_basis": "on the fly (Legendre recurrence, fp32)" if args.sh_basis == "recur" else
                            "stored split-bf16",
                "value": mesh.N * args.steps / fel, "ms_per_step": 1e3 * fel / args.steps,
                "kernel_ms": {"face": ftm["face_ms"], "cell": ftm["cell_ms"], "sh_analysis": ftm["sh_analysis_ms"],
                              "sh_coef": ftm["sh_coef_ms"], "sh_synth": ftm["sh_synth_ms"]},
                "rel_err_vs_fp64_after_%d_steps" % (args.steps + args.warmup): {
                    "eta": rel_l2(fs["eta"], state["eta"]), "v": rel_l2(fs["v"], state["v"]),
                    "dissipation": abs(float(fd[0]) - float(diss[0])) / abs(float(diss[0]))},
            }
            fctx.close()
            # FP32 stencils with the SH passes as split-BF16 products on the
            # matrix cores (one instance: each pass reads the stored basis once)
            mctx = G.Context(mesh, p, precision=G.FP32)
            apply_launch_options(G, mctx, args)
            mctx.set_sh(SH_DEGREE, beta, mode=G.SH_MFMA)
            m_steps = args.steps
            mel, mtm = timed_run(mctx, m_steps, args.warmup, args.timing_stride)
            ms_ = mctx.get_state()
            md = mctx.dissipation()
            out["config"]["fp32_mfma_form"] = {
                "sh": "split-bf16 matrix products (stored basis in fragment order, three products per term, fp32 "
                      "accumulation, fp64 partial sums and coupling)",
                "value": mesh.N * m_steps / mel, "ms_per_step": 1e3 * mel / m_steps, "steps": m_steps,
                "kernel_ms": {"step": mtm["step_ms"], "sh_analysis": mtm["sh_analysis_ms"],
                              "sh_coef": mtm["sh_coef_ms"], "sh_synth": mtm["sh_synth_ms"]},
                "mfma": mfma_utilisation(mtm, mesh.N, SH_DEGREE, 1),
                "basis_bytes_per_pass": ((K + 31) // 32 * 32) * ((mesh.N + 31) // 32 * 32) * 4,
                "rel_err_vs_fp64_after_%d_steps" % (args.steps + args.warmup): {
                    "eta": rel_l2(ms_["eta"], state["eta"]), "v": rel_l2(ms_["v"], state["v"]),
                    "dissipation": abs(float(md[0]) - float(diss[0])) / abs(float(diss[0]))},
            }
            mctx.close()
        if shell:
            # the stencils alone: free-surface ocean at its own stable step
            pf, _ = G.make_params(time_step=6.5)
            fmesh = G.Mesh(grid, pf.radius)
            for name, fprec in (("free_surface_fp64", G.FP64), ("free_surface_fp32", G.FP32)):
                c2 = G.Context(fmesh, pf, precision=fprec)
                apply_launch_options(G, c2, args)
                el2, tm2 = timed_run(c2, args.steps, args.warmup, args.timing_stride)
                one2 = c2.option(G.OPT_ONE_KERNEL) == 1
                tn2 = ("double" if fprec == G.FP64 else "float") + ",compact"
                kn, kb, km, _ = stencil_roofline(G, c2, fmesh, tm2, 8 if fprec == G.FP64 else 4, one2)
                out["config"][name] = {
                    "value": fmesh.N * args.steps / el2, "ms_per_step": 1e3 * el2 / args.steps,
                    "kernel": "%s<%s>" % (kn, tn2), "kernel_ms": km, "alg_bytes": kb,
                    "hbm_frac": (kb / (km * 1e-3) / 1e9 / HBM_PEAK_GBS) if km > 0 else None}
                c2.close()

    if not args.no_extra and not args.no_sweep:
        out["config"]["sweep_fp32"] = sweep_run(G, args)
    if not args.no_extra and not args.no_uneven:
        out["config"]["uneven_fp32"] = uneven_run(G, args)

    if not args.no_extra and not args.no_converged:
        out["config"]["time_to_converged"] = converged_leg(G, args)

    if not args.no_cpu_baseline:
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        os.environ.setdefault("OMP_PROC_BIND", "close")  # read when the OpenMP runtime loads with the oracle
        arrs = grid.arrays()
        om = O.Mesh(arrs, p.radius)
        sim = O.Sim(om, O.make_params(g=p.g, h=p.h, alpha=p.alpha, omega=p.omega, radius=p.radius,
                                      love_reduct=p.love_reduct, ecc=p.ecc, dt=p.dt, tide=p.tide,
                                      surface=p.surface, shell_thickness=p.shell_thickness))
        setup = 0.0
        if shell:
            ts = time.perf_counter()
            Y = O.real_sh_basis(SH_DEGREE, arrs["lat"], arrs["lon"])
            Pc = np.linalg.solve(Y.T @ Y, Y.T)
            dfac = np.zeros(K)
            for l in range(2, SH_DEGREE + 1):
                dfac[l * l:(l + 1) ** 2] = beta[l] - 1.0
            sim.set_projector(Y * dfac[None, :], Pc)
            del Y, Pc
            setup = time.perf_counter() - ts
        sim.step(2)
        batch = 2 if shell else 10

        def cpu_sample(seconds):
            n, t_cpu = 0, 0.0
            while t_cpu < seconds and n < 5000:
                ts = time.perf_counter()
                sim.step(batch)
                t_cpu += time.perf_counter() - ts
                n += batch
            return n, t_cpu

        used = int(os.environ.get("OMP_NUM_THREADS", threads))
        O.set_threads(used)
        n, t_cpu = cpu_sample(args.cpu_seconds)
        O.set_threads(1)
        n1, t1 = cpu_sample(args.cpu_seconds / 3.0)
        O.set_threads(used)
        out["cpu_baseline"] = {
            "value": mesh.N * n / t_cpu,
            "unit": "cell-steps/s",
            "cores": used,
            "kind": "port",
            "sample": ("FP64 C oracle (OpenMP rows%s), same %d-cell grid and parameters, %d steps after 2 warm-up "
                       "steps%s" % (", SH projection as dense LSQ analysis and synthesis products" if shell else "",
                                    mesh.N, n, " (projector set-up %.1f s not timed)" % setup if shell else "")),
            "implementation": "untuned restatement (the CPU oracle), not a tuned CPU port: OpenMP covers its row "
                              "loops only, its reductions are serial",
            "host_cpu": host["model"], "physical_cores": host["physical_cores"],
            "logical_cpus": host["logical_cpus"], "cpus_allowed": host["cpus_allowed"],
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "unset"),
            "one_thread": {"value": mesh.N * n1 / t1, "steps": n1},
        }
    if args.cpu_sizes:
        out["cpu_baseline_sizes"] = cpu_sizes_run(G, args)
    print(json.dumps(out), flush=True)


def host_cpu_info():
    """CPU model, physical cores (distinct (package, core) pairs), logical
    CPUs and the CPUs this process may run on."""
    model, pairs, phys, core = "", set(), None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and not model:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    pairs.add((phys, core))
                    phys = core = None
        if phys is not None:
            pairs.add((phys, core))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = None
    return {"model": model, "physical_cores": len(pairs) or None, "logical_cpus": os.cpu_count(),
            "cpus_allowed": allowed}


def cpu_sizes_run(G, args):
    """The FP64 C oracle on the plain-method free-surface step (ECC, 38 km,
    alpha 1e-8, dt below 0.35 / omega_max) at 10242, 40962 and 163842 cells,
    with OMP_NUM_THREADS threads and with one: 2 warm-up steps, then steps
    until args.cpu_size_seconds of stepping."""
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", min(16, os.cpu_count() or 1)))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    rows = []
    for level in (6, 7, 8):
        grid = generated_grid(G, level)
        p, _ = G.make_params(time_step=0.35 / (0.0193 * (grid.n / 10242.0) ** 0.5))
        sim = O.Sim(O.Mesh(grid.arrays(), p.radius),
                    O.make_params(g=p.g, h=p.h, alpha=p.alpha, omega=p.omega, radius=p.radius,
                                  love_reduct=p.love_reduct, ecc=p.ecc, dt=p.dt, tide=p.tide))
        for th in (threads, 1):
            O.set_threads(th)
            sim.step(2)
            n, t = 0, 0.0
            while t < args.cpu_size_seconds:
                t0 = time.perf_counter()
                sim.step(2)
                t += time.perf_counter() - t0
                n += 2
            rows.append({"cells": grid.n, "threads": th, "value": grid.n * n / t, "steps": n})
            print("bench: cpu %d cells, %d threads: %.3e cell-steps/s" % (grid.n, th, rows[-1]["value"]),
                  file=sys.stderr, flush=True)
    O.set_threads(threads)
    return {"workload": "plain method, free surface, ECC, FP64 C oracle (OpenMP rows)", "unit": "cell-steps/s",
            "host_cpu": model, "logical_cpus": os.cpu_count(), "rows": rows}


if __name__ == "__main__":
    main()
