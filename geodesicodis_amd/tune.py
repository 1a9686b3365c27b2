"""Launch-shape autotuner for a context's step.

The step kernels have launch shapes that only the device can rank: the
one-kernel step or the face + cell pair, cells per patch, lanes per cell,
the interleaved batch layout with the faces / (cell, instance) pairs each
lane keeps in flight, its workgroup order and its operand records, and
HIP-graph replay.  Every candidate gives the same
state bit for bit (dissipation sums differ by rounding), so tuning never
changes results.

tune() times each candidate on a scratch context of the same mesh, batch,
precision and SH coupling (the caller's context is not stepped), keeps the
fastest, and caches the choice per (cells, instances, precision, SH mode,
degree, device) in memory and in a JSON file (ODIS_TUNE_CACHE, else
~/.cache/geodesicodis_amd/tune.json).  A candidate replaces the default only
when it is at least `margin` faster.
"""
import itertools
import json
import os
import time

from . import (FP64, OPT_GRAPHS, OPT_IL_ORDER, OPT_IL_RECORDS, OPT_IL_UNROLL, OPT_INTERLEAVE,
               OPT_INTERLEAVE_ACTIVE, OPT_ONE_KERNEL, OPT_PATCH_CELLS, OPT_STEP_LANES, SH_OFF, Context, OdisError)

TUNED = (OPT_ONE_KERNEL, OPT_PATCH_CELLS, OPT_STEP_LANES, OPT_INTERLEAVE, OPT_IL_UNROLL, OPT_IL_ORDER, OPT_IL_RECORDS,
         OPT_GRAPHS)
NAMES = {OPT_ONE_KERNEL: "one_kernel", OPT_PATCH_CELLS: "patch_cells", OPT_STEP_LANES: "step_lanes",
         OPT_INTERLEAVE: "interleave", OPT_IL_UNROLL: "il_unroll", OPT_IL_ORDER: "il_order",
         OPT_IL_RECORDS: "il_records", OPT_GRAPHS: "graphs"}

_MEMO = {}


def _cache_path():
    return os.environ.get("ODIS_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache",
                                                              "geodesicodis_amd", "tune.json")


def _load_cache():
    try:
        with open(_cache_path()) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_cache(key, entry):
    path = _cache_path()
    data = _load_cache()
    data[key] = entry
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: the in-memory cache still holds it


def _device_name():
    try:
        with open("/sys/class/kfd/kfd/topology/nodes/1/name") as fh:
            return fh.read().strip()
    except OSError:
        return "gpu"


def candidates(ctx, graphs=True):
    """Option settings (dicts) valid for the context; graphs=False leaves
    out graph replay (for runs that stamp kernels with events, which replay
    no graphs)."""
    out = []
    defaults = {o: ctx.option(o) for o in TUNED}
    one_ok = defaults[OPT_ONE_KERNEL] == 1 or _accepts(ctx, OPT_ONE_KERNEL, 1)
    for graphs in ((0, 1) if graphs else (0,)):
        out.append({OPT_ONE_KERNEL: 0, OPT_GRAPHS: graphs, OPT_INTERLEAVE: 0})
        if not one_ok:
            continue
        for patch, lanes in itertools.product((64, 128, 256), (4, 2)):
            base = {OPT_ONE_KERNEL: 1, OPT_PATCH_CELLS: patch, OPT_STEP_LANES: lanes, OPT_GRAPHS: graphs,
                    OPT_INTERLEAVE: 0}
            out.append(dict(base))
            if patch == 64 and lanes == 4 and ctx.B >= 2:  # lanes per cell do not apply to it
                for unroll, order, rec in itertools.product((0, 2), (0, 1, 2, 4), (1, 0)):
                    out.append({**base, OPT_INTERLEAVE: 1, OPT_IL_UNROLL: unroll, OPT_IL_ORDER: order,
                                OPT_IL_RECORDS: rec})
    return out


def _accepts(ctx, opt, value):
    old = ctx.option(opt)
    try:
        ctx.set_option(opt, value)
    except OdisError:
        return False
    ctx.set_option(opt, old)
    return True


def apply(ctx, options):
    """Set option values (a dict option -> value) on a context; options the
    context refuses are left as they are.  Returns the options set."""
    done = {}
    order = (OPT_ONE_KERNEL, OPT_PATCH_CELLS, OPT_STEP_LANES, OPT_INTERLEAVE, OPT_IL_UNROLL, OPT_IL_ORDER,
             OPT_IL_RECORDS, OPT_GRAPHS)
    for o in order:
        if o in options:
            try:
                ctx.set_option(o, int(options[o]))
                done[o] = int(options[o])
            except OdisError:
                pass
    return done


def _time(ctx, steps, warmup, repeats):
    ctx.step(warmup)
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        ctx.step(steps)
        best = min(best, time.perf_counter() - t0)
    return best / steps


def tune(mesh, params, precision=FP64, sh=None, steps=100, warmup=20, repeats=3, margin=0.02, use_cache=True,
         log=None, graphs=True):
    """The fastest launch options for stepping this batch; sh = (l_max, beta,
    mode) of the SH coupling or None; graphs=False leaves graph replay out.
    Returns a dict: "options" (name -> value, to pass to apply() through
    option_ids()), "ms_per_step" tuned and "default_ms_per_step", "speedup",
    "candidates" timed, "cached"."""
    params = list(params)
    key = "%d/%d/%d/%s/%d/%s%s" % (mesh.N, len(params), precision, sh[2] if sh else SH_OFF,
                                   sh[0] if sh else 0, _device_name(), "" if graphs else "/no-graphs")
    if use_cache:
        if key in _MEMO:
            return dict(_MEMO[key], cached=True)
        disk = _load_cache()
        if key in disk:
            _MEMO[key] = disk[key]
            return dict(disk[key], cached=True)
    ctx = Context(mesh, params, precision=precision)
    if sh:
        ctx.set_sh(*sh[:2], mode=sh[2])
    default = {o: ctx.option(o) for o in TUNED}
    t_default = _time(ctx, steps, warmup, repeats)
    rows = []
    best, t_best = dict(default), t_default
    for cand in candidates(ctx, graphs):
        apply(ctx, default)
        set_ = apply(ctx, cand)
        if set_ != {o: int(v) for o, v in cand.items()}:
            continue  # the context refused part of it
        if cand.get(OPT_INTERLEAVE) == 1 and ctx.option(OPT_INTERLEAVE_ACTIVE) != 1:
            continue
        t = _time(ctx, steps, warmup, repeats)
        rows.append(({NAMES[o]: v for o, v in cand.items()}, t * 1e3))
        if t < t_best * (1.0 - margin):
            best, t_best = {**default, **cand}, t
    ctx.close()
    entry = {"options": {NAMES[o]: int(v) for o, v in best.items()}, "ms_per_step": t_best * 1e3,
             "default_ms_per_step": t_default * 1e3, "speedup": t_default / t_best,
             "default": {NAMES[o]: int(v) for o, v in default.items()}, "candidates": len(rows),
             "timings_ms": sorted(rows, key=lambda r: r[1])[:8]}
    if log:
        log("tune %s: %d candidates, default %.4f ms, best %.4f ms %s" % (key, len(rows), t_default * 1e3,
                                                                       t_best * 1e3, entry["options"]))
    _MEMO[key] = entry
    if use_cache:
        _store_cache(key, entry)
    return dict(entry, cached=False)


def option_ids(named):
    """Option name -> value dict of a tune() result as option id -> value."""
    ids = {v: k for k, v in NAMES.items()}
    return {ids[k]: v for k, v in named.items()}
