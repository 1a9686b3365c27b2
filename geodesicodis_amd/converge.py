"""Batched runs to converged orbit-averaged dissipation.

The reference runs one ocean for a fixed number of orbits and logs the
dissipation at its output cadence (src/timeIntegrator.cpp:280-304,
src/energy.cpp:13-62).  A parameter sweep instead needs every instance run
until its orbit-averaged dissipation is converged.  This driver steps a batch
of instances in one context and, at each instance's orbit ends:

- takes the FP64 orbit mean D_k of its per-step dissipation and its FP64
  orbit-periodicity residuals ||x(t) - x(t - P)|| / ||x(t)|| on the device
  (odis_orbit_stats);
- applies the stop rule: stationary (|D_k - D_{k-1}| / D_k < tol for
  `stable` consecutive orbits), or, when a reference D* is given, within
  `band` of D* for `stable` consecutive orbits; or the orbit cap;
- retires finished instances from the batch (odis_retire), so that later
  steps cost only the instances still running.

Each instance's orbit is its own quantised N_s steps (multiples of 100), so
orbit ends fall on multiples of the 100-step chunk the driver steps by.
"""
import time

import numpy as np


class InstanceRecord:
    """Orbit history and outcome of one instance."""

    def __init__(self, steps_per_orbit):
        self.steps_per_orbit = int(steps_per_orbit)
        self.D = []            # orbit means, W/m^2
        self.resid_v = []      # periodicity residuals of v and eta (NaN for the first orbit)
        self.resid_eta = []
        self.t = []            # wall seconds at each orbit end
        self.orbits = 0
        self.stop = None       # "stationary", "band" or "cap"
        self.wall_s = None
        self.run = 0           # consecutive orbits meeting the stop rule

    @property
    def D_star(self):
        return self.D[-1] if self.D else None

    def as_dict(self):
        return {"orbits": self.orbits, "stop": self.stop, "wall_s": self.wall_s, "D_star": self.D_star,
                "steps_per_orbit": self.steps_per_orbit,
                "resid_v": self.resid_v[-1] if self.resid_v else None,
                "resid_eta": self.resid_eta[-1] if self.resid_eta else None}


def run_to_convergence(ctx, steps_per_orbit, tol=1e-4, stable=3, max_orbits=2000, chunk=100, retire=True,
                       d_ref=None, band=0.01, log=None, on_stop=None):
    """Step the batch of ctx until every instance has stopped.

    steps_per_orbit: N_s of each instance (multiples of chunk).  d_ref (optional,
    one value per instance): stop an instance once its orbit means stay within
    `band` of d_ref for `stable` orbits instead of at stationarity.  The
    context's state is the start (odis_set_state or rest) at step 0.
    on_stop(slot, b), if given, runs when instance b (context slot `slot`)
    stops, before it is retired (to read its final state).  Returns one
    InstanceRecord per instance, in the context's original order; ctx ends
    holding the instances that stopped at the last orbit end (a context keeps
    at least one), or all of them when retire is False."""
    ns = np.asarray(steps_per_orbit, np.int64)
    B = len(ns)
    if ctx.B != B:
        raise ValueError("steps_per_orbit must hold one entry per instance")
    if np.any(ns % chunk):
        raise ValueError("every orbit length must be a multiple of the chunk")
    recs = [InstanceRecord(n) for n in ns]
    live = list(range(B))   # context slot j holds original instance live[j]
    done = np.zeros(B, bool)
    ctx.sync()
    t0 = time.perf_counter()
    step = 0
    while not done.all():
        ctx.step(chunk, sync=True)
        step += chunk
        slots = [j for j, b in enumerate(live) if not done[b] and step % ns[b] == 0]
        if not slots:
            continue
        st = ctx.orbit_stats(slots, [int(ns[live[j]]) for j in slots])
        now = time.perf_counter() - t0
        for (j, (d, rv, re)) in zip(slots, st):
            b = live[j]
            r = recs[b]
            r.D.append(float(d))
            r.resid_v.append(float(rv))
            r.resid_eta.append(float(re))
            r.t.append(now)
            r.orbits += 1
            if d_ref is not None:
                ok = abs(d - d_ref[b]) <= band * abs(d_ref[b])
            else:
                ok = len(r.D) > 1 and abs(r.D[-1] - r.D[-2]) <= tol * abs(r.D[-1])
            r.run = r.run + 1 if ok else 0
            if r.run >= stable:
                r.stop = "band" if d_ref is not None else "stationary"
            elif r.orbits >= max_orbits:
                r.stop = "cap"
            if r.stop:
                r.wall_s = now
                done[b] = True
                if on_stop is not None:
                    on_stop(j, b)
        finished = [j for j in slots if done[live[j]]]
        if finished and log:
            log("converge: step %d, %d of %d instances stopped (%.1f s)" % (step, int(done.sum()), B, now))
        if finished and retire:
            keep = [j for j in range(len(live)) if not done[live[j]]]
            if keep:
                ctx.retire(keep)
                live = [live[j] for j in keep]
    return recs


def time_to_band(rec, d_star, band=0.01, stable=3):
    """(orbits, wall seconds) until the orbit means of rec stay within band of
    d_star for `stable` orbits (the first of them), or (None, None)."""
    ok = [abs(d - d_star) <= band * abs(d_star) for d in rec.D]
    for k in range(len(ok) - stable + 1):
        if all(ok[k:]):
            return k + 1, rec.t[k]
    return None, None
