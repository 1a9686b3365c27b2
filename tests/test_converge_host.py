"""The run-to-convergence driver's host logic without a GPU: stop rules
(stationary, band, orbit cap), orbit ends of instances with different orbit
lengths, and the bookkeeping of retired instances.  A stand-in context
serves scripted orbit means per instance, checks that every orbit_stats
request names the instance's own orbit length at one of its orbit ends, and
applies retire() as odis_retire does (the kept slots, in order)."""
import numpy as np
import pytest

from geodesicodis_amd import converge as C


class FakeBatch:
    def __init__(self, ns, means):
        self.ns = list(ns)
        self.means = [list(m) for m in means]     # orbit means of each original instance
        self.live = list(range(len(ns)))          # slot j holds original instance live[j]
        self.step_n = 0
        self.retired = []

    @property
    def B(self):
        return len(self.live)

    def sync(self):
        pass

    def step(self, n, sync=False):
        self.step_n += n

    def orbit_stats(self, slots, ns):
        out = []
        for j, n in zip(slots, ns):
            b = self.live[j]
            assert n == self.ns[b] and self.step_n % n == 0, (j, b, n, self.step_n)
            k = self.step_n // n - 1
            out.append((self.means[b][k], 0.0 if k else np.nan, 0.0 if k else np.nan))
        return out

    def retire(self, keep):
        assert list(keep) == sorted(set(keep)) and all(0 <= j < len(self.live) for j in keep)
        self.retired.append([self.live[j] for j in range(len(self.live)) if j not in keep])
        self.live = [self.live[j] for j in keep]


def _settling(d_inf, rate, n):
    """Orbit means approaching d_inf geometrically."""
    return [d_inf * (1.0 + rate ** (k + 1)) for k in range(n)]


def test_stationary_stop_orbit_lengths_and_retirement():
    ns = [300, 200, 500]
    means = [_settling(1.0, 0.1, 50), _settling(2.0, 0.5, 50), _settling(3.0, 0.01, 50)]
    ctx = FakeBatch(ns, means)
    stops = []
    recs = C.run_to_convergence(ctx, ns, tol=1e-4, stable=3, chunk=100,
                                on_stop=lambda slot, b: stops.append((slot, b, ctx.live[slot])))
    for b, r in enumerate(recs):
        assert r.stop == "stationary"
        rel = [abs(means[b][k] - means[b][k - 1]) / means[b][k] for k in range(1, len(means[b]))]
        # the first orbit k with 3 consecutive relative changes <= 1e-4 ends the run
        k = next(k for k in range(2, len(rel)) if all(x <= 1e-4 for x in rel[k - 2:k + 1]))
        assert r.orbits == k + 2 and r.D == means[b][:k + 2]
        assert r.D_star == means[b][k + 1]
        assert np.isnan(r.resid_v[0]) and r.resid_v[-1] == 0.0
    # every on_stop saw the instance it names in that slot, before retirement
    assert all(b == held for _, b, held in stops)
    assert sorted(b for _, b, _ in stops) == [0, 1, 2]
    # the instances that stopped at the last orbit end stay in the context
    # (odis_retire keeps at least one); the others were retired as they stopped
    last = max(r.wall_s for r in recs)
    assert ctx.live == [b for b in range(3) if recs[b].wall_s == last]
    assert sorted(sum(ctx.retired, []) + ctx.live) == [0, 1, 2]


def test_band_rule_and_cap():
    ns = [100, 100]
    means = [_settling(1.0, 0.5, 40), [1.0 + 0.5 * (-1) ** k for k in range(40)]]
    ctx = FakeBatch(ns, means)
    recs = C.run_to_convergence(ctx, ns, stable=3, max_orbits=25, chunk=100, d_ref=[1.0, 1.0], band=0.01)
    # instance 0: within 1 % from orbit 7 on (0.5^7 < 0.01), stops 3 orbits later
    first = next(k for k in range(40) if abs(means[0][k] - 1.0) <= 0.01)
    assert recs[0].stop == "band" and recs[0].orbits == first + 3
    # instance 1 never settles: stopped by the cap
    assert recs[1].stop == "cap" and recs[1].orbits == 25
    assert recs[0].wall_s <= recs[1].wall_s


def test_without_retirement_every_slot_stays():
    ns = [100, 200]
    ctx = FakeBatch(ns, [_settling(1.0, 0.01, 20), _settling(1.0, 0.01, 20)])
    recs = C.run_to_convergence(ctx, ns, chunk=100, retire=False)
    assert all(r.stop == "stationary" for r in recs) and ctx.live == [0, 1] and ctx.retired == []


def test_refuses_bad_orbit_lengths():
    with pytest.raises(ValueError):
        C.run_to_convergence(FakeBatch([150], [[1.0]]), [150], chunk=100)
    with pytest.raises(ValueError):
        C.run_to_convergence(FakeBatch([100, 100], [[1.0], [1.0]]), [100], chunk=100)


def test_time_to_band_needs_the_means_to_stay():
    r = C.InstanceRecord(100)
    r.D = [2.0, 1.005, 1.2, 1.004, 0.995, 1.001, 1.0]
    r.t = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0]
    # orbit 2 is inside the band but orbit 3 leaves it: the band is reached
    # for good at orbit 4
    assert C.time_to_band(r, 1.0, band=0.01, stable=3) == (4, 4.0)
    assert C.time_to_band(r, 1.5, band=0.01, stable=3) == (None, None)
    r.D = r.D[:5]
    r.t = r.t[:5]
    # two orbits in the band at the end are not enough for stable = 3
    assert C.time_to_band(r, 1.0, band=0.01, stable=3) == (None, None)
