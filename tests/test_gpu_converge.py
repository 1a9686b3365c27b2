"""Orbit statistics, batch retirement and the run-to-convergence driver on
the GPU, through the C ABI."""
import numpy as np
import pytest

from conftest import generated_grid, rel_l2

pytestmark = pytest.mark.gpu


def _batch(G, level, specs, precision):
    g, _ = generated_grid(level)
    ps, ns = [], []
    for h, a, ts in specs:
        p, n = G.make_params(h=h, alpha=a, time_step=ts)
        ps.append(p)
        ns.append(n)
    mesh = G.Mesh(g, ps[0].radius)
    return mesh, ps, ns


def test_orbit_stats_match_series_and_states(product):
    """odis_orbit_stats: the orbit mean equals the mean of the per-step
    dissipation series over the orbit, and the periodicity residuals equal
    ||x(t) - x(t - P)|| / ||x(t)|| of the states read at the two orbit ends
    (NaN at an instance's first orbit).  Instances have their own orbit
    lengths (1200 and 2400 steps)."""
    G = product
    mesh, ps, ns = _batch(G, 4, [(38e3, 1e-7, 100), (5e3, 1e-8, 50), (20e3, 3e-7, 100)], G.FP64)
    assert ns == [1200, 2400, 1200]
    ctx = G.Context(mesh, ps)
    ctx.step(1200)
    s1 = [ctx.get_state(b) for b in range(3)]
    st = ctx.orbit_stats([0, 2], [1200, 1200])
    for j, b in enumerate((0, 2)):
        want = np.mean(ctx.dissipation_series(b, 0, 1200))
        assert abs(st[j, 0] - want) <= 1e-13 * abs(want)
        assert np.isnan(st[j, 1]) and np.isnan(st[j, 2])
    ctx.step(1200)
    st = ctx.orbit_stats([0, 1, 2], [1200, 2400, 1200])
    for j in range(3):
        s2 = ctx.get_state(j)
        want = np.mean(ctx.dissipation_series(j, 2400 - ns[j], ns[j]))
        assert abs(st[j, 0] - want) <= 1e-13 * abs(want)
        if j != 1:
            rv = np.linalg.norm(s2["v"] - s1[j]["v"]) / np.linalg.norm(s2["v"])
            re = np.linalg.norm(s2["eta"] - s1[j]["eta"]) / np.linalg.norm(s2["eta"])
            assert abs(st[j, 1] - rv) <= 1e-12 * rv and abs(st[j, 2] - re) <= 1e-12 * re
        else:
            assert np.isnan(st[j, 1])
    with pytest.raises(G.OdisError):
        ctx.orbit_stats([0], [4800])   # longer than the series held
    with pytest.raises(G.OdisError):
        ctx.orbit_stats([0, 0], [1200, 1200])
    ctx.close()


@pytest.mark.parametrize("level,n,precision", [(4, 5, 0), (7, 20, 1)])
def test_retire_keeps_instances_bit_for_bit(product, level, n, precision):
    """odis_retire drops instances from a batch: the kept ones continue
    exactly as in a batch that kept everyone (state and history bit for bit;
    dissipation series and orbit statistics to rounding), including a
    20-instance FP32 batch held interleaved that drops below the 16
    instances of one interleaved group."""
    G = product
    rng = np.random.default_rng(4)
    ts = 40 if level == 4 else 5   # below 0.35 / omega_max of a 100 km ocean on the grid
    specs = [(10 ** rng.uniform(3.5, 5), 10 ** rng.uniform(-9, -7), ts) for _ in range(n)]
    mesh, ps, ns = _batch(G, level, specs, precision)
    full = G.Context(mesh, ps, precision=precision)
    part = G.Context(mesh, ps, precision=precision)
    if level == 7:
        assert part.option(G.OPT_INTERLEAVE_ACTIVE) == 1
    for c in (full, part):
        c.step(60)
    keep = [b for b in range(n) if b % 3 != 1]
    part.retire(keep)
    assert part.B == len(keep)
    for c in (full, part):
        c.step(140)
    for j, b in enumerate(keep):
        a, r = part.get_state(j), full.get_state(b)
        for k in ("v", "eta", "dvdt", "detadt"):
            assert np.array_equal(a[k], r[k]), (j, b, k)
        # the series: the same sums, except that a batch below 16 instances steps
        # instance-major, which groups the per-face terms of the FP64 sums another way
        da, db = part.dissipation_series(j, 0, 200), full.dissipation_series(b, 0, 200)
        assert np.allclose(da, db, rtol=1e-12, atol=0), np.abs(da - db).max()
    sa = part.orbit_stats(list(range(len(keep))), [150] * len(keep))
    sb = full.orbit_stats(keep, [150] * len(keep))
    assert np.allclose(sa[:, 0], sb[:, 0], rtol=1e-12, atol=0)
    with pytest.raises(G.OdisError):
        part.retire([1, 0])
    full.close()
    part.close()


def test_run_to_convergence_retires_and_matches_single_runs(product):
    """The driver on a batch of strongly damped oceans (642 cells): every
    instance reaches stationary orbit means (|D_k - D_k-1| / D_k < 1e-4 for 3
    orbits) and is retired; each D* equals the orbit mean of the same
    instance run alone for the same number of orbits, and the periodicity
    residual of a stationary instance is small."""
    from geodesicodis_amd.converge import run_to_convergence
    G = product
    specs = [(38e3, 3e-6, 100), (10e3, 1e-5, 50), (60e3, 5e-6, 50), (20e3, 2e-6, 100)]
    mesh, ps, ns = _batch(G, 4, specs, G.FP64)
    ctx = G.Context(mesh, ps)
    recs = run_to_convergence(ctx, ns, tol=1e-4, stable=3, max_orbits=200)
    for b, r in enumerate(recs):
        assert r.stop == "stationary", (b, r.as_dict())
        assert r.resid_v[-1] < 1e-3 and r.resid_eta[-1] < 1e-3, r.as_dict()
        solo = G.Context(mesh, ps[b])
        solo.step(r.orbits * ns[b])
        d = np.mean(solo.dissipation_series(0, (r.orbits - 1) * ns[b], ns[b]))
        assert abs(d - r.D_star) <= 1e-12 * abs(d), (b, d, r.D_star)
        solo.close()
    ctx.close()
