"""The benchmark configurations at their real shapes on the GPU against the
CPU FP64 oracle, through the C ABI: the seeded 256-instance sweep on 40962
cells, the seeded 512-instance uneven batch over four grids (its shell
groups coupled by the batched split-BF16 products), one full orbit of the
plain method on the 163842-cell grid, and one full orbit of a shell-coupled
step on 10242 cells."""
import numpy as np
import pytest

from conftest import generated_grid, rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _oracle_params(O, p):
    return O.make_params(**{k: getattr(p, k) for k in
                            ("g", "h", "alpha", "omega", "radius", "love_reduct", "ecc", "obliquity", "dt",
                             "shell_thickness", "semimajor", "tide", "surface", "friction", "init_load")},
                         quad_drag_momentum=p.quad_drag)


def _errs(a, b):
    return {k: rel_l2(a[k], b[k]) for k in ("v", "eta", "dvdt", "detadt")}


def _projector(O, arrs, L, beta):
    """The oracle's SH coupling as dense products: eta_eff = eta + Pb (Pc eta)
    with Pc the least-squares analysis on scipy's harmonics and Pb the
    synthesis weighted by beta_l - 1 (2 <= l <= L)."""
    Y = O.real_sh_basis(L, arrs["lat"], arrs["lon"])
    Pc = np.linalg.solve(Y.T @ Y, Y.T)
    d = np.zeros((L + 1) ** 2)
    for l in range(2, L + 1):
        d[l * l:(l + 1) ** 2] = beta[l] - 1.0
    return Y * d[None, :], Pc


def test_sweep_256_instances_40962_cells(product, oracle):
    """The batched sweep as benchmarked: 256 instances on 40962 cells,
    thickness 10^U(3,5) m and drag 10^U(-11,-7) s^-1 from
    default_rng(20261203), per-instance dt, ECC tide, rest start, 300 steps.
    Six instances (thinnest, median and thickest ocean; weakest, median and
    strongest drag) against single oracle runs: the FP64 batch within 1e-10,
    the FP32 batch (stepping interleaved) within 1e-4.  Every instance's
    dissipation is finite, and an instance stepped alone in FP32 equals its
    copy in the batch bit for bit."""
    import bench
    G, O = product, oracle
    ps = bench.sweep_params(G, 256)
    g, arrs = generated_grid(7)
    mesh = G.Mesh(g, ps[0].radius)
    assert mesh.N == 40962
    n = 300
    c32 = G.Context(mesh, ps, precision=G.FP32)
    assert c32.option(G.OPT_INTERLEAVE_ACTIVE) == 1
    c32.step(n)
    d32 = c32.dissipation()
    assert np.all(np.isfinite(d32)) and np.all(d32 > 0)
    c64 = G.Context(mesh, ps, precision=G.FP64)
    c64.step(n)
    d64 = c64.dissipation()
    hs = np.array([p.h for p in ps])
    al = np.array([p.alpha for p in ps])
    picks = sorted({int(np.argmin(hs)), int(np.argsort(hs)[128]), int(np.argmax(hs)),
                    int(np.argmin(al)), int(np.argsort(al)[128]), int(np.argmax(al))})
    om = O.Mesh(arrs, ps[0].radius)
    for b in picks:
        sim = O.Sim(om, _oracle_params(O, ps[b]))
        sim.step(n)
        want = sim.get_state()
        e64 = _errs(c64.get_state(b), want)
        e32 = _errs(c32.get_state(b), want)
        print("sweep instance %d (h %.3g m, alpha %.3g): fp64 %s fp32 %s" % (b, hs[b], al[b], e64, e32))
        assert max(e64.values()) <= TOL64, e64
        assert max(e32[k] for k in ("v", "eta", "dvdt")) <= 1e-4, e32
        dor = sim.dissipation()
        assert abs(d64[b] - dor) <= TOL64 * abs(dor)
        assert abs(d32[b] - dor) <= 1e-4 * abs(dor)
    b = picks[0]
    solo = G.Context(mesh, ps[b], precision=G.FP32)
    solo.step(n)
    s1, s2 = solo.get_state(), c32.get_state(b)
    for k in ("v", "eta", "dvdt", "detadt"):
        assert np.array_equal(s1[k], s2[k]), k
    for c in (solo, c32, c64):
        c.close()


def test_uneven_512_instances(product, oracle):
    """The uneven batch as benchmarked: 512 instances from
    default_rng(20261204) over the 2562- to 163842-cell grids, FULL tide,
    quadratic drag in the momentum equation, half under a 20 km LID_MEMBR
    shell coupled to degree 16 through the batched split-BF16 products, FP32
    state, 8 contexts (one per grid and surface).  The 8 contexts step
    together, one stream each; every instance's dissipation is finite; each
    context's state then equals the same context stepped alone, bit for bit;
    and two instances per context match single oracle runs (free surface
    1e-4, shell 2e-3 relative: the split-BF16 products keep about 16 bits)."""
    import bench
    G, O = product, oracle
    groups, beta = bench.uneven_groups(G, 512)
    assert sum(len(v) for v in groups.values()) == 512 and len(groups) == 8
    L = bench.UNEVEN_SH_DEGREE
    n = 120
    built = {}
    for key, ps in sorted(groups.items()):
        g, arrs = generated_grid(key[0] + 1)
        mesh = G.Mesh(g, ps[0].radius)
        ctx = G.Context(mesh, ps, precision=G.FP32)
        if key[1]:
            ctx.set_sh(L, np.tile(beta, (len(ps), 1)), mode=G.SH_MFMA)
        built[key] = (g, arrs, mesh, ctx)
    for _, _, _, ctx in built.values():
        ctx.sync()
    for _, _, _, ctx in built.values():
        ctx.step(n, sync=False)
    for _, _, _, ctx in built.values():
        ctx.sync()
    for key, (g, arrs, mesh, ctx) in sorted(built.items()):
        ps = groups[key]
        d = ctx.dissipation()
        assert np.all(np.isfinite(d)) and np.all(d > 0), key
        alone = G.Context(mesh, ps, precision=G.FP32)
        if key[1]:
            alone.set_sh(L, np.tile(beta, (len(ps), 1)), mode=G.SH_MFMA)
        alone.step(n)
        picks = (0, len(ps) - 1)
        om = O.Mesh(arrs, ps[0].radius)
        proj = _projector(O, arrs, L, beta) if key[1] else None
        for b in picks:
            sim = O.Sim(om, _oracle_params(O, ps[b]))
            if proj is not None:
                sim.set_projector(*proj)
            sim.step(n)
            want = sim.get_state()
            s1, s2 = alone.get_state(b), ctx.get_state(b)
            e, e_alone = _errs(s2, want), _errs(s1, want)
            print("uneven group %s instance %d: concurrent %s alone %s" % (key, b, e, e_alone))
            tol = 2e-3 if key[1] else 1e-4
            assert max(e[k] for k in ("v", "eta", "dvdt")) <= tol, (key, b, e)
            for k in ("v", "eta", "dvdt", "detadt"):
                assert np.array_equal(s1[k], s2[k]), (key, b, k)
        alone.close()
        ctx.close()


def test_full_size_one_orbit_fp64(product, oracle):
    """The plain method on the 163842-cell grid for one full orbit: free
    surface, ECC, h = 38 km, alpha = 1e-8, dt = P / 18300 (the quantised
    6.5 s target), rest start, FP64.  Every state array and the per-step
    dissipation series within 1e-10 of the oracle after 18300 steps."""
    G, O = product, oracle
    g, arrs = generated_grid(8)
    p, n = G.make_params(time_step=6.5)
    assert n == 18300
    mesh = G.Mesh(g, p.radius)
    ctx = G.Context(mesh, p)
    ctx.step(n)
    sim = O.Sim(O.Mesh(arrs, p.radius), _oracle_params(O, p))
    d_orc = sim.step(n, record=True)
    e = _errs(ctx.get_state(), sim.get_state())
    print("163842 cells, one orbit (18300 steps), fp64:", e)
    assert max(e.values()) <= TOL64, e
    series = ctx.dissipation_series(0, 1, n - 1)
    assert rel_l2(series, d_orc[:-1]) <= TOL64
    ctx.close()


def test_shell_one_orbit_10242_cells(product, oracle):
    """One full orbit of the shell-coupled step: 10242 cells, Enceladus ocean
    (38 km) under a 5 km LID_MEMBR shell, self-gravity / shell coupling by
    the SH projection to degree 16 with the basis evaluated on the fly, ECC,
    FP64, dt below 0.35 / omega_max of the grid and shell modes.  State
    within 1e-10 of the oracle's dense least-squares projector after the
    orbit."""
    G, O = product, oracle
    L, hs, ho = 16, 5e3, 38e3
    beta = np.array(G.surface_beta(G.LID_MEMBR, L, 252.1e3, 0.113, hs, ho))
    nu, _, ro, go = O.membrane_nu_beta(252.1e3, hs, ho, L)
    ls = np.arange(L + 1)
    w = max(0.0193, float(np.max(np.sqrt(go * ho * np.maximum(beta, 1.0) * ls * (ls + 1)) / ro)))
    p, n = G.make_params(radius=ro, g=go, surface=G.LID_MEMBR, shell_thickness=hs, love_reduct=nu[2],
                         time_step=0.35 / w)
    g, arrs = generated_grid(6)
    mesh = G.Mesh(g, p.radius)
    ctx = G.Context(mesh, p)
    ctx.set_sh(L, beta, mode=G.SH_RECUR)
    ctx.step(n)
    sim = O.Sim(O.Mesh(arrs, p.radius), _oracle_params(O, p))
    sim.set_projector(*_projector(O, arrs, L, beta))
    sim.step(n)
    e = _errs(ctx.get_state(), sim.get_state())
    print("shell 5 km, 10242 cells, degree 16, one orbit (%d steps of %.3f s):" % (n, p.dt), e)
    assert max(e.values()) <= TOL64, e
    ctx.close()
