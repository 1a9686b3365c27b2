"""The launch-shape autotuner on the GPU: it times the candidates on a
scratch context, its choice changes no result, and it caches the choice."""
import numpy as np
import pytest

from conftest import generated_grid

pytestmark = pytest.mark.gpu


def test_tuner_choice_is_bit_identical_and_cached(product, tmp_path, monkeypatch):
    """A 24-instance FP32 batch on 40962 cells: the tuner times the one-kernel
    step (patch sizes, lanes per cell), the face + cell pair, the interleaved
    layout with its unrolls and graph replay; the chosen options step the
    batch to the same state bit for bit as the defaults, no slower than the
    defaults by the tuner's own timing, and a second call is served from
    the cache file."""
    from geodesicodis_amd import tune
    G = product
    monkeypatch.setenv("ODIS_TUNE_CACHE", str(tmp_path / "tune.json"))
    tune._MEMO.clear()
    g, _ = generated_grid(7)
    rng = np.random.default_rng(9)
    ps = [G.make_params(h=10 ** rng.uniform(3.5, 5), alpha=10 ** rng.uniform(-10, -7), time_step=5)[0]
          for _ in range(24)]
    mesh = G.Mesh(g, ps[0].radius)
    res = tune.tune(mesh, ps, precision=G.FP32, steps=40, warmup=5, repeats=2)
    print(res)
    assert res["candidates"] >= 10 and not res["cached"]
    assert res["ms_per_step"] <= res["default_ms_per_step"]
    a = G.Context(mesh, ps, precision=G.FP32)
    b = G.Context(mesh, ps, precision=G.FP32)
    tune.apply(b, tune.option_ids(res["options"]))
    for c in (a, b):
        c.step(120)
    for k in (0, 23):
        sa, sb = a.get_state(k), b.get_state(k)
        for key in ("v", "eta", "dvdt", "detadt"):
            assert np.array_equal(sa[key], sb[key]), (k, key)
    tune._MEMO.clear()
    again = tune.tune(mesh, ps, precision=G.FP32)
    assert again["cached"] and again["options"] == res["options"]


def test_every_tuner_candidate_steps_alike(product):
    """Each launch shape the tuner may pick (face + cell pair, one-kernel
    step at every patch size and lane count, interleaved layout with every
    unroll and workgroup order, graph replay) steps a 20-instance FP32 batch
    on 40962 cells to the same state bit for bit."""
    from geodesicodis_amd import tune
    G = product
    g, _ = generated_grid(7)
    rng = np.random.default_rng(10)
    ps = [G.make_params(h=10 ** rng.uniform(3.5, 5), alpha=10 ** rng.uniform(-10, -7), time_step=5)[0]
          for _ in range(20)]
    mesh = G.Mesh(g, ps[0].radius)
    ref = G.Context(mesh, ps, precision=G.FP32)
    ref.step(70)
    want = [ref.get_state(k) for k in (0, 19)]
    ref.close()
    probe = G.Context(mesh, ps, precision=G.FP32)
    cands = tune.candidates(probe)
    probe.close()
    assert len(cands) >= 30
    for cand in cands:
        c = G.Context(mesh, ps, precision=G.FP32)
        tune.apply(c, cand)
        c.step(70)
        for k, w in zip((0, 19), want):
            s = c.get_state(k)
            for key in ("v", "eta", "dvdt", "detadt"):
                assert np.array_equal(s[key], w[key]), (cand, k, key)
        c.close()
