"""The autotuner's host logic without a GPU: which launch options it
proposes for a context, how it applies a choice, that option names map back
to the C ABI's ids, and how tune() ranks candidates (the margin, inactive
interleaving skipped) and caches its choice in memory and on disk.  A
stand-in context records the options set on it and refuses the ones a real
context would refuse; a stand-in timer gives each option set its time."""
import json

import pytest

import geodesicodis_amd as G
from geodesicodis_amd import tune as T


class FakeContext:
    """Options as odis_set_option keeps them; `refuse` holds (option, value)
    pairs the context rejects (as a context without the one-kernel step
    rejects ODIS_OPT_ONE_KERNEL 1)."""

    def __init__(self, B, refuse=()):
        self.B = B
        self.refuse = set(refuse)
        self.opts = {o: 0 for o in T.TUNED}
        self.opts.update({G.OPT_ONE_KERNEL: 1, G.OPT_PATCH_CELLS: 64, G.OPT_STEP_LANES: 4,
                          G.OPT_INTERLEAVE: 2, G.OPT_IL_RECORDS: 1})

    def option(self, o):
        return self.opts[o]

    def set_option(self, o, v):
        if (o, v) in self.refuse:
            raise G.OdisError(-4, "refused")  # ODIS_E_UNSUPPORTED
        self.opts[o] = v


def test_candidates_cover_the_launch_shapes():
    cands = T.candidates(FakeContext(B=256))
    # per graph setting: the face + cell pair, 6 one-kernel shapes, and for
    # the 64-cell patch with 4 lanes 2 unrolls x 4 orders x records on / off
    assert len(cands) == 2 * (1 + 6 + 2 * 4 * 2)
    il = [c for c in cands if c.get(G.OPT_INTERLEAVE) == 1]
    assert all(c[G.OPT_PATCH_CELLS] == 64 and c[G.OPT_STEP_LANES] == 4 for c in il)
    assert {c[G.OPT_IL_ORDER] for c in il} == {0, 1, 2, 4}
    assert {c[G.OPT_IL_RECORDS] for c in il} == {0, 1}
    assert len({tuple(sorted(c.items())) for c in cands}) == len(cands)  # no duplicates


def test_candidates_without_graphs_or_one_kernel():
    assert all(c[G.OPT_GRAPHS] == 0 for c in T.candidates(FakeContext(B=64), graphs=False))
    single = T.candidates(FakeContext(B=1))
    assert not any(c.get(G.OPT_INTERLEAVE) == 1 for c in single)  # one instance: no interleaving
    ctx = FakeContext(B=64, refuse={(G.OPT_ONE_KERNEL, 1)})
    ctx.opts[G.OPT_ONE_KERNEL] = 0
    assert T.candidates(ctx) == [{G.OPT_ONE_KERNEL: 0, G.OPT_GRAPHS: g, G.OPT_INTERLEAVE: 0} for g in (0, 1)]


def test_apply_skips_refused_options_and_names_round_trip():
    ctx = FakeContext(B=32, refuse={(G.OPT_IL_ORDER, 4)})
    done = T.apply(ctx, {G.OPT_INTERLEAVE: 1, G.OPT_IL_ORDER: 4, G.OPT_IL_UNROLL: 2})
    assert done == {G.OPT_INTERLEAVE: 1, G.OPT_IL_UNROLL: 2}
    assert ctx.option(G.OPT_IL_ORDER) == 0
    named = {T.NAMES[o]: v for o, v in done.items()}
    assert T.option_ids(named) == done
    assert set(T.NAMES) == set(T.TUNED)


@pytest.mark.parametrize("name", ["il_order", "il_records", "graphs"])
def test_option_names_are_abi_options(name):
    ids = T.option_ids({name: 0})
    (o,) = ids
    assert o in T.TUNED and 1 <= o <= 12


class FakeTuneContext(FakeContext):
    """A scratch context as tune() builds one: (mesh, params, precision),
    set_sh, the interleaved layout active when asked for with 16 instances
    or more, close.  Counts the contexts made."""
    made = 0

    def __init__(self, mesh, params, precision=0):
        super().__init__(B=len(params))
        FakeTuneContext.made += 1
        self.sh = None

    def set_sh(self, L, beta, mode=0):
        self.sh = (L, mode)

    def option(self, o):
        if o == G.OPT_INTERLEAVE_ACTIVE:
            return int(self.opts[G.OPT_INTERLEAVE] == 1 and self.B >= 16)
        return self.opts[o]

    def close(self):
        pass


class _Mesh:
    N = 40962


def _timer(table):
    """ms per step of a context's options: the first matching rule wins."""
    def time_fn(ctx, steps, warmup, repeats):
        for cond, ms in table:
            if all(ctx.opts.get(o) == v for o, v in cond.items()):
                return ms * 1e-3
        return 1e-3
    return time_fn


@pytest.fixture
def fake_tuner(monkeypatch, tmp_path):
    monkeypatch.setattr(T, "Context", FakeTuneContext)
    monkeypatch.setattr(T, "_MEMO", {})
    monkeypatch.setenv("ODIS_TUNE_CACHE", str(tmp_path / "tune.json"))
    FakeTuneContext.made = 0
    return tmp_path / "tune.json"


def test_tune_keeps_the_default_within_the_margin(fake_tuner, monkeypatch):
    # graphs 1% faster: under the 2% margin, so the default stays
    monkeypatch.setattr(T, "_time", _timer([({G.OPT_GRAPHS: 1}, 0.99)]))
    r = T.tune(_Mesh, [object()] * 4, margin=0.02, use_cache=False)
    assert r["options"] == r["default"] and r["speedup"] == 1.0 and not r["cached"]
    # 4 instances: the interleaved candidates are not active and not timed
    assert r["candidates"] == 2 * (1 + 6)


def test_tune_picks_the_fastest_valid_candidate(fake_tuner, monkeypatch):
    table = [({G.OPT_INTERLEAVE: 1, G.OPT_IL_ORDER: 2, G.OPT_IL_RECORDS: 1, G.OPT_GRAPHS: 0}, 0.80),
             ({G.OPT_PATCH_CELLS: 128, G.OPT_STEP_LANES: 2}, 0.90)]
    monkeypatch.setattr(T, "_time", _timer(table))
    # 32 instances: the interleaved layout is active and wins
    r = T.tune(_Mesh, [object()] * 32, use_cache=False)
    assert r["options"]["interleave"] == 1 and r["options"]["il_order"] == 2
    assert r["options"]["il_records"] == 1 and r["options"]["graphs"] == 0
    assert abs(r["speedup"] - 1.25) < 1e-12
    # 8 instances: the interleaved layout would not be active, so those
    # candidates are skipped and the 128-cell, 2-lane step wins
    r = T.tune(_Mesh, [object()] * 8, use_cache=False)
    assert r["options"]["interleave"] == 0
    assert (r["options"]["patch_cells"], r["options"]["step_lanes"]) == (128, 2)
    assert r["candidates"] == 2 * (1 + 6)


def test_tune_cache_in_memory_and_on_disk(fake_tuner, monkeypatch):
    monkeypatch.setattr(T, "_time", _timer([({G.OPT_ONE_KERNEL: 0, G.OPT_GRAPHS: 1}, 0.5)]))
    r1 = T.tune(_Mesh, [object()] * 2, graphs=False)
    assert not r1["cached"] and r1["options"]["graphs"] == 0  # graph replay left out
    made = FakeTuneContext.made
    r2 = T.tune(_Mesh, [object()] * 2, graphs=False)
    assert r2["cached"] and r2["options"] == r1["options"] and FakeTuneContext.made == made
    disk = json.loads(fake_tuner.read_text())
    (key,) = disk
    assert key.startswith("40962/2/") and key.endswith("/no-graphs")
    # a fresh process (empty memo) reads the file; with graphs the key differs
    monkeypatch.setattr(T, "_MEMO", {})
    assert T.tune(_Mesh, [object()] * 2, graphs=False)["cached"]
    r3 = T.tune(_Mesh, [object()] * 2)
    assert not r3["cached"] and r3["options"]["one_kernel"] == 0 and r3["options"]["graphs"] == 1
    assert len(json.loads(fake_tuner.read_text())) == 2


def test_tune_cache_survives_a_corrupt_file(fake_tuner, monkeypatch):
    fake_tuner.write_text("{not json")
    monkeypatch.setattr(T, "_time", _timer([]))
    r = T.tune(_Mesh, [object()])
    assert not r["cached"]
    assert len(json.loads(fake_tuner.read_text())) == 1
