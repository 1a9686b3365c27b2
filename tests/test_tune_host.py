"""The autotuner's host logic without a GPU: which launch options it
proposes for a context, how it applies a choice, and that option names map
back to the C ABI's ids.  A stand-in context records the options set on it
and refuses the ones a real context would refuse."""
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
