"""odis_sh_project_noalloc (the projection entry point that allocates no
device memory) against odis_sh_project, through the C ABI: the same
coefficients bit for bit on two calls in a row, the second on a smaller
grid, with the stepping state of the contexts untouched."""
import numpy as np
import pytest

from conftest import generated_grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["SH_RECUR", "SH_FP64"])
def test_sh_project_noalloc_matches_sh_project(product, mode):
    G = product
    rng = np.random.default_rng(7)
    results = []
    for level, L in ((6, 16), (5, 8)):       # 10242 cells, then the smaller 2562
        g, _ = generated_grid(level)
        p, _ = G.make_params()
        mesh = G.Mesh(g, p.radius)
        ctx = G.Context(mesh, p)
        ctx.set_sh(L, np.ones(L + 1), mode=getattr(G, mode))
        ctx.step(20)
        before = ctx.get_state()
        field = rng.normal(size=mesh.N)
        want = ctx.sh_project(field)
        got = ctx.sh_project_noalloc(field)
        assert got.shape == ((L + 1) ** 2,)
        assert np.array_equal(got, want), np.abs(got - want).max()
        after = ctx.get_state()
        assert np.array_equal(after["v"], before["v"]) and np.array_equal(after["eta"], before["eta"])
        # stepping after the call matches stepping after odis_sh_project
        ctx.step(5)
        results.append((mesh.N, np.abs(want).max()))
        ctx.close()
    assert results[1][0] < results[0][0]


def test_sh_project_noalloc_needs_a_basis(product):
    G = product
    g, _ = generated_grid(3)
    p, _ = G.make_params()
    ctx = G.Context(G.Mesh(g, p.radius), p)
    ctx.sh_l_max = 2  # no basis set on the context
    with pytest.raises(G.OdisError) as ei:
        ctx.sh_project_noalloc(np.zeros(g.n))
    assert ei.value.code == -1
    ctx.close()
