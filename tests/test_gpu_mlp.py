"""The warm start's network on the GPU (odis_mlp_*, FP32 matrix-core
products) against the FP64 oracle network: forward pass, loss and gradient
at the real shape (5 inputs, 4 x 256 tanh, 866 outputs), and training."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SHAPE = (5, 256, 4, 866)


def _data(rng, n, shape=SHAPE):
    x = rng.uniform(-1, 1, size=(n, shape[0])).astype(np.float32)
    t = rng.normal(size=(n, shape[3])).astype(np.float32)
    return x, t


def test_mlp_forward_and_gradient_match_oracle(product, oracle):
    """Forward pass within 1e-5 and loss / gradient within 1e-5 / 1e-4
    (relative L2, per layer) of the FP64 oracle on the same FP32 weights,
    for 300 rows (not a multiple of the 64-row tiles) and for the rows of a
    shuffled index set."""
    G, O = product, oracle
    rng = np.random.default_rng(0)
    m = G.Mlp(*SHAPE, seed=11)
    p = m.params()
    assert p.size == 5 * 256 + 256 + 3 * (256 * 256 + 256) + 256 * 866 + 866
    x, t = _data(rng, 300)
    y = m.forward(x)
    want = O.mlp_forward(SHAPE, p.astype(np.float64), x.astype(np.float64))
    assert rel_l2(y, want) < 1e-5, rel_l2(y, want)
    m.set_data(x, t)
    idx = rng.permutation(300)[:217].astype(np.int32)
    loss, g = m.loss_grad(idx)
    lw, gw = O.mlp_loss_grad(SHAPE, p.astype(np.float64), x[idx].astype(np.float64), t[idx].astype(np.float64))
    assert abs(loss - lw) < 1e-5 * lw, (loss, lw)
    off = 0
    for l in range(SHAPE[2] + 1):
        ni = SHAPE[0] if l == 0 else SHAPE[1]
        no = SHAPE[3] if l == SHAPE[2] else SHAPE[1]
        for nm, n in (("W", ni * no), ("b", no)):
            e = rel_l2(g[off:off + n], gw[off:off + n])
            assert e < 1e-4, (l, nm, e)
            off += n
    m.close()


def test_mlp_training_fits_and_is_deterministic(product):
    """Adam on the GPU fits a smooth map (the loss falls by 100x), and two
    runs from the same seed end with the same weights bit for bit."""
    G = product
    rng = np.random.default_rng(2)
    shape = (5, 64, 2, 40)
    x = rng.uniform(-1, 1, size=(500, 5)).astype(np.float32)
    A = rng.normal(size=(5, 40))
    t = np.tanh(x @ A).astype(np.float32)
    runs = []
    for _ in range(2):
        m = G.Mlp(*shape, seed=3)
        m.set_data(x, t)
        loss = m.train(epochs=300, batch=100, lr_start=3e-3, lr_end=1e-5, seed=5)
        runs.append((loss, m.params()))
        m.close()
    assert runs[0][0][-1] < runs[0][0][0] / 100, (runs[0][0][0], runs[0][0][-1])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][0], runs[1][0])


def test_mlp_argument_errors(product):
    G = product
    with pytest.raises(G.OdisError):
        G.Mlp(0, 8, 1, 1)
    m = G.Mlp(2, 8, 1, 1)
    with pytest.raises(G.OdisError):
        m.loss_grad(np.array([0], np.int32))   # no data set
    m.set_data(np.zeros((3, 2), np.float32), np.zeros((3, 1), np.float32))
    with pytest.raises(G.OdisError):
        m.loss_grad(np.array([3], np.int32))
    m.close()
