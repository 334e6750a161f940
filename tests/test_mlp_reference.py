"""The CPU reference of the scorer's ranking loss against finite differences."""
import numpy as np

from test_gpu_mlp import rank_loss


def test_rank_loss_gradient_by_finite_differences():
    rng = np.random.default_rng(4)
    G = 8
    y = rng.normal(0, 2, 4 * G)
    t = rng.integers(0, 5, 4 * G).astype(float)
    _, dy = rank_loss(y, t, G)
    for i in range(len(y)):
        e = np.zeros(len(y))
        e[i] = 1e-6
        fd = (rank_loss(y + e, t, G)[0] - rank_loss(y - e, t, G)[0]) / 2e-6
        assert abs(fd - dy[i]) <= 1e-6 + 1e-5 * abs(dy[i])


def test_rank_loss_ignores_ties_and_orders_pairs():
    # Two rows, the first closer: loss softplus(y0 - y1); swapping the scores raises it.
    t = np.array([1.0, 2.0])
    low, _ = rank_loss(np.array([-3.0, 3.0]), t, 2)
    high, _ = rank_loss(np.array([3.0, -3.0]), t, 2)
    assert low < 0.01 < 5.9 < high
    tied, dy = rank_loss(np.array([0.3, -0.2]), np.array([2.0, 2.0]), 2)
    assert tied == 0 and not dy.any()
