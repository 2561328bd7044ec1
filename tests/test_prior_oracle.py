"""CPU pins of the learned-prior oracle (oracle/mlp_oracle.py): the backprop
algebra against central finite differences, BF16 rounding against its
definition, and the elite-tour labels against a direct restatement."""
import numpy as np

import mlp_oracle as M


def _params(rng, scale=0.3):
    return (rng.standard_normal(M.NPAR) * scale).astype(np.float64)


def test_bf16_rounding_is_nearest_even():
    # BF16 keeps 7 stored mantissa bits: the spacing at 1.0 is 2^-7
    vals = np.array([1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 3.0e-20],
                    np.float32)
    got = M.bf16(vals)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0);
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to 1 + 2^-6
    want = np.array([1.0, 1.0 + 2 ** -7, 1.0, 1.0 + 2 ** -6, -2.5, 0.0], np.float32)
    assert np.array_equal(got[:6], want)
    rel = abs(float(got[6]) - float(vals[6])) / float(vals[6])
    assert rel <= 2 ** -8


def test_gradient_matches_finite_differences():
    """exact=True drops the BF16 rounding; the analytic gradient of the mean
    BCE must match central differences (pins the backward formulas)."""
    rng = np.random.default_rng(3)
    p = _params(rng)
    x = rng.uniform(-1, 1, size=(24, M.F))
    y = (rng.random(24) < 0.3).astype(np.uint8)
    loss, g = M.loss_and_grad(p, x, y, exact=True)
    idx = np.concatenate([rng.choice(np.arange(M.PW1, M.PB1), 12),
                          rng.choice(np.arange(M.PB1, M.PW2), 6),
                          rng.choice(np.arange(M.PW2, M.PB2), 20),
                          rng.choice(np.arange(M.PB2, M.PW3), 6),
                          rng.choice(np.arange(M.PW3, M.PB3), 6), [M.PB3]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_and_grad(pp, x, y, exact=True)[0] - M.loss_and_grad(pm, x, y, exact=True)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_labels_from_tour():
    # nodes 0..5, k = 3; tour 0-1-2-0-3-0-0 (an empty route at the end)
    cand = np.array([[1, 3, 5], [2, 0, 4], [1, 3, 0], [0, 2, 5], [5, 1, 2], [4, 3, 0]])
    tour = [0, 1, 2, 0, 3, 0, 0]
    y = M.labels_from_tour(cand, 6, 3, tour).reshape(6, 3)
    want = np.array([[1, 1, 0],    # depot: 1 and 3 start/end routes, (0,0) is an edge too
                     [1, 1, 0],    # 1: neighbours 0 and 2
                     [1, 0, 1],    # 2: neighbours 1 and 0
                     [1, 0, 0],    # 3: neighbour 0 (twice)
                     [0, 0, 0],    # 4: unvisited
                     [0, 0, 0]])
    assert np.array_equal(y, want)
