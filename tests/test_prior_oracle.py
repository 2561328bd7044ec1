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


def test_features_hand_computed_rows():
    """A 3-node instance worked by hand: depot (10,20), customer 1 (13,24)
    with demand 3, customer 2 (10,30) with demand 5, Q = 8, k = 2, so the
    bounding-box scale is L = max(13-10, 30-20) = 10.  Candidate rows:
    depot -> [1, 2] (nint 5, 10), 1 -> [2, pad 0] (nint 7, 0),
    2 -> [1, pad 0] (nint 7, 0).  Expected features (before BF16 rounding):
      slot (1,0), j=2: (0.3, 0.4, 0, 1, 5/8, 0.7, cos, 0) with cos of the
        angle at 1 between (-3,6) and (-3,-4) = -15/(sqrt(45)*5) = -1/sqrt(5);
      slot (0,1), j=2: (0, 0, 0, 1, 5/8, 1.0, 0 (i is the depot), 1/2);
      slot (2,0), j=1: (0, 1, 0.3, 0.4, 3/8, 0.7, 60/(sqrt(45)*10) = 2/sqrt(5), 0)."""
    x, y, dem, Q, k = [10, 13, 10], [20, 24, 30], [0, 3, 5], 8, 2
    cand = [[1, 2], [2, 0], [1, 0]]
    dist = [[5, 10], [7, 0], [7, 0]]
    f = M.bf16_float(M.features(x, y, dem, Q, cand, dist, k)).reshape(3, k, M.F)
    r5 = 5.0 ** 0.5
    want = {(1, 0): [0.3, 0.4, 0.0, 1.0, 0.625, 0.7, -1.0 / r5, 0.0],
            (0, 1): [0.0, 0.0, 0.0, 1.0, 0.625, 1.0, 0.0, 0.5],
            (2, 0): [0.0, 1.0, 0.3, 0.4, 0.375, 0.7, 2.0 / r5, 0.0]}
    for (i, t), w in want.items():
        w = np.array(w)
        got = f[i, t].astype(np.float64)
        # BF16 keeps 8 significant bits: round to nearest is within 2^-8 relative
        assert np.all(np.abs(got - w) <= np.abs(w) * 2.0 ** -8), (i, t, got, w)
        exact = np.isin(w, [0.0, 0.375, 0.5, 0.625, 1.0])
        assert np.array_equal(got[exact], w[exact]), (i, t)
