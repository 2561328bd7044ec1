"""CPU oracle of the learned edge prior.  TEST INFRASTRUCTURE: only tests use
it, as the checker of the GPU network kernels (aco_prior.hip), never as the
product path.

The network has no reference counterpart (it replaces eta_beta at the seam of
aco_cuda_kernels.cu:295).  This restates its arithmetic:

* features of candidate slot (i, t) in float32 operation order, rounded to
  BF16 (round to nearest even);
* forward: z1 = x W1^T + b1, h1 = bf16(relu(z1)), z2 = h1 W2^T + b2,
  h2 = relu(z2), z3 = h2 . w3 + b3 with W1, W2 rounded to BF16;
* gradient of mean BCE(sigmoid(z3), y): the products that run on the
  matrix cores take BF16 operands (dZ2 and H1 for dW2, dZ2 and W2 for dH1,
  dZ1 and X for dW1), everything else stays in float.

Sums are taken in float64 here (the GPU sums in float32 in another order), so
comparisons carry a tolerance.  `exact=True` drops every BF16 rounding: that
model is what the finite-difference test differentiates.
"""
import numpy as np

F, H = 8, 64
PW1, PB1, PW2, PB2, PW3, PB3, NPAR = 0, 512, 576, 4672, 4736, 4800, 4801


def bf16_bits(a):
    """float32 -> BF16 bits (uint16), round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_float(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16(a):
    return bf16_float(bf16_bits(a))


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[PW1:PB1].reshape(H, F), p[PB1:PW2], p[PW2:PB2].reshape(H, H), p[PB2:PW3],
            p[PW3:PB3], p[PB3])


def features(x, y, dem, Q, cand, dist, k):
    """BF16 feature bits [(n+1)*k, 8] of every candidate slot (float32 steps
    as the device kernel; the cosine in float64, rounded once)."""
    x = np.asarray(x, np.int64)
    y = np.asarray(y, np.int64)
    dem = np.asarray(dem, np.int64)
    n1 = len(x)
    L = np.float32(max(1, int(max(x.max() - x.min(), y.max() - y.min()))))
    cand = np.asarray(cand, np.int64).reshape(n1, k)
    dist = np.asarray(dist, np.int64).reshape(n1, k)
    i = np.repeat(np.arange(n1), k)
    t = np.tile(np.arange(k), n1)
    j = cand.reshape(-1)
    f = np.zeros((n1 * k, F), np.float32)
    f[:, 0] = (x[i] - x[0]).astype(np.float32) / L
    f[:, 1] = (y[i] - y[0]).astype(np.float32) / L
    f[:, 2] = (x[j] - x[0]).astype(np.float32) / L
    f[:, 3] = (y[j] - y[0]).astype(np.float32) / L
    f[:, 4] = dem[j].astype(np.float32) / np.float32(Q)
    f[:, 5] = dist.reshape(-1).astype(np.float32) / L
    ux, uy = x[j] - x[i], y[j] - y[i]
    vx, vy = x[0] - x[i], y[0] - y[i]
    nu, nv = ux * ux + uy * uy, vx * vx + vy * vy
    ok = (nu > 0) & (nv > 0)
    cosv = np.zeros(n1 * k, np.float64)
    cosv[ok] = (ux * vx + uy * vy)[ok].astype(np.float64) / (
        np.sqrt(nu[ok].astype(np.float64)) * np.sqrt(nv[ok].astype(np.float64)))
    f[:, 6] = cosv.astype(np.float32)
    f[:, 7] = t.astype(np.float32) / np.float32(k)
    return bf16_bits(f)


def _r(a, exact):
    return a if exact else bf16(np.asarray(a, np.float32)).astype(np.float64)


def forward(params, xbits, exact=False, return_all=False):
    W1, b1, W2, b2, w3, b3 = unpack(params)
    x = xbits.astype(np.float64) if exact else bf16_float(xbits).astype(np.float64)
    W1e, W2e = _r(W1, exact), _r(W2, exact)
    z1 = x @ W1e.T + b1
    h1 = _r(np.maximum(z1, 0.0), exact)
    z2 = h1 @ W2e.T + b2
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ w3 + b3
    if return_all:
        return x, W1e, W2e, z1, h1, z2, h2, z3
    return z3


def loss_and_grad(params, xbits, labels, exact=False):
    """Mean BCE and its gradient (NPAR,) in the device's rounding model."""
    x, W1e, W2e, z1, h1, z2, h2, z3 = forward(params, xbits, exact, return_all=True)
    W1, b1, W2, b2, w3, b3 = unpack(params)
    y = np.asarray(labels, np.float64)
    E = len(y)
    loss = np.mean(np.maximum(z3, 0) - z3 * y + np.log1p(np.exp(-np.abs(z3))))
    dz3 = (1.0 / (1.0 + np.exp(-z3)) - y) / E
    gw3 = h2.T @ dz3
    gb3 = dz3.sum()
    dz2 = np.outer(dz3, w3) * (z2 > 0)
    gb2 = dz2.sum(0)
    dz2r = _r(dz2, exact)
    gW2 = dz2r.T @ h1
    dh1 = dz2r @ W2e
    dz1 = dh1 * (z1 > 0)
    gb1 = dz1.sum(0)
    gW1 = _r(dz1, exact).T @ x
    g = np.zeros(NPAR, np.float64)
    g[PW1:PB1] = gW1.reshape(-1)
    g[PB1:PW2] = gb1
    g[PW2:PB2] = gW2.reshape(-1)
    g[PB2:PW3] = gb2
    g[PW3:PB3] = gw3
    g[PB3] = gb3
    return loss, g


def labels_from_tour(cand, n1, k, tour):
    """1 where candidate j of node i is consecutive to i in the tour."""
    cand = np.asarray(cand, np.int64).reshape(n1, k)
    adj = [set() for _ in range(n1)]
    for a, b in zip(tour[:-1], tour[1:]):
        adj[a].add(b)
        adj[b].add(a)
    return np.array([[1 if cand[i, t] in adj[i] else 0 for t in range(k)] for i in range(n1)],
                    np.uint8).reshape(-1)
