"""Pins of the numpy network used as the checker of the GPU learned guess:
its analytic gradient against central finite differences (smooth variant,
no bf16 rounding), and the neighbour-mean transpose against the adjoint
identity."""
import numpy as np

from oracle import mlp


def random_graph(rng, n=40):
    # symmetric neighbour lists of a ring lattice with a few chords
    nb = [set() for _ in range(n)]
    for i in range(n):
        for d in (1, 2, 7):
            j = (i + d) % n
            nb[i].add(j)
            nb[j].add(i)
    off = np.zeros(n + 1, np.int64)
    idx = []
    for i in range(n):
        s = sorted(nb[i])
        idx += s
        off[i + 1] = off[i] + len(s)
    return off, np.array(idx)


def init_params(rng):
    p = np.zeros(mlp.NPARAMS)
    for k, (o, s) in mlp.LAYOUT.items():
        if k.startswith("w"):
            p[o:o + s] = 0.3 * rng.standard_normal(s)
        else:
            p[o:o + s] = 0.1 * rng.standard_normal(s)
    o, s = mlp.LAYOUT["w3"]
    p[o + 5 * 64:o + s] = 0
    o, s = mlp.LAYOUT["b3"]
    p[o + 5:o + s] = 0
    return p


def test_agg_transpose_is_adjoint():
    rng = np.random.default_rng(0)
    off, idx = random_graph(rng)
    h = rng.standard_normal((40, 64))
    d = rng.standard_normal((40, 64))
    lhs = (mlp.agg(h, off, idx) * d).sum()
    rhs = (h * mlp.agg_t(d, off, idx)).sum()
    assert np.isclose(lhs, rhs, rtol=1e-12)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    off, idx = random_graph(rng)
    x = np.zeros((40, 32))
    x[:, :19] = rng.standard_normal((40, 19))
    y = rng.standard_normal((40, 8))
    p = init_params(rng)
    loss, grad, _ = mlp.loss_and_grad(p, x, y, off, idx, round_bf16=False)
    picks = []
    for k, (o, s) in mlp.LAYOUT.items():
        n = s if k != "w3" and k != "b3" else (5 * 64 if k == "w3" else 5)
        picks += list(o + rng.choice(n, size=min(6, n), replace=False))
    for i in picks:
        h = 1e-6
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        lp, _, _ = mlp.loss_and_grad(pp, x, y, off, idx, round_bf16=False)
        lm, _, _ = mlp.loss_and_grad(pm, x, y, off, idx, round_bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(grad[i])) + 1e-9, (i, fd, grad[i])


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], np.float32)
    r = mlp.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7        # above the tie rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2 ** -8 * 3.0e-3  # half a bf16 ulp
