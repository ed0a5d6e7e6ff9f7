"""Pins the CPU oracle of the warm-start network: its hand-written backward
pass against central finite differences (float64, small random graphs), and
the mean aggregation against a per-cell loop."""
import numpy as np
import pytest

import mpnn


def small_problem(n=30, w=4, hidden=6, layers=2, seed=3):
    rng = np.random.default_rng(seed)
    p = mpnn.init_params(hidden=hidden, layers=layers, seed=seed)
    p = {k: v.astype(np.float64) for k, v in p.items()}
    for k in ("b0", "c", "bo"):
        p[k] = rng.standard_normal(p[k].shape) * 0.1
    # symmetric neighbour lists with ragged degree (some cells isolated)
    pairs = set()
    while len(pairs) < n:
        i, j = rng.integers(0, n, 2)
        if i != j:
            pairs.add((min(i, j), max(i, j)))
    nb = [[] for _ in range(n)]
    for i, j in sorted(pairs):
        if len(nb[i]) < w and len(nb[j]) < w:
            nb[i].append(j)
            nb[j].append(i)
    nbr = -np.ones((n, w), np.int64)
    for i, lst in enumerate(nb):
        nbr[i, :len(lst)] = lst
    X = rng.standard_normal((n, mpnn.F_IN))
    t = rng.standard_normal((n, mpnn.F_OUT)) * 0.3
    return p, X, nbr, t


def test_neighbour_mean_matches_loop():
    p, X, nbr, t = small_problem()
    h = np.random.default_rng(1).standard_normal((X.shape[0], 5))
    m, deg = mpnn.neighbour_mean(h, nbr)
    for i in range(X.shape[0]):
        js = [j for j in nbr[i] if j >= 0]
        ref = np.mean(h[js], axis=0) if js else np.zeros(5)
        assert np.allclose(m[i], ref, rtol=1e-14, atol=1e-15)
        assert deg[i] == len(js)


@pytest.mark.parametrize("key", ["W0", "b0", "A", "B", "c", "Wo", "bo"])
def test_backward_matches_finite_differences(key):
    p, X, nbr, t = small_problem()
    _, g = mpnn.backward(p, X, nbr, t)
    rng = np.random.default_rng(7)
    flat = p[key].reshape(-1)
    picks = rng.choice(flat.size, size=min(12, flat.size), replace=False)
    eps = 1e-6
    for i in picks:
        old = flat[i]
        flat[i] = old + eps
        lp = mpnn.loss(mpnn.forward(p, X, nbr), t)
        flat[i] = old - eps
        lm = mpnn.loss(mpnn.forward(p, X, nbr), t)
        flat[i] = old
        fd = (lp - lm) / (2 * eps)
        an = g[key].reshape(-1)[i]
        assert abs(fd - an) <= 1e-7 + 1e-5 * abs(fd), (key, i, fd, an)


def test_adam_reduces_the_loss():
    p, X, nbr, t = small_problem(n=40, hidden=16, layers=2)
    st = {}
    l0, _ = mpnn.backward(p, X, nbr, t)
    for _ in range(200):
        _, g = mpnn.backward(p, X, nbr, t)
        mpnn.adam_step(p, g, st, 1e-2)
    l1, _ = mpnn.backward(p, X, nbr, t)
    assert l1 < 0.5 * l0


def _scatter_transpose(g, nbr):
    """Explicit transpose of the neighbour mean: out[j] += g[i] / deg[i] for
    every j in i's list (any lists, symmetric or not)."""
    valid = nbr >= 0
    deg = np.maximum(valid.sum(axis=1), 1)
    out = np.zeros_like(g)
    for i in range(nbr.shape[0]):
        for j in nbr[i]:
            if j >= 0:
                out[j] += g[i] / deg[i]
    return out


def test_mean_transpose_is_the_adjoint_on_symmetric_lists():
    """<mean(h), g> == <h, mean^T(g)>, and the gather form of the transpose
    (the one the device kernel uses) equals the explicit scatter."""
    _, X, nbr, _ = small_problem()
    rng = np.random.default_rng(4)
    h = rng.standard_normal((X.shape[0], 7))
    g = rng.standard_normal((X.shape[0], 7))
    m, _ = mpnn.neighbour_mean(h, nbr)
    mt = mpnn.neighbour_mean_transpose(g, nbr)
    assert np.allclose(mt, _scatter_transpose(g, nbr), rtol=0, atol=1e-13)
    assert abs(np.sum(m * g) - np.sum(h * mt)) <= 1e-12 * np.sum(np.abs(m * g))


def test_gather_transpose_needs_symmetric_lists():
    """With one asymmetric entry (cell 0 lists 5, cell 5 does not list 0) the
    gather form is no longer the transpose: the reason phl_mpnn_set_graph
    rejects lists that are not symmetric."""
    nbr = -np.ones((6, 2), np.int64)
    nbr[0, 0] = 5
    for i in range(1, 5):
        nbr[i, 0], nbr[i + 1, 1] = i + 1, i
    g = np.random.default_rng(2).standard_normal((6, 3))
    assert not np.allclose(mpnn.neighbour_mean_transpose(g, nbr), _scatter_transpose(g, nbr))


def test_bf16_rounding_is_nearest_even():
    """mpnn.bf16 against a brute-force nearest search over the bfloat16
    values around each input (8 significant bits; ties to the even
    mantissa), on values spread over many binades, exact bfloat16 values and
    exact ties."""
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-6, 6, 2000),
                        [1.0, -2.0, 0.375, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(3.0 + 2.0 ** -7), 0.0]])
    x = x.astype(np.float32).astype(np.float64)
    got = mpnn.bf16(x)
    for v, g in zip(x, got):
        if v == 0.0:
            assert g == 0.0
            continue
        e = np.floor(np.log2(abs(v)))
        ulp = 2.0 ** (e - 7)
        lo = np.floor(v / ulp) * ulp
        cands = [lo, lo + ulp]
        d = [abs(v - c) for c in cands]
        if d[0] == d[1]:  # a tie: the candidate with the even last mantissa bit
            want = cands[0] if int(round(cands[0] / ulp)) % 2 == 0 else cands[1]
        else:
            want = cands[int(np.argmin(d))]
        assert g == want, (v, g, want)
