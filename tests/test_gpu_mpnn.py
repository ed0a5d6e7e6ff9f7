"""GPU: the warm-start network's HIP forward, backward and Adam kernels
against the CPU oracle (oracle/mpnn.py, itself pinned by finite
differences).  FP32 on the device with MFMA accumulation against FP64 on the
host: agreement to FP32 rounding, stated per test."""
import numpy as np
import pytest

import mpnn
import phenlei_amd as pa

pytestmark = pytest.mark.gpu


def mesh_neighbours(mesh):
    c = mesh.connectivity()
    nb = mesh.n_bface
    lists = [[] for _ in range(mesh.n_cell)]
    for a, b in zip(c["left"][nb:], c["right"][nb:]):
        lists[a].append(b)
        lists[b].append(a)
    w = max(len(x) for x in lists)
    out = -np.ones((mesh.n_cell, w), np.int64)
    for i, x in enumerate(lists):
        out[i, :len(x)] = x
    return out


def problem(seed=0):
    mesh = pa.Mesh.bump_hex(10, 6, 4, 0.05)  # 240 cells: two full 128-row tiles
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((mesh.n_cell, 8)).astype(np.float32)
    T = (0.3 * rng.standard_normal((mesh.n_cell, 5))).astype(np.float32)
    net = pa.Mpnn(layers=6, seed=seed)
    p = net.get_params()
    for k in ("b0", "c", "bo"):  # non-zero biases exercise the bias paths
        p[k] = (0.05 * rng.standard_normal(p[k].shape)).astype(np.float32)
    net.set_params(p)
    net.set_graph(mesh=mesh)
    net.set_data(X, T)
    return mesh, net, {k: v.astype(np.float64) for k, v in p.items()}, X.astype(np.float64), T.astype(np.float64)


def test_forward_matches_oracle():
    mesh, net, p, X, T = problem()
    nbr = mesh_neighbours(mesh)
    y = net.forward()
    yr = mpnn.forward(p, X, nbr)
    err = np.max(np.abs(y - yr)) / np.max(np.abs(yr))
    print("forward rel err", err)
    assert err <= 1e-5


def test_bf16_forward_matches_oracle():
    """The BF16 inference form (hidden-layer products from BF16 operands,
    FP32 accumulation): within 2e-3 of the largest output of the oracle with
    the same operands rounded to bfloat16 (the rest is the FP32 activations'
    rounding before the BF16 rounding), and within 1e-2 of the FP64 network;
    the FP32 form is unchanged by the switch."""
    mesh, net, p, X, T = problem(seed=2)
    nbr = mesh_neighbours(mesh)
    y32 = net.forward()
    net.set_forward_precision(True)
    y16 = net.forward()
    net.set_forward_precision(False)
    assert np.array_equal(net.forward(), y32)
    yb = mpnn.forward(p, X, nbr, bf16_hidden=True)
    yr = mpnn.forward(p, X, nbr)
    sc = np.max(np.abs(yr))
    eb, er = np.max(np.abs(y16 - yb)) / sc, np.max(np.abs(y16 - yr)) / sc
    print("bf16 forward: vs bf16 oracle", eb, "vs fp64 oracle", er)
    assert eb <= 2e-3 and er <= 1e-2
    assert not np.array_equal(y16, y32)


def torus(nx, ny):
    """Every node has exactly four neighbours (periodic grid), so a
    neighbour mean divides by 4 exactly."""
    i = np.arange(nx * ny)
    x, y = i % nx, i // nx
    return np.stack([(x + 1) % nx + nx * y, (x - 1) % nx + nx * y, x + nx * ((y + 1) % ny),
                     x + nx * ((y - 1) % ny)], axis=1)


def test_bf16_forward_is_exact_on_dyadic_values():
    """Pins the BF16 MFMA operand layout: with features in {0, 1} and
    weights in {0, +-1/4, +-1/2, +-1}, every activation is a short dyadic
    number that BF16 holds exactly and every sum is exact in FP32, so one
    hidden layer's BF16 form, its FP32 form and the FP64 oracle agree bit for
    bit; a misplaced row, column or k of either operand would not.  2,010
    nodes: 15 full 128-row tiles and a partial one."""
    nx, ny = 30, 67
    nbr = torus(nx, ny)
    n = nx * ny
    rng = np.random.default_rng(7)
    net = pa.Mpnn(layers=1, seed=0)
    p = net.get_params()
    p["W0"] = rng.choice([0.0, 0.5, -0.5, 1.0, -1.0], p["W0"].shape).astype(np.float32)
    p["A"] = rng.choice([0.0, 0.25, -0.25, 0.5, -0.5], p["A"].shape).astype(np.float32)
    p["B"] = rng.choice([0.0, 0.25, -0.25, 0.5, -0.5], p["B"].shape).astype(np.float32)
    p["Wo"] = rng.choice([0.0, 0.5, -0.5], p["Wo"].shape).astype(np.float32)
    for k in ("b0", "c", "bo"):
        p[k] = np.zeros_like(p[k])
    net.set_params(p)
    net.set_graph(nbr.astype(np.int32))
    X = rng.integers(0, 2, (n, 8)).astype(np.float32)
    net.set_data(X)
    y32 = net.forward()
    net.set_forward_precision(True)
    y16 = net.forward()
    yr = mpnn.forward({k: v.astype(np.float64) for k, v in p.items()}, X.astype(np.float64), nbr)
    assert np.abs(yr).max() > 10.0  # the sums are not trivial
    assert np.array_equal(y32.astype(np.float64), yr)
    assert np.array_equal(y16, y32)


def test_gradient_matches_oracle():
    mesh, net, p, X, T = problem(seed=1)
    nbr = mesh_neighbours(mesh)
    loss, g = net.gradient()
    lr, gr = mpnn.backward(p, X, nbr, T)
    assert loss == pytest.approx(lr, rel=1e-5)
    for k in gr:
        err = np.max(np.abs(g[k] - gr[k])) / max(np.max(np.abs(gr[k])), 1e-30)
        print("grad", k, "rel err", err)
        assert err <= 1e-4, k


@pytest.mark.parametrize("w,pairs", [(5, 500), (12, 1200)])
def test_ragged_explicit_graph(w, pairs):
    """Explicit neighbour lists with isolated cells and uneven degree; width
    12 takes the neighbour kernels' path for lists wider than 8."""
    rng = np.random.default_rng(4)
    n = 300
    nbr = -np.ones((n, w), np.int64)
    deg = np.zeros(n, int)
    for _ in range(pairs):
        i, j = rng.integers(0, n, 2)
        if i != j and deg[i] < w and deg[j] < w and j not in nbr[i]:
            nbr[i, deg[i]] = j
            nbr[j, deg[j]] = i
            deg[i] += 1
            deg[j] += 1
    net = pa.Mpnn(layers=2, seed=3)
    net.set_graph(nbr=nbr)
    X = rng.standard_normal((n, 8)).astype(np.float32)
    T = rng.standard_normal((n, 5)).astype(np.float32)
    net.set_data(X, T)
    p = {k: v.astype(np.float64) for k, v in net.get_params().items()}
    loss, g = net.gradient()
    lr, gr = mpnn.backward(p, X.astype(np.float64), nbr, T.astype(np.float64))
    assert loss == pytest.approx(lr, rel=1e-5)
    for k in gr:
        assert np.max(np.abs(g[k] - gr[k])) <= 1e-4 * max(np.max(np.abs(gr[k])), 1e-30), k


def test_adam_training_tracks_oracle_and_learns():
    mesh, net, p, X, T = problem(seed=2)
    nbr = mesh_neighbours(mesh)
    hist = net.train(3, 1e-3)
    st = {}
    q = {k: v.copy() for k, v in p.items()}
    ref = []
    for _ in range(3):
        l, g = mpnn.backward(q, X, nbr, T)
        ref.append(l)
        mpnn.adam_step(q, g, st, 1e-3)
        q = {k: v.astype(np.float64) for k, v in q.items()}
    assert np.allclose(hist, ref, rtol=1e-5)
    # Adam steps by about lr * sign(g); components whose gradient is at FP32
    # noise level may step the other way, so compare the updates in norm
    got = net.get_params()
    for k in q:
        d_gpu, d_ref = got[k] - p[k], q[k] - p[k]
        assert np.linalg.norm(d_gpu - d_ref) <= 0.02 * np.linalg.norm(d_ref) + 1e-7, k
    more = net.train(300, 1e-3)
    print("loss", hist[0], "->", more[-1])
    assert more[-1] < 0.3 * hist[0]


def test_set_graph_rejects_asymmetric_lists():
    """The backward pass transposes the neighbour mean as a gather over each
    node's own list, so lists that are not symmetric are an argument error."""
    net = pa.Mpnn(layers=1, seed=0, device=0)
    nbr = -np.ones((6, 2), np.int32)
    for i in range(1, 5):
        nbr[i, 0], nbr[i + 1, 1] = i + 1, i
    net.set_graph(nbr)  # a symmetric chain
    nbr[0, 0] = 5  # 0 lists 5, 5 does not list 0
    with pytest.raises(pa.PhlError) as e:
        net.set_graph(nbr)
    assert e.value.status == pa.PHL_ERR_ARG
    net.close()
