"""CPU oracle of the warm-start message-passing network (test infrastructure).

Only tests/ and bench.py's CPU baseline use this file: it is the checker of
the HIP training and inference kernels in phenlei_amd/csrc/phl_mpnn.hip, never
part of the product path.  The reference has no learned component, so there
is no reference file to cite; the network is this project's own:

    h0   = relu(X W0 + b0)                      X: [N, 8] cell features
    for l = 1..L:
        m    = mean over face neighbours j of h_j  (interior faces only)
        z    = h A_l + m B_l + c_l
        h    = h + relu(z)
    y    = h Wo + bo                            y: [N, 5] predicted state

and the loss is the mean over cells and the 5 outputs of (y - t)^2.

The hand-written backward pass below is pinned by central finite differences
on small graphs (tests/test_mpnn_oracle.py).
"""
import numpy as np

F_IN, F_OUT = 8, 5


def init_params(hidden=128, layers=6, seed=0):
    """He-style initialisation, FP32, as the device model initialises."""
    rng = np.random.default_rng(seed)
    f = np.float32
    p = {
        "W0": (rng.standard_normal((F_IN, hidden)) * np.sqrt(2.0 / F_IN)).astype(f),
        "b0": np.zeros(hidden, f),
        "A": (rng.standard_normal((layers, hidden, hidden)) * np.sqrt(1.0 / hidden) * 0.5).astype(f),
        "B": (rng.standard_normal((layers, hidden, hidden)) * np.sqrt(1.0 / hidden) * 0.5).astype(f),
        "c": np.zeros((layers, hidden), f),
        "Wo": (rng.standard_normal((hidden, F_OUT)) * np.sqrt(1.0 / hidden)).astype(f),
        "bo": np.zeros(F_OUT, f),
    }
    return p


def neighbour_mean(h, nbr):
    """nbr: [N, W] neighbour indices, -1 padded.  Mean of h over the valid
    neighbours (zero for a cell without any)."""
    valid = nbr >= 0
    deg = valid.sum(axis=1)
    idx = np.where(valid, nbr, 0)
    s = (h[idx] * valid[..., None]).sum(axis=1)
    return s / np.maximum(deg, 1)[:, None], deg


def neighbour_mean_transpose(g, nbr):
    """Transpose of neighbour_mean applied to g: cell j receives g_i / deg_i
    from every cell i that lists j.  Written as a gather over j's own list,
    as the device kernel does, which equals the scatter only when the lists
    are symmetric with equal multiplicity (every face relation is)."""
    valid = nbr >= 0
    deg = np.maximum(valid.sum(axis=1), 1)
    idx = np.where(valid, nbr, 0)
    w = g / deg[:, None]
    return (w[idx] * valid[..., None]).sum(axis=1)


def bf16(a):
    """Round to bfloat16 (8 significant bits, round to nearest even) through
    FP32, as the device's v_cvt_pk_bf16_f32 does; returned as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, X, nbr, keep=False, bf16_hidden=False):
    """Returns y (and the activations needed by backward when keep).
    bf16_hidden: the hidden layers' products take h, m and A_l, B_l rounded
    to bfloat16 (the device's BF16 inference form)."""
    z0 = X @ p["W0"] + p["b0"]
    h = np.maximum(z0, 0)
    hs, ms, zs = [h], [], []
    rnd = bf16 if bf16_hidden else (lambda a: a)
    for l in range(p["A"].shape[0]):
        m, _ = neighbour_mean(h, nbr)
        z = rnd(h) @ rnd(p["A"][l]) + rnd(m) @ rnd(p["B"][l]) + p["c"][l]
        h = h + np.maximum(z, 0)
        ms.append(m)
        zs.append(z)
        hs.append(h)
    y = h @ p["Wo"] + p["bo"]
    if keep:
        return y, dict(z0=z0, hs=hs, ms=ms, zs=zs)
    return y


def loss(y, t):
    return float(np.mean((y - t) ** 2))


def backward(p, X, nbr, t):
    """Loss and gradients of every parameter (same keys as p)."""
    y, a = forward(p, X, nbr, keep=True)
    n = y.shape[0]
    dy = 2.0 * (y - t) / (n * y.shape[1])
    g = {}
    h = a["hs"][-1]
    g["Wo"] = h.T @ dy
    g["bo"] = dy.sum(axis=0)
    dh = dy @ p["Wo"].T
    L = p["A"].shape[0]
    gA, gB, gc = np.zeros_like(p["A"]), np.zeros_like(p["B"]), np.zeros_like(p["c"])
    for l in reversed(range(L)):
        z, m, hin = a["zs"][l], a["ms"][l], a["hs"][l]
        dz = dh * (z > 0)
        gA[l] = hin.T @ dz
        gB[l] = m.T @ dz
        gc[l] = dz.sum(axis=0)
        dm = dz @ p["B"][l].T
        dh = dh + dz @ p["A"][l].T + neighbour_mean_transpose(dm, nbr)
    g["A"], g["B"], g["c"] = gA, gB, gc
    dz0 = dh * (a["z0"] > 0)
    g["W0"] = X.T @ dz0
    g["b0"] = dz0.sum(axis=0)
    return loss(y, t), g


def adam_step(p, g, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """In-place Adam update with bias correction; state holds m, v and step."""
    state["t"] = state.get("t", 0) + 1
    tt = state["t"]
    for k in p:
        m = state.setdefault("m_" + k, np.zeros_like(p[k]))
        v = state.setdefault("v_" + k, np.zeros_like(p[k]))
        m[...] = beta1 * m + (1 - beta1) * g[k]
        v[...] = beta2 * v + (1 - beta2) * g[k] * g[k]
        mh = m / (1 - beta1 ** tt)
        vh = v / (1 - beta2 ** tt)
        p[k] = (p[k] - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
