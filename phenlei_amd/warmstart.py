"""Learned warm start: cell features, training targets from converged fields,
and the prediction loaded as a solver's initial state.

The solver generates its own training data (a batched solve of seeded
bump-channel cases), the device network (phenlei_amd.Mpnn) is trained on it
with the hand-written HIP kernels, and for a new case the prediction replaces
the uniform freestream as the initial field.  The plain freestream start
remains the fallback, and the solver's residual judges the prediction.
"""
import numpy as np

from . import Mesh, Mpnn, Solver, default_params

Y_TOP, X_HALF, Z_SPAN = 0.8, 1.5, 0.2


def wall_height(x, bump):
    return bump * np.exp(-25.0 * x * x)


def features(mesh, bump, mach, alpha_deg):
    """[n_cell, 8] features: position (x, y, z scaled to the channel), the
    wall-normal fraction (y - y_w) / (0.8 - y_w), Mach, sin/cos of the angle
    of attack and the bump height / 0.08."""
    met = mesh.metrics()
    n = mesh.n_cell
    x, y, z = met["xcc"][:n], met["ycc"][:n], met["zcc"][:n]
    yw = wall_height(x, bump)
    a = np.radians(alpha_deg)
    X = np.empty((n, 8), np.float32)
    X[:, 0] = x / X_HALF
    X[:, 1] = y / Y_TOP
    X[:, 2] = z / Z_SPAN
    X[:, 3] = (y - yw) / (Y_TOP - yw)
    X[:, 4] = mach
    X[:, 5] = np.sin(a)
    X[:, 6] = np.cos(a)
    X[:, 7] = bump / 0.08
    return X


def state_scale(q_inf):
    """Per-variable scale of the targets: density and velocity as they are,
    pressure relative to the freestream pressure."""
    return np.array([1.0, 1.0, 1.0, 1.0, q_inf[4]])


def targets(q_conv, q_inf, n_cell):
    """[n_cell, 5] (q - q_inf) / scale of a converged primitive field."""
    s = state_scale(q_inf)
    return ((q_conv[:, :n_cell] - q_inf[:, None]) / s[:, None]).T.astype(np.float32)


def prediction_to_state(y, q_inf, n_total):
    """[5, n_total] primitive state from the network output; ghost cells get
    the freestream (the first boundary evaluation overwrites them)."""
    s = state_scale(q_inf)
    q = np.tile(q_inf[:, None], (1, n_total))
    n = y.shape[0]
    q[:, :n] = q_inf[:, None] + (y.astype(np.float64) * s[None, :]).T
    q[0, :n] = np.maximum(q[0, :n], 1e-3 * q_inf[0])
    q[4, :n] = np.maximum(q[4, :n], 1e-3 * q_inf[4])
    return q


def freestream(mach, alpha_deg):
    a = np.radians(alpha_deg)
    return np.array([1.0, np.cos(a), np.sin(a), 0.0, 1.0 / (1.4 * mach * mach)])


def sample_cases(n, seed, mach=(0.3, 0.8), alpha=(-2.0, 6.0), bump=(0.02, 0.08)):
    rng = np.random.default_rng(seed)
    return [dict(mach=float(rng.uniform(*mach)), alpha=float(rng.uniform(*alpha)), bump=float(rng.uniform(*bump)))
            for _ in range(n)]


def solve_cases(cases, dims, drop, max_steps, cfl=10.0, device=0):
    """Batched solve of the cases on (nx, ny, nz) bump channels; returns the
    meshes, the converged primitive fields [5, n_total] and per-case steps."""
    meshes = [Mesh.bump_hex(*dims, c["bump"]) for c in cases]
    params = [default_params(mach=c["mach"], alpha_deg=c["alpha"], cfl_start=cfl, cfl_end=cfl) for c in cases]
    union = Mesh.concat(meshes)
    s = Solver(union, params, device=device)
    s.init_freestream()
    steps, _ = s.solve_batch(max_steps, drop)
    q = s.get_state()
    fields = []
    for k, m in enumerate(meshes):
        cs, gs = union.case_slices(k)
        fields.append(np.concatenate([q[:, cs], q[:, gs]], axis=1))
    s.close()
    return meshes, fields, steps


def train_network(meshes, cases, fields, steps, lr=1e-3, layers=6, seed=0, device=0):
    """Train the device network on the union graph of the training cases."""
    union = Mesh.concat(meshes)
    X = np.concatenate([features(m, c["bump"], c["mach"], c["alpha"]) for m, c in zip(meshes, cases)])
    T = np.concatenate([targets(f, freestream(c["mach"], c["alpha"]), m.n_cell)
                        for m, c, f in zip(meshes, cases, fields)])
    net = Mpnn(layers=layers, seed=seed, device=device)
    net.set_graph(mesh=union)
    net.set_data(X, T)
    hist = net.train(steps, lr)
    return net, hist


def predict_state(net, mesh, case):
    """Initial primitive state [5, n_total] of a case from the network."""
    net.set_graph(mesh=mesh)
    net.set_data(features(mesh, case["bump"], case["mach"], case["alpha"]))
    y = net.forward()
    return prediction_to_state(y, freestream(case["mach"], case["alpha"]), mesh.n_total)


def steps_to(solver, target, max_steps, chunk=64):
    """Iterate until the density-residual RMS is at or below `target`; returns
    (steps, first residual) or (-steps, first residual) if not reached."""
    done, r1 = 0, None
    while done < max_steps:
        h = solver.iterate(1 + done, min(chunk, max_steps - done))
        if r1 is None:
            r1 = float(h[0])
        hit = np.nonzero(h <= target)[0]
        if hit.size:
            return done + int(hit[0]) + 1, r1
        done += h.size
    return -done, r1
