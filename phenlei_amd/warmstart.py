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
    """Train the device network on the union graph of the training cases.
    steps/lr: one Adam run, or lists of equal length (a step-wise learning-rate
    schedule; the optimiser state carries over)."""
    union = Mesh.concat(meshes)
    X = np.concatenate([features(m, c["bump"], c["mach"], c["alpha"]) for m, c in zip(meshes, cases)])
    T = np.concatenate([targets(f, freestream(c["mach"], c["alpha"]), m.n_cell)
                        for m, c, f in zip(meshes, cases, fields)])
    net = Mpnn(layers=layers, seed=seed, device=device)
    net.set_graph(mesh=union)
    net.set_data(X, T)
    steps = list(np.atleast_1d(steps))
    lrs = list(np.atleast_1d(lr)) if np.ndim(lr) else [lr] * len(steps)
    hist = np.concatenate([net.train(int(n), float(r)) for n, r in zip(steps, lrs)])
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


def time_to_target(solver, target, max_steps, switch_stall=0, chunk=64):
    """Iterations and wall seconds until the density-residual RMS is at or
    below `target`; with switch_stall > 0 a mixed-precision solver switches
    to FP64 once its residual has not halved for that many steps."""
    import time
    solver.synchronize()
    t0 = time.perf_counter()
    done, best, mark, since = 0, np.inf, np.inf, 0
    while done < max_steps:
        h = solver.iterate(1 + done, min(chunk, max_steps - done))
        hit = np.nonzero(h <= target)[0]
        if hit.size:
            solver.synchronize()
            return done + int(hit[0]) + 1, time.perf_counter() - t0
        for i, v in enumerate(h):
            best = min(best, v)
            if best < 0.5 * mark:
                mark, since = best, done + i
        done += h.size
        if switch_stall and solver.precision() == "mixed" and done - since >= switch_stall:
            solver.set_precision("fp64")
    solver.synchronize()
    return -done, time.perf_counter() - t0


def compare(n_train=32, n_held=4, dims=(24, 12, 4), drop=1e-8, max_steps=12000,
            train_steps=(1500, 1500, 1500, 1500), train_lr=(2e-3, 1e-3, 3e-4, 1e-4), log=None):
    """The learned warm start end to end on bump-channel cases: converged
    fields of n_train seeded cases (one batched solve), the network trained
    on them, then for n_held held-out cases the wall time to drop * R_1 (R_1:
    the first residual from freestream) of (a) the plain FP64 form from
    freestream and (b) the mixed form started from the prediction, switched
    to FP64 when it stalls.  Returns a dict of per-case and total figures."""
    import time
    say = log or (lambda *a: None)
    t0 = time.perf_counter()
    train = sample_cases(n_train, seed=11)
    meshes, fields, steps = solve_cases(train, dims, drop=drop, max_steps=max_steps)
    t_data = time.perf_counter() - t0
    say(f"training data: {n_train} cases of {meshes[0].n_cell} cells, median {int(np.median(np.abs(steps)))} "
        f"iterations, {t_data:.1f} s")
    t0 = time.perf_counter()
    net, hist = train_network(meshes, train, fields, list(train_steps), list(train_lr))
    t_train = time.perf_counter() - t0
    say(f"training: loss {hist[0]:.3e} -> {hist[-1]:.3e} in {len(hist)} Adam steps, {t_train:.1f} s")
    out = dict(n_train=n_train, cells_per_case=meshes[0].n_cell, drop=drop, data_seconds=round(t_data, 2),
               train_seconds=round(t_train, 2), final_loss=float(hist[-1]), max_steps=max_steps, cases=[])
    tot_a = tot_b = 0.0
    reached_a = reached_b = 0
    for case in sample_cases(n_held, seed=99):
        mesh = Mesh.bump_hex(*dims, case["bump"])
        base = dict(mach=case["mach"], alpha_deg=case["alpha"], cfl_start=10.0, cfl_end=10.0)
        a = Solver(mesh, default_params(**base))
        a.init_freestream()
        r1 = float(a.iterate(1, 1)[0])
        a.init_freestream()
        na, ta = time_to_target(a, drop * r1, max_steps)
        b = Solver(mesh, default_params(precision=1, **base))
        t1 = time.perf_counter()
        b.set_state(predict_state(net, mesh, case))
        t_pred = time.perf_counter() - t1
        nb, tb = time_to_target(b, drop * r1, max_steps, switch_stall=512)
        tb += t_pred
        a.close()
        b.close()
        rec = dict(case, fp64_iterations=na, fp64_seconds=round(ta, 3), warm_mixed_iterations=nb,
                   warm_mixed_seconds=round(tb, 3), predict_seconds=round(t_pred, 3))
        say(f"held-out {rec}")
        out["cases"].append(rec)
        # a run that missed the target within max_steps counts at its full
        # wall time (a lower bound of its time to the target), so a failed
        # warm start is never dropped from the ratio
        reached_a += na > 0
        reached_b += nb > 0
        tot_a += ta
        tot_b += tb
    out["fp64_reached"], out["warm_mixed_reached"] = int(reached_a), int(reached_b)
    out["speedup"] = round(tot_a / tot_b, 3) if tot_b > 0 else None
    out["speedup_note"] = ("total wall time over all held-out cases; a case that missed the target counts at its "
                           "max_steps wall time")
    return out


# ---------------------------------------------------------------- swept wing
# The RANS configuration's wing: root chord 1 on the symmetry plane z = 0,
# semi-span 1.5, farfield 20 root chords away.
WING_SEMI_SPAN, WING_FARFIELD = 1.5, 20.0


def wing_features(mesh, mach, alpha_deg, wall_distance):
    """[n_cell, 8] features of a swept-wing cell: position scaled by the
    semi-span (x, y normal to the wing plane, z spanwise), the wall distance
    on a log scale (0 at 1e-6, 1 at 10 chords), the distance from the root
    quarter chord relative to the farfield, Mach, sin/cos of the angle of
    attack."""
    met = mesh.metrics()
    n = mesh.n_cell
    x, y, z = met["xcc"][:n], met["ycc"][:n], met["zcc"][:n]
    d = np.asarray(wall_distance[:n], dtype=np.float64)
    a = np.radians(alpha_deg)
    X = np.empty((n, 8), np.float32)
    X[:, 0] = x / WING_SEMI_SPAN
    X[:, 1] = y / WING_SEMI_SPAN
    X[:, 2] = z / WING_SEMI_SPAN
    X[:, 3] = (np.log10(d + 1e-6) + 6.0) / 7.0
    X[:, 4] = np.sqrt((x - 0.25) ** 2 + y ** 2 + z ** 2) / WING_FARFIELD
    X[:, 5] = mach
    X[:, 6] = np.sin(a)
    X[:, 7] = np.cos(a)
    return X


def sample_wing_cases(n, seed, mach=(0.5, 0.7), alpha=(0.0, 4.0)):
    rng = np.random.default_rng(seed)
    return [dict(mach=float(rng.uniform(*mach)), alpha=float(rng.uniform(*alpha))) for _ in range(n)]


def wing_params(case, **over):
    """The RANS wing configuration's numerics at a case's Mach and angle."""
    p = dict(mach=case["mach"], alpha_deg=case["alpha"], viscous=3, reynolds=3e6, cfl_start=0.01, cfl_end=5.0,
             cfl_vary_step=500, force_ref_area=1.125, force_ref_length=0.7778, force_ref_span=1.5)
    p.update(over)
    return default_params(**p)


def wing_training_set(cases, dims, steps, log=None):
    """Late-time FP64 fields of seeded wing cases on a coarse member of the
    wing mesh family (one context per case): (mesh, [(X, T)], per-case
    final residual drops)."""
    say = log or (lambda *a: None)
    mesh = Mesh.swept_wing(*dims)
    data, drops = [], []
    for k, case in enumerate(cases):
        s = Solver(mesh, wing_params(case))
        s.init_freestream()
        h = s.iterate(1, steps)
        wd = s.get_field("wdist")
        q = s.get_state()
        data.append((wing_features(mesh, case["mach"], case["alpha"], wd),
                     targets(q, freestream(case["mach"], case["alpha"]), mesh.n_cell)))
        drops.append(float(h[-1] / h[0]))
        say(f"wing case {k}: M{case['mach']:.3f} a{case['alpha']:.2f}, {steps} steps, drop {drops[-1]:.3e}")
        s.close()
    return mesh, data, drops


def train_wing_network(mesh, data, steps, lr, layers=6, seed=0, device=0):
    """The device network trained on the union graph of the training cases
    (the same coarse mesh once per case)."""
    union = Mesh.concat([mesh] * len(data))
    net = Mpnn(layers=layers, seed=seed, device=device)
    net.set_graph(mesh=union)
    net.set_data(np.concatenate([x for x, _ in data]), np.concatenate([t for _, t in data]))
    hist = np.concatenate([net.train(int(n), float(r)) for n, r in zip(steps, lr)])
    return net, hist


def wing_candidates(net, mesh, case, wall_distance, windows=(None, 0.5)):
    """Initial states of a wing case from the network: the prediction
    itself, and the prediction faded to the freestream beyond a wall
    distance (exp(-d / width)), which keeps the large far-field cells at the
    freestream they start from."""
    q_inf = freestream(case["mach"], case["alpha"])
    net.set_graph(mesh=mesh)
    net.set_data(wing_features(mesh, case["mach"], case["alpha"], wall_distance))
    y = net.forward()
    out = []
    d = np.asarray(wall_distance[:mesh.n_cell], dtype=np.float64)
    for w in windows:
        yy = y if w is None else (y * np.exp(-d / w)[:, None]).astype(np.float32)
        out.append((w, prediction_to_state(yy, q_inf, mesh.n_total)))
    return out


def judged_start(solver, candidates, r1_freestream):
    """The solver's residual judges the predicted starts: the candidate with
    the smallest first density-residual RMS is used if it is below the
    freestream start's, else the freestream start (the fallback).  Returns
    (label, its first residual, all candidates' residuals)."""
    seen = []
    for w, q in candidates:
        solver.set_state(q)
        solver.residual()
        seen.append((w, float(np.sqrt(solver.residual_norm()[0] / solver.mesh.n_cell))))
    w, r = min(seen, key=lambda t: t[1])
    if r < r1_freestream:
        solver.set_state(dict(candidates)[w])
        return ("prediction" if w is None else f"prediction faded beyond d={w}"), r, seen
    solver.init_freestream()
    return "freestream (fallback)", r1_freestream, seen


WING_TRAIN_DIMS = (24, 12, 10, 1e-4, 8)


def save_network(net, path, **meta):
    """The parameters as one flat FP32 array in the network's layout order,
    with plain numeric metadata (nothing that needs unpickling)."""
    p = net.get_params()
    flat = np.concatenate([np.asarray(p[k], np.float32).reshape(-1) for k, _ in net.layout()])
    np.savez(path, params=flat, layers=np.int32(net.layers), **meta)


def load_network(path, device=0):
    """A network saved by save_network."""
    with np.load(path, allow_pickle=False) as z:
        net = Mpnn(layers=int(z["layers"]), seed=0, device=device)
        net.set_params(net.unflatten(np.asarray(z["params"], np.float32)))
        meta = {k: z[k] for k in z.files if k not in ("params", "layers")}
    return net, meta


def main(argv=None):
    """python -m phenlei_amd.warmstart train-wing OUT.npz [n_cases solve_steps adam_steps]:
    generate the wing training set on the GPU, train the network on it and
    save its parameters."""
    import sys
    import time
    args = list(sys.argv[1:] if argv is None else argv)
    if not args or args[0] != "train-wing" or len(args) < 2:
        raise SystemExit(main.__doc__)
    out = args[1]
    n_cases = int(args[2]) if len(args) > 2 else 16
    solve_steps = int(args[3]) if len(args) > 3 else 3000
    adam = int(args[4]) if len(args) > 4 else 2400

    def say(m):
        print("[warmstart] " + m, flush=True)

    t0 = time.perf_counter()
    cases = sample_wing_cases(n_cases, seed=2026)
    mesh, data, drops = wing_training_set(cases, WING_TRAIN_DIMS, solve_steps, log=say)
    t_data = time.perf_counter() - t0
    say(f"training set: {n_cases} cases of {mesh.n_cell} cells in {t_data:.1f} s")
    t0 = time.perf_counter()
    stages = [adam // 2, adam // 4, adam - adam // 2 - adam // 4]
    net, hist = train_wing_network(mesh, data, stages, [1e-3, 3e-4, 1e-4])
    t_train = time.perf_counter() - t0
    say(f"training: loss {hist[0]:.3e} -> {hist[-1]:.3e} in {len(hist)} Adam steps, {t_train:.1f} s")
    save_network(net, out, train_dims=np.array(WING_TRAIN_DIMS[:3], np.int32), train_h1=WING_TRAIN_DIMS[3],
                 n_cases=np.int32(n_cases), solve_steps=np.int32(solve_steps), adam_steps=np.int32(len(hist)),
                 loss=np.array([hist[0], hist[-1]]), data_seconds=t_data, train_seconds=t_train,
                 cases=np.array([[c["mach"], c["alpha"]] for c in cases]))
    say(f"saved {out}")


if __name__ == "__main__":
    main()
