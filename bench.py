"""Benchmark of the residual + LU-SGS path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]

A step is one outer iteration of the LU-SGS branch of Controller::Relaxation
(boundary conditions, gradients, limiter, fluxes, spectral radius, time step,
forward/backward sweeps, update) over the whole mesh, with the state already
resident in HBM.  value = cells * steps / wall seconds (cell-updates/s).

Besides the contract fields the JSON line carries:
  roofline      the face-flux kernel's algorithmic bytes per launch over its
                mean duration (HIP events on the solver stream, a second pass
                of the same steps), against the 8 TB/s HBM peak
  cpu_baseline  the CPU oracle (sequential FP64 restatement, one core) on a
                bounded sample of the same workload
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import phenlei_amd as pa  # noqa: E402

METRIC = "cell-updates/s and time-to-1e-8 residual drop: mixed+warm-start vs plain FP64"
HBM_PEAK_GBS = 8000.0


# Workloads.  'c1' is the reference's CPU-runnable case (64x32x8 hex bump
# channel, inviscid, M0.5, CFL 10).
CONFIGS = {
    "c1": dict(workload="bump channel, 64x32x8 hex (16,384 cells), inviscid Euler, M0.5, AoA 0, "
                        "Roe + GG-cell + Venkat(K=5), LU-SGS CFL 10, FP64",
               mesh=("bump_hex", 64, 32, 8, 0.05),
               params=dict(mach=0.5, cfl_start=10.0, cfl_end=10.0), cpu_sample=(5, 100, 5, 100)),
    "c2": dict(workload="bump channel, hybrid 1,007,616 cells (15 prism layers, first height 1e-4, over a "
                        "128x16 wall; tet fill of a seeded jittered lattice), laminar NS, M0.3, Re 5e5/unit "
                        "length, Sutherland, Roe + GG-cell + Venkat(K=5), LU-SGS CFL 10, FP64",
               mesh=("bump_hybrid", 128, 16, 15, 1e-4, 77, 0.1, 1, 0.05),
               params=dict(mach=0.3, viscous=1, reynolds=5e5, cfl_start=10.0, cfl_end=10.0),
               cpu_sample=(5, 20, 1, 5)),
    "hex1m": dict(workload="bump channel, 256x64x64 hex (1,048,576 cells), inviscid Euler, M0.5, "
                           "Roe + GG-cell + Venkat(K=5), LU-SGS CFL 10, FP64",
                  mesh=("bump_hex", 256, 64, 64, 0.05),
                  params=dict(mach=0.5, cfl_start=10.0, cfl_end=10.0), cpu_sample=(2, 10, 1, 4)),
}
CONFIGS["c3"] = dict(workload="swept wing (root chord 1, taper 0.5, AR 4, quarter-chord sweep 25 deg, symmetric "
                              "cubic 12% section), hybrid 3,983,616 cells (30 prism layers each side of the wing plane, "
                              "first height 1e-5, growth 1.2; tetrahedral rows to a farfield at 20 root chords), "
                              "Spalart-Allmaras RANS, M0.6, AoA 2 deg, Re 3e6 per root chord, Roe + GG-cell + "
                              "Venkat(K=5), LU-SGS with a CFL ramp 0.01 -> 5 over 500 steps (the reference's ramp "
                              "to 10 diverges behind the tip near step 450)",
                     mesh=("swept_wing", 80, 40, 30, 1e-5, 29),
                     params=dict(mach=0.6, alpha_deg=2.0, viscous=3, reynolds=3e6, cfl_start=0.01, cfl_end=5.0,
                                 cfl_vary_step=500, force_ref_area=1.125, force_ref_length=0.7778,
                                 force_ref_span=1.5),
                     cpu_sample=(2, 8, 1, 2),
                     # the headline leg: time to a residual drop, plain FP64 from freestream against the mixed
                     # form started from the learned warm start (judged by the solver's residual)
                     converge=dict(drop=0.25, max_steps=5000))
# The reference's own example case on its own grid: the RAE2822 2-D hybrid
# grid it ships (4 zones, joined) with cfd_para_subsonic.hypara and the
# cfd_para.hypara defaults (ggnode gradient, 100 first-order Van Leer steps).
CONFIGS["rae2822"] = dict(workload="RAE2822 airfoil, the reference's example grid rae2822_hybrid2d__4_0.fts (2-D hybrid, "
                                   "36,036 cells, 4 zones joined), its subsonic example case: Spalart-Allmaras RANS, "
                                   "M0.2, AoA 2.79 deg, Re 6.5e6, Roe + GG-node + Venkat(K=50), LU-SGS 5 sweep pairs "
                                   "(tolerance 0.01), CFL 0.01 -> 100 over 500 steps, 100 first-order Van Leer steps",
                          mesh=("fts", os.path.join("tests", "golden", "rae2822_hybrid2d__4_0.fts")),
                          params=dict(mach=0.2, alpha_deg=2.79, reynolds=6.5e6, viscous=3, gradient=2,
                                      venkat_k=50.0, cfl_start=0.01, cfl_end=100.0, cfl_vary_step=500,
                                      n_lusgs_sweeps=5, flow_init_step=100),
                          cpu_sample=(5, 30, 2, 10), data="the reference's RAE2822 example grid (tests/golden)")
DEFAULT_CONFIG = "c3"
WING_WEIGHTS = os.path.join(ROOT, "phenlei_amd", "data", "wing_warmstart.npz")

# Batched sweep: 64 inviscid bump-channel cases in one context, sizes and
# freestream drawn from a fixed seed.
C4_CASES, C4_SEED = 64, 20241202


def c4_cases(n_cases=C4_CASES, seed=C4_SEED):
    """(nx, ny, nz, bump, mach, alpha) per case: cells ~ U[50k, 400k] with the
    channel's 8:4:1 aspect, Mach ~ U[0.3, 0.8], AoA ~ U[-2, 6] deg, bump
    height ~ U[0.02, 0.08]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_cases):
        n = rng.uniform(50e3, 400e3)
        nz = max(2, int(round((n / 32.0) ** (1.0 / 3.0))))
        ny = 4 * nz
        nx = max(4, int(round(n / (ny * nz))))
        out.append((nx, ny, nz, float(rng.uniform(0.02, 0.08)), float(rng.uniform(0.3, 0.8)),
                    float(rng.uniform(-2.0, 6.0))))
    return out


CONFIGS["c4"] = dict(workload="batched sweep: 64 inviscid bump-channel hex cases in one context, cells ~ U[50k,400k] "
                              "(seeded), Mach ~ U[0.3,0.8], AoA ~ U[-2,6] deg, bump ~ U[0.02,0.08]; Roe + GG-cell + "
                              "Venkat(K=5), LU-SGS CFL 10; FP32 face fluxes + FP64 accumulation; autotuned tiles; hipGraph",
                     batch=True, params=dict(cfl_start=10.0, cfl_end=10.0, precision=1), cpu_sample=(2, 20, 1, 10))


def build_batch(cfg):
    cases = c4_cases()
    meshes = [pa.Mesh.bump_hex(nx, ny, nz, h) for nx, ny, nz, h, _, _ in cases]
    params = [pa.default_params(mach=mach, alpha_deg=alpha, **cfg["params"]) for *_, mach, alpha in cases]
    return meshes, params


def build_mesh(spec):
    kind = spec[0]
    if kind == "bump_hex":
        return pa.Mesh.bump_hex(*spec[1:])
    if kind == "bump_hybrid":
        return pa.Mesh.bump_hybrid(*spec[1:])
    if kind == "swept_wing":
        return pa.Mesh.swept_wing(*spec[1:])
    if kind == "fts":
        return pa.Mesh.from_fts(os.path.join(ROOT, spec[1]))
    raise ValueError(kind)


def flux_algorithmic_bytes(n_cell, n_face, viscous, mixed=False):
    """Compulsory HBM bytes of one face-flux launch: read q 40N, gradients
    120N (60N stored in FP32 by the mixed form), limiter 8N, centroids 24N,
    face centres 24F, normal+area 32F, left/right 8F; write residual 40N and
    spectral radius 8N; viscous adds the temperature gradient (24N, 12N in
    FP32) and the viscous spectral radius 8N."""
    g = 60 if mixed else 120
    b = (240 - 120 + g) * n_cell + 64 * n_face
    if viscous:
        b += (12 if mixed else 24) * n_cell + 8 * n_cell
    return b


def sweep_algorithmic_bytes(n_cell, n_face, mixed=False, viscous=False, levels_form=False):
    """Compulsory HBM bytes of one LU-SGS sweep pair.  Every face side is used
    once over the two directions.  Level form: the side's stored coefficient
    row (128 B, 64 B in the mixed form) and the neighbour's dq (40 B); per
    cell, forward reads res and the diagonal and writes dq and the
    off-diagonal sum (128 B), backward also reads that sum (168 B).  Dataflow
    form (the default): the sweep evaluates the coefficients, so a face side
    reads the face record (64 B), the neighbour's state and dq (80 B), and
    when viscous its eddy viscosity and centroid (32 B); a cell also reads
    its own state (40 B; viscous +32 B) in each direction."""
    if levels_form:
        row = 64 if mixed else 128
        return 296 * n_cell + 2 * n_face * (row + 40)
    side = 144 + (32 if viscous else 0)
    cell = 376 + (64 if viscous else 0)
    return cell * n_cell + 2 * n_face * side


def host_cpu():
    """(logical CPUs, CPU model) of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count(), model


def _oracle_rate(orc, conn, met, params, n_cell, steps, warmup, wall_distance, omp):
    cpu = orc.OracleSolver(conn, met, params, omp=omp)
    if wall_distance is not None:
        cpu.set_wall_distance(wall_distance)  # preprocessing, outside the timed iterations
    cpu.init_freestream()
    if warmup:
        cpu.iterate(1, warmup)
    t0 = time.perf_counter()
    cpu.iterate(1 + warmup, steps)
    dt = time.perf_counter() - t0
    cpu.close()
    return n_cell * steps / dt, dt


def cpu_baseline(mesh, params, sample, wall_distance=None):
    """The CPU oracle on this host: the OpenMP build over the host's cores
    (face loops split by cell owner, LU-SGS by dependency level; the same
    bits as the sequential build) and the sequential build on one core, each
    on a bounded sample of iterations of the same mesh from freestream.
    sample = (omp warm-up, omp timed, one-core warm-up, one-core timed)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc  # the checker, timed here as the CPU baseline only
    conn, met = mesh.connectivity(), mesh.metrics()
    nproc, model = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", nproc or 1))
    w_omp, n_omp, w_seq, n_seq = sample
    print(f"[bench] CPU oracle, OpenMP on {threads} threads: {w_omp} + {n_omp} iterations", file=sys.stderr, flush=True)
    v_omp, t_omp = _oracle_rate(orc, conn, met, params, mesh.n_cell, n_omp, w_omp, wall_distance, True)
    print(f"[bench] CPU oracle, one core: {w_seq} + {n_seq} iterations", file=sys.stderr, flush=True)
    v_seq, t_seq = _oracle_rate(orc, conn, met, params, mesh.n_cell, n_seq, w_seq, wall_distance, False)
    return dict(value=v_omp, unit="cell-updates/s", cores=threads, kind="port", nproc=nproc, cpu_model=model,
                sample=f"{n_omp} FP64 iterations after {w_omp} warm-up of the same mesh from freestream "
                       f"({mesh.n_cell} cells, {t_omp:.1f} s timed), the C oracle built with OpenMP on {threads} "
                       f"threads (same results as the sequential oracle)",
                single_core=dict(value=v_seq, cores=1,
                                 sample=f"{n_seq} iteration(s) after {w_seq} warm-up, {t_seq:.1f} s timed, "
                                        f"sequential C oracle"))


def time_to_drop(mesh, pover, drop, max_steps, chunk=256, stall=2048):
    """Wall time and iterations to reach drop * R_1 (R_1: the first density-
    residual RMS of the FP64 run from freestream) for (a) the plain FP64 form
    and (b) the mixed form, switched to FP64 on the live context once its
    residual has not halved for `stall` iterations.  Both start from
    freestream; the clock covers the iterations and the host's residual checks
    every `chunk` iterations."""
    out = {}
    target = None
    for form in ("fp64", "mixed_then_fp64"):
        s = pa.Solver(mesh, pa.default_params(**dict(pover, precision=0 if form == "fp64" else 1)))
        s.init_freestream()
        s.synchronize()
        t0 = time.perf_counter()
        n, hit, best, since, switched = 0, -1, np.inf, 0, None
        while n < max_steps and hit < 0:
            h = s.iterate(1 + n, min(chunk, max_steps - n))
            if target is None:
                target = drop * float(h[0])
            below = np.nonzero(h <= target)[0]
            if below.size:
                hit = n + int(below[0]) + 1
            if form != "fp64" and switched is None:
                if h.min() < 0.5 * best:
                    best, since = float(h.min()), 0
                else:
                    since += h.size
                if since >= stall:
                    s.set_precision("fp64")
                    switched = n + h.size
            n += h.size
        s.synchronize()
        wall = time.perf_counter() - t0
        out[form] = dict(iterations=hit if hit > 0 else -n, seconds=round(wall, 3),
                         reached=hit > 0, switched_to_fp64_at=switched)
        print(f"[bench] time to {drop:g} drop, {form}: {out[form]}", file=sys.stderr, flush=True)
        s.close()
    a, b = out["fp64"], out["mixed_then_fp64"]
    out["speedup"] = round(a["seconds"] / b["seconds"], 3) if a["reached"] and b["reached"] else None
    out["drop"] = drop
    out["target_abs"] = target
    out["warm_start"] = None  # the warm-start network is trained per family of cases (tests, warmstart.py)
    return out


COPY_BYTES = 2 * 1024**3  # the calibration copy: 1 GiB read and 1 GiB written per launch


def traffic_child(cfg):
    """The workload of a counter pass: two warm-up and two eager steps,
    then the calibration copy."""
    mesh = build_mesh(cfg["mesh"])
    s = pa.Solver(mesh, pa.default_params(**cfg["params"]))
    s.set_launch_mode("eager")
    s.init_freestream()
    s.iterate(1, 4, history=False)
    s.synchronize()
    s.close()
    pa.hbm_copy_gbps(0, COPY_BYTES, 2)


def _counter_means(out_dir):
    """Mean counter value per dispatch, by (kernel, counter), from a
    rocprofv3 csv counter collection."""
    import csv
    import glob
    sums = {}
    for path in glob.glob(os.path.join(out_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                m = re.search(r"\b(k_\w+)(<[^()]*>)?", row["Kernel_Name"])
                if not m:
                    continue
                key = (m.group(0), row["Counter_Name"])
                v = sums.setdefault(key, [0.0, 0])
                v[0] += float(row["Counter_Value"])
                v[1] += 1
    return {k: v[0] / v[1] for k, v in sums.items() if v[1]}


def measure_traffic(config, prefixes, counters=(("FETCH_SIZE",), ("WRITE_SIZE",)), timeout=240):
    """HBM bytes per dispatch of the kernels whose names start with one of
    `prefixes`, measured in this run: one rocprofv3 counter pass per counter
    set over a short child run of the same workload, each counter scaled by
    what it reads for the copy kernel of known size in the same pass (1 GiB
    read, 1 GiB written).  Returns {kernel: bytes} and the scale factors, or
    None if the profiler is not available."""
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    per_kernel, factors = {}, {}
    work = tempfile.mkdtemp(prefix="phl_pmc_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        for group in counters:
            out = os.path.join(work, group[0].lower())
            cmd = [exe, "--pmc", *group, "--output-format", "csv", "-d", out, "-o", "pass", "--",
                   sys.executable, os.path.join(ROOT, "bench.py"), "--traffic-child", "--config", config]
            print(f"[bench] counter pass {' '.join(group)}", file=sys.stderr, flush=True)
            r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=timeout)
            if r.returncode != 0:
                print(f"[bench] counter pass failed ({r.returncode}): {r.stderr[-400:]}", file=sys.stderr, flush=True)
                return None
            means = _counter_means(out)
            for cname in group:
                copy = means.get(("k_copy", cname))
                if not copy:
                    return None
                # a copy launch reads and writes COPY_BYTES / 2 each
                factor = (COPY_BYTES / 2) / (copy * 1024.0)
                factors[cname] = factor
                for (kname, c2), v in means.items():
                    if c2 == cname and any(kname.startswith(p) for p in prefixes):
                        per_kernel.setdefault(kname, {})[cname] = v * 1024.0 * factor
    except (OSError, subprocess.TimeoutExpired) as e:
        print(f"[bench] counter pass not run: {e}", file=sys.stderr, flush=True)
        return None
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return dict(kernels=per_kernel, factors=factors)


def wing_time_to_drop(mesh, pover, drop, max_steps, weights, chunk=256, stall=1024):
    """The headline comparison on the wing: wall time until the density-
    residual RMS falls to drop * R_1, R_1 being the first residual of the
    plain FP64 run from freestream (one absolute target for both runs), for
    (a) plain FP64 from freestream and (b) the mixed form started from the
    learned warm start: the network (trained on coarse wing cases, loaded
    from `weights`) predicts the field, the solver's first residual judges
    the prediction against the freestream start (the fallback), and the run
    switches to FP64 if it stalls for `stall` steps.  (b)'s clock covers the
    network load, the feature assembly, the forward pass and the judging.
    A run that misses the target counts at its full wall time."""
    from phenlei_amd import warmstart as ws
    case = dict(mach=pover["mach"], alpha=pover["alpha_deg"])
    out = dict(drop=drop, max_steps=max_steps, weights=os.path.relpath(weights, ROOT) if weights else None)
    fields = {}
    r1 = None
    for form in ("fp64", "mixed_warm_start"):
        s = pa.Solver(mesh, pa.default_params(**dict(pover, precision=0 if form == "fp64" else 1)))
        s.init_freestream()
        s.synchronize()
        rec = {}
        t0 = time.perf_counter()
        if form != "fp64":
            if weights and os.path.exists(weights):
                net, meta = ws.load_network(weights)
                cands = ws.wing_candidates(net, mesh, case, s.get_field("wdist"))
                label, r_start, seen = ws.judged_start(s, cands, r1)
                net.close()
                rec.update(start=label, start_residual=r_start, candidates=[
                    dict(fade=w, first_residual=r, ratio_to_freestream=r / r1) for w, r in seen],
                    network=dict(train_cases=int(meta["n_cases"]), train_cells=int(np.prod(meta["train_dims"])),
                                 solve_steps=int(meta["solve_steps"]), adam_steps=int(meta["adam_steps"])))
            else:
                rec.update(start="freestream (no trained network)")
        n, hit, best, mark, since, switched = 0, -1, np.inf, np.inf, 0, None
        while n < max_steps and hit < 0:
            h = s.iterate(1 + n, min(chunk, max_steps - n))
            if r1 is None:
                r1 = float(h[0])
            below = np.nonzero(h <= drop * r1)[0]
            if below.size:
                hit = n + int(below[0]) + 1
            for i, v in enumerate(h):
                best = min(best, float(v))
                if best < 0.5 * mark:
                    mark, since = best, n + i
            n += h.size
            if form != "fp64" and switched is None and s.precision() == "mixed" and n - since >= stall:
                s.set_precision("fp64")
                switched = n
            print(f"[bench]   {form}: step {n}, residual {h[-1]:.3e} ({h[-1] / r1:.3e} of R_1)", file=sys.stderr,
                  flush=True)
        s.synchronize()
        wall = time.perf_counter() - t0
        f = s.forces()
        rec.update(iterations=hit if hit > 0 else -n, seconds=round(wall, 3), reached=hit > 0,
                   switched_to_fp64_at=switched, cl=f["CL"], cd=f["CD"])
        fields[form] = s.get_state()[:, :mesh.n_cell]
        out[form] = rec
        print(f"[bench] time to {drop:g} x R_1, {form}: {rec}", file=sys.stderr, flush=True)
        s.close()
    out["r1"] = r1
    a, b = out["fp64"], out["mixed_warm_start"]
    out["speedup"] = round(a["seconds"] / b["seconds"], 3)
    out["speedup_note"] = ("wall time of plain FP64 over mixed + warm start to the same absolute residual; a run "
                           "that missed it counts at its full wall time")
    out["dcl"], out["dcd"] = abs(a["cl"] - b["cl"]), abs(a["cd"] - b["cd"])
    qa, qb = fields["fp64"], fields["mixed_warm_start"]
    dev = qa - np.mean(qa, axis=1, keepdims=True)
    out["field_l2_rel"] = [float(np.linalg.norm(qa[m] - qb[m]) / max(np.linalg.norm(dev[m]), 1e-300))
                           for m in range(5)]
    return out


def precision_compare(mesh, pover, steps, chunk=250):
    """The same number of iterations from freestream in the FP64 and the
    mixed form: lift and drag of both, and the relative L2 difference of the
    fields (each primitive variable's deviation from freestream, over real
    cells)."""
    out = {"steps": steps}
    fields = {}
    for form in ("fp64", "mixed"):
        s = pa.Solver(mesh, pa.default_params(**dict(pover, precision=0 if form == "fp64" else 1)))
        s.init_freestream()
        done, last = 0, None
        while done < steps:
            h = s.iterate(1 + done, min(chunk, steps - done))
            done += h.size
            last = float(h[-1])
            f = s.forces()
            print(f"[bench]   {form}: step {done}, density residual {last:.3e}, CL {f['CL']:.6f} CD {f['CD']:.6f}",
                  file=sys.stderr, flush=True)
        f = s.forces()
        out[f"cl_{form}"], out[f"cd_{form}"], out[f"rms_{form}"] = f["CL"], f["CD"], last
        q = s.get_state()[:, :mesh.n_cell]
        if pover.get("viscous") == 3:
            fields[form + "_nut"] = s.get_field("nut")[:mesh.n_cell]
        fields[form] = q
        s.close()
    out["dcl"] = abs(out["cl_fp64"] - out["cl_mixed"])
    out["dcd"] = abs(out["cd_fp64"] - out["cd_mixed"])
    a, b = fields["fp64"], fields["mixed"]
    dev = a - np.mean(a, axis=1, keepdims=True)
    out["field_l2_rel"] = [float(np.linalg.norm(a[m] - b[m]) / max(np.linalg.norm(dev[m]), 1e-300)) for m in range(5)]
    if "fp64_nut" in fields:
        na, nb = fields["fp64_nut"], fields["mixed_nut"]
        out["nut_l2_rel"] = float(np.linalg.norm(na - nb) / max(np.linalg.norm(na - np.mean(na)), 1e-300))
    print(f"[bench] precision comparison: {out}", file=sys.stderr, flush=True)
    return out


def batch_solve_logged(s, max_steps, drop, stall, label, chunk=256):
    """phl_solve_batch's loop (per-case drop, freezing, precision switch)
    driven from here so that each chunk of iterations reports its progress."""
    n = s.n_cases
    r1 = None
    hit = np.full(n, -1)
    best = np.full(n, np.inf)
    mark = np.full(n, np.inf)
    since = np.zeros(n, int)
    done = 0
    while done < max_steps and np.any(hit < 0):
        h = s.iterate(1 + done, min(chunk, max_steps - done))
        if r1 is None:
            r1 = h[0].copy()
        for k in np.nonzero(hit < 0)[0]:
            for i in range(h.shape[0]):
                v = h[i, k]
                best[k] = min(best[k], v)
                if best[k] < 0.5 * mark[k]:
                    mark[k], since[k] = best[k], done + i
                if v <= drop * r1[k]:
                    hit[k] = done + i + 1
                    break
        done += h.shape[0]
        s.set_frozen(hit > 0)
        left = hit < 0
        if stall and s.precision() == "mixed" and left.any() and np.all(done - since[left] >= stall):
            s.set_precision("fp64")
            print(f"[bench]   {label}: switched to fp64 at step {done}", file=sys.stderr, flush=True)
        print(f"[bench]   {label}: step {done}, {int(np.sum(hit > 0))}/{n} cases at the drop, "
              f"worst remaining drop {np.max((best / r1)[left]) if left.any() else 0:.2e}",
              file=sys.stderr, flush=True)
    return np.where(hit > 0, hit, -done)


def batch_time_to_drop(mesh, params_list, drop, max_steps, stall=2048, forms=("fp64", "mixed_then_fp64")):
    """Batched form: every case of the batch to drop * its own R_1, (a) FP64
    throughout and (b) mixed with the switch to FP64 once every unconverged
    case has stalled (phl_set_precision_switch).  Converged cases freeze."""
    out = {}
    for form in forms:
        plist = []
        for p in params_list:
            q = pa.Params.from_buffer_copy(p)
            q.precision = 0 if form == "fp64" else 1
            plist.append(q)
        s = pa.Solver(mesh, plist, device=0)
        s.autotune()
        s.init_freestream()
        if form != "fp64":
            s.set_precision_switch(stall)
        s.synchronize()
        t0 = time.perf_counter()
        steps = batch_solve_logged(s, max_steps, drop, stall if form != "fp64" else 0, form)
        s.synchronize()
        wall = time.perf_counter() - t0
        out[form] = dict(seconds=round(wall, 3), cases_reached=int(np.sum(steps > 0)),
                         max_iterations=int(np.max(np.abs(steps))), median_iterations=int(np.median(np.abs(steps))),
                         final_precision=s.precision())
        print(f"[bench] batch time to {drop:g} drop, {form}: {out[form]}", file=sys.stderr, flush=True)
        s.close()
    if "fp64" in out and "mixed_then_fp64" in out:
        out["speedup"] = round(out["fp64"]["seconds"] / out["mixed_then_fp64"]["seconds"], 3)
    out["drop"] = drop
    return out


def mpnn_dense_flops(layers, train):
    """Dense-layer FLOPs per cell of the warm-start network (hidden width
    128, 8 inputs, 5 outputs): the input layer, per layer h A + m B, the
    output layer; a training step adds, per layer, the weight gradients of A
    and B and the products dZ A^T and dZ B^T, plus the input- and
    output-layer gradients.  Neighbour means, ReLU, loss and Adam are not
    counted (they are HBM-bound and small)."""
    H, I, O = 128, 8, 5
    fwd = 2 * I * H + layers * 2 * (2 * H * H) + 2 * H * O
    if not train:
        return fwd
    return fwd + layers * 4 * (2 * H * H) + 2 * I * H + 2 * (2 * H * O)


def mpnn_throughput(n_cases, dims, steps, layers=6):
    """The warm-start network at inference and training scale: n_cases
    bump-channel hex cases of dims cells as one union graph, synthetic
    features and targets.  Times `steps` forward passes (each copies the
    prediction y [cells, 5] to the host) and `steps` Adam steps, and prices
    both by their dense-layer FLOPs against the FP32 MFMA peak."""
    print(f"[bench] mpnn: building {n_cases} cases of {dims}", file=sys.stderr, flush=True)
    meshes = [pa.Mesh.bump_hex(*dims, 0.02 + 0.06 * i / max(n_cases - 1, 1)) for i in range(n_cases)]
    union = pa.Mesh.concat(meshes)
    n = union.n_cell
    rng = np.random.default_rng(5)
    net = pa.Mpnn(layers=layers, seed=0, device=0)
    net.set_graph(mesh=union)
    net.set_data(rng.standard_normal((n, net.N_IN)).astype(np.float32),
                 0.1 * rng.standard_normal((n, net.N_OUT)).astype(np.float32))
    print(f"[bench] mpnn: {n} cells; forward x {steps}", file=sys.stderr, flush=True)
    net.forward()
    t0 = time.perf_counter()
    for _ in range(steps):
        net.forward()
    t_fwd = (time.perf_counter() - t0) / steps
    net.set_forward_precision(True)
    net.forward()
    t0 = time.perf_counter()
    for _ in range(steps):
        net.forward()
    t_fwd16 = (time.perf_counter() - t0) / steps
    net.set_forward_precision(False)
    print(f"[bench] mpnn: forward {t_fwd * 1e3:.1f} ms (BF16 hidden layers {t_fwd16 * 1e3:.1f} ms); "
          f"training x {steps}", file=sys.stderr, flush=True)
    net.train(1, 1e-4)
    t0 = time.perf_counter()
    net.train(steps, 1e-4)
    t_train = (time.perf_counter() - t0) / steps
    print(f"[bench] mpnn: training step {t_train * 1e3:.1f} ms", file=sys.stderr, flush=True)
    net.close()
    peak = 157.3e12
    out = dict(workload=f"{n_cases} bump-channel hex cases of {dims[0]}x{dims[1]}x{dims[2]} as one graph, "
                        f"{layers} layers, width 128, FP32 (MFMA 32x32x2 f32), synthetic features",
               cells=n, steps=steps)
    for key, t, train, pk in (("forward", t_fwd, False, peak), ("train_step", t_train, True, peak),
                              ("forward_bf16", t_fwd16, False, 2.5e15)):
        fl = mpnn_dense_flops(layers, train) * n
        out[key] = dict(ms=round(t * 1e3, 3), cells_per_s=n / t, dense_tflop=fl / 1e12,
                        achieved_tflops=round(fl / t / 1e12, 2), peak_tflops=pk / 1e12,
                        frac=round(fl / t / pk, 4))
    out["forward"]["note"] = "wall time per call, including the copy of y to the host"
    out["forward_bf16"]["note"] = ("hidden layers on v_mfma_f32_32x32x16_bf16 (BF16 operands, FP32 accumulation), "
                                   "priced against the dense BF16 peak; input/output layers FP32; wall time per call "
                                   "including the copy of y to the host")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default=DEFAULT_CONFIG, choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", action="store_true", help="(the HBM copy probe always runs)")
    ap.add_argument("--autotune", action="store_true", help="pick the flux/gradient tiles by timing first")
    ap.add_argument("--launch", default="graph", choices=["graph", "eager", "eager_overlap"],
                    help="replay captured hipGraphs (default), launch every kernel eagerly on one stream, or "
                         "eagerly with the turbulence sweeps on a second stream")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"],
                    help="override the configuration's precision")
    ap.add_argument("--warm-start", action="store_true",
                    help="also run the learned warm start end to end on bump-channel cases (32 training cases, "
                         "4 held out) and time mixed + warm start against plain FP64 to a 1e-8 drop")
    ap.add_argument("--ws-cases", default="32,4", metavar="TRAIN,HELD",
                    help="warm start: training and held-out case counts")
    ap.add_argument("--ws-dims", default="24,12,4", metavar="NX,NY,NZ",
                    help="warm start: bump-channel grid of every case")
    ap.add_argument("--ws-steps", default="1500,1500,1500,1500", metavar="N,...",
                    help="warm start: Adam steps per learning-rate stage (2e-3, 1e-3, 3e-4, 1e-4)")
    ap.add_argument("--mpnn", default=None, metavar="CASES,NX,NY,NZ",
                    help="also time the warm-start network at this scale (e.g. 64,64,56,56: 64 x 200,704 cells)")
    ap.add_argument("--mpnn-steps", type=int, default=5, help="forward passes and Adam steps timed by --mpnn")
    ap.add_argument("--precision-compare", type=int, default=None, metavar="STEPS",
                    help="also run STEPS iterations in FP64 and in the mixed form and compare CL/CD and fields")
    ap.add_argument("--face-loop", default=None, choices=["tile", "split"],
                    help="face-loop form (default: the library's, split)")
    ap.add_argument("--sweep-form", default=None, choices=["dataflow", "levels", "ticket", "xcd"],
                    help="LU-SGS sweep form (default: the library's, dataflow in launch order)")
    ap.add_argument("--storage-order", default=None, choices=["locality", "mesh"],
                    help="device storage order of the cells (default: the library's, locality)")
    ap.add_argument("--converge", type=float, default=None, metavar="DROP",
                    help="also time the FP64 and mixed-then-FP64 forms to this residual drop from freestream")
    ap.add_argument("--max-steps", type=int, default=20000, help="iteration cap of --converge")
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the configuration's default time-to-drop leg (the wing's headline comparison)")
    ap.add_argument("--forms", default="fp64,mixed_then_fp64",
                    help="forms timed by --converge on a batch (comma separated)")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-traffic", action="store_true", help="skip the in-run HBM counter passes")
    args = ap.parse_args()
    if args.traffic_child:
        traffic_child(CONFIGS[args.config])
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    cfg = CONFIGS[args.config]
    print(f"[bench] config {args.config}: building mesh", file=sys.stderr, flush=True)
    so = {"locality": 0, "mesh": 1}.get(args.storage_order)
    fl = {"split": 0, "tile": 1}.get(args.face_loop)
    sf = {"dataflow": 0, "levels": 1, "ticket": 2, "xcd": 3}.get(args.sweep_form)
    if cfg.get("batch"):
        parts, params_list = build_batch(cfg)
        for p in params_list:
            if so is not None:
                p.storage_order = so
            if fl is not None:
                p.face_loop = fl
            if sf is not None:
                p.sweep_form = sf
        mesh = pa.Mesh.concat(parts)
        params = params_list[0]
        solver = pa.Solver(mesh, params_list, device=0)
    else:
        parts = [build_mesh(cfg["mesh"])]
        mesh = parts[0]
        pover = dict(cfg["params"])
        if args.precision:
            pover["precision"] = 1 if args.precision == "mixed" else 0
        if so is not None:
            pover["storage_order"] = so
        if fl is not None:
            pover["face_loop"] = fl
        if sf is not None:
            pover["sweep_form"] = sf
        params = pa.default_params(**pover)
        print(f"[bench] {mesh.n_cell} cells: creating the context", file=sys.stderr, flush=True)
        solver = pa.Solver(mesh, params, device=0)
    solver.set_launch_mode(args.launch)
    # the device's exact wall distance, for the CPU baseline's SA model (preprocessing)
    wdist = solver.get_field("wdist")[:mesh.n_cell].copy() if params.viscous == 3 else None
    solver.init_freestream()
    tuned = None
    if args.autotune or cfg.get("batch"):
        tuned = solver.autotune()
        solver.init_freestream()
        print(f"[bench] autotuned tiles {tuned}", file=sys.stderr, flush=True)
    print(f"[bench] {mesh.n_cell} cells, {mesh.n_face} faces; warmup {args.warmup}", file=sys.stderr, flush=True)
    solver.iterate(1, args.warmup, history=False)
    solver.synchronize()

    t0 = time.perf_counter()
    solver.iterate(1 + args.warmup, args.steps, history=False)
    solver.synchronize()
    wall = time.perf_counter() - t0
    value = mesh.n_cell * args.steps / wall
    print(f"[bench] timed {args.steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)

    # roofline of the face-flux kernel: same steps again with HIP events
    solver.set_kernel_timing(True)
    solver.iterate(1 + args.warmup + args.steps, args.steps, history=False)
    times, n_timed = solver.kernel_timing()
    solver.set_kernel_timing(False)
    flux_ms = times["flux"] / max(n_timed, 1)
    mixed = params.precision == 1
    fb = flux_algorithmic_bytes(mesh.n_cell, mesh.n_face, params.viscous, mixed)
    achieved = fb / (flux_ms * 1e-3) / 1e9
    roof_face = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                     frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                     kernel="k_face_flux + k_cell_gather", kernel_ms=round(flux_ms, 5), algorithmic_bytes=fb,
                     bytes_note="240N + 64F (+32N viscous; FP32 gradients: 180N + 64F)")
    per_class = {k: round(v / max(n_timed, 1), 5) for k, v in times.items()}
    sweep_ms = (times["sweep_fwd"] + times["sweep_bwd"]) / max(n_timed, 1)
    levels_form = params.sweep_form == 1
    # the LU-SGS pair's compulsory bytes as the hot-path survey prices them:
    # per sweep read res, q, dq, the off-diagonal sums (160N), the diagonal
    # (8N), normal and area (32F), connectivity (16F); write dq and the sums
    # (80N): 496N + 96F for both sweeps
    sb = 496 * mesh.n_cell + 96 * mesh.n_face
    sweep_gbs = sb / (sweep_ms * 1e-3) / 1e9
    roof_sweep = dict(bound="hbm", achieved=round(sweep_gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(sweep_gbs / HBM_PEAK_GBS, 4), traffic=None,
                      kernel="k_sweep_level_so + k_sweep_block_so (forward + backward)" if levels_form
                      else "k_sweep_flow (forward + backward launch)",
                      kernel_ms=round(sweep_ms, 5), algorithmic_bytes=sb, bytes_note="496N + 96F",
                      note="latency bound: each dependency level is a launch with two dependent memory rounds"
                      if levels_form else "latency bound: one launch per direction; each dependency level costs "
                      "a flag round trip, the neighbours' dq loads and the store acknowledgement")
    print(f"[bench] mean ms per step by kernel class: {per_class}", file=sys.stderr, flush=True)
    # the dominant kernel of the step carries the contract's roofline object
    roof, other = (roof_sweep, roof_face) if sweep_ms >= flux_ms else (roof_face, roof_sweep)
    copy_gbs = round(pa.hbm_copy_gbps(0, COPY_BYTES, 10), 1)
    roof["copy_probe_gbs"] = copy_gbs
    if not args.no_traffic and not cfg.get("batch"):
        tr = measure_traffic(args.config, ("k_sweep_flow", "k_face_flux", "k_cell_gather"))
        if tr:
            def total(prefixes):
                ks = [k for k in tr["kernels"] if any(k.startswith(p) for p in prefixes)]
                return sum(sum(tr["kernels"][k].values()) for k in ks) if ks else None
            roof_sweep["traffic"] = total(("k_sweep_flow",))
            roof_face["traffic"] = total(("k_face_flux", "k_cell_gather"))
            roof["traffic_per_kernel"] = {k: {c: round(v) for c, v in d.items()} for k, d in tr["kernels"].items()}
            roof["traffic_calibration"] = {c: round(f, 4) for c, f in tr["factors"].items()}

    line = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 face fluxes + f64" if mixed else "f64",
        "data": cfg.get("data", "synthetic"),
        "config": {"workload": cfg["workload"], "cells": mesh.n_cell, "faces": mesh.n_face,
                   "cases": mesh.n_cases, "precision": "mixed" if mixed else "fp64",
                   "parallelism": "single GPU", "launch": args.launch,
                   "sweep_form": ["dataflow", "levels", "ticket", "xcd"][params.sweep_form],
                   "tiles": None if tuned is None else {"flux": tuned["flux_tile"], "grad": tuned["grad_tile"]}},
        "roofline": roof,
        "roofline_other": other,
        "kernel_ms_per_step": per_class,
    }
    if args.warm_start:
        from phenlei_amd import warmstart as ws
        n_train, n_held = (int(v) for v in args.ws_cases.split(","))
        dims = tuple(int(v) for v in args.ws_dims.split(","))
        stages = [int(v) for v in args.ws_steps.split(",")]
        lrs = [2e-3, 1e-3, 3e-4, 1e-4][:len(stages)]
        line["warm_start"] = ws.compare(n_train=n_train, n_held=n_held, dims=dims, train_steps=stages, train_lr=lrs,
                                        log=lambda m: print("[bench] warm start: " + m, file=sys.stderr, flush=True))
    if args.mpnn:
        solver.close()
        c, nx, ny, nz = (int(v) for v in args.mpnn.split(","))
        line["mpnn"] = mpnn_throughput(c, (nx, ny, nz), args.mpnn_steps)
    if args.precision_compare is not None and not cfg.get("batch"):
        solver.close()
        line["precision_compare"] = precision_compare(mesh, pover, args.precision_compare)
    conv = cfg.get("converge")
    if conv and args.converge is None and not args.no_converge:
        solver.close()
        line["time_to_drop"] = wing_time_to_drop(mesh, pover, conv["drop"], conv["max_steps"], WING_WEIGHTS)
        line["time_to_drop"]["target_note"] = (
            f"drop {conv['drop']} x R_1 within {conv['max_steps']} steps: on this mesh the plain FP64 run needs "
            "about 4,400 steps for a 0.25 drop (the residual RMS is dominated by the large far-field cells, whose "
            "relaxation sets the rate), so a 1e-8 drop does not fit the run")
    if args.converge is not None:
        solver.close()
        if cfg.get("batch"):
            line["time_to_drop"] = batch_time_to_drop(mesh, params_list, args.converge, args.max_steps,
                                                      forms=tuple(args.forms.split(",")))
        else:
            line["time_to_drop"] = time_to_drop(mesh, pover, args.converge, args.max_steps)
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(parts[0], params, cfg["cpu_sample"], wdist)
    solver.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
