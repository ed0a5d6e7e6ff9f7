"""GPU parity: every hook of the path against the CPU oracle on identical
meshes and states.  FP64 bar: per-call residuals within 1e-12 relative to the
local flux magnitude; the kernels are written to round
like the oracle, so most fields agree exactly."""
import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc
from conftest import perturbed_state

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel_err(a, b, floor):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


def make_pair(nx, ny, nz, bump=0.05, **pover):
    m = pa.Mesh.bump_hex(nx, ny, nz, bump)
    conn, met = m.connectivity(), m.metrics()
    p = pa.default_params(**pover)
    gpu = pa.Solver(m, p)
    cpu = orc.OracleSolver(conn, met, p)
    return m, gpu, cpu


def seeded(m, cpu, seed):
    q = perturbed_state(m.n_total, m.n_cell, cpu.freestream(), seed)
    return q


@pytest.mark.parametrize("limiter", [0, 1, 2])
def test_residual_parity(limiter):
    m, gpu, cpu = make_pair(16, 8, 4, limiter=limiter)
    q = seeded(m, cpu, 1 + limiter)
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.residual()
    cpu.residual()
    nc = m.n_cell
    qg, qc = gpu.get_state(), cpu.get("q")
    assert np.max(np.abs(qg - qc)) <= 1e-14  # ghost cells from the BCs
    for name in ("dqdx", "dqdy", "dqdz"):
        g, c = gpu.get_field(name)[:, :nc], cpu.get(name)[:, :nc]
        assert rel_err(g, c, 1e-12 * np.max(np.abs(c))) <= TOL, name
    lg, lc = gpu.get_field("limit"), cpu.get("limit")
    assert np.max(np.abs(lg - lc)) <= 1e-13
    rg, rc = gpu.get_residual()[:, :nc], cpu.get("res")[:, :nc]
    floor = 1e-3 * np.max(np.abs(rc), axis=1, keepdims=True)
    assert rel_err(rg, rc, floor) <= TOL
    cpu.spectrum()
    sg, sc = gpu.get_field("spec_inv")[:nc], cpu.get("spec_inv")[:nc]
    assert rel_err(sg, sc, 1e-300) <= TOL
    print("limiter", limiter, "max |res diff|", np.max(np.abs(rg - rc)), "bitwise equal:",
          np.array_equal(rg, rc), "grad bitwise:", np.array_equal(gpu.get_field("dqdx")[:, :nc],
                                                                 cpu.get("dqdx")[:, :nc]))


def test_time_step_and_lusgs_parity():
    m, gpu, cpu = make_pair(16, 8, 4, cfl_start=10.0, cfl_end=10.0)
    q = seeded(m, cpu, 7)
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.residual()
    cpu.residual()
    cpu.spectrum()
    gpu.time_step(3)
    cpu.time_step(3)
    nc = m.n_cell
    for name in ("dt",):
        g, c = gpu.get_field(name)[:nc], cpu.get(name)[:nc]
        assert rel_err(g, c, 1e-300) <= TOL, name
    nfix = gpu.lusgs()
    nfix_c = cpu.lusgs()
    assert nfix == nfix_c
    for name in ("diag",):
        g, c = gpu.get_field(name)[:nc], cpu.get(name)[:nc]
        assert rel_err(g, c, 1e-300) <= TOL, name
    dg, dc = gpu.get_field("dq"), cpu.get("dq")
    floor = 1e-3 * np.max(np.abs(dc[:, :nc]), axis=1, keepdims=True)
    assert rel_err(dg[:, :nc], dc[:, :nc], floor) <= TOL
    qg, qc = gpu.get_state()[:, :nc], cpu.get("q")[:, :nc]
    assert rel_err(qg, qc, 1e-300) <= TOL


@pytest.mark.parametrize("dims", [(16, 8, 4), (64, 32, 8)])
def test_iterations_parity(dims):
    m, gpu, cpu = make_pair(*dims, cfl_start=10.0, cfl_end=10.0)
    gpu.init_freestream()
    cpu.init_freestream()
    n = 10
    hg = gpu.iterate(1, n)
    hc, _ = cpu.iterate(1, n)
    assert np.max(np.abs(hg - hc) / hc) <= 1e-11
    nc = m.n_cell
    qg, qc = gpu.get_state()[:, :nc], cpu.get("q")[:, :nc]
    dq = qc - cpu.freestream()[:, None]
    err = np.max(np.abs(qg - qc) / np.maximum(np.max(np.abs(dq), axis=1, keepdims=True), 1e-300))
    print("rms hist", hg[:3], "state rel err", err)
    assert err <= 1e-10


def test_deterministic():
    m, gpu, cpu = make_pair(32, 16, 8, cfl_start=10.0, cfl_end=10.0)
    gpu.init_freestream()
    h1 = gpu.iterate(1, 5)
    q1 = gpu.get_state()
    gpu.init_freestream()
    h2 = gpu.iterate(1, 5)
    q2 = gpu.get_state()
    assert np.array_equal(h1, h2)
    assert np.array_equal(q1, q2)


def make_hybrid_pair(**pover):
    m = pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)
    conn, met = m.connectivity(), m.metrics()
    p = pa.default_params(**pover)
    return m, pa.Solver(m, p), orc.OracleSolver(conn, met, p)


@pytest.mark.parametrize("viscous", [0, 1])
def test_hybrid_mesh_residual_parity(viscous):
    m, gpu, cpu = make_hybrid_pair(mach=0.3, viscous=viscous, reynolds=5e5, cfl_start=10.0, cfl_end=10.0)
    q = seeded(m, cpu, 21 + viscous)
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.residual()
    cpu.residual()
    cpu.spectrum()
    nc = m.n_cell
    rg, rc = gpu.get_residual()[:, :nc], cpu.get("res")[:, :nc]
    floor = 1e-3 * np.max(np.abs(rc), axis=1, keepdims=True)
    assert rel_err(rg, rc, floor) <= TOL
    for name in ("spec_inv", "spec_vis"):
        g, c = gpu.get_field(name)[:nc], cpu.get(name)[:nc]
        assert rel_err(g, c, 1e-300) <= TOL, name
    if viscous:
        for name in ("dtdx", "dtdy", "dtdz"):
            g, c = gpu.get_field(name)[:nc], cpu.get(name)[:nc]
            assert rel_err(g, c, 1e-12 * np.max(np.abs(c))) <= TOL, name
    gpu.time_step(2)
    cpu.time_step(2)
    assert gpu.lusgs() == cpu.lusgs()
    qg, qc = gpu.get_state()[:, :nc], cpu.get("q")[:, :nc]
    assert rel_err(qg, qc, 1e-300) <= TOL


@pytest.mark.parametrize("viscous", [0, 1])
def test_hybrid_mesh_iterations_parity(viscous):
    m, gpu, cpu = make_hybrid_pair(mach=0.3, viscous=viscous, reynolds=5e5, cfl_start=5.0, cfl_end=5.0)
    gpu.init_freestream()
    cpu.init_freestream()
    hg = gpu.iterate(1, 8)
    hc, _ = cpu.iterate(1, 8)
    assert np.max(np.abs(hg - hc) / hc) <= 1e-10
    nc = m.n_cell
    qg, qc = gpu.get_state()[:, :nc], cpu.get("q")[:, :nc]
    dq = qc - cpu.freestream()[:, None]
    err = np.max(np.abs(qg - qc) / np.maximum(np.max(np.abs(dq), axis=1, keepdims=True), 1e-300))
    print("hybrid viscous", viscous, "state rel err", err)
    assert err <= 1e-9


@pytest.mark.parametrize("viscous", [0, 1])
def test_forces_parity(viscous):
    """Wall force and moment coefficients after a few iterations on the hybrid
    bump, against the oracle's AirForceCoefParts + CompForce.  The GPU sums
    the wall faces in a fixed tree order, the oracle in face order, so the
    sums agree to rounding (1e-12 of the largest coefficient)."""
    m, gpu, cpu = make_hybrid_pair(mach=0.3, viscous=viscous, reynolds=5e5, cfl_start=5.0, cfl_end=5.0,
                                   alpha_deg=2.0, force_ref_x=0.25, force_ref_y=0.1, force_ref_z=0.05)
    gpu.init_freestream()
    cpu.init_freestream()
    gpu.iterate(1, 5)
    cpu.iterate(1, 5)
    fg = gpu.forces()
    fc = cpu.forces()
    vals = np.array([fg[k] for k in pa.FORCE_NAMES])
    scale = np.max(np.abs(fc[:15]))
    assert scale > 0
    print("forces viscous", viscous, "CL", fg["CL"], "CD", fg["CD"], "max diff", np.max(np.abs(vals - fc)))
    assert np.max(np.abs(vals[:21] - fc[:21])) <= 1e-12 * scale
    assert vals[21] == pytest.approx(fc[21], rel=1e-9, abs=1e-12)
    # the forces call refreshes derived fields only: the state is unchanged and
    # the next iterations still match
    hg = gpu.iterate(6, 2)
    hc, _ = cpu.iterate(6, 2)
    assert np.max(np.abs(hg - hc) / hc) <= 1e-10


def test_graph_replay_equals_eager_launch():
    """phl_iterate replays captured hipGraphs of 8 steps by default, one per
    assignment of the double-buffered state.  Calls of 9, 8 and 11 steps start
    on both buffer parities, so both graphs are captured and replayed and the
    eager tails run too.  Every replay equals the eager launches bit for bit,
    and the launch counters show the graphs really ran."""
    m = pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)
    p = pa.default_params(mach=0.3, viscous=1, reynolds=5e5, cfl_start=5.0, cfl_end=5.0)
    a, b = pa.Solver(m, p), pa.Solver(m, p)
    a.set_launch_mode("eager")
    a.init_freestream()
    b.init_freestream()
    for first, n in ((1, 9), (10, 8), (18, 11)):
        ha = a.iterate(first, n)
        hb = b.iterate(first, n)
        assert np.array_equal(ha, hb)
        assert np.array_equal(a.get_state(), b.get_state())
    assert a.launch_counts() == (0, 28)
    assert b.launch_counts() == (3, 4)  # 8+1, 8, 8+3 steps


@pytest.mark.parametrize("viscous", [0, 1])
def test_mixed_precision_residual(viscous):
    """Mixed form (FP32 gradients, reconstruction, Roe and viscous fluxes,
    FP64 accumulation) against the FP64 oracle on a perturbed state: the
    residual agrees to FP32 rounding of the face fluxes, and the spectral
    radii and LU-SGS inputs, which stay FP64, agree exactly as in FP64."""
    m, gpu, cpu = make_hybrid_pair(mach=0.3, viscous=viscous, reynolds=5e5, precision=1)
    q = seeded(m, cpu, 31 + viscous)
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.residual()
    cpu.residual()
    cpu.spectrum()
    nc = m.n_cell
    rg, rc = gpu.get_residual()[:, :nc], cpu.get("res")[:, :nc]
    scale = np.max(np.abs(rc), axis=1, keepdims=True)
    err = np.max(np.abs(rg - rc) / scale)
    print("mixed residual, viscous", viscous, "max err / max |res| =", err)
    assert err <= 1e-4
    assert err > 0.0  # the FP32 path really ran
    sg, sc = gpu.get_field("spec_inv")[:nc], cpu.get("spec_inv")[:nc]
    assert rel_err(sg, sc, 1e-300) <= TOL
    gg, gc = gpu.get_field("dqdx")[:, :nc], cpu.get("dqdx")[:, :nc]
    assert np.max(np.abs(gg - gc)) <= 1e-6 * np.max(np.abs(gc))  # FP32 storage


def test_mixed_precision_forces_track_fp64():
    """50 iterations from freestream in the FP64 and the mixed form: lift and
    drag agree within 1e-4 (absolute coefficient), the density residual
    histories within 1e-3 relative; switching precision on a live context
    continues the run."""
    m = pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)
    base = dict(mach=0.3, viscous=1, reynolds=5e5, cfl_start=5.0, cfl_end=5.0)
    a = pa.Solver(m, pa.default_params(**base))
    b = pa.Solver(m, pa.default_params(precision=1, **base))
    a.init_freestream()
    b.init_freestream()
    ha = a.iterate(1, 50)
    hb = b.iterate(1, 50)
    fa, fb = a.forces(), b.forces()
    print("CL", fa["CL"], fb["CL"], "CD", fa["CD"], fb["CD"], "hist rel", np.max(np.abs(hb - ha) / ha))
    assert abs(fa["CL"] - fb["CL"]) <= 1e-4 and abs(fa["CD"] - fb["CD"]) <= 1e-4
    assert np.max(np.abs(hb - ha) / ha) <= 1e-3
    b.set_precision("fp64")
    hb2 = b.iterate(51, 4)
    assert np.all(np.isfinite(hb2))


def _batch_cases():
    meshes = [pa.Mesh.bump_hex(12, 6, 4, 0.05), pa.Mesh.bump_hex(16, 5, 3, 0.03), pa.Mesh.bump_hex(8, 8, 2, 0.07)]
    params = [pa.default_params(mach=0.5, alpha_deg=0.0, cfl_start=5.0, cfl_end=5.0),
              pa.default_params(mach=0.35, alpha_deg=3.0, cfl_start=5.0, cfl_end=5.0),
              pa.default_params(mach=0.75, alpha_deg=-1.5, cfl_start=5.0, cfl_end=5.0)]
    return meshes, params


def test_batch_equals_separate_contexts():
    """Three cases of different size, Mach, angle and bump height in one
    batched context: each case's state after 12 iterations equals its own
    context's bit for bit, residual histories and forces agree to rounding."""
    meshes, params = _batch_cases()
    u = pa.Mesh.concat(meshes)
    batch = pa.Solver(u, params)
    batch.init_freestream()
    hb = batch.iterate(1, 12)
    qb = batch.get_state()
    for k, (m, p) in enumerate(zip(meshes, params)):
        s = pa.Solver(m, p)
        s.init_freestream()
        hs = s.iterate(1, 12)
        q = s.get_state()
        cs, gs = u.case_slices(k)
        assert np.array_equal(qb[:, cs], q[:, :m.n_cell]), k
        assert np.array_equal(qb[:, gs], q[:, m.n_cell:]), k
        assert np.max(np.abs(hb[:, k] - hs) / hs) <= 1e-12, k
        fb, fs = batch.forces_case(k), s.forces()
        assert abs(fb["CL"] - fs["CL"]) <= 1e-13 and abs(fb["CD"] - fs["CD"]) <= 1e-13, k


def test_batch_solve_freezes_converged_cases():
    meshes, params = _batch_cases()
    u = pa.Mesh.concat(meshes)
    batch = pa.Solver(u, params)
    batch.init_freestream()
    steps, hist = batch.solve_batch(320, 0.2)
    print("steps to 0.2 drop per case", steps)
    assert np.all(steps > 0)
    qb = batch.get_state()
    for k, (m, p) in enumerate(zip(meshes, params)):
        # frozen at the end of the 32-step chunk in which it converged
        n_run = int(np.ceil(steps[k] / 32.0)) * 32
        s = pa.Solver(m, p)
        s.init_freestream()
        h = s.iterate(1, n_run)
        assert np.argmax(h / h[0] <= 0.2) + 1 == steps[k]
        cs, _ = u.case_slices(k)
        assert np.array_equal(qb[:, cs], s.get_state()[:, :m.n_cell]), k


def test_autotuned_tiles_keep_the_state():
    """The autotuner's workgroup tiles regroup only the residual partial sums:
    with the tiles it picked, the state is bit-identical to the default
    tiles and the history agrees to rounding."""
    m = pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)
    p = pa.default_params(mach=0.3, viscous=1, reynolds=5e5, cfl_start=5.0, cfl_end=5.0)
    a, b = pa.Solver(m, p), pa.Solver(m, p)
    a.init_freestream()
    b.init_freestream()
    tuned = b.autotune()
    print("autotune", tuned)
    assert tuned["flux_tile"] in (64, 128, 256) and tuned["grad_tile"] in (64, 128, 256)
    b.init_freestream()
    ha, hb = a.iterate(1, 6), b.iterate(1, 6)
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.max(np.abs(ha - hb) / ha) <= 1e-12


@pytest.mark.parametrize("viscous", [1, 3])
def test_storage_order_does_not_change_results(viscous):
    """Cells stored in locality order (default) or in the mesh's order give
    the same states bit for bit: every accumulation and sweep follows the
    mesh's numbering; only the residual-norm partial sums regroup."""
    m = pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)
    base = dict(mach=0.3, viscous=viscous, reynolds=5e5, cfl_start=5.0, cfl_end=5.0)
    a = pa.Solver(m, pa.default_params(storage_order=0, **base))
    b = pa.Solver(m, pa.default_params(storage_order=1, **base))
    for s in (a, b):
        s.init_freestream()
    ha, hb = a.iterate(1, 12), b.iterate(1, 12)
    assert np.max(np.abs(ha - hb) / hb) <= 1e-14
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(a.get_field("dqdx"), b.get_field("dqdx"))
    if viscous == 3:
        assert np.array_equal(a.get_field("nut"), b.get_field("nut"))
        assert np.array_equal(a.get_field("wdist"), b.get_field("wdist"))


def test_solve_switches_a_stalled_mixed_context_to_fp64():
    """phl_set_precision_switch: the mixed form stalls near its FP32 floor;
    phl_solve then switches the context to FP64, which carries the residual
    below that floor."""
    m = pa.Mesh.bump_hex(24, 12, 4, 0.05)
    s = pa.Solver(m, pa.default_params(precision=1, cfl_start=10.0, cfl_end=10.0))
    s.set_precision_switch(256)
    s.init_freestream()
    n, h = s.solve(6000, 1e-8)
    print("steps", n, "final drop", h[-1] / h[0], "precision", s.precision())
    assert s.precision() == "fp64"
    assert h[-1] / h[0] <= 1e-7


def test_frozen_case_keeps_its_state():
    a, b = pa.Mesh.bump_hex(8, 4, 2, 0.05), pa.Mesh.bump_hex(6, 4, 2, 0.03)
    u = pa.Mesh.concat([a, b])
    s = pa.Solver(u, [pa.default_params(mach=0.5), pa.default_params(mach=0.6)])
    s.init_freestream()
    s.iterate(1, 5)
    q0 = s.get_state()
    s.set_frozen([True, False])
    s.iterate(6, 5)
    q1 = s.get_state()
    cs0, gs0 = u.case_slices(0)
    cs1, _ = u.case_slices(1)
    assert np.array_equal(q0[:, cs0], q1[:, cs0])
    assert not np.array_equal(q0[:, cs1], q1[:, cs1])


@pytest.mark.parametrize("viscous", [0, 1])
def test_face_loop_forms_agree(viscous):
    """The tile-fused face loop (face values in LDS) and the split one (face
    kernel + cell gather through HBM) give identical residuals and states;
    only the residual norm's partial sums are grouped differently."""
    m = pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)
    base = dict(mach=0.3, viscous=viscous, reynolds=5e5, cfl_start=5.0, cfl_end=5.0)
    a = pa.Solver(m, pa.default_params(face_loop=1, **base))
    b = pa.Solver(m, pa.default_params(face_loop=0, **base))
    for s in (a, b):
        s.init_freestream()
    ha, hb = a.iterate(1, 10), b.iterate(1, 10)
    assert np.max(np.abs(ha - hb) / hb) <= 1e-14  # the norm's block partials group 128 vs 256 cells
    assert np.array_equal(a.get_state(), b.get_state())
    assert np.array_equal(a.get_residual(), b.get_residual())


@pytest.mark.parametrize("kind,precision,form", [("hex", 0, 0), ("hybrid", 0, 0), ("hybrid", 1, 0), ("wing", 0, 0),
                                                 ("hybrid", 0, 2), ("wing", 1, 2), ("hex", 0, 3), ("hybrid", 1, 3),
                                                 ("wing", 0, 3)])
def test_dataflow_sweep_equals_level_sweep(kind, precision, form):
    """The dataflow sweeps (one launch per direction, per-chunk flags; chunks
    in launch order, form 0, ticket order, form 2, or split over the XCDs,
    form 3) and the level-by-level
    launches (form 1) apply the same arithmetic in the same order: states,
    residual histories and dq are bitwise equal, over several calls (the
    flags carry a new epoch each sweep) and in both launch modes."""
    if kind == "hex":
        m = pa.Mesh.bump_hex(48, 16, 6, 0.08)
    elif kind == "wing":
        m = pa.Mesh.swept_wing(12, 6, 5, 1e-3, 3)
    else:
        m = pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)
    base = dict(mach=0.3, viscous=0 if kind == "hex" else 1, reynolds=5e5, cfl_start=5.0, cfl_end=5.0,
                precision=precision)
    a = pa.Solver(m, pa.default_params(sweep_form=form, **base))
    b = pa.Solver(m, pa.default_params(sweep_form=1, **base))
    for s in (a, b):
        s.init_freestream()
    for mode, first, n in (("eager", 1, 5), ("graph", 6, 17)):
        a.set_launch_mode(mode)
        b.set_launch_mode(mode)
        ha, hb = a.iterate(first, n), b.iterate(first, n)
        assert np.array_equal(ha, hb), mode
        assert np.array_equal(a.get_state(), b.get_state()), mode
        assert np.array_equal(a.get_field("dq"), b.get_field("dq")), mode


def test_dataflow_sweep_batch_equals_level_sweep():
    """A batched context (levels merged over the cases) in the dataflow form."""
    parts = [pa.Mesh.bump_hex(24 + 8 * i, 8, 4, 0.08) for i in range(3)]
    u = pa.Mesh.concat(parts)
    out = []
    for form in (0, 1, 3):
        s = pa.Solver(u, [pa.default_params(cfl_start=5.0, cfl_end=5.0, sweep_form=form) for _ in parts])
        s.init_freestream()
        out.append((s.iterate(1, 12), s.get_state()))
    for k in (1, 2):
        assert np.array_equal(out[0][0], out[k][0])
        assert np.array_equal(out[0][1], out[k][1])


@pytest.mark.parametrize("kind,viscous", [("hex", 0), ("hybrid", 0), ("hybrid", 1)])
def test_least_squares_gradient_parity(kind, viscous):
    """gradientName "lsq": per-call gradients, temperature gradients,
    limiter, residual and one LU-SGS step against the oracle (ghosts take
    their owner's gradient, which the boundary faces' reconstruction and
    viscous terms read); then iterations."""
    pover = dict(mach=0.3, viscous=viscous, reynolds=5e5, cfl_start=5.0, cfl_end=5.0, gradient=1)
    if kind == "hex":
        m, gpu, cpu = make_pair(16, 8, 4, bump=0.08, **pover)
    else:
        m, gpu, cpu = make_hybrid_pair(**pover)
    q = seeded(m, cpu, 41 + viscous)
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.residual()
    cpu.residual()
    nc = m.n_cell
    for name in ("dqdx", "dqdy", "dqdz"):
        g, c = gpu.get_field(name)[:, :nc], cpu.get(name)[:, :nc]
        assert rel_err(g, c, 1e-12 * np.max(np.abs(c))) <= TOL, name
    if viscous:
        for name in ("dtdx", "dtdy", "dtdz"):
            g, c = gpu.get_field(name)[:nc], cpu.get(name)[:nc]
            assert rel_err(g, c, 1e-12 * np.max(np.abs(c))) <= TOL, name
    assert np.max(np.abs(gpu.get_field("limit") - cpu.get("limit"))) <= 1e-13
    rg, rc = gpu.get_residual()[:, :nc], cpu.get("res")[:, :nc]
    floor = 1e-3 * np.max(np.abs(rc), axis=1, keepdims=True)
    err_r = rel_err(rg, rc, floor)
    print(kind, viscous, "lsq residual rel err", err_r, "bitwise:", np.array_equal(rg, rc))
    assert err_r <= TOL
    gpu.time_step(2)
    cpu.spectrum()
    cpu.time_step(2)
    assert gpu.lusgs() == cpu.lusgs()
    qg, qc = gpu.get_state()[:, :nc], cpu.get("q")[:, :nc]
    assert rel_err(qg, qc, 1e-300) <= TOL
    gpu.init_freestream()
    cpu.init_freestream()
    hg = gpu.iterate(1, 8)
    hc, _ = cpu.iterate(1, 8)
    assert np.max(np.abs(hg - hc) / hc) <= 1e-10


@pytest.mark.parametrize("dims,viscous,form", [((1, 1, 1), 0, 0), ((2, 1, 1), 1, 0), ((1, 2, 3), 0, 0),
                                               ((1, 1, 1), 0, 1), ((1, 2, 3), 1, 2), ((1, 1, 1), 0, 3),
                                               ((1, 2, 3), 1, 3)])
def test_tiny_meshes_match_oracle(dims, viscous, form):
    """Edge sizes: one cell with only boundary faces (no interior face, one
    level, one chunk), two and six cells; every sweep form; against the
    oracle over a few iterations."""
    m, gpu, cpu = make_pair(*dims, mach=0.4, viscous=viscous, reynolds=1e4, cfl_start=2.0, cfl_end=2.0,
                            sweep_form=form)
    gpu.init_freestream()
    cpu.init_freestream()
    q = seeded(m, cpu, 7)
    gpu.set_state(q)
    cpu.set("q", q)
    hg = gpu.iterate(1, 4)
    hc, _ = cpu.iterate(1, 4)
    nc = m.n_cell
    assert np.all(np.isfinite(hg))
    assert np.max(np.abs(hg - hc) / np.maximum(hc, 1e-300)) <= 1e-10
    qg, qc = gpu.get_state()[:, :nc], cpu.get("q")[:, :nc]
    assert np.max(np.abs(qg - qc) / np.maximum(np.abs(qc), 1e-300)) <= 1e-12


@pytest.mark.parametrize("mach,alpha,spread", [(0.5, 3.0, 0.8), (2.0, 10.0, 0.3)])
def test_farfield_ghosts_all_regimes(mach, alpha, spread):
    """Farfield ghost states on the GPU equal the oracle's (pinned by the
    Riemann-invariant construction) in every regime: subsonic inflow and
    outflow (reversed interior flow at the outlet), supersonic inflow and
    outflow."""
    m, gpu, cpu = make_pair(6, 3, 2, mach=mach, alpha_deg=alpha)
    qo = cpu.freestream()
    rng = np.random.default_rng(17)
    q = np.tile(qo[:, None], (1, m.n_total))
    nc = m.n_cell
    q[0, :nc] *= rng.uniform(0.9, 1.1, nc)
    q[1:4, :nc] += rng.uniform(-spread, spread, (3, nc)) * np.linalg.norm(qo[1:4])
    q[4, :nc] *= rng.uniform(0.9, 1.1, nc)
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.boundary()
    cpu.boundary()
    qg, qc = gpu.get_state()[:, nc:], cpu.get("q")[:, nc:]
    assert np.max(np.abs(qg - qc) / np.maximum(np.abs(qc), 1e-300)) <= 1e-13


@pytest.mark.parametrize("limiter", [0, 1])
def test_residual_parity_with_first_order_fallback(limiter):
    """A rough state (rho and p over two decades) makes many reconstructed face
    states non-positive, so the MUSCL first-order fallback
    (PositiveCheckForDensityAndPressure, FaceProxy.h:209) is taken on many
    sides; the residual still equals the oracle's."""
    m, gpu, cpu = make_pair(8, 6, 3, mach=0.5, limiter=limiter)
    rng = np.random.default_rng(4)
    q = np.tile(cpu.freestream()[:, None], (1, m.n_total))
    nc = m.n_cell
    q[0, :nc] = rng.uniform(0.02, 2.0, nc)
    q[4, :nc] *= rng.uniform(0.02, 2.0, nc)
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.residual()
    cpu.residual()
    rg, rc = gpu.get_residual()[:, :nc], cpu.get("res")[:, :nc]
    scale = np.max(np.abs(rc), axis=1, keepdims=True)
    assert np.max(np.abs(rg - rc) / scale) <= TOL


@pytest.mark.parametrize("t_wall", [350.0, 20.0])
def test_isothermal_wall_matches_oracle(t_wall):
    """wallTemperature > 0: ghost states, residual and iterations on the
    hybrid mesh with a laminar isothermal wall (and one cold enough to hit
    the ghost-temperature floor)."""
    m, gpu, cpu = make_hybrid_pair(mach=0.3, viscous=1, reynolds=5e5, cfl_start=5.0, cfl_end=5.0, t_wall=t_wall)
    q = seeded(m, cpu, 33)
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.residual()
    cpu.residual()
    nc = m.n_cell
    qg, qc = gpu.get_state(), cpu.get("q")
    assert np.max(np.abs(qg[:, nc:] - qc[:, nc:]) / np.abs(qc[:, nc:]).max(axis=1, keepdims=True)) <= 1e-15
    rg, rc = gpu.get_residual()[:, :nc], cpu.get("res")[:, :nc]
    assert rel_err(rg, rc, 1e-3 * np.max(np.abs(rc), axis=1, keepdims=True)) <= TOL
    gpu.init_freestream()
    cpu.init_freestream()
    hg = gpu.iterate(1, 10)
    hc, _ = cpu.iterate(1, 10)
    assert np.max(np.abs(hg - hc) / hc) <= 1e-12
