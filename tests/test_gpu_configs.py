"""GPU parity at the configured sizes and forms, against the CPU oracle:

- the batched form (seeded cases of the 64-case sweep: uneven sizes, Mach,
  angle and bump height) in FP64 and in the mixed form, with autotuned tiles
  and graph replay, each case against an oracle run of that case alone;
- one full step of the path (residual, time step, LU-SGS, and for RANS the
  turbulence relaxation) on the full 1,007,616-cell laminar hybrid channel
  and on the full 3,983,616-cell swept wing at M0.6, AoA 2 deg, Re 3e6.

FP64 bar: 1e-12 relative per call (floors at 1e-3 of each field's largest
value), 1e-10 over iterations.  Mixed form (FP32 face fluxes): the per-case
density-residual histories within 1e-3 relative of the FP64 oracle, state
deviations from freestream within 1e-3 of their largest value, lift and
drag within 1e-4."""
import numpy as np
import pytest

import bench
import phenlei_amd as pa
import oracle as orc
from conftest import perturbed_state

pytestmark = pytest.mark.gpu

TOL = 1e-12
CL, CD = pa.FORCE_NAMES.index("CL"), pa.FORCE_NAMES.index("CD")


def rel_err(a, b, floor):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


def _sweep_subset(n_cases, shrink):
    """The first n_cases of the seeded 64-case sweep with every grid
    dimension divided by `shrink` (sizes stay uneven)."""
    meshes, params = [], []
    for nx, ny, nz, h, mach, alpha in bench.c4_cases()[:n_cases]:
        dims = (max(4, nx // shrink), max(4, ny // shrink), max(2, nz // shrink))
        meshes.append(pa.Mesh.bump_hex(*dims, h))
        params.append(dict(mach=mach, alpha_deg=alpha, cfl_start=10.0, cfl_end=10.0))
    return meshes, params


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_sweep_matches_oracle_per_case(precision):
    """Six seeded cases in one batched context (autotuned tiles, 16 steps as
    two replayed 8-step graphs); every case against the FP64 oracle of that
    case alone."""
    meshes, pover = _sweep_subset(6, 4)
    u = pa.Mesh.concat(meshes)
    batch = pa.Solver(u, [pa.default_params(precision=precision, **p) for p in pover])
    batch.init_freestream()
    batch.autotune()
    batch.init_freestream()
    n = 16
    hb = batch.iterate(1, n)
    graphs, _ = batch.launch_counts()
    assert graphs >= 2
    qb = batch.get_state()
    for k, (m, p) in enumerate(zip(meshes, pover)):
        cpu = orc.OracleSolver(m.connectivity(), m.metrics(), pa.default_params(**p))
        cpu.init_freestream()
        hc, _ = cpu.iterate(1, n)
        cs, _ = u.case_slices(k)
        qc = cpu.get("q")[:, :m.n_cell]
        dev = np.max(np.abs(qc - cpu.freestream()[:, None]), axis=1, keepdims=True)
        herr = np.max(np.abs(hb[:, k] - hc) / hc)
        qerr = np.max(np.abs(qb[:, cs] - qc) / dev)
        fb = batch.forces_case(k)
        fc = cpu.forces()
        dcl, dcd = abs(fb["CL"] - fc[CL]), abs(fb["CD"] - fc[CD])
        print(f"case {k} ({m.n_cell} cells, M{p['mach']:.3f}, a{p['alpha_deg']:.2f}) precision {precision}: "
              f"history {herr:.2e}, state {qerr:.2e}, dCL {dcl:.2e}, dCD {dcd:.2e}")
        if precision == 0:
            assert herr <= 1e-10 and qerr <= 1e-10, k
            assert dcl <= 1e-12 and dcd <= 1e-12, k
        else:
            assert 0.0 < herr <= 1e-3 and qerr <= 1e-3, k
            assert dcl <= 1e-4 and dcd <= 1e-4, k
        cpu.close()


def _one_step(name, seed):
    cfg = bench.CONFIGS[name]
    m = bench.build_mesh(cfg["mesh"])
    p = pa.default_params(**cfg["params"])
    nc, nt = m.n_cell, m.n_total
    print(f"{name}: {nc} cells, {m.n_face} faces", flush=True)
    gpu = pa.Solver(m, p)
    cpu = orc.OracleSolver(m.connectivity(), m.metrics(), p)
    q = perturbed_state(nt, nc, cpu.freestream(), seed, amp=0.01)
    gpu.set_state(q)
    cpu.set("q", q)
    sa = p.viscous == 3
    if sa:
        # the oracle takes the device's exact wall distance (bitwise equal to
        # the oracle's own on the smaller meshes, see test_gpu_sa.py)
        cpu.set_wall_distance(gpu.get_field("wdist")[:nc])
        rng = np.random.default_rng(seed)
        nut = 3.0 * (1.0 + rng.random(nt))
        vist = 1e-3 + 0.1 * rng.random(nt)
        for k, v in (("nut", nut), ("vist", vist)):
            gpu.set_field(k, v)
            cpu.set(k, v)
    gpu.residual()
    cpu.residual()
    rg, rc = gpu.get_residual()[:, :nc], cpu.get("res")[:, :nc]
    err = rel_err(rg, rc, 1e-3 * np.max(np.abs(rc), axis=1, keepdims=True))
    print(f"{name}: residual rel err {err:.2e}, bitwise equal: {np.array_equal(rg, rc)}", flush=True)
    assert err <= TOL
    assert np.array_equal(gpu.residual_norm(), cpu.residual_norm()) or \
        np.max(np.abs(gpu.residual_norm() - cpu.residual_norm()) / cpu.residual_norm()) <= 1e-12
    step = 2 * p.cfl_vary_step  # the end of the CFL ramp
    gpu.spectrum_radius()
    cpu.spectrum()
    gpu.time_step(step)
    cpu.time_step(step)
    assert rel_err(gpu.get_field("dt")[:nc], cpu.get("dt")[:nc], 1e-300) <= TOL
    assert gpu.lusgs() == cpu.lusgs()
    dg, dc = gpu.get_field("dq")[:, :nc], cpu.get("dq")[:, :nc]
    err = rel_err(dg, dc, 1e-3 * np.max(np.abs(dc), axis=1, keepdims=True))
    print(f"{name}: dq rel err {err:.2e}, bitwise equal: {np.array_equal(dg, dc)}", flush=True)
    assert err <= TOL
    qg, qc = gpu.get_state()[:, :nc], cpu.get("q")[:, :nc]
    assert rel_err(qg, qc, 1e-3 * np.max(np.abs(qc), axis=1, keepdims=True)) <= TOL
    if sa:
        # one turbulence relaxation on the updated state (the oracle's
        # TurbSolver relaxation: boundary, laminar and eddy viscosity, nu~
        # gradient, residual, spectrum, scalar LU-SGS)
        gpu.turb_relax()
        cpu.sa_relax()
        tg, tc = gpu.get_field("nut")[:nc], cpu.get("nut")[:nc]
        err = rel_err(tg, tc, 1e-300)
        print(f"{name}: nu~ after the turbulence relaxation rel err {err:.2e}", flush=True)
        assert err <= 1e-11
        vg, vc = gpu.get_field("vist")[:nc], cpu.get("vist")[:nc]
        assert rel_err(vg, vc, 1e-300) <= 1e-11
    cpu.close()
    gpu.close()


@pytest.mark.timeout(400)
def test_full_size_laminar_hybrid_step_matches_oracle():
    _one_step("c2", 21)


@pytest.mark.timeout(600)
def test_full_size_swept_wing_rans_step_matches_oracle():
    _one_step("c3", 22)
