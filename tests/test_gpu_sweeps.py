"""Several LU-SGS sweep pairs per step (nLUSGSSweeps > 1 with
LUSGSTolerance; Controller::LUSGS, Controller.cpp:1878-1961) on the device
against the oracle: the same number of pairs per call, the pairs' squared dq
changes (summed in another order: 1e-12 relative), and the states and
residual histories (1e-12; 1e-9 with the SA model, whose source calls pow),
for the NS equations and the SA model, every sweep form, and the reference's
own RAE2822 example (cfd_para_subsonic.hypara: nLUSGSSweeps = 5;
cfd_para_transonic.hypara: 3)."""
import os

import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def mesh_of(kind):
    if kind == "hex":
        return pa.Mesh.bump_hex(24, 12, 4, 0.08)
    if kind == "wing":
        return pa.Mesh.swept_wing(10, 5, 4, 1e-3, 3)
    if kind == "rae":
        return pa.Mesh.from_fts(os.path.join(GOLD, "rae2822_hybrid2d__4_0.fts"))
    return pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)


@pytest.mark.parametrize("kind,viscous,sweeps,tol,form", [("hex", 0, 3, 0.01, 0), ("hybrid", 1, 4, 0.0, 0),
                                                          ("hybrid", 1, 3, 0.01, 1), ("hex", 0, 5, 0.01, 3),
                                                          ("wing", 3, 3, 0.01, 0), ("rae", 3, 5, 0.01, 0),
                                                          ("rae", 1, 3, 0.01, 2)])
def test_sweep_pairs_match_oracle(kind, viscous, sweeps, tol, form):
    m = mesh_of(kind)
    conn, met = m.connectivity(), m.metrics()
    mach = 0.2 if kind == "rae" else 0.5
    p = pa.default_params(mach=mach, alpha_deg=2.79 if kind == "rae" else 0.0, viscous=viscous,
                          reynolds=6.5e6 if kind == "rae" else 1e5, cfl_start=2.0, cfl_end=5.0, cfl_vary_step=20,
                          n_lusgs_sweeps=sweeps, lusgs_tol=tol, sweep_form=form)
    gpu = pa.Solver(m, p)
    cpu = orc.OracleSolver(conn, met, p)
    if viscous == 3:
        cpu.set_wall_distance(orc.wall_distance(conn, met))
    gpu.init_freestream()
    cpu.init_freestream()
    bar = 1e-9 if viscous == 3 else 1e-12
    for first in (1, 4):
        hg = gpu.iterate(first, 3)
        hc, _ = cpu.iterate(first, 3)
        err = np.max(np.abs(hg - hc) / hc)
        ng, nc_ = gpu.sweep_normals(), cpu.sweep_normals()
        print(kind, viscous, sweeps, tol, form, "history rel err", err, "pairs", len(ng), len(nc_))
        assert err <= bar
        assert len(ng) == len(nc_) and len(ng) >= (2 if tol > 0 else sweeps)
        np.testing.assert_allclose(ng, nc_, rtol=1e-9 if viscous == 3 else 1e-11)
        if viscous == 3:
            tg, tc = gpu.sweep_normals(turb=True), cpu.sweep_normals(turb=True)
            assert len(tg) == len(tc)
            np.testing.assert_allclose(tg, tc, rtol=1e-8)
    n = m.n_cell
    qg, qc = gpu.get_state()[:, :n], cpu.get("q")[:, :n]
    floor = 1e-3 * np.max(np.abs(qc), axis=1, keepdims=True) + 1e-300
    assert np.max(np.abs(qg - qc) / np.maximum(np.abs(qc), floor)) <= bar


def test_lusgs_hook_with_sweep_pairs():
    """phl_lusgs on a seeded state: the pairs, their normals and the update."""
    m = mesh_of("hybrid")
    conn, met = m.connectivity(), m.metrics()
    p = pa.default_params(mach=0.3, viscous=1, reynolds=5e5, cfl_start=5.0, cfl_end=5.0, n_lusgs_sweeps=6,
                          lusgs_tol=0.001)
    gpu = pa.Solver(m, p)
    cpu = orc.OracleSolver(conn, met, p)
    rng = np.random.default_rng(3)
    q = np.tile(cpu.freestream()[:, None], (1, m.n_total))
    q[:, :m.n_cell] *= 1 + rng.uniform(-0.03, 0.03, size=(5, m.n_cell))
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.residual()
    cpu.residual()
    gpu.time_step(2)
    cpu.spectrum()
    cpu.time_step(2)
    assert gpu.lusgs() == cpu.lusgs()
    ng, nc_ = gpu.sweep_normals(), cpu.sweep_normals()
    assert len(ng) == len(nc_) >= 2
    np.testing.assert_allclose(ng, nc_, rtol=1e-11)
    n = m.n_cell
    dg, dc = gpu.get_field("dq")[:, :n], cpu.get("dq")[:, :n]
    floor = 1e-3 * np.max(np.abs(dc), axis=1, keepdims=True)
    assert np.max(np.abs(dg - dc) / np.maximum(np.abs(dc), floor)) <= 1e-12
