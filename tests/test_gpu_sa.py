"""GPU parity of the Spalart-Allmaras path (viscous = 3) against the oracle:
the wall distance, the NS residual and LU-SGS with an eddy viscosity, every
turbulence hook on the same state, whole iterations on the hex and hybrid
bump channels, and graph replay against eager launches.  FP64 bar: 1e-12
relative per call (the source calls pow, whose last bit may differ between
the device and host libraries), 1e-9 over a dozen iterations."""
import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc
from conftest import perturbed_state

pytestmark = pytest.mark.gpu

TOL = 1e-12
SA = 3


def rel_err(a, b, floor):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


def mesh_of(kind):
    if kind == "hex":
        return pa.Mesh.bump_hex(24, 12, 4, 0.08)
    if kind == "wing":
        return pa.Mesh.swept_wing(10, 5, 4, 1e-3, 3)
    return pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)


def pair(m, **pover):
    conn, met = m.connectivity(), m.metrics()
    p = pa.default_params(viscous=SA, reynolds=1e5, **pover)
    gpu = pa.Solver(m, p)
    cpu = orc.OracleSolver(conn, met, p)
    wd = orc.wall_distance(conn, met)
    cpu.set_wall_distance(wd)
    return gpu, cpu, wd


@pytest.mark.parametrize("kind", ["hex", "hybrid", "wing"])
def test_wall_distance_matches_oracle(kind):
    m = mesh_of(kind)
    gpu, cpu, wd = pair(m)
    g = gpu.get_field("wdist")[:m.n_cell]
    print(kind, "wall distance bitwise equal:", np.array_equal(g, wd), "max rel", np.max(np.abs(g - wd) / wd))
    assert np.max(np.abs(g - wd) / wd) <= 1e-14


@pytest.mark.parametrize("kind", ["hex", "hybrid", "wing"])
def test_turbulence_hooks_match_oracle(kind):
    m = mesh_of(kind)
    nc, nt = m.n_cell, m.n_total
    gpu, cpu, wd = pair(m, cfl_start=10.0, cfl_end=10.0)
    rng = np.random.default_rng(11)
    q = perturbed_state(nt, nc, cpu.freestream(), 5)
    nut = 3.0 * (1.0 + rng.random(nt))
    neg = rng.random(nt) < 0.1
    nut[neg] *= -0.2  # the negative-nu~ branches of the source and diffusion
    vist = 1e-3 + 0.5 * rng.random(nt)
    gpu.set_field("q", q)
    cpu.set("q", q)
    gpu.set_field("nut", nut)
    gpu.set_field("vist", vist)
    cpu.set("nut", nut)
    cpu.set("vist", vist)

    # the NS residual, spectrum, time step and LU-SGS see the eddy viscosity
    gpu.residual()
    cpu.residual()
    rg, rc = gpu.get_residual()[:, :nc], cpu.get("res")[:, :nc]
    assert rel_err(rg, rc, 1e-3 * np.max(np.abs(rc), axis=1, keepdims=True)) <= TOL
    cpu.spectrum()
    sg, sc = gpu.get_field("spec_vis")[:nc], cpu.get("spec_vis")[:nc]
    assert rel_err(sg, sc, 1e-300) <= TOL
    gpu.time_step(2)
    cpu.time_step(2)
    assert rel_err(gpu.get_field("dt")[:nc], cpu.get("dt")[:nc], 1e-300) <= TOL
    gpu.gamma_t()  # laminar viscosity of the state (the oracle's residual call made it)
    assert rel_err(gpu.get_field("visl"), cpu.get("visl"), 1e-300) <= 1e-15

    # turbulence hooks, in the order of one relaxation
    gpu.turb_boundary()
    cpu.sa_boundary()
    assert np.array_equal(gpu.get_field("nut"), cpu.get("nut"))
    assert np.array_equal(gpu.get_field("vist"), cpu.get("vist"))
    gpu.turb_viscosity()
    cpu.sa_viscosity()
    assert rel_err(gpu.get_field("vist")[:nc], cpu.get("vist")[:nc], 1e-300) <= 1e-15
    gpu.turb_gradient()
    cpu.sa_gradient()
    for d in ("dnudx", "dnudy", "dnudz"):
        g, c = gpu.get_field(d)[:nc], cpu.get(d)[:nc]
        assert rel_err(g, c, 1e-12 * np.max(np.abs(c))) <= TOL, d
    gpu.turb_update_residuals()
    cpu.sa_update_residuals()
    tg, tc = gpu.get_field("res_t")[:nc], cpu.get("res_t")[:nc]
    err_res = rel_err(tg, tc, 1e-3 * np.max(np.abs(tc)))
    print(kind, "res_t rel err", err_res, "bitwise equal:", np.array_equal(tg, tc))
    assert err_res <= TOL
    sg, sc = gpu.get_field("spec_t")[:nc], cpu.get("spec_t")[:nc]
    assert rel_err(sg, sc, 1e-300) <= TOL
    gpu.turb_spectrum()
    cpu.sa_spectrum()
    sg, sc = gpu.get_field("spec_t")[:nc], cpu.get("spec_t")[:nc]
    assert rel_err(sg, sc, 1e-300) <= TOL
    gpu.turb_lusgs()
    cpu.sa_lusgs()
    dg, dc = gpu.get_field("dq_t")[:nc], cpu.get("dq_t")[:nc]
    err_dq = rel_err(dg, dc, 1e-3 * np.max(np.abs(dc)))
    print(kind, "dq_t rel err", err_dq)
    assert err_dq <= 1e-11  # a sweep chains a dozen divisions per level
    ng, nc_ = gpu.get_field("nut")[:nc], cpu.get("nut")[:nc]
    assert rel_err(ng, nc_, 1e-300) <= 1e-12


@pytest.mark.parametrize("kind", ["hex", "hybrid", "wing"])
def test_sa_iterations_match_oracle(kind):
    m = mesh_of(kind)
    nc = m.n_cell
    gpu, cpu, wd = pair(m, cfl_start=5.0, cfl_end=5.0)
    gpu.init_freestream()
    cpu.init_freestream()
    n = 12
    hg = gpu.iterate(1, n)
    hc, _ = cpu.iterate(1, n)
    err_h = np.max(np.abs(hg - hc) / hc)
    qg, qc = gpu.get_state()[:, :nc], cpu.get("q")[:, :nc]
    dq = qc - cpu.freestream()[:, None]
    err_q = np.max(np.abs(qg - qc) / np.maximum(np.max(np.abs(dq), axis=1, keepdims=True), 1e-300))
    ng, ncpu = gpu.get_field("nut")[:nc], cpu.get("nut")[:nc]
    err_n = np.max(np.abs(ng - ncpu)) / np.max(np.abs(ncpu - 3.0))
    vg, vc = gpu.get_field("vist")[:nc], cpu.get("vist")[:nc]
    err_v = np.max(np.abs(vg - vc)) / np.max(np.abs(vc))
    print(kind, "rms", hg[:3], "errs hist", err_h, "q", err_q, "nut", err_n, "vist", err_v,
          "nut range", ncpu.min(), ncpu.max())
    assert err_h <= 1e-10
    assert err_q <= 1e-9
    assert err_n <= 1e-9
    assert err_v <= 1e-9


@pytest.mark.parametrize("mode", ["graph", "eager_overlap"])
def test_sa_overlapped_sweeps_equal_serial_launches(mode):
    """The turbulence sweeps of a step run on a second stream beside the next
    step's NS kernels (graph replay, or eager with overlap): the results equal
    the serial eager launches bit for bit, across several calls."""
    m = mesh_of("hybrid")
    p = pa.default_params(viscous=SA, reynolds=1e5, cfl_start=5.0, cfl_end=5.0)
    a, b = pa.Solver(m, p), pa.Solver(m, p)
    a.set_launch_mode(mode)
    b.set_launch_mode("eager")
    for s in (a, b):
        s.init_freestream()
    for first, n in ((1, 19), (20, 16), (36, 3)):
        ha, hb = a.iterate(first, n), b.iterate(first, n)
        assert np.array_equal(ha, hb)
        assert np.array_equal(a.get_state(), b.get_state())
        assert np.array_equal(a.get_field("nut"), b.get_field("nut"))
        assert np.array_equal(a.get_field("vist"), b.get_field("vist"))


def test_sa_mixed_precision_tracks_fp64():
    m = mesh_of("hybrid")
    nc = m.n_cell
    p64 = pa.default_params(viscous=SA, reynolds=1e5, cfl_start=5.0, cfl_end=5.0)
    p32 = pa.default_params(viscous=SA, reynolds=1e5, cfl_start=5.0, cfl_end=5.0, precision=1)
    a, b = pa.Solver(m, p64), pa.Solver(m, p32)
    for s in (a, b):
        s.init_freestream()
    ha, hb = a.iterate(1, 20), b.iterate(1, 20)
    na, nb = a.get_field("nut")[:nc], b.get_field("nut")[:nc]
    err_h = np.max(np.abs(ha - hb) / ha)
    err_n = np.max(np.abs(na - nb)) / np.max(np.abs(na - 3.0))
    print("mixed vs fp64: hist", err_h, "nut", err_n)
    assert err_h <= 1e-3
    assert err_n <= 1e-3


@pytest.mark.parametrize("kind,mode,form", [("hybrid", "graph", 0), ("wing", "eager_overlap", 0), ("wing", "eager", 0),
                                            ("wing", "graph", 3)])
def test_sa_dataflow_sweeps_equal_level_sweeps(kind, mode, form):
    """NS sweeps in the dataflow form (the default) against one launch per
    level, with the SA model: bitwise equal states, nu~ and mu_t, also with
    the (level-launched) SA sweeps beside the next NS step's dataflow sweeps
    on the second stream (graph and eager_overlap)."""
    m = mesh_of(kind)
    base = dict(viscous=SA, reynolds=1e5, cfl_start=5.0, cfl_end=5.0)
    a = pa.Solver(m, pa.default_params(sweep_form=form, **base))
    b = pa.Solver(m, pa.default_params(sweep_form=1, **base))
    a.set_launch_mode(mode)
    b.set_launch_mode("eager")
    for s in (a, b):
        s.init_freestream()
    for first, n in ((1, 11), (12, 9)):
        ha, hb = a.iterate(first, n), b.iterate(first, n)
        assert np.array_equal(ha, hb)
        assert np.array_equal(a.get_state(), b.get_state())
        assert np.array_equal(a.get_field("nut"), b.get_field("nut"))
        assert np.array_equal(a.get_field("vist"), b.get_field("vist"))
