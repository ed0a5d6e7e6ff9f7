"""GPU parity on the reference's own example grid: the RAE2822 2-D hybrid
grid (PHengLEI/examples/grid/rae2822_hybrid2d__4_0.fts, 36,036 cells in 4
zones joined into one, copied to tests/golden/) at the flow of the
reference's example cases (cfd_para_subsonic.hypara: M 0.2, alpha 2.79 deg,
Re 6.5e6, Spalart-Allmaras; cfd_para_transonic.hypara: M 0.73).  2-D metrics
and the 2-D wall distance (the reference's bisection) on the device against
the oracle and against the reference's own .wdt distances; residuals and
iterations against the oracle.  FP64 bar: 1e-12 relative per call, 1e-9 over
a few iterations (the SA source and farfield BC call pow)."""
import os

import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc
from conftest import perturbed_state

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FTS = os.path.join(GOLD, "rae2822_hybrid2d__4_0.fts")
WDT = os.path.join(GOLD, "rae2822_hybrid2d__4_0.wdt")
TOL = 1e-12


@pytest.fixture(scope="module")
def rae():
    m = pa.Mesh.from_fts(FTS)
    c, g = m.connectivity(), m.metrics()
    return m, c, g, orc.wall_distance(c, g)


def rel_err(a, b, floor):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


def test_wall_distance_2d_on_device(rae):
    m, c, g, wd = rae
    p = pa.default_params(mach=0.2, alpha_deg=2.79, reynolds=6.5e6, viscous=3)
    s = pa.Solver(m, p)
    dev = s.get_field("wdist")[:m.n_cell]
    print("2-D wall distance bitwise equal to the oracle:", np.array_equal(dev, wd))
    assert np.max(np.abs(dev - wd) / wd) <= 1e-14
    ref = m.read_wdt(WDT)
    assert np.max(np.abs(dev - ref) / ref) < 1e-7


@pytest.mark.parametrize("mach,viscous,gradient,precision", [(0.2, 3, 0, 0), (0.73, 3, 0, 0), (0.73, 1, 2, 0),
                                                             (0.2, 0, 2, 0), (0.73, 3, 0, 1), (0.2, 3, 2, 0)])
def test_rae2822_iterations_match_oracle(rae, mach, viscous, gradient, precision):
    """From freestream: the first residual per call, then iterations, against
    the oracle (mixed precision: the residual norms to 1e-4)."""
    m, c, g, wd = rae
    p = pa.default_params(mach=mach, alpha_deg=2.79, reynolds=6.5e6, viscous=viscous, gradient=gradient,
                          cfl_start=1.0, cfl_end=10.0, cfl_vary_step=50, precision=precision)
    gpu = pa.Solver(m, p)
    cpu = orc.OracleSolver(c, g, p)
    if viscous == 3:
        cpu.set_wall_distance(wd)
    nc = m.n_cell
    if precision == 0:
        # a seeded state: at freestream the residual is rounding noise around
        # the cancelling face fluxes, where one-ulp pow differences of the
        # farfield ghosts show
        gpu.init_freestream()  # nu~ and mu_t of the SA model at their freestream values
        cpu.init_freestream()
        q = perturbed_state(m.n_total, nc, cpu.freestream(), 17)
        q[3] = 0.0  # 2-D: no w
        gpu.set_state(q)
        cpu.set("q", q)
        gpu.residual()
        cpu.residual()
        rg, rc = gpu.get_residual()[:, :nc], cpu.get("res")[:, :nc]
        # a 2-D grid has no z-momentum residual at all: exactly zero on both
        assert not np.any(rg[3]) and not np.any(rc[3])
        assert rel_err(rg[[0, 1, 2, 4]], rc[[0, 1, 2, 4]],
                       1e-3 * np.max(np.abs(rc[[0, 1, 2, 4]]), axis=1, keepdims=True)) <= TOL
    gpu.init_freestream()
    cpu.init_freestream()
    hg = gpu.iterate(1, 6)
    hc, nfix = cpu.iterate(1, 6)
    err = np.max(np.abs(hg - hc) / hc)
    print(f"M{mach} viscous {viscous} gradient {gradient} precision {precision}: history rel err {err:.2e}")
    assert err <= (1e-4 if precision else 1e-9)
    if precision == 0:
        qg, qc = gpu.get_state()[:, :nc], cpu.get("q")[:, :nc]
        assert not np.any(qg[3]) and not np.any(qc[3])  # w stays exactly zero in 2-D
        keep = [0, 1, 2, 4]
        assert rel_err(qg[keep], qc[keep], 1e-3 * np.max(np.abs(qc[keep]), axis=1, keepdims=True)) <= 1e-9
        cl_g, cl_c = gpu.forces()["CL"], cpu.forces()[pa.FORCE_NAMES.index("CL")]
        assert abs(cl_g - cl_c) <= 1e-9 * max(abs(cl_c), 1e-3)


@pytest.mark.parametrize("case", ["subsonic", "transonic"])
def test_reference_example_configuration(rae, case):
    """The reference's own example cases on its own grid: cfd_para_subsonic.
    hypara (M 0.2, venkatCoeff 50, CFLEnd 100, nLUSGSSweeps 5) and
    cfd_para_transonic.hypara (M 0.73, venkatCoeff 5, CFLEnd 30,
    nLUSGSSweeps 3), both alpha 2.79, Re 6.5e6, SA, with the defaults of
    cfd_para.hypara (gradientName ggnode, flowInitStep 100 Van Leer
    iterations, CFLStart 0.01 over CFLVaryStep 500).  The Roe scheme of the
    transonic file's structured part is the unstructured default here (its
    "gks" unstructured scheme is not on this path).  12 iterations against
    the oracle."""
    m, c, g, wd = rae
    sub = case == "subsonic"
    p = pa.default_params(mach=0.2 if sub else 0.73, alpha_deg=2.79, reynolds=6.5e6, viscous=3, gradient=2,
                          venkat_k=50.0 if sub else 5.0, cfl_start=0.01, cfl_end=100.0 if sub else 30.0,
                          cfl_vary_step=500, n_lusgs_sweeps=5 if sub else 3, flow_init_step=100)
    gpu = pa.Solver(m, p)
    cpu = orc.OracleSolver(c, g, p)
    cpu.set_wall_distance(wd)
    gpu.init_freestream()
    cpu.init_freestream()
    hg = gpu.iterate(1, 12)
    hc, _ = cpu.iterate(1, 12)
    err = np.max(np.abs(hg - hc) / hc)
    print(case, "history rel err", err, "pairs", gpu.sweep_normals().size, cpu.sweep_normals().size)
    assert err <= 1e-9
    assert gpu.sweep_normals().size == cpu.sweep_normals().size
    n = m.n_cell
    nut_g, nut_c = gpu.get_field("nut")[:n], cpu.get("nut")[:n]
    assert np.max(np.abs(nut_g - nut_c) / np.maximum(np.abs(nut_c), 1e-3 * np.max(np.abs(nut_c)))) <= 1e-9
