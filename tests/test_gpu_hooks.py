"""The fine-grained hook sequence of Controller::Relaxation against phl_iterate
and the oracle, SolutionFix and NaN handling, and the dataflow sweep's wait
bound.

Hook order per iteration (CFD/Solver/src/Controller.cpp:567-608, 1002-1083,
1176, 1715): LoadResiduals, Boundary, ComputeGamaAndTField/ViscousCoeff,
ComputeGradient, ComputeLimiter, NSSolver::UpdateResiduals (InviscidFlux,
ViscousFlux), SpectrumRadius, TimeStep, LUSGS; then the turbulence relaxation.
"""
import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc
from conftest import perturbed_state

pytestmark = pytest.mark.gpu


def _pair(mesh, **pover):
    p = pa.default_params(**pover)
    gpu = pa.Solver(mesh, p)
    cpu = orc.OracleSolver(mesh.connectivity(), mesh.metrics(), p)
    return gpu, cpu


def _gpu_hooks(s, step, split=False, turb=False, ggnode=False):
    s.load_residuals()
    s.boundary()
    if ggnode:  # ComputeNodeValue of the previous PostSolve (Controller.cpp:5151-5198)
        s.compute_node_value()
    s.gamma_t()
    s.pressure_factor()
    s.gradient()
    s.limiter()
    if split:
        s.inviscid_flux()
        s.viscous_flux()
    else:
        s.update_residuals()
    s.spectrum_radius()
    s.time_step(step)
    n = s.lusgs()
    if turb:
        s.turb_relax()
    return n


def _cpu_hooks(c, step, viscous, ggnode=False):
    c.zero_residual()
    c.boundary()
    if ggnode:
        c.compute_node_value()
    c.gamma_t()
    c.pressure_factor()
    if viscous:
        c.viscosity()
    c.gradient()
    c.limiter()
    c.update_residuals()
    c.spectrum()
    c.time_step(step)
    n = c.lusgs()
    if viscous == 3:
        c.sa_relax()
    return n


CASES = {
    "hex_inviscid": (lambda: pa.Mesh.bump_hex(16, 8, 4, 0.05), dict(mach=0.5, cfl_start=10.0, cfl_end=10.0)),
    "hybrid_laminar": (lambda: pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7),
                       dict(mach=0.3, viscous=1, reynolds=5e5, cfl_start=5.0, cfl_end=5.0)),
    "hex_inviscid_ggnode": (lambda: pa.Mesh.bump_hex(16, 8, 4, 0.05),
                            dict(mach=0.5, cfl_start=10.0, cfl_end=10.0, gradient=2)),
    "hybrid_laminar_ggnode": (lambda: pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7),
                              dict(mach=0.3, viscous=1, reynolds=5e5, cfl_start=5.0, cfl_end=5.0, gradient=2)),
    "hybrid_laminar_entropy6": (lambda: pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7),
                                dict(mach=0.8, viscous=1, reynolds=5e5, cfl_start=5.0, cfl_end=5.0, entropy_fix=6)),
    "wing_sa": (lambda: pa.Mesh.swept_wing(10, 5, 4, 1e-3, 3),
                dict(mach=0.6, alpha_deg=2.0, viscous=3, reynolds=3e6, cfl_start=1.0, cfl_end=5.0,
                     cfl_vary_step=50)),
}


@pytest.mark.parametrize("kind", sorted(CASES))
@pytest.mark.parametrize("split", [False, True])
def test_hook_sequence_equals_iterate_and_oracle(kind, split):
    """Three iterations driven hook by hook equal phl_iterate bit for bit
    (state and eddy viscosity) and the oracle's own hook sequence to 1e-12."""
    make, pover = CASES[kind]
    m = make()
    viscous = pover.get("viscous", 0)
    gpu, cpu = _pair(m, **pover)
    ref = pa.Solver(m, pa.default_params(**pover))
    for s in (gpu, ref):
        s.init_freestream()
    cpu.init_freestream()
    if viscous == 3:
        wd = gpu.get_field("wdist")[:m.n_cell]
        cpu.set_wall_distance(wd)
    gg = pover.get("gradient", 0) == 2
    for step in (1, 2, 3):
        ng = _gpu_hooks(gpu, step, split=split, turb=viscous == 3, ggnode=gg)
        nc = _cpu_hooks(cpu, step, viscous, ggnode=gg)
        ref.iterate(step, 1)
        assert ng == nc == 0
        qg, qr = gpu.get_state(), ref.get_state()
        assert np.array_equal(qg, qr), (kind, step)
        qc = cpu.get("q")
        scale = np.max(np.abs(qc[:, :m.n_cell]), axis=1, keepdims=True)
        err = np.max(np.abs(qg[:, :m.n_cell] - qc[:, :m.n_cell]) / scale)
        assert err <= 1e-12, (kind, step, err)
        if viscous == 3:
            assert np.array_equal(gpu.get_field("nut"), ref.get_field("nut"))


def test_residual_norm_after_hooks_equals_fused_residual():
    """phl_residual_norm of the hook-assembled residual equals the sums the
    fused phl_residual computes (same blocks, same order) and the oracle's."""
    m = pa.Mesh.bump_hybrid(16, 4, 15, 1e-4, 10, 0.1, 7)
    pover = dict(mach=0.3, viscous=1, reynolds=5e5)
    gpu, cpu = _pair(m, **pover)
    fused = pa.Solver(m, pa.default_params(**pover))
    q = perturbed_state(m.n_total, m.n_cell, cpu.freestream(), 5)
    for s in (gpu, fused):
        s.set_state(q)
    cpu.set("q", q)
    gpu.load_residuals()
    gpu.boundary()
    gpu.gradient()
    gpu.limiter()
    gpu.inviscid_flux()
    gpu.viscous_flux()
    fused.residual()
    cpu.residual()
    assert np.array_equal(gpu.get_residual(), fused.get_residual())
    a, b, c = gpu.residual_norm(), fused.residual_norm(), cpu.residual_norm()
    assert np.array_equal(a, b)
    assert np.max(np.abs(a - c) / c) <= 1e-12


def test_limiter_hook_after_gradient_is_free_and_correct():
    """phl_limiter straight after phl_gradient keeps the limiter in place; on
    a state the gradient has not seen it evaluates it."""
    m = pa.Mesh.bump_hex(16, 8, 4, 0.05)
    gpu, cpu = _pair(m, mach=0.5)
    q = perturbed_state(m.n_total, m.n_cell, cpu.freestream(), 8)
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.boundary()
    cpu.boundary()
    gpu.limiter()  # the gradient has not seen this state: runs it
    cpu.gradient()
    cpu.limiter()
    lg, lc = gpu.get_field("limit"), cpu.get("limit")
    assert np.max(np.abs(lg - lc)) <= 1e-13
    gpu.gradient()
    gpu.limiter()
    assert np.array_equal(gpu.get_field("limit"), lg)


@pytest.mark.parametrize("inject,cells,expect_listed", [
    ((1e-3, 1e-7), [100, 101, 116, 300], True),
    ((1e-8, 1e-8), [100, 101, 116, 300], True),
    ((1e-3, 1e-7), "every3", False),
])
def test_solution_fix_matches_oracle(inject, cells, expect_listed):
    """A state with cells of tiny density and pressure (scaled by `inject`)
    makes the update produce rho or p below 1e-6 of freestream; SolutionFix
    (NSSolverUnstruct.cpp:14509-14585) replaces them by the average of their
    admissible neighbours, in increasing cell order.  The fix count and the
    repaired state equal the oracle's.  Up to 4,096 flagged cells are sorted
    in LDS; more take the walk over the mesh order (third case)."""
    m = pa.Mesh.bump_hex(16, 8, 4, 0.05) if expect_listed else pa.Mesh.bump_hex(64, 32, 8, 0.05)
    gpu, cpu = _pair(m, mach=0.5, cfl_start=10.0, cfl_end=10.0)
    cpu.init_freestream()
    q = cpu.get("q").copy()
    idx = np.arange(0, m.n_cell, 3) if cells == "every3" else np.array(cells)
    q[0, idx] *= inject[0]
    q[4, idx] *= inject[1]
    gpu.set_state(q)
    cpu.set("q", q)
    gpu.residual()
    cpu.residual()
    gpu.spectrum_radius()
    cpu.spectrum()
    gpu.time_step(1)
    cpu.time_step(1)
    ng, nc = gpu.lusgs(), cpu.lusgs()
    print("SolutionFix count gpu", ng, "oracle", nc)
    assert ng == nc > 0
    assert (ng <= 4096) == expect_listed
    qg, qc = gpu.get_state()[:, :m.n_cell], cpu.get("q")[:, :m.n_cell]
    assert np.all(np.isfinite(qg))
    err = np.max(np.abs(qg - qc) / np.maximum(np.abs(qc), 1e-300))
    assert err <= 1e-10, err
    # the flags were cleared: a following ordinary step fixes the same way
    # the oracle does
    for s in (gpu,):
        s.residual()
        s.spectrum_radius()
        s.time_step(2)
    cpu.residual()
    cpu.spectrum()
    cpu.time_step(2)
    assert gpu.lusgs() == cpu.lusgs()


def test_nan_state_is_a_numeric_error_and_the_context_recovers():
    """A NaN in the state makes the density residual non-finite: phl_iterate
    returns PHL_ERR_NUMERIC (the reference stops, Residual.cpp:579-589);
    after a fresh freestream the context iterates as a new one does."""
    m = pa.Mesh.bump_hex(16, 8, 4, 0.05)
    p = pa.default_params(mach=0.5, cfl_start=10.0, cfl_end=10.0)
    s, ref = pa.Solver(m, p), pa.Solver(m, p)
    s.init_freestream()
    q = s.get_state()
    q[0, 37] = np.nan
    s.set_state(q)
    with pytest.raises(pa.PhlError) as e:
        s.iterate(1, 4)
    assert e.value.status == pa.PHL_ERR_NUMERIC
    s.init_freestream()
    ref.init_freestream()
    assert np.array_equal(s.iterate(1, 10), ref.iterate(1, 10))
    assert np.array_equal(s.get_state(), ref.get_state())


@pytest.mark.parametrize("form", [0, 2, 3])
def test_sweep_wait_bound_gives_up_and_the_context_recovers(form):
    """With the dataflow sweep's wait bound set to 0 from the host, every
    wavefront that would wait gives up at once (every kernel still exits);
    phl_iterate and phl_lusgs report PHL_ERR_NUMERIC.  After restoring the
    bound and the state, the context iterates exactly like a fresh one."""
    m = pa.Mesh.bump_hex(16, 8, 4, 0.05)
    p = pa.default_params(mach=0.5, cfl_start=10.0, cfl_end=10.0, sweep_form=form)
    s, ref = pa.Solver(m, p), pa.Solver(m, p)
    fwd, bwd = s.sweep_levels()
    assert fwd > 1 and bwd > 1  # there is something to wait for
    s.init_freestream()
    s.set_sweep_wait(0)
    with pytest.raises(pa.PhlError) as e:
        s.iterate(1, 2)
    assert e.value.status == pa.PHL_ERR_NUMERIC
    s.residual()
    s.spectrum_radius()
    s.time_step(1)
    with pytest.raises(pa.PhlError) as e:
        s.lusgs()
    assert e.value.status == pa.PHL_ERR_NUMERIC
    s.set_sweep_wait(20000000)
    s.init_freestream()
    ref.init_freestream()
    assert np.array_equal(s.iterate(1, 12), ref.iterate(1, 12))
    assert np.array_equal(s.get_state(), ref.get_state())
