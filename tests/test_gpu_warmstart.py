"""GPU: the learned warm start end to end at small scale.  The solver
generates converged fields for seeded cases, the device network is trained
on them, and for held-out cases the prediction is judged by the solver's own
residual: it must reach the same absolute residual level as the freestream
start in fewer iterations.  The freestream start is the fallback."""
import numpy as np
import pytest

import phenlei_amd as pa
from phenlei_amd import warmstart as ws

pytestmark = pytest.mark.gpu

DIMS = (24, 12, 4)


def test_warm_start_beats_freestream_on_held_out_cases():
    train = ws.sample_cases(10, seed=11)
    held = ws.sample_cases(2, seed=99)
    meshes, fields, steps = ws.solve_cases(train, DIMS, drop=1e-6, max_steps=6000)
    print("training cases converged in", steps)
    assert np.all(steps > 0)
    net, hist = ws.train_network(meshes, train, fields, steps=600, lr=2e-3, layers=6)
    print("training loss", hist[0], "->", hist[-1])
    assert hist[-1] < 0.1 * hist[0]
    for case in held:
        mesh = pa.Mesh.bump_hex(*DIMS, case["bump"])
        p = pa.default_params(mach=case["mach"], alpha_deg=case["alpha"], cfl_start=10.0, cfl_end=10.0)
        cold = pa.Solver(mesh, p)
        cold.init_freestream()
        r_cold = cold.iterate(1, 1)[0]
        target = 1e-6 * r_cold
        cold.init_freestream()
        n_cold, _ = ws.steps_to(cold, target, 6000)
        warm = pa.Solver(mesh, p)
        warm.set_state(ws.predict_state(net, mesh, case))
        n_warm, r_warm = ws.steps_to(warm, target, 6000)
        print("case", case, "freestream: R1", r_cold, "steps", n_cold, "| predicted: R1", r_warm, "steps", n_warm)
        assert n_cold > 0 and n_warm > 0
        assert n_warm < n_cold
