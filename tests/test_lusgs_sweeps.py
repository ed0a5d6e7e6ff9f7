"""Pins of the oracle's LU-SGS with several sweep pairs (nLUSGSSweeps > 1,
LUSGSTolerance; Controller::LUSGS, Controller.cpp:1878-1961, with the
deltaFlux bookkeeping of SolveLUSGSForward/Backward, NSSolverUnstruct.cpp:
15118-15764, and TurbSolverUnstr.cpp:1400-1637).

With k pairs and no early stop, dq is the k-th symmetric Gauss-Seidel iterate
of (D + L + U) dq = res from dq = 0 (one pair starts from dq = res, which
only its sweep normal sees): the forward sweep solves
(D + L) x* = res - U x_(k-1), the backward sweep (D + U) x_k = res - L x*.
The blocks of L and U are MXDQ_STD at each neighbour (the NS system) or the
SA face matrices (the scalar system), built densely here; the sweep normals
are the squared changes of each half-sweep; the pairs stop once a pair's
normal falls below LUSGSTolerance times the first pair's."""
import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc

GAMMA = 1.4


def farfield_box(nx=3, ny=2, nz=2):
    base = pa.Mesh.bump_hex(nx, ny, nz, 0.05)
    c = base.connectivity()
    m = pa.Mesh.from_arrays(c["x"], c["y"], c["z"], c["face_node_ptr"], c["face_nodes"], c["left"], c["right"],
                            c["n_bface"], np.full(c["n_bface"], 4, np.int32), c["cell_node_ptr"], c["cell_nodes"])
    return m, m.connectivity(), m.metrics()


def ns_setup(n_sweeps, tol):
    m, conn, met = farfield_box()
    p = pa.default_params(mach=0.5, cfl_start=5.0, cfl_end=5.0, n_lusgs_sweeps=n_sweeps, lusgs_tol=tol)
    s = orc.OracleSolver(conn, met, p)
    rng = np.random.default_rng(5)
    q = np.tile(s.freestream()[:, None], (1, m.n_total))
    q[:, :m.n_cell] *= 1 + rng.uniform(-0.05, 0.05, size=(5, m.n_cell))
    s.set("q", q)
    s.residual()
    s.spectrum()
    s.time_step(1)
    return m, conn, met, s


def ns_dense(m, conn, met, s):
    qb = s.get("q")
    nc = m.n_cell
    L = np.zeros((5 * nc, 5 * nc))
    U = np.zeros((5 * nc, 5 * nc))
    for f in range(m.n_bface, m.n_face):
        a, b = conn["left"][f], conn["right"][f]
        n = np.array([met["xfn"][f], met["yfn"][f], met["zfn"][f]])
        for i, j, nn in ((a, b, n), (b, a, -n)):
            rad = orc.lusgs_rad(qb[:, i], qb[:, j], nn, met["area"][f])
            B = np.column_stack([orc.mxdq(qb[:, j], nn, met["area"][f], GAMMA, e, rad, -1) for e in np.eye(5)])
            (L if j < i else U)[5 * i:5 * i + 5, 5 * j:5 * j + 5] += B
    return L, U


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_ns_sweep_pairs_are_symmetric_gauss_seidel(k):
    m, conn, met, s = ns_setup(k, 0.0)  # tolerance 0: every pair runs
    res = s.get("res")[:, :m.n_cell].T.reshape(-1).copy()
    L, U = ns_dense(m, conn, met, s)
    s.lusgs()
    nc = m.n_cell
    dq = s.get("dq")[:, :nc].T.reshape(-1)
    D = np.kron(np.diag(s.get("diag")[:nc]), np.eye(5))
    x = np.zeros_like(res)
    normals = []
    for i in range(k):
        xs = np.linalg.solve(D + L, res - U @ x)
        xn = np.linalg.solve(D + U, res - L @ xs)
        # one pair starts dq from the residual, several from zero: only the
        # first normal sees the difference
        old = res if (k == 1 and i == 0) else x
        normals.append(np.sum((xs - old) ** 2) + np.sum((xn - xs) ** 2))
        x = xn
    assert np.max(np.abs(dq - x)) <= 1e-12 * np.max(np.abs(x))
    np.testing.assert_allclose(s.sweep_normals(), normals, rtol=1e-10)


def test_ns_sweeps_converge_to_the_unfactored_system_and_stop_at_the_tolerance():
    m, conn, met, s = ns_setup(60, 1e-30)
    res = s.get("res")[:, :m.n_cell].T.reshape(-1).copy()
    L, U = ns_dense(m, conn, met, s)
    s.lusgs()
    nc = m.n_cell
    dq = s.get("dq")[:, :nc].T.reshape(-1)
    D = np.kron(np.diag(s.get("diag")[:nc]), np.eye(5))
    exact = np.linalg.solve(D + L + U, res)
    assert np.max(np.abs(dq - exact)) <= 1e-9 * np.max(np.abs(exact))
    # a loose tolerance stops after the first pair whose normal drops below
    # tol times the first pair's (never after the first pair itself)
    m, conn, met, s2 = ns_setup(8, 0.05)
    s2.lusgs()
    nrm = s2.sweep_normals()
    assert len(nrm) >= 2
    assert nrm[-1] / nrm[0] < 0.05 and all(v / nrm[0] >= 0.05 for v in nrm[1:-1])


@pytest.mark.parametrize("k", [1, 2, 4])
def test_sa_sweep_pairs_are_symmetric_gauss_seidel(k):
    """The scalar SA system: L, U from the face matrices (matrixTurbulenceLeft
    for a neighbour on the left of the face, Right otherwise), D the spectrum.
    On a farfield box each ghost copies its owner's dq after the forward
    sweep (SetGhostDQLUSGS), so the boundary faces B enter the backward sweep
    with the forward iterate xs and the next forward sweep (through the stored
    deltaFlux) with the previous pair's xs:
        (D + L) xs_k = res - U x_(k-1) - B xs_(k-1),
        (D + U) x_k  = res - L xs_k - B xs_k."""
    m, conn, met = farfield_box(4, 3, 2)
    p = pa.default_params(mach=0.5, viscous=3, reynolds=1e5, cfl_start=5.0, cfl_end=5.0, n_lusgs_sweeps=k,
                          lusgs_tol=0.0)
    s = orc.OracleSolver(conn, met, p)
    s.set_wall_distance(np.full(m.n_cell, 0.3))
    s.init_freestream()
    rng = np.random.default_rng(9)
    nut = s.get("nut")
    nut[:m.n_cell] *= 1 + rng.uniform(-0.3, 0.3, m.n_cell)
    s.set("nut", nut)
    s.residual()
    s.spectrum()
    s.time_step(1)
    s.sa_boundary()
    s.sa_viscosity()
    s.sa_gradient()
    s.sa_update_residuals()
    s.sa_spectrum()
    nc, nb = m.n_cell, m.n_bface
    nut0 = s.get("nut")[:nc].copy()
    r = s.get("res_t")[:nc].copy()
    D = np.diag(s.get("spec_t")[:nc])
    matL, matR = s.sa_face_matrices(m.n_face)
    Lm = np.zeros((nc, nc))
    Um = np.zeros((nc, nc))
    B = np.zeros((nc, nc))
    for f in range(m.n_face):
        a, b = conn["left"][f], conn["right"][f]
        if f < nb:
            B[a, a] += matR[f]
            continue
        Lm[b, a] += matL[f]  # cell b sees its lower neighbour a (the face's left) through matL
        Um[a, b] += matR[f]  # cell a sees its upper neighbour b through matR
    s.sa_lusgs()
    dq = s.get("nut")[:nc] - nut0
    x = np.zeros(nc)
    xs = np.zeros(nc)
    normals = []
    for i in range(k):
        xs_new = np.linalg.solve(D + Lm, r - Um @ x - B @ xs)
        x_new = np.linalg.solve(D + Um, r - Lm @ xs_new - B @ xs_new)
        old = r if (k == 1 and i == 0) else x
        normals.append(np.sum((xs_new - old) ** 2) + np.sum((x_new - xs_new) ** 2))
        x, xs = x_new, xs_new
    # dq is the update nu~_new - nu~_old: rounding of the addition aside
    assert np.max(np.abs(dq - x)) <= 1e-12 * np.max(np.abs(nut0))
    np.testing.assert_allclose(s.sweep_normals(turb=True), normals, rtol=1e-8)
