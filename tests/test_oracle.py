"""Pins of the CPU oracle (runs on the CPU only).

The reference ships no golden vectors for this path, so each oracle function
is pinned by checks that an independent computation or the mathematics
fixes: an eigen-decomposition of the complex-step flux Jacobian for the Roe
flux and MXDQ_STD, closed forms (Harten, VenFun, entropy coefficient,
spectral radius on a box), exactness of Green-Gauss for linear fields,
freestream preservation, and a dense solve of (D+L)D^-1(D+U) for the LU-SGS
sweep structure.
"""
import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc

GAMMA = 1.4


def cons(q, g=GAMMA):
    r, u, v, w, p = q
    return np.array([r, r * u, r * v, r * w, p / (g - 1.0) + 0.5 * r * (u * u + v * v + w * w)])


def flux_n(U, n, g=GAMMA):
    """Physical Euler flux F(U).n from conservative variables (works on complex U)."""
    r = U[0]
    u, v, w = U[1] / r, U[2] / r, U[3] / r
    p = (g - 1.0) * (U[4] - 0.5 * r * (u * u + v * v + w * w))
    vn = u * n[0] + v * n[1] + w * n[2]
    return np.array([r * vn, r * u * vn + p * n[0], r * v * vn + p * n[1], r * w * vn + p * n[2],
                     (U[4] + p) * vn])


def jacobian(U, n, g=GAMMA):
    """dF.n/dU by complex step (exact to rounding for this rational function)."""
    h = 1e-30
    A = np.zeros((5, 5))
    for k in range(5):
        Uc = U.astype(complex)
        Uc[k] += 1j * h
        A[:, k] = flux_n(Uc, n, g).imag / h
    return A


def random_state(rng):
    return np.array([rng.uniform(0.5, 2.0), *rng.uniform(-1.0, 1.0, 3), rng.uniform(0.5, 3.0)])


def random_normal(rng):
    n = rng.normal(size=3)
    return n / np.linalg.norm(n)


def textbook_roe(qL, qR, n, eps, g=GAMMA):
    """Roe flux from an eigen-decomposition of A at the reference's averaged
    state (rho = sqrt(rL rR), velocity and pressure sqrt(rho)-weighted)."""
    r = np.sqrt(qR[0] / qL[0])
    c = 1.0 / (1.0 + r)
    rho = np.sqrt(qL[0] * qR[0])
    vel = (qL[1:4] + qR[1:4] * r) * c
    p = (qL[4] + qR[4] * r) * c
    Ua = cons(np.array([rho, *vel, p]), g)
    A = jacobian(Ua, n, g)
    vn = vel @ n
    H = g / (g - 1.0) * p / rho + 0.5 * vel @ vel
    cm = np.sqrt((g - 1.0) * (H - 0.5 * vel @ vel))
    # textbook right eigenvectors (acoustic, entropy, two shear, acoustic)
    t1 = np.cross(n, [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    R = np.column_stack([
        [1.0, *(vel - cm * n), H - cm * vn],
        [1.0, *vel, 0.5 * vel @ vel],
        [0.0, *t1, vel @ t1],
        [0.0, *t2, vel @ t2],
        [1.0, *(vel + cm * n), H + cm * vn]])
    lam = np.array([vn - cm, vn, vn, vn, vn + cm])
    assert np.allclose(A @ R, R * lam, rtol=1e-12, atol=1e-12)  # R diagonalises the Jacobian
    delta = eps * (abs(vn) + cm)
    alam = np.where(np.abs(lam) < delta, 0.5 * (lam * lam + delta * delta) / delta, np.abs(lam))
    absA = R @ np.diag(alam) @ np.linalg.inv(R)
    UL, UR = cons(qL, g), cons(qR, g)
    return 0.5 * (flux_n(UL, n, g) + flux_n(UR, n, g)) - 0.5 * absA @ (UR - UL)


def test_roe_matches_eigen_decomposition():
    rng = np.random.default_rng(11)
    for _ in range(200):
        qL, qR, n = random_state(rng), random_state(rng), random_normal(rng)
        eps = rng.choice([1e-4, 0.05, 0.3])  # small and large entropy-fix widths
        f = orc.roe_flux(qL, qR, n, GAMMA, 3, eps)
        ref = textbook_roe(qL, qR, n, eps)
        scale = max(np.max(np.abs(ref)), 1.0)
        assert np.max(np.abs(f - ref)) <= 1e-12 * scale


def test_roe_consistency_and_conservation():
    rng = np.random.default_rng(12)
    for _ in range(100):
        q, q2, n = random_state(rng), random_state(rng), random_normal(rng)
        f = orc.roe_flux(q, q, n)
        assert np.allclose(f, flux_n(cons(q), n), rtol=1e-14, atol=1e-14)
        a = orc.roe_flux(q, q2, n)
        b = orc.roe_flux(q2, q, -n)
        assert np.allclose(a, -b, rtol=1e-13, atol=1e-13)


def test_roe_supersonic_is_jacobian_upwind():
    """All eigenvalues positive: |A| = A, so F = (FL+FR)/2 - A(U~)(UR-UL)/2.
    (The reference averages pressure rather than enthalpy, so A(U~)dU is not
    exactly FR-FL and F differs slightly from FL in the energy component.)"""
    rng = np.random.default_rng(13)
    for _ in range(50):
        qL = np.array([1.0, 3.0, 0.1, -0.2, 1.0])  # M ~ 2.5 along +x
        qR = qL * (1 + rng.uniform(-0.02, 0.02, 5))
        n = np.array([1.0, 0.0, 0.0])
        f = orc.roe_flux(qL, qR, n)
        r = np.sqrt(qR[0] / qL[0])
        c = 1.0 / (1.0 + r)
        qa = np.array([np.sqrt(qL[0] * qR[0]), *((qL[1:4] + qR[1:4] * r) * c), (qL[4] + qR[4] * r) * c])
        A = jacobian(cons(qa), n)
        UL, UR = cons(qL), cons(qR)
        ref = 0.5 * (flux_n(UL, n) + flux_n(UR, n)) - 0.5 * A @ (UR - UL)
        assert np.allclose(f, ref, rtol=1e-13, atol=1e-13)
        assert np.allclose(f[:4], flux_n(UL, n)[:4], rtol=1e-12, atol=1e-12)  # Roe property for mass, momentum


def test_harten_cases():
    assert orc.load().orc_harten(0.5, 0.1) == 0.5
    assert orc.load().orc_harten(-0.5, 0.1) == 0.5
    assert orc.load().orc_harten(0.05, 0.1) == pytest.approx(0.5 * (0.0025 + 0.01) / 0.1, rel=1e-15)
    # continuous at the switch, never below lim/2
    lim = 0.3
    assert orc.load().orc_harten(lim * (1 - 1e-12), lim) == pytest.approx(lim, rel=1e-11)
    assert orc.load().orc_harten(0.0, lim) == pytest.approx(lim / 2, rel=1e-15)


def test_venfun_closed_forms():
    v = orc.load().orc_venfun
    assert v(2.0, 0.0, 0.0) == 1.0
    assert v(1.0, 1.0, 0.0) == pytest.approx(0.75, rel=1e-15)
    assert v(3.0, 1.0, 0.0) == pytest.approx((9 + 6) / (9 + 2 + 3), rel=1e-15)
    assert v(1e8, 1.0, 0.0) == pytest.approx(1.0, rel=1e-7)
    assert v(0.0, 1.0, 1.0) == pytest.approx(1.0 / 3.0, rel=1e-15)


def test_entropy_coefficient_schedule():
    ec = orc.load().orc_entropy_coef
    assert ec(0.5, 1.0) == 1e-4
    assert ec(1.3, 1.0) == pytest.approx(0.5 * 1e-4 + 0.5 * 0.01, rel=1e-14)
    assert ec(2.0, 1.0) == pytest.approx(0.01, rel=1e-14)
    assert ec(3.5, 2.0) == pytest.approx(2 * (0.5 * 0.01 + 0.5 * 0.5), rel=1e-14)
    assert ec(7.0, 1.0) == 1.0


def test_mxdq_is_half_jacobian_plus_radius():
    rng = np.random.default_rng(14)
    for _ in range(100):
        q, n = random_state(rng), random_normal(rng)
        area, rad = rng.uniform(0.1, 2.0), rng.uniform(0.1, 5.0)
        dq = rng.normal(size=5)
        for ipn in (-1, 1):
            f = orc.mxdq(q, n, area, GAMMA, dq, rad, ipn)
            A = jacobian(cons(q), n)
            ref = 0.5 * (A @ dq + ipn * rad / area * dq) * area
            assert np.max(np.abs(f - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_lusgs_radius_equal_states():
    rng = np.random.default_rng(15)
    for _ in range(20):
        q, n = random_state(rng), random_normal(rng)
        area = rng.uniform(0.1, 2.0)
        ref = (abs(q[1:4] @ n) + np.sqrt(GAMMA * q[4] / q[0])) * area
        assert orc.lusgs_rad(q, q, n, area) == pytest.approx(ref, rel=1e-14)


def box_mesh(nx=6, ny=4, nz=3):
    return pa.Mesh.bump_hex(nx, ny, nz, 0.0)  # uniform Cartesian box


def oracle_for(mesh, **p):
    return orc.OracleSolver(mesh.connectivity(), mesh.metrics(), pa.default_params(**p))


def test_green_gauss_exact_for_linear_fields():
    m = box_mesh()
    met = m.metrics()
    s = oracle_for(m)
    coef = np.array([[1.0, 0.3, -0.2, 0.1], [0.2, -0.5, 0.7, 0.3], [0.1, 0.2, 0.4, -0.6],
                     [-0.3, 0.1, 0.2, 0.9], [2.0, 0.4, -0.1, 0.25]])
    q = coef[:, :1] + coef[:, 1:2] * met["xcc"] + coef[:, 2:3] * met["ycc"] + coef[:, 3:4] * met["zcc"]
    s.set("q", q)
    s.gradient()
    nc = m.n_cell
    for name, col in (("dqdx", 1), ("dqdy", 2), ("dqdz", 3)):
        g = s.get(name)[:, :nc]
        assert np.max(np.abs(g - coef[:, col:col + 1])) <= 1e-12, name


@pytest.mark.parametrize("kind", ["bump", "hybrid"])
def test_least_squares_exact_for_linear_fields_in_every_cell(kind):
    """Least squares (gradientName "lsq") recovers a linear field's gradient in
    every real cell of a distorted mesh, boundary cells included, when the
    ghosts carry the same linear field; ghosts take their owner's gradient."""
    m = pa.Mesh.bump_hex(12, 6, 4, 0.08) if kind == "bump" else pa.Mesh.bump_hybrid(8, 3, 6, 1e-3, 6, 0.08, 5)
    met = m.metrics()
    s = oracle_for(m, gradient=1)
    coef = np.array([[1.0, 0.3, -0.2, 0.1], [0.2, -0.5, 0.7, 0.3], [0.1, 0.2, 0.4, -0.6],
                     [-0.3, 0.1, 0.2, 0.9], [2.0, 0.4, -0.1, 0.25]])
    q = coef[:, :1] + coef[:, 1:2] * met["xcc"] + coef[:, 2:3] * met["ycc"] + coef[:, 3:4] * met["zcc"]
    s.set("q", q)
    s.gradient()
    nc = m.n_cell
    owner = m.connectivity()["left"][:m.n_bface]
    for name, col in (("dqdx", 1), ("dqdy", 2), ("dqdz", 3)):
        g = s.get(name)
        assert np.max(np.abs(g[:, :nc] - coef[:, col:col + 1])) <= 1e-11, name
        assert np.array_equal(g[:, nc:], g[:, owner]), name


def test_least_squares_solves_the_weighted_normal_equations():
    """Each cell's lsq gradient minimises sum_f w_f (g . d_f - dq_f)^2 over its
    faces, w_f = 1/|d_f|^2 (d_f: centroid difference to the neighbour or
    ghost): checked against numpy's solve of the 3x3 normal equations on a
    random field."""
    m = pa.Mesh.bump_hybrid(8, 3, 6, 1e-3, 6, 0.08, 5)
    conn, met = m.connectivity(), m.metrics()
    s = oracle_for(m, gradient=1)
    rng = np.random.default_rng(3)
    q = 1.0 + 0.1 * rng.random((5, m.n_total))
    s.set("q", q)
    s.gradient()
    g = np.stack([s.get("dqdx")[0], s.get("dqdy")[0], s.get("dqdz")[0]])
    xc = np.stack([met["xcc"], met["ycc"], met["zcc"]])
    left, right = conn["left"], conn["right"]
    A = np.zeros((m.n_cell, 3, 3))
    b = np.zeros((m.n_cell, 3))
    for f in range(m.n_face):
        le, re = left[f], right[f]
        d = xc[:, re] - xc[:, le]
        w = 1.0 / (d @ d)
        for c, sgn in ((le, 1.0), (re, 1.0)):
            if c >= m.n_cell:
                continue
            A[c] += w * np.outer(d, d)
            b[c] += w * d * (q[0, re] - q[0, le]) * sgn
    ref = np.linalg.solve(A, b[..., None])[..., 0].T
    err = np.max(np.abs(g[:, :m.n_cell] - ref)) / np.max(np.abs(ref))
    assert err <= 1e-10, err


def test_metrics_closure_and_volume():
    m = pa.Mesh.bump_hex(16, 8, 4, 0.05)
    conn, met = m.connectivity(), m.metrics()
    nc = m.n_cell
    acc = np.zeros((nc, 3))
    for k, comp in enumerate(("xfn", "yfn", "zfn")):
        sa = met[comp] * met["area"]
        np.add.at(acc[:, k], conn["left"], sa)
        inner = np.arange(m.n_bface, m.n_face)
        np.add.at(acc[:, k], conn["right"][inner], -sa[inner])
    assert np.max(np.abs(acc)) <= 1e-14
    b = box_mesh(8, 5, 4)
    assert b.metrics()["vol"][:b.n_cell].sum() == pytest.approx(3.0 * 0.8 * 0.2, rel=1e-13)


def test_oracle_metrics_restate_library_metrics():
    m = pa.Mesh.bump_hex(16, 8, 4, 0.05)
    conn, met = m.connectivity(), m.metrics()
    fm = orc.face_metrics(conn)
    cm = orc.cell_metrics(conn, fm)
    for k, v in list(fm.items()) + list(cm.items()):
        assert np.array_equal(v, met[k]), k


def test_freestream_preserved_in_flat_channel():
    m = box_mesh(8, 4, 3)
    s = oracle_for(m, mach=0.5)
    s.init_freestream()
    s.residual()
    res = s.get("res")[:, :m.n_cell]
    pinf = s.freestream()[4]
    assert np.max(np.abs(res)) <= 1e-13 * pinf * m.metrics()["area"].max()
    # an update from a zero residual leaves the state unchanged
    s.spectrum()
    s.time_step(1)
    s.lusgs()
    q = s.get("q")[:, :m.n_cell]
    assert np.max(np.abs(q - s.freestream()[:, None])) <= 1e-13


def test_spectral_radius_and_time_step_on_box():
    nx, ny, nz = 6, 4, 3
    m = box_mesh(nx, ny, nz)
    s = oracle_for(m, mach=0.5, cfl_start=10.0, cfl_end=10.0)
    s.init_freestream()
    s.residual()
    s.spectrum()
    s.time_step(1)
    dx, dy, dz = 3.0 / nx, 0.8 / ny, 0.2 / nz
    ax, ay, az = dy * dz, dx * dz, dx * dy
    c = 2.0  # sound speed at M = 0.5 with |V| = 1
    spec = 0.5 * (2 * ax * (1.0 + c) + 2 * ay * c + 2 * az * c)
    nc = m.n_cell
    assert np.allclose(s.get("spec_inv")[:nc], spec, rtol=1e-13)
    assert np.allclose(s.get("dt")[:nc], 10.0 / spec, rtol=1e-13)


def test_residual_norm_by_hand():
    m = box_mesh(4, 2, 2)
    s = oracle_for(m)
    rng = np.random.default_rng(3)
    r = rng.normal(size=(5, m.n_total))
    s.set("res", r)
    assert np.allclose(s.residual_norm(), (r[:, :m.n_cell] ** 2).sum(axis=1), rtol=1e-14)


def test_lusgs_equals_dense_factorised_solve():
    """One forward + backward sweep solves (D+L) D^-1 (D+U) dq = res, with the
    blocks of L and U given by MXDQ_STD at each neighbour (all-farfield mesh,
    so the ghost dq of the backward sweep is zero)."""
    base = pa.Mesh.bump_hex(3, 2, 2, 0.05)
    conn = base.connectivity()
    m = pa.Mesh.from_arrays(conn["x"], conn["y"], conn["z"], conn["face_node_ptr"], conn["face_nodes"],
                            conn["left"], conn["right"], conn["n_bface"],
                            np.full(conn["n_bface"], 4, np.int32), conn["cell_node_ptr"], conn["cell_nodes"])
    conn, met = m.connectivity(), m.metrics()
    p = pa.default_params(mach=0.5, cfl_start=5.0, cfl_end=5.0)
    s = orc.OracleSolver(conn, met, p)
    rng = np.random.default_rng(5)
    q = np.tile(s.freestream()[:, None], (1, m.n_total))
    q[:, :m.n_cell] *= 1 + rng.uniform(-0.05, 0.05, size=(5, m.n_cell))
    s.set("q", q)
    s.residual()
    s.spectrum()
    s.time_step(1)
    res = s.get("res")[:, :m.n_cell].copy()
    qb = s.get("q")
    s.lusgs()
    dq = s.get("dq")[:, :m.n_cell]
    diag = s.get("diag")[:m.n_cell]
    nc = m.n_cell
    L = np.zeros((5 * nc, 5 * nc))
    U = np.zeros((5 * nc, 5 * nc))
    for f in range(m.n_bface, m.n_face):
        a, b = conn["left"][f], conn["right"][f]
        n = np.array([met["xfn"][f], met["yfn"][f], met["zfn"][f]])
        for i, j, nn in ((a, b, n), (b, a, -n)):
            rad = orc.lusgs_rad(qb[:, i], qb[:, j], nn, met["area"][f])
            B = np.column_stack([orc.mxdq(qb[:, j], nn, met["area"][f], GAMMA, e, rad, -1) for e in np.eye(5)])
            tgt = L if j < i else U
            tgt[5 * i:5 * i + 5, 5 * j:5 * j + 5] += B
    D = np.kron(np.diag(diag), np.eye(5))
    r = res.T.reshape(-1)
    dstar = np.linalg.solve(D + L, r)
    x = np.linalg.solve(D + U, D @ dstar)
    assert np.max(np.abs(x.reshape(nc, 5).T - dq)) <= 1e-12 * np.max(np.abs(dq))


def test_viscous_couette_residual_closed_form():
    """Linear shear u = a*y, uniform rho, p on a uniform box: in cells away
    from the boundary the momentum residual vanishes and the energy residual
    is the viscous dissipation mu*a^2*V/Re (inviscid part cancels exactly)."""
    nx, ny, nz = 6, 6, 4
    m = box_mesh(nx, ny, nz)
    met = m.metrics()
    re_num = 1000.0
    s = oracle_for(m, mach=0.5, viscous=1, reynolds=re_num, limiter=1)
    fs = s.freestream()
    a = 0.7
    q = np.tile(fs[:, None], (1, m.n_total))
    q[1] = 0.3 + a * met["ycc"]
    q[2] = 0.0
    q[3] = 0.0
    s.set("q", q)
    # keep the linear field in the ghosts instead of applying the BCs
    s.gamma_t()
    s.viscosity()
    s.gradient()
    s.limiter()
    s.zero_residual()
    s.inviscid_flux()
    s.viscous_flux()
    res = s.get("res")
    xc, yc, zc = met["xcc"], met["ycc"], met["zcc"]
    dx, dy, dz = 3.0 / nx, 0.8 / ny, 0.2 / nz
    inner = np.where((np.abs(xc[:m.n_cell]) < 1.5 - dx) & (yc[:m.n_cell] > dy) & (yc[:m.n_cell] < 0.8 - dy) &
                     (zc[:m.n_cell] > dz) & (zc[:m.n_cell] < 0.2 - dz))[0]
    assert inner.size > 0
    t = fs[4] / (fs[0] / (1.4 * 0.25))  # T = p / (R rho) with R = 1/(gamma M^2)
    ts = 110.4 / 288.15
    mu = t * np.sqrt(t) * (1 + ts) / (t + ts)
    vol = met["vol"][inner]
    scale = fs[4] * dy * dz
    assert np.max(np.abs(res[:4, inner])) <= 1e-12 * scale
    assert np.allclose(res[4, inner], mu * a * a * vol / re_num, rtol=1e-9)


def _wall_box(nx, ny, nz):
    """Uniform box with the lower wall y=0 a solid surface and every other
    boundary an extrapolation (copy) boundary."""
    base = box_mesh(nx, ny, nz)
    c = base.connectivity()
    bc = np.where(c["bc_type"] == 2, 2, 1).astype(np.int32)
    return pa.Mesh.from_arrays(c["x"], c["y"], c["z"], c["face_node_ptr"], c["face_nodes"], c["left"], c["right"],
                               c["n_bface"], bc, c["cell_node_ptr"], c["cell_nodes"])


def test_forces_closed_form_on_a_sheared_flat_wall():
    """Uniform overpressure delta and linear shear u = a*y over the flat wall
    y = 0 of a 3 x 0.8 x 0.2 box: the wall integrals of AirForceCoefParts have
    closed forms (pressure -2*delta*A in y, friction 2*mu*a*A/Re in x with the
    wall-corrected gradient, moment arms about a chosen reference point), and
    CompForce's scalings and wind-axis projections are fixed by alpha."""
    nx, ny, nz = 8, 4, 5
    m = _wall_box(nx, ny, nz)
    met = m.metrics()
    re_num, a, delta, alpha = 2000.0, 0.6, 0.05, 30.0
    x0, y0, z0 = 0.5, 1.0, 0.05
    aref, lref, span = 2.0, 0.5, 4.0
    s = oracle_for(m, mach=0.5, viscous=1, reynolds=re_num, alpha_deg=alpha, force_ref_x=x0, force_ref_y=y0,
                   force_ref_z=z0, force_ref_area=aref, force_ref_length=lref, force_ref_span=span)
    fs = s.freestream()
    q = np.tile(fs[:, None], (1, m.n_total))
    q[1] = a * met["ycc"]
    q[2] = 0.0
    q[3] = 0.0
    q[4] = fs[4] + delta
    s.set("q", q)
    out = s.forces()
    lx, lz, aw = 3.0, 0.2, 3.0 * 0.2
    t = q[4, 0] / (q[0, 0] / (1.4 * 0.25))
    ts = 110.4 / 288.15
    mu = t * np.sqrt(t) * (1 + ts) / (t + ts)
    fx_f = 2.0 * mu * a * aw / re_num          # raw sum of fvsx
    fy_p = -2.0 * delta * aw                   # raw sum of n_y * area * cp
    scf, sb, vcm = 1 / aref, 1 / (aref * span), 1 / (aref * lref)
    expect = {
        0: fx_f * scf, 1: 0.0, 2: 0.0, 3: fy_p * scf, 4: 0.0, 5: 0.0, 6: 0.0, 7: fy_p * scf, 8: 0.0,
        9: 0.0,                                             # Cmx_f: fvsy = fvsz = 0
        10: -(-2.0 * delta) * lx * (lz * lz / 2 - z0 * lz) * sb,   # -sum (z - z0) dpy
        11: fx_f / aw * lx * (lz * lz / 2 - z0 * lz) * sb,        # sum (z - z0) fvsx
        12: 0.0,
        13: y0 * fx_f * vcm,                                 # -(0 - y0) fvsx
        14: (-2.0 * delta) * lz * (0.0 - 3.0 * x0) * vcm,   # sum (x - x0) dpy
    }
    scale = max(abs(v) for v in expect.values())
    for i, v in expect.items():
        assert abs(out[i] - v) <= 1e-12 * scale, (i, out[i], v)
    al = np.radians(alpha)
    cfx, cfy = out[0] + out[1], out[2] + out[3]
    assert out[15] == pytest.approx(-np.sin(al) * cfx + np.cos(al) * cfy, rel=1e-13)
    assert out[16] == pytest.approx(np.cos(al) * cfx + np.sin(al) * cfy, rel=1e-13)
    assert out[18] == pytest.approx(np.cos(al) * out[6] + np.sin(al) * out[7], rel=1e-13)
    assert out[19] == pytest.approx(out[16] - out[18], rel=1e-13)
    assert out[20] == pytest.approx(out[16] - out[15] ** 2 / (np.pi * 9.5), rel=1e-13)
    assert out[21] == pytest.approx(np.sign(cfy) * (out[13] + out[14]) / abs(cfy), rel=1e-13)


def test_forces_inviscid_ignore_shear():
    m = _wall_box(6, 3, 3)
    met = m.metrics()
    s = oracle_for(m, mach=0.5, viscous=0)
    fs = s.freestream()
    q = np.tile(fs[:, None], (1, m.n_total))
    q[1] = 0.8 * met["ycc"]
    s.set("q", q)
    out = s.forces()
    assert np.all(out[[0, 2, 4, 9, 11, 13]] == 0.0)
    assert np.max(np.abs(out[:15])) <= 1e-15  # p = p_inf: no pressure force either
