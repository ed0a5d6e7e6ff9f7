"""Pins of the oracle's entropy fixes 2, 4 and 6 of the Roe flux
(Flux_RoeEntropyCorrection, API/Physics/src/Flux_RoeEntropyCorrection.cpp:
93-166, the build without USE_GMRESSOLVER) and of the pressure factor that
methods 2 and 6 read (Controller::PressureFactor, Controller.cpp:1202-1234;
NSSolverUnstruct::CalPressureFactor, NSSolverUnstruct.cpp:6976-7045).

The flux pins build the Roe flux from an eigen-decomposition of the flux
Jacobian at the Roe-averaged state, with the three eigenvalue magnitudes
(|vn|, |vn + c|, |vn - c|) modified as each method states, so a mistake in
how the oracle assigns or assembles the modified eigenvalues fails them; the
reductions to methods 1 and 3 and the symmetry in the two cells' sensors pin
the max/min choices.  The pressure factor is checked against closed forms
(uniform pressure, a jump beyond and below the ratio thresholds) and a
brute force over the face list."""
import math

import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc
from test_oracle import GAMMA, cons, flux_n, jacobian, random_normal, random_state


def harten(e, lim):
    e = abs(e)
    return 0.5 * (e * e + lim * lim) / lim if e < lim else e


def roe_average(qL, qR):
    r = np.sqrt(qR[0] / qL[0])
    c = 1.0 / (1.0 + r)
    rho = np.sqrt(qL[0] * qR[0])
    vel = (qL[1:4] + qR[1:4] * r) * c
    p = (qL[4] + qR[4] * r) * c
    return rho, vel, p


def textbook_flux(qL, qR, n, modify, g=GAMMA):
    """(FL + FR)/2 - R |L| R^-1 (UR - UL)/2 with |L| from modify(e1, e2, e3,
    vn, c, |V|), e1 the convective and e2, e3 the acoustic magnitudes."""
    rho, vel, p = roe_average(qL, qR)
    A = jacobian(cons(np.array([rho, *vel, p]), g), n, g)
    vn = vel @ n
    H = g / (g - 1.0) * p / rho + 0.5 * vel @ vel
    cm = np.sqrt((g - 1.0) * (H - 0.5 * vel @ vel))
    t1 = np.cross(n, [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    R = np.column_stack([[1.0, *(vel - cm * n), H - cm * vn], [1.0, *vel, 0.5 * vel @ vel],
                         [0.0, *t1, vel @ t1], [0.0, *t2, vel @ t2], [1.0, *(vel + cm * n), H + cm * vn]])
    lam = np.array([vn - cm, vn, vn, vn, vn + cm])
    assert np.allclose(A @ R, R * lam, rtol=1e-12, atol=1e-12)
    e1, e2, e3 = modify(abs(vn), abs(vn + cm), abs(vn - cm), vn, cm, np.linalg.norm(vel))
    absA = R @ np.diag([e3, e1, e1, e1, e2]) @ np.linalg.inv(R)
    UL, UR = cons(qL, g), cons(qR, g)
    return 0.5 * (flux_n(UL, n, g) + flux_n(UR, n, g)) - 0.5 * absA @ (UR - UL)


def method2(pcL, pcR):
    def f(e1, e2, e3, vn, cm, vel):
        ps = max(pcL, pcR)
        l1 = (abs(vn) + cm) * 15.0 * ps
        l2 = (abs(vn) + cm) * (0.25 + 5.0 * ps)
        return harten(e1, l1), harten(e2, l2), harten(e3, l2)
    return f


def method4(eps):
    def f(e1, e2, e3, vn, cm, vel):
        mx = max(e2, e3)
        t = max(mx * eps * min(1.0, abs(vn) / cm), mx * eps / 100.0)
        return max(t, e1), max(t, e2), max(t, e3)
    return f


def method6(pcL, pcR):
    def f(e1, e2, e3, vn, cm, vel):
        lim = (1.0 - min(pcL, pcR)) * (vel + cm) + 1e-4 * cm
        return harten(e1, lim), harten(e2, lim), harten(e3, lim)
    return f


@pytest.mark.parametrize("method", [2, 4, 6])
def test_fixes_match_eigen_decomposition(method):
    rng = np.random.default_rng(40 + method)
    for _ in range(150):
        qL, qR, n = random_state(rng), random_state(rng), random_normal(rng)
        a, b = rng.uniform(0.0, 1.0, 2)
        eps = rng.choice([1e-4, 0.05, 0.3])
        f = orc.roe_flux(qL, qR, n, GAMMA, method, eps, pc=(a, b))
        mod = {2: method2(a, b), 4: method4(eps), 6: method6(a, b)}[method]
        ref = textbook_flux(qL, qR, n, mod)
        assert np.max(np.abs(f - ref)) <= 1e-12 * max(np.max(np.abs(ref)), 1.0)


def test_method2_without_pressure_jump_is_harten_on_the_acoustic_waves():
    """A zero sensor leaves the convective eigenvalue alone and smooths the
    acoustic ones over 0.25 (|vn| + c): method 3 with those coefficients."""
    rng = np.random.default_rng(50)
    for _ in range(50):
        qL, qR, n = random_state(rng), random_state(rng), random_normal(rng)
        a = orc.roe_flux(qL, qR, n, GAMMA, 2, 1e-4, pc=(0.0, 0.0))
        b = orc.roe_flux(qL, qR, n, GAMMA, 3, 0.25, convect=0.0)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("method,pick", [(2, max), (6, min)])
def test_sensor_of_the_face_is_the_larger_or_smaller_cell_value(method, pick):
    rng = np.random.default_rng(51 + method)
    for _ in range(30):
        qL, qR, n = random_state(rng), random_state(rng), random_normal(rng)
        a, b = rng.uniform(0.0, 1.0, 2)
        f1 = orc.roe_flux(qL, qR, n, GAMMA, method, 1e-4, pc=(a, b))
        f2 = orc.roe_flux(qL, qR, n, GAMMA, method, 1e-4, pc=(b, a))
        s = pick(a, b)
        f3 = orc.roe_flux(qL, qR, n, GAMMA, method, 1e-4, pc=(s, s))
        assert np.array_equal(f1, f2) and np.array_equal(f1, f3)


def test_method4_reduces_to_method1_at_sonic_normal_speeds_and_floors_at_a_hundredth():
    rng = np.random.default_rng(52)
    n = np.array([1.0, 0.0, 0.0])
    for _ in range(30):
        eps = rng.choice([0.05, 0.3])
        # |vn| >= c: ratio 1, the limit is method 1's
        qL = np.array([1.0, 2.5, *rng.uniform(-0.3, 0.3, 2), 1.0])
        qR = qL * (1 + rng.uniform(-0.05, 0.05, 5))
        assert np.array_equal(orc.roe_flux(qL, qR, n, GAMMA, 4, eps), orc.roe_flux(qL, qR, n, GAMMA, 1, eps))
        # vn = 0: the floor of a hundredth of method 1's limit
        qL = np.array([1.0, 0.0, *rng.uniform(-0.3, 0.3, 2), 1.0])
        qR = qL.copy()
        qR[[0, 4]] *= 1 + rng.uniform(-0.05, 0.05, 2)
        qR[1] = -qL[1]
        f4 = orc.roe_flux(qL, qR, n, GAMMA, 4, eps)
        f1 = orc.roe_flux(qL, qR, n, GAMMA, 1, eps / 100.0)
        assert np.max(np.abs(f4 - f1)) <= 1e-14 * np.max(np.abs(f1))


def brute_pressure_factor(conn, q, method):
    nc, nb = conn["n_cell"], conn["n_bface"]
    left, right = conn["left"], conn["right"]
    faces = [[] for _ in range(nc)]
    for f in range(len(left)):
        faces[left[f]].append(f)
        if f >= nb:
            faces[right[f]].append(f)
    p = q[4]
    out = np.zeros(q.shape[1])
    for c in range(nc):
        if method == 2:
            sl = sr = 0.0
            for f in faces[c]:
                sl += p[left[f]]
                sr += p[right[f]]
            out[c] = abs(sr - sl) / abs(sr + sl) / len(faces[c])
        else:
            ps = [p[c]] + [p[right[f] if left[f] == c else left[f]] for f in faces[c] if f >= nb]
            ratio = max(ps) / min(ps)
            fn = min(1.0, max(0.0, (2.6 - ratio) / (2.6 - 1.3)))
            out[c] = 0.5 * (1.0 - max(-1.0, min(math.cos(fn * 3.141592653589), 1.0)))
    for f in range(nb):
        out[right[f]] = out[left[f]]
    return out


def solver_with_state(method, q_fn):
    m = pa.Mesh.bump_hybrid(6, 2, 3, 1e-3, 3, 0.1, 5)
    conn, met = m.connectivity(), m.metrics()
    s = orc.OracleSolver(conn, met, pa.default_params(mach=0.5, entropy_fix=method))
    s.init_freestream()
    q = s.get("q")
    q[:, :m.n_cell] = q_fn(m, met, q[:, :m.n_cell])
    s.set("q", q)
    s.boundary()
    s.gamma_t()
    s.pressure_factor()
    conn = dict(conn, n_cell=m.n_cell, n_bface=m.n_bface)
    return m, conn, s


@pytest.mark.parametrize("method", [2, 6])
def test_pressure_factor_brute_force(method):
    rng = np.random.default_rng(60 + method)

    def q_fn(m, met, q):
        q = q.copy()
        q[4] *= np.exp(rng.uniform(-0.6, 0.6, m.n_cell))  # neighbour ratios across 1.3 .. 2.6
        return q
    m, conn, s = solver_with_state(method, q_fn)
    got = s.get("pressure_factor")
    ref = brute_pressure_factor(conn, s.get("q"), method)
    assert np.array_equal(got, ref)
    if method == 6:  # the sensor really spans its range
        assert got.min() == 0.0 and got.max() == 1.0 and np.any((got > 0.05) & (got < 0.95))


def test_pressure_factor_closed_forms():
    # uniform pressure: method 2 sees no jump, method 6 a ratio of 1 (f = 1)
    m, conn, s = solver_with_state(2, lambda m, met, q: q)
    assert np.all(s.get("pressure_factor") == 0.0)
    m, conn, s = solver_with_state(6, lambda m, met, q: q)
    assert np.all(s.get("pressure_factor") == 1.0)

    # a pressure jump of 3x at x = 0.5: method 6 gives 0 on both sides of it
    def jump(m, met, q):
        q = q.copy()
        q[4] = np.where(met["xcc"][:m.n_cell] > 0.5, 3.0 * q[4], q[4])
        return q
    m, conn, s = solver_with_state(6, jump)
    rt = s.get("pressure_factor")[:m.n_cell]
    left, right, nb = conn["left"], conn["right"], m.n_bface
    hi = s.get("q")[4][:m.n_cell] > 2.0 * s.freestream()[4]
    touching = np.zeros(m.n_cell, bool)
    for f in range(nb, len(left)):
        if hi[left[f]] != hi[right[f]]:
            touching[left[f]] = touching[right[f]] = True
    assert touching.any() and np.all(rt[touching] == 0.0) and np.all(rt[~touching] == 1.0)
    # method 2: zero in interior cells away from the jump (boundary cells see
    # their ghosts' pressures), positive at the jump
    interior = np.ones(m.n_cell, bool)
    interior[left[:nb]] = False
    m, conn, s = solver_with_state(2, jump)
    rt2 = s.get("pressure_factor")[:m.n_cell]
    assert np.all(rt2[~touching & interior] == 0.0) and np.all(rt2[touching] > 0.0)


@pytest.mark.parametrize("method", [2, 4, 6])
def test_residual_uses_the_fix(method):
    """The residual of a perturbed state differs from method 3's."""
    m = pa.Mesh.bump_hybrid(6, 2, 3, 1e-3, 3, 0.1, 5)
    conn, met = m.connectivity(), m.metrics()
    scale = 1 + np.random.default_rng(70).uniform(-0.3, 0.3, size=(5, m.n_cell))
    res = {}
    for meth in (3, method):
        s = orc.OracleSolver(conn, met, pa.default_params(mach=0.5, entropy_fix=meth))
        s.init_freestream()
        q = s.get("q")
        q[:, :m.n_cell] *= scale
        s.set("q", q)
        s.residual()
        res[meth] = s.get("res")[:, :m.n_cell].copy()
    assert not np.array_equal(res[3], res[method])
