"""Pins of the oracle functions that the restatement alone left open (runs on
the CPU only).  Each check is computed here from the mathematics or the
reference's stated rule, independently of the oracle's code:

- farfield ghost states: Riemann invariants, entropy and tangential velocity
  of the upwind side (NSSolverUnstruct.cpp:21195-21330), subsonic and
  supersonic, inflow and outflow;
- the SA farfield ghost: in/outflow by the boundary normal velocity of the NS
  ghost (TurbSolverUnstr.cpp:6546);
- SolutionFix on a small mesh: the update and the neighbour averages redone
  here in increasing cell order (NSSolverUnstruct.cpp:14509-14716);
- ComputePTotal (PerfectGas.cpp:1609-1641) against the Rayleigh pitot
  formula (NACA 1135 eq. 100), and the low-pressure CFL branch of LimitCFL
  (:6052-6117) as a piecewise-linear law of the pressure;
- the Venkatakrishnan limiter as a whole: brute force over 2-8 cell meshes
  written from Limiter.cpp:1091-1254 and SetLimitBoundary
  (Geo_UnstructGrid.cpp:10744-10777), bounds, zero beside farfield faces, one
  on linear fields away from boundaries;
- MUSCL reconstruction: exact for linear fields, and the first-order fallback
  of PositiveCheckForDensityAndPressure (FaceProxy.h:209);
- the SA implicit source linearisation (dsdnu, dfwdg, drdnu in
  TurbSolverUnstr.cpp:2702-2916) against central finite differences of the
  source.
"""
import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc

G = 1.4
BC_WALL, BC_SYMM, BC_FAR = 2, 3, 4


def oracle_for(mesh, **p):
    return orc.OracleSolver(mesh.connectivity(), mesh.metrics(), pa.default_params(**p))


def freestream(mach, alpha_deg):
    a = np.radians(alpha_deg)
    return np.array([1.0, np.cos(a), np.sin(a), 0.0, 1.0 / (G * mach * mach)])


# ----------------------------------------------------------- farfield BC
def riemann_ghost(qi, qo, n, g=G):
    """Characteristic farfield state at a boundary face of outward normal n
    (Blazek, Computational Fluid Dynamics, 3rd ed., sec. 8.2.1, with the
    Riemann invariants R+- = vn +- 2c/(g-1))."""
    ri, vi, pi = qi[0], np.asarray(qi[1:4]), qi[4]
    ro, vo, po = qo[0], np.asarray(qo[1:4]), qo[4]
    ci, co = np.sqrt(g * pi / ri), np.sqrt(g * po / ro)
    vni, vno = vi @ n, vo @ n
    if np.linalg.norm(vi) > ci:  # supersonic: everything from the upwind side
        return np.array(qi, float) if vni >= 0.0 else np.array(qo, float)
    r_out = vni + 2.0 * ci / (g - 1.0)   # leaves the domain
    r_in = vno - 2.0 * co / (g - 1.0)    # enters from the farfield
    vnb = 0.5 * (r_out + r_in)
    cb = 0.25 * (g - 1.0) * (r_out - r_in)
    up_r, up_v, up_p = (ri, vi, pi) if vnb >= 0.0 else (ro, vo, po)
    entropy = up_p / up_r ** g
    vt = up_v - (up_v @ n) * n
    rb = (cb * cb / (g * entropy)) ** (1.0 / (g - 1.0))
    pb = rb * cb * cb / g
    return np.array([rb, *(vt + vnb * n), pb])


@pytest.mark.parametrize("mach,alpha,spread", [(0.5, 0.0, 0.05), (0.5, 3.0, 0.8), (2.0, 0.0, 0.05),
                                               (2.0, 10.0, 0.3)])
def test_farfield_ghost_is_the_riemann_invariant_state(mach, alpha, spread):
    """Every farfield ghost equals the characteristic state computed here,
    and it carries the interior's outgoing and the freestream's incoming
    Riemann invariant (subsonic), or the upwind state (supersonic).  A large
    `spread` of the interior velocity turns some outlet faces into inflow."""
    m = pa.Mesh.bump_hex(6, 3, 2, 0.05)
    conn, met = m.connectivity(), m.metrics()
    cpu = orc.OracleSolver(conn, met, pa.default_params(mach=mach, alpha_deg=alpha))
    qo = freestream(mach, alpha)
    rng = np.random.default_rng(int(mach * 10 + alpha))
    q = np.tile(qo[:, None], (1, m.n_total))
    nc = m.n_cell
    q[0, :nc] *= rng.uniform(0.9, 1.1, nc)
    q[1:4, :nc] += rng.uniform(-spread, spread, (3, nc)) * np.linalg.norm(qo[1:4])
    q[4, :nc] *= rng.uniform(0.9, 1.1, nc)
    cpu.set("q", q)
    cpu.boundary()
    out = cpu.get("q")
    seen = {"sub_in": 0, "sub_out": 0, "sup_in": 0, "sup_out": 0}
    for f in range(m.n_bface):
        if conn["bc_type"][f] != BC_FAR:
            continue
        le, gh = conn["left"][f], nc + f
        n = np.array([met["xfn"][f], met["yfn"][f], met["zfn"][f]])
        qi = q[:, le]
        expect = riemann_ghost(qi, qo, n)
        got = out[:, gh]
        assert np.allclose(got, expect, rtol=1e-13, atol=1e-13), (f, got, expect)
        ci = np.sqrt(G * qi[4] / qi[0])
        if np.linalg.norm(qi[1:4]) > ci:
            seen["sup_out" if qi[1:4] @ n >= 0 else "sup_in"] += 1
        else:
            cb = np.sqrt(G * got[4] / got[0])
            vnb = got[1:4] @ n
            assert abs((vnb + 2 * cb / (G - 1)) - (qi[1:4] @ n + 2 * ci / (G - 1))) <= 1e-12
            co = np.sqrt(G * qo[4] / qo[0])
            assert abs((vnb - 2 * cb / (G - 1)) - (qo[1:4] @ n - 2 * co / (G - 1))) <= 1e-12
            seen["sub_out" if vnb >= 0 else "sub_in"] += 1
    kinds = [k for k, v in seen.items() if v]
    print(mach, alpha, seen)
    if mach < 1:
        assert seen["sub_in"] and seen["sub_out"]
    else:
        assert seen["sup_in"] and seen["sup_out"]
    assert kinds


def test_sa_farfield_ghost_follows_the_boundary_normal_velocity():
    """SA farfield ghost: an outflow face (boundary normal velocity >= 0)
    copies nu~ and mu_t from the interior, an inflow face takes the
    freestream values (muoo, freeStreamViscosity)."""
    m = pa.Mesh.bump_hex(6, 3, 2, 0.05)
    conn, met = m.connectivity(), m.metrics()
    p = pa.default_params(mach=0.5, alpha_deg=3.0, viscous=3, reynolds=1e6, sa_nu_inf=3.0, sa_vist_inf=1e-3)
    cpu = orc.OracleSolver(conn, met, p)
    cpu.init_freestream()
    qo = freestream(0.5, 3.0)
    rng = np.random.default_rng(3)
    q = cpu.get("q").copy()
    nc = m.n_cell
    q[1:4, :nc] += rng.uniform(-0.8, 0.8, (3, nc))
    cpu.set("q", q)
    nut = np.full(m.n_total, 3.0)
    nut[:nc] = rng.uniform(5.0, 50.0, nc)
    vist = np.zeros(m.n_total)
    vist[:nc] = rng.uniform(0.1, 2.0, nc)
    cpu.set("nut", nut)
    cpu.set("vist", vist)
    cpu.boundary()
    cpu.sa_boundary()
    qg, ng, vg = cpu.get("q"), cpu.get("nut"), cpu.get("vist")
    kinds = set()
    for f in range(m.n_bface):
        if conn["bc_type"][f] != BC_FAR:
            continue
        le, gh = conn["left"][f], nc + f
        n = np.array([met["xfn"][f], met["yfn"][f], met["zfn"][f]])
        vnb = riemann_ghost(q[:, le], qo, n)[1:4] @ n
        if vnb >= 0.0:
            assert ng[gh] == nut[le] and vg[gh] == vist[le]
            kinds.add("out")
        else:
            assert ng[gh] == 3.0 and vg[gh] == 1e-3
            kinds.add("in")
    assert kinds == {"in", "out"}


# ------------------------------------------------------------ SolutionFix
def prim_to_cons(q, g=G):
    r, u, v, w, p = q
    return np.array([r, r * u, r * v, r * w, p / (g - 1.0) + 0.5 * r * (u * u + v * v + w * w)])


def cons_to_prim(U, g=G):
    r = U[0]
    u, v, w = U[1] / r, U[2] / r, U[3] / r
    return np.array([r, u, v, w, (g - 1.0) * (U[4] - 0.5 * r * (u * u + v * v + w * w))])


@pytest.mark.parametrize("inject", [(1e-3, 1e-7), (1e-8, 1e-8)])
def test_solution_fix_by_hand(inject):
    """Redo UpdateFlowField + SolutionFix here: Q += dq in every real cell;
    a cell whose rho or p falls below 1e-6 of freestream gets the average of
    its admissible neighbours (|rho|, u, v, w, |p|), reading a neighbour of
    lower index from the new field and one of higher index (or a ghost) from
    the old one, exactly as the reference's in-place loop sees them."""
    m = pa.Mesh.bump_hex(16, 8, 4, 0.05)
    conn = m.connectivity()
    cpu = oracle_for(m, mach=0.5, cfl_start=10.0, cfl_end=10.0)
    cpu.init_freestream()
    q = cpu.get("q").copy()
    cells = [100, 101, 116, 117, 300]
    q[0, cells] *= inject[0]
    q[4, cells] *= inject[1]
    cpu.set("q", q)
    cpu.residual()
    cpu.spectrum()
    cpu.time_step(1)
    q_old = cpu.get("q").copy()  # real cells as set, ghosts from this step's Boundary
    nfix = cpu.lusgs()
    dq, q_new = cpu.get("dq"), cpu.get("q")
    nc = m.n_cell
    qinf = freestream(0.5, 0.0)
    rlim, plim = 1e-6 * qinf[0], 1e-6 * qinf[4]
    nbrs = [[] for _ in range(nc)]
    for f, (le, re) in enumerate(zip(conn["left"], conn["right"])):
        nbrs[le].append(re)
        if f >= m.n_bface:
            nbrs[re].append(le)
    expect = q_old.copy()
    flagged = 0
    for c in range(nc):
        upd = cons_to_prim(prim_to_cons(q_old[:, c]) + dq[:, c])
        if upd[0] < rlim or upd[4] < plim:
            flagged += 1
            acc, k = np.zeros(5), 0
            for nb in nbrs[c]:
                v = expect[:, nb] if nb < c else q_old[:, nb]
                if v[0] < rlim or v[4] < plim:
                    continue
                acc += np.array([abs(v[0]), v[1], v[2], v[3], abs(v[4])])
                k += 1
            upd = acc / k if k else qinf
        expect[:, c] = upd
    print("flagged", flagged, "oracle", nfix)
    assert nfix == flagged > 0
    scale = np.max(np.abs(expect[:, :nc]), axis=1, keepdims=True)
    assert np.max(np.abs(q_new[:, :nc] - expect[:, :nc]) / scale) <= 1e-12


# --------------------------------------------------- pTotal and LimitCFL
def rayleigh_pitot(mach, g=G):
    """Pitot pressure behind a normal shock over the static pressure
    (NACA 1135 eq. 100)."""
    m2 = mach * mach
    return ((g + 1.0) * m2 / 2.0) ** (g / (g - 1.0)) * ((g + 1.0) / (2.0 * g * m2 - (g - 1.0))) ** (1.0 / (g - 1.0))


def test_ptotal_is_the_rayleigh_pitot_pressure():
    """ComputePTotal is the pitot pressure behind a normal shock (c1/c2 is the
    Mach number squared behind the shock).  Below Mach 1 it keeps that form
    without the static-pressure jump, so pTotal times the jump equals the
    pitot formula's continuation; below M = sqrt((g-1)/(2g)) = 0.378 its base
    is negative and pTotal is NaN (the low-pressure CFL guard is then off,
    since every comparison with NaN is false)."""
    for mach in (1.1, 1.5, 2.0, 3.0, 5.0):
        p_inf = 1.0 / (G * mach * mach)
        assert np.isclose(orc.load().orc_ptotal(mach, G), p_inf * rayleigh_pitot(mach), rtol=1e-13)
    for mach in (0.4, 0.5, 0.6, 0.8, 0.95):
        p_inf = 1.0 / (G * mach * mach)
        jump = (2.0 * G * mach * mach - (G - 1.0)) / (G + 1.0)
        assert np.isclose(orc.load().orc_ptotal(mach, G) * jump, p_inf * rayleigh_pitot(mach), rtol=1e-13)
    for mach in (0.1, 0.3, 0.37):
        assert np.isnan(orc.load().orc_ptotal(mach, G))


def test_low_pressure_cfl_branch_is_piecewise_linear():
    """LimitCFL: the local CFL is CFLStart below 0.001 pTotal, CFL above
    0.02 pTotal, and linear in p in between (continuous at both ends); the
    time step is CFL_cell V / spectral radius (undivided skewness on a box)."""
    mach, cfl_start, cfl_end = 0.5, 0.5, 10.0
    m = pa.Mesh.bump_hex(8, 6, 4, 0.0)
    cpu = oracle_for(m, mach=mach, cfl_start=cfl_start, cfl_end=cfl_end, cfl_vary_step=10)
    cpu.init_freestream()
    nc = m.n_cell
    p_inf = 1.0 / (G * mach * mach)
    jump = (2.0 * G * mach * mach - (G - 1.0)) / (G + 1.0)
    p_total = p_inf * rayleigh_pitot(mach) / jump
    lo, hi = 0.001 * p_total, 0.02 * p_total
    q = cpu.get("q").copy()
    q[4, :nc] = np.geomspace(0.2 * lo, 3.0 * hi, nc)
    cpu.set("q", q)
    cpu.boundary()
    cpu.spectrum()
    cpu.time_step(100)
    cfl = cpu.get("dt")[:nc] * cpu.get("spec_inv")[:nc]  # dt is stored divided by the volume
    p = q[4, :nc]
    assert np.allclose(cfl[p < lo], cfl_start, rtol=1e-13)
    assert np.allclose(cfl[p >= hi], cfl_end, rtol=1e-13)
    mid = (p >= lo) & (p < hi)
    assert mid.sum() > 10
    assert np.allclose(cfl[mid], cfl_start + (cfl_end - cfl_start) * (p[mid] - lo) / (hi - lo), rtol=1e-11)


# ---------------------------------------------------------- Venkat limiter
def venfun(x, y, eps):
    """Math_Limiter.h:174: (x^2 + eps + 2xy) / (x^2 + 2y^2 + xy + eps)."""
    return (x * x + eps + 2.0 * x * y) / (x * x + 2.0 * y * y + x * y + eps)


def brute_force_limiter(conn, met, q, grads, K=5.0):
    """Limiter.cpp:47-145 (scalar limiter over rho and p), VencatLimiter7
    (:1091-1199), MinMaxDiff (:1202-1254: ghosts of physical boundaries do
    not enter the min/max) and SetLimitBoundary (zero in cells and ghosts of
    every boundary that is not symmetry, wall or outflow), cell by cell."""
    nc = len(conn["cell_node_ptr"]) - 1
    nb = conn["n_bface"]
    left, right = conn["left"], conn["right"]
    vol = met["vol"][:nc]
    avg = vol.sum() / nc
    faces_of = [[] for _ in range(nc)]
    for f in range(len(left)):
        faces_of[left[f]].append(f)
        if f >= nb:
            faces_of[right[f]].append(f)
    lim = np.ones(nc + nb)
    for var in (0, 4):
        qq = q[var]
        gx, gy, gz = grads[0][var], grads[1][var], grads[2][var]
        for c in range(nc):
            nbv = [qq[right[f] if left[f] == c else left[f]] for f in faces_of[c] if f >= nb]
            dmax = max([qq[c]] + nbv) - qq[c]
            dmin = min([qq[c]] + nbv) - qq[c]
            eps = K ** 3 / (avg ** (1.0 / 3.0)) ** 3 * qq[c] ** 2 * (vol[c] ** (1.0 / 3.0)) ** 3
            for f in faces_of[c]:
                d = gx[c] * (met["xfc"][f] - met["xcc"][c]) + gy[c] * (met["yfc"][f] - met["ycc"][c]) + \
                    gz[c] * (met["zfc"][f] - met["zcc"][c])
                if d > 1e-40:
                    lim[c] = min(lim[c], venfun(dmax, d, eps))
                elif d < -1e-40:
                    lim[c] = min(lim[c], venfun(dmin, d, eps))
    for f in range(nb):
        if conn["bc_type"][f] not in (BC_SYMM, BC_WALL, 6):
            lim[left[f]] = 0.0
            lim[nc + f] = 0.0
    return lim


@pytest.mark.parametrize("dims", [(2, 1, 1), (2, 2, 1), (2, 2, 2), (3, 2, 1), (4, 2, 1)])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_venkat_limiter_brute_force_on_small_meshes(dims, seed):
    """The oracle's limiter equals the cell-by-cell brute force on meshes of
    2 to 8 cells (the ones of 2 to 4 cells in x have cells away from the
    farfield ends on the larger ones), with rough random states."""
    m = pa.Mesh.bump_hex(*dims, 0.05)
    conn, met = m.connectivity(), m.metrics()
    cpu = orc.OracleSolver(conn, met, pa.default_params(mach=0.5))
    rng = np.random.default_rng(seed)
    q = np.tile(freestream(0.5, 0.0)[:, None], (1, m.n_total))
    q[:, :m.n_cell] *= rng.uniform(0.5, 1.5, (5, m.n_cell))
    cpu.set("q", q)
    cpu.boundary()
    cpu.gradient()
    cpu.limiter()
    qb = cpu.get("q")
    grads = (cpu.get("dqdx"), cpu.get("dqdy"), cpu.get("dqdz"))
    expect = brute_force_limiter(conn, met, qb, grads)
    got = cpu.get("limit")
    assert np.allclose(got, expect, rtol=1e-14, atol=1e-15)


def test_venkat_limiter_on_a_hybrid_mesh_brute_force_and_bounds():
    """On a prism/tet mesh with a rough state: equal to the brute force,
    inside [0, 1], zero in every cell next to a farfield face, and below 1
    somewhere (the limiter is active)."""
    m = pa.Mesh.bump_hybrid(6, 2, 3, 1e-2, 3, 0.1, 5)
    conn, met = m.connectivity(), m.metrics()
    cpu = orc.OracleSolver(conn, met, pa.default_params(mach=0.5))
    rng = np.random.default_rng(9)
    q = np.tile(freestream(0.5, 0.0)[:, None], (1, m.n_total))
    q[:, :m.n_cell] *= rng.uniform(0.7, 1.3, (5, m.n_cell))
    cpu.set("q", q)
    cpu.boundary()
    cpu.gradient()
    cpu.limiter()
    got = cpu.get("limit")
    expect = brute_force_limiter(conn, met, cpu.get("q"), (cpu.get("dqdx"), cpu.get("dqdy"), cpu.get("dqdz")))
    assert np.allclose(got, expect, rtol=1e-14, atol=1e-15)
    nc = m.n_cell
    assert np.all(got >= 0.0) and np.all(got <= 1.0)
    far = conn["left"][:m.n_bface][conn["bc_type"] == BC_FAR]
    assert far.size and np.all(got[far] == 0.0)
    assert np.any((got[:nc] > 0.0) & (got[:nc] < 0.999))


def test_venkat_limiter_is_one_on_linear_fields_away_from_boundaries():
    """Linear rho and p on a uniform box: in cells not touching a boundary the
    Green-Gauss gradient is exact, the face increment is half the neighbour
    difference, VenFun(2y, y, eps) = 1, so the limiter is 1 (to rounding)."""
    m = pa.Mesh.bump_hex(10, 6, 5, 0.0)
    conn, met = m.connectivity(), m.metrics()
    cpu = orc.OracleSolver(conn, met, pa.default_params(mach=0.5))
    x, y, z = met["xcc"], met["ycc"], met["zcc"]
    q = np.tile(freestream(0.5, 0.0)[:, None], (1, m.n_total))
    q[0] = 1.0 + 0.1 * x + 0.05 * y - 0.2 * z
    q[4] = q[4] * (1.0 - 0.05 * x + 0.1 * y + 0.1 * z)
    cpu.set("q", q)
    cpu.gradient()
    cpu.limiter()
    lim = cpu.get("limit")
    inner = np.ones(m.n_cell, bool)
    inner[conn["left"][:m.n_bface]] = False
    assert inner.sum() > 20
    assert np.all(np.abs(lim[:m.n_cell][inner] - 1.0) <= 1e-12)


# ------------------------------------------------------------------ MUSCL
def test_muscl_is_exact_for_linear_fields():
    """With no limiter (phi = 1), the reconstructed face state of an interior
    face between two interior cells is the linear field at the face centre."""
    m = pa.Mesh.bump_hex(10, 6, 5, 0.0)
    conn, met = m.connectivity(), m.metrics()
    cpu = orc.OracleSolver(conn, met, pa.default_params(mach=0.5, limiter=1))
    a = np.array([[1.0, 0.1, 0.05, -0.2], [0.8, 0.02, 0.0, 0.01], [0.1, 0.0, 0.03, 0.0],
                  [0.0, 0.01, -0.02, 0.05], [2.8, -0.1, 0.2, 0.1]])

    def lin(x, y, z):
        return np.array([r[0] + r[1] * x + r[2] * y + r[3] * z for r in a])

    q = lin(met["xcc"], met["ycc"], met["zcc"])
    cpu.set("q", q)
    cpu.gradient()
    cpu.limiter()  # nolim: phi = 1
    inner = np.ones(m.n_cell, bool)
    inner[conn["left"][:m.n_bface]] = False
    checked = 0
    for f in range(m.n_bface, m.n_face):
        if inner[conn["left"][f]] and inner[conn["right"][f]]:
            qL, qR = cpu.face_states(f)
            exact = lin(met["xfc"][f], met["yfc"][f], met["zfc"][f])
            assert np.allclose(qL, exact, rtol=0, atol=1e-13) and np.allclose(qR, exact, rtol=0, atol=1e-13)
            checked += 1
    assert checked > 50


def test_muscl_falls_back_to_first_order_on_negative_density_or_pressure():
    """PositiveCheckForDensityAndPressure: a side whose reconstructed rho or
    p is not positive keeps its cell value (first order); the other side is
    still reconstructed.  A rough state makes both cases occur."""
    m = pa.Mesh.bump_hex(8, 6, 3, 0.05)
    conn, met = m.connectivity(), m.metrics()
    cpu = orc.OracleSolver(conn, met, pa.default_params(mach=0.5, limiter=1))
    rng = np.random.default_rng(4)
    q = np.tile(freestream(0.5, 0.0)[:, None], (1, m.n_total))
    nc = m.n_cell
    q[0, :nc] = rng.uniform(0.02, 2.0, nc)
    q[4, :nc] *= rng.uniform(0.02, 2.0, nc)
    cpu.set("q", q)
    cpu.boundary()
    cpu.gradient()
    cpu.limiter()  # nolim: phi = 1
    assert np.all(cpu.get("limit") == 1.0)
    qb = cpu.get("q")
    gx, gy, gz = cpu.get("dqdx"), cpu.get("dqdy"), cpu.get("dqdz")
    first = recon = 0
    for f in range(m.n_face):
        if f < m.n_bface and conn["bc_type"][f] in (BC_WALL, BC_SYMM):
            continue  # BoundaryQlQrFix replaces both sides there
        qL, qR = cpu.face_states(f)
        for side, got in ((conn["left"][f], qL), (conn["right"][f], qR)):
            d = np.array([met["xfc"][f], met["yfc"][f], met["zfc"][f]]) - \
                np.array([met["xcc"][side], met["ycc"][side], met["zcc"][side]])
            g = (gx[:, side] * d[0] + gy[:, side] * d[1] + gz[:, side] * d[2]) if side < nc else np.zeros(5)
            trial = qb[:, side] + 1.0 * g
            if trial[0] <= 0.0 or trial[4] <= 0.0:
                assert np.array_equal(got, qb[:, side])
                first += 1
            else:
                assert np.allclose(got, trial, rtol=1e-14, atol=1e-15)
                recon += 1
    print("first-order sides", first, "reconstructed sides", recon)
    assert first > 10 and recon > 100


# ---------------------------------------------------- SA linearisation
@pytest.mark.parametrize("nut,omega,d", [(3.0, 50.0, 1e-3), (30.0, 5.0, 1e-2), (300.0, 2.0, 0.05),
                                         (1.0, 200.0, 2e-4), (80.0, 0.5, 0.2)])
def test_sa_source_linearisation_matches_finite_differences(nut, omega, d):
    """On the main branch (nu~ > 0, S~ = Omega + S-bar, r below its clip at
    10) the reference's linearisation prdep (built from dsdnu, dfwdg, dgdr,
    drdnu) is d/dnu~ of (P - D)/nu~: central finite differences of the
    oracle's own source agree to 1e-6, and (P - D)/nu~ times nu~ is the
    source."""
    rho, mu, re = 1.0, 1.0, 3e6
    s, prde, prdep, _ = orc.sa_source_point(nut, omega, d, rho, mu, re)
    assert np.isclose(prde * nut, s, rtol=1e-12)
    h = 1e-5 * nut
    fp = orc.sa_source_point(nut + h, omega, d, rho, mu, re)[1]
    fm = orc.sa_source_point(nut - h, omega, d, rho, mu, re)[1]
    fd = (fp - fm) / (2.0 * h)
    kap2 = 0.41 ** 2
    rs = nut / (kap2 * d * d * re)
    assert rs / (omega + rs) < 10.0  # below the clip of r
    print(nut, omega, d, "prdep", prdep, "fd", fd)
    assert abs(prdep - fd) <= 1e-6 * max(abs(fd), abs(prdep))


def test_sa_implicit_term_is_the_negative_part_of_the_linearised_source():
    """The diagonal term per volume is -min(prde, 0) - min(prdep, 0) nu~: only
    the parts that make the update more stable enter the diagonal."""
    for args in [(3.0, 50.0, 1e-3), (30.0, 5.0, 1e-2), (300.0, 2.0, 0.05), (-2.0, 10.0, 1e-2)]:
        s, prde, prdep, diag = orc.sa_source_point(*args, 1.0, 1.0, 3e6)
        assert np.isclose(diag, -min(prde, 0.0) - min(prdep, 0.0) * args[0], rtol=1e-14, atol=1e-300)


@pytest.mark.parametrize("t_wall", [200.0, 400.0, 20.0])
def test_isothermal_wall_ghost(t_wall):
    """wallTemperature > 0 (ViscousIsotropicWallBCRegion): the ghost of a wall
    face keeps the pressure, reverses the velocity, and carries the
    temperature mirrored about T_w (T_ghost = 2 T_w - T, floored at T_w / 10);
    the face mean of the temperature is then T_w unless floored; the density
    follows from p = rho R T."""
    m = pa.Mesh.bump_hex(8, 4, 2, 0.05)
    conn, met = m.connectivity(), m.metrics()
    p = pa.default_params(mach=0.5, viscous=1, reynolds=1e5, t_wall=t_wall)
    s = orc.OracleSolver(conn, met, p)
    rng = np.random.default_rng(3)
    q = np.tile(s.freestream()[:, None], (1, m.n_total))
    q[:, :m.n_cell] *= 1 + rng.uniform(-0.1, 0.1, size=(5, m.n_cell))
    s.set("q", q)
    s.boundary()
    qb = s.get("q")
    R = 1.0 / (1.4 * 0.5 ** 2)
    tw = t_wall / 288.15
    walls = np.flatnonzero(conn["bc_type"] == 2)
    le, gh = conn["left"][walls], m.n_cell + walls
    np.testing.assert_allclose(qb[4, gh], qb[4, le], rtol=0)
    np.testing.assert_allclose(qb[1:4, gh], -qb[1:4, le], rtol=0)
    t_in = qb[4, le] / (R * qb[0, le])
    t_gh = qb[4, gh] / (R * qb[0, gh])
    expect = np.maximum(2 * tw - t_in, 0.1 * tw)
    np.testing.assert_allclose(t_gh, expect, rtol=1e-14)
    unfloored = 2 * tw - t_in > 0.1 * tw
    np.testing.assert_allclose(0.5 * (t_in + t_gh)[unfloored], tw, rtol=1e-14)
    if t_wall == 20.0:
        assert not unfloored.any()  # a wall far colder than the gas: the floor everywhere
    else:
        assert unfloored.all()
