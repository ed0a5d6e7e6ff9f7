"""Pins the oracle's Spalart-Allmaras restatement (NSMB formation) with checks
that do not retype its code: exact wall distances (flat wall, dense sampling
of a bump), the source term against the published SA formulation (Allmaras,
Johnson & Spalart 2012) on fields where convection and diffusion vanish, the
diffusion operator against the continuous SA diffusion on a linear field,
the upwind convection of a uniform stream, the wall ghost rule, and the
scalar LU-SGS against a dense factorised solve."""
import numpy as np
import pytest

import oracle as orc
import phenlei_amd as pa

SA = 3


def make(mesh, **p):
    conn, met = mesh.connectivity(), mesh.metrics()
    params = pa.default_params(viscous=SA, reynolds=1e5, mach=0.5, **p)
    s = orc.OracleSolver(conn, met, params)
    wd = orc.wall_distance(conn, met)
    s.set_wall_distance(wd)
    return s, conn, met, wd


def test_wall_distance_flat_wall_is_height():
    m = pa.Mesh.bump_hex(8, 6, 3, 0.0)  # flat wall at y = 0
    conn, met = m.connectivity(), m.metrics()
    wd = orc.wall_distance(conn, met)
    assert np.allclose(wd, met["ycc"][:m.n_cell], rtol=0, atol=1e-14)


def test_wall_distance_bump_against_dense_sampling():
    m = pa.Mesh.bump_hex(12, 6, 2, 0.08)
    conn, met = m.connectivity(), m.metrics()
    wd = orc.wall_distance(conn, met)
    # sample the piecewise-planar wall densely through its face triangles
    nb = m.n_bface
    pts = []
    fp = conn["face_node_ptr"]
    for f in range(nb):
        if conn["bc_type"][f] != 2:
            continue
        c = np.array([met["xfc"][f], met["yfc"][f], met["zfc"][f]])
        nodes = conn["face_nodes"][fp[f]:fp[f + 1]]
        P = np.stack([conn["x"][nodes], conn["y"][nodes], conn["z"][nodes]], axis=1)
        for i in range(len(nodes)):
            a, b = P[i], P[(i + 1) % len(nodes)]
            for u in np.linspace(0, 1, 13):
                for v in np.linspace(0, 1 - u, max(2, int(13 * (1 - u)))):
                    pts.append(c + u * (a - c) + v * (b - c))
    pts = np.array(pts)
    cc = np.stack([met["xcc"], met["ycc"], met["zcc"]], axis=1)[:m.n_cell]
    for i in range(0, m.n_cell, 7):
        dmin = np.min(np.linalg.norm(pts - cc[i], axis=1))
        assert wd[i] <= dmin + 1e-13          # exact distance is never above a sample
        assert dmin - wd[i] <= 2e-3           # and the sampling converges to it


def textbook_sa_source(nut, omega, d, mu, rho, re):
    """SA source per unit volume from Allmaras, Johnson & Spalart (ICCFD7,
    2012), nondimensional with nu~ scaled by the reference viscosity: the
    nu~ terms carry 1/Re.  Negative branch with the reference's
    destruction factor 50 (TurbSolverUnstr.cpp:2866-2871)."""
    kap, cb1, cb2, cv1, cw2, cw3, sig, c2, c3, ct3 = 0.41, 0.1355, 0.622, 7.1, 0.3, 2.0, 2.0 / 3.0, 0.7, 0.9, 1.2
    cw1 = cb1 / kap ** 2 + (1 + cb2) / sig
    chi = nut * rho / mu
    fv1 = chi ** 3 / (chi ** 3 + cv1 ** 3)
    fv2 = 1 - chi / (1 + chi * fv1)
    sbar = nut * fv2 / (kap ** 2 * d ** 2 * re)
    stilde = np.where(sbar >= -c2 * omega, omega + sbar,
                      omega + omega * (c2 ** 2 * omega + c3 * sbar) / ((c3 - 2 * c2) * omega - sbar))
    r = np.minimum(nut / (stilde * kap ** 2 * d ** 2 * re), 10.0)
    g = r + cw2 * (r ** 6 - r)
    fw = g * ((1 + cw3 ** 6) / (g ** 6 + cw3 ** 6)) ** (1 / 6)
    pos = cb1 * stilde * nut - cw1 * fw * (nut / d) ** 2 / re
    neg = cb1 * (1 - ct3) * omega * nut + 50.0 * cw1 * (nut / d) ** 2 / re
    return np.where(nut >= 0, pos, neg)


@pytest.mark.parametrize("nut_value,shear", [(3.0, 2.0), (50.0, 2.0), (50.0, 1e-3), (-2.0, 2.0)])
def test_source_matches_textbook_sa(nut_value, shear):
    """Uniform nu~, linear shear u = a y on a box: in interior cells the
    convection and diffusion vanish exactly and the GG vorticity is |a|, so
    res / vol is the SA source.  Covers both branches of S~ (a small shear
    with a large nu~ makes S-bar < -c2 Omega) and the negative branch."""
    nx, ny, nz = 6, 8, 4
    m = pa.Mesh.bump_hex(nx, ny, nz, 0.0)
    s, conn, met, wd = make(m)
    q = np.tile(s.freestream()[:, None], (1, m.n_total))
    q[1] = shear * met["ycc"]
    q[2:4] = 0.0
    s.set("q", q)
    s.gamma_t()
    s.viscosity()
    s.gradient()
    s.set("nut", np.full(m.n_total, nut_value))
    s.set("dt", np.full(m.n_total, 1e30))
    s.sa_update_residuals()
    res = s.get("res_t")
    xc, yc, zc = met["xcc"][:m.n_cell], met["ycc"][:m.n_cell], met["zcc"][:m.n_cell]
    dx, dy, dz = 3.0 / nx, 0.8 / ny, 0.2 / nz
    inner = np.where((np.abs(xc) < 1.5 - dx) & (yc > dy) & (yc < 0.8 - dy) & (zc > dz) & (zc < 0.2 - dz))[0]
    assert inner.size > 0
    mu = s.get("visl")[inner]
    rho = q[0, inner]
    expect = textbook_sa_source(nut_value, abs(shear), wd[inner], mu, rho, 1e5)
    got = res[inner] / met["vol"][inner]
    assert np.allclose(got, expect, rtol=1e-10, atol=1e-14 * np.max(np.abs(expect)))
    assert np.all(s.get("spec_t")[inner] >= 0.0)


def test_diffusion_matches_continuous_operator_on_linear_field():
    """nu~ = a*y on a uniform box, uniform rho and laminar nu: in interior
    cells the NSMB diffusion residual equals (1/sigma)(1 + cb2) a^2 V / Re."""
    nx, ny, nz = 6, 8, 4
    m = pa.Mesh.bump_hex(nx, ny, nz, 0.0)
    s, conn, met, wd = make(m)
    q = np.tile(s.freestream()[:, None], (1, m.n_total))
    q[1:4, :] = 0.0  # no convection
    s.set("q", q)
    s.gamma_t()
    s.viscosity()
    s.gradient()
    a = 2.5
    nut = 10.0 + a * met["ycc"]
    s.set("nut", nut)
    s.set("wdist", np.full(m.n_total, 1e30))  # no destruction ...
    s.set("dt", np.full(m.n_total, 1e30))
    s.sa_update_residuals()
    res = s.get("res_t")
    # ... and no production (zero vorticity); ghosts were overwritten by the
    # boundary rule, so look at cells away from the boundary
    xc, yc, zc = met["xcc"][:m.n_cell], met["ycc"][:m.n_cell], met["zcc"][:m.n_cell]
    dx, dy, dz = 3.0 / nx, 0.8 / ny, 0.2 / nz
    inner = np.where((np.abs(xc) < 1.5 - dx) & (yc > dy) & (yc < 0.8 - dy) & (zc > dz) & (zc < 0.2 - dz))[0]
    assert inner.size > 0
    vol = met["vol"][inner]
    expect = (1.0 / (2.0 / 3.0)) * (1.0 + 0.622) * a * a * vol / 1e5
    assert np.allclose(res[inner], expect, rtol=1e-9)


def test_convection_of_uniform_field_vanishes():
    m = pa.Mesh.bump_hex(6, 4, 3, 0.05)
    s, conn, met, wd = make(m)
    s.init_freestream()
    s.boundary()
    s.gamma_t()
    s.viscosity()
    s.gradient()
    s.set("wdist", np.full(m.n_total, 1e30))
    s.set("dt", np.full(m.n_total, 1e30))
    s.sa_update_residuals()
    res = s.get("res_t")[:m.n_cell]
    # uniform nu~, uniform stream: convective fluxes cancel in every cell not
    # touching a wall (where the ghost nu~ flips sign) or the farfield
    c, nb = conn, m.n_bface
    touch = np.zeros(m.n_cell, bool)
    touch[c["left"][:nb]] = True
    assert np.max(np.abs(res[~touch])) <= 1e-13


def test_wall_ghost_is_antisymmetric():
    m = pa.Mesh.bump_hex(6, 4, 3, 0.05)
    s, conn, met, wd = make(m)
    rng = np.random.default_rng(2)
    s.init_freestream()
    nut = 1 + rng.random(m.n_total)
    s.set("nut", nut)
    s.sa_boundary()
    out = s.get("nut")
    nb = m.n_bface
    for f in range(nb):
        le, re = conn["left"][f], conn["right"][f]
        if conn["bc_type"][f] == 2:
            assert out[re] == -nut[le]
        elif conn["bc_type"][f] == 3:
            assert out[re] == nut[le]


def test_scalar_lusgs_equals_dense_solve():
    """One forward + backward sweep solves (D + L) D^-1 (D + U) dq = res with
    L and U from the face matrices (all-farfield mesh, ghosts copy)."""
    base = pa.Mesh.bump_hex(3, 3, 2, 0.05)
    c = base.connectivity()
    m = pa.Mesh.from_arrays(c["x"], c["y"], c["z"], c["face_node_ptr"], c["face_nodes"], c["left"], c["right"],
                            c["n_bface"], np.full(c["n_bface"], 4, np.int32), c["cell_node_ptr"], c["cell_nodes"])
    s, conn, met, wd = make(m)
    rng = np.random.default_rng(4)
    q = np.tile(s.freestream()[:, None], (1, m.n_total))
    q[:, :m.n_cell] *= 1 + rng.uniform(-0.05, 0.05, size=(5, m.n_cell))
    s.set("q", q)
    s.gamma_t()
    s.viscosity()
    s.gradient()
    s.set("nut", 3.0 + rng.random(m.n_total))
    s.set("wdist", np.full(m.n_total, 0.3))
    s.set("dt", np.full(m.n_total, 1e-3))
    s.sa_update_residuals()
    s.sa_spectrum()
    res = s.get("res_t")[:m.n_cell].copy()
    spec = s.get("spec_t")[:m.n_cell].copy()
    matL, matR = s.sa_face_matrices(m.n_face)
    nut0 = s.get("nut")[:m.n_cell].copy()
    # the ghost dq of an all-farfield mesh copies the owner: fold it into D
    nc = m.n_cell
    L = np.zeros((nc, nc))
    U = np.zeros((nc, nc))
    for f in range(m.n_bface, m.n_face):
        a, b = conn["left"][f], conn["right"][f]
        # row a (left cell) takes matR * dq[b]; row b takes matL * dq[a]
        (L if b < a else U)[a, b] += matR[f]
        (L if a < b else U)[b, a] += matL[f]
    Dg = np.zeros(nc)
    for f in range(m.n_bface):
        Dg[conn["left"][f]] += matR[f]  # backward sweep: ghost dq = dq[le]
    D = np.diag(spec)
    dstar = np.linalg.solve(D + L, res)
    x = np.linalg.solve(D + U + np.diag(Dg), D @ dstar)
    s.sa_lusgs()
    dq = s.get("nut")[:nc] - nut0
    assert np.max(np.abs(dq - x)) <= 1e-8 * np.max(np.abs(x))  # dense solve rounding
