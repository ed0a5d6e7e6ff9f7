"""The reference's own grid files: the .fts reader, 2-D metrics and the 2-D
wall distance, against the RAE2822 example grid the reference ships
(PHengLEI/examples/grid/rae2822_hybrid2d__4_0.fts, 4 zones, 2-D hybrid) and
the reference's own wall distances for it (.wdt), copied as data into
tests/golden/ by tests/golden/copy_rae2822.py."""
import os

import numpy as np
import pytest

import phenlei_amd as pa
import oracle as orc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FTS = os.path.join(GOLD, "rae2822_hybrid2d__4_0.fts")
WDT = os.path.join(GOLD, "rae2822_hybrid2d__4_0.wdt")


@pytest.fixture(scope="module")
def rae():
    m = pa.Mesh.from_fts(FTS)
    return m, m.connectivity(), m.metrics()


def test_reader_sizes_and_boundaries(rae):
    m, c, g = rae
    # the unpartitioned grid (rae2822_hybrid2d.cas) has 23,334 nodes, 59,370
    # faces and 36,036 cells: the joined zones must give exactly these
    assert (m.n_node, m.n_face, m.n_cell, m.dim, m.n_zones) == (23334, 59370, 36036, 2, 4)
    bc, cnt = np.unique(c["bc_type"], return_counts=True)
    assert dict(zip(bc.tolist(), cnt.tolist())) == {2: 304, 4: 296}  # wall, farfield; interfaces joined
    assert np.all(np.diff(c["face_node_ptr"]) == 2)
    # every node is used, every interior face joins two different cells
    assert np.array_equal(np.unique(c["face_nodes"]), np.arange(m.n_node))
    nb = m.n_bface
    assert np.all(c["right"][:nb] == m.n_cell + np.arange(nb))
    assert np.all(c["left"][nb:] != c["right"][nb:])
    # each cell is closed by its faces: 3 or 4 edges, as many as its nodes
    deg = np.bincount(np.concatenate([c["left"], c["right"][nb:]]), minlength=m.n_cell)
    assert np.array_equal(deg, np.diff(c["cell_node_ptr"]))


def test_2d_metrics_equal_oracle_and_close(rae):
    m, c, g = rae
    fm = orc.face_metrics(c)
    cm = orc.cell_metrics(c, fm)
    for k, v in list(fm.items()) + list(cm.items()):
        assert np.array_equal(v, g[k]), k
    assert np.all(g["zfn"] == 0.0) and np.all(g["zcc"] == 0.0)
    # closure: sum over a cell's faces of n A is zero
    nb, nc = m.n_bface, m.n_cell
    s = np.zeros((nc, 2))
    w = np.stack([g["xfn"] * g["area"], g["yfn"] * g["area"]], 1)
    np.add.at(s, c["left"], w)
    np.add.at(s, c["right"][nb:], -w[nb:])
    per = np.zeros(nc)
    np.add.at(per, c["left"], g["area"])
    np.add.at(per, c["right"][nb:], g["area"][nb:])
    assert np.max(np.abs(s).max(1) / per) < 1e-14


def test_2d_metrics_pinned_by_polygon_formulas(rae):
    """Cell areas by the shoelace formula over the cell's nodes taken in
    boundary order (the face loop around the cell), centroids by the
    polygon centroid formula."""
    m, c, g = rae
    nb, nc = m.n_bface, m.n_cell
    x, y = c["x"], c["y"]
    fp, fnodes = c["face_node_ptr"], c["face_nodes"]
    # directed edges with the cell on their left (counter-clockwise): the
    # normal (y2 - y1, x1 - x2) of a face a->b points to the right of a->b,
    # towards the right cell, so a->b runs counter-clockwise round the left
    # cell and b->a round the right cell
    edges = [[] for _ in range(nc)]
    for f in range(m.n_face):
        a, b = fnodes[fp[f]], fnodes[fp[f] + 1]
        edges[c["left"][f]].append((a, b))
        if f >= nb:
            edges[c["right"][f]].append((b, a))
    area = np.empty(nc)
    cx = np.empty(nc)
    cy = np.empty(nc)
    cond = np.empty(nc)
    for k in range(nc):
        e = edges[k]
        xa, ya = x[[u for u, _ in e]], y[[u for u, _ in e]]
        xb, yb = x[[v for _, v in e]], y[[v for _, v in e]]
        cr = xa * yb - xb * ya
        a = 0.5 * cr.sum()
        area[k] = a
        cx[k] = ((xa + xb) * cr).sum() / (6 * a)
        cy[k] = ((ya + yb) * cr).sum() / (6 * a)
        # the divergence-theorem sums are not translation invariant: rounding
        # grows like eps (R + L)^2 L / A (R the distance from the origin, L
        # the perimeter), which matters for the thin wall cells
        per = np.hypot(xb - xa, yb - ya).sum()
        cond[k] = np.finfo(float).eps * (np.hypot(xa, ya).max() + per) ** 2 * per / a
    assert np.all(area > 0)
    np.testing.assert_allclose(g["vol"][:nc], area, rtol=1e-9)
    err = np.hypot(g["xcc"][:nc] - cx, g["ycc"][:nc] - cy)
    assert np.all(err <= 16 * cond)
    # ghosts: mirror images of their owners across the boundary edge
    le = c["left"][:nb]
    mid = np.stack([g["xfc"][:nb], g["yfc"][:nb]], 1)
    own = np.stack([g["xcc"][le], g["ycc"][le]], 1)
    gh = np.stack([g["xcc"][nc:], g["ycc"][nc:]], 1)
    n = np.stack([g["xfn"][:nb], g["yfn"][:nb]], 1)
    assert np.max(np.abs((((gh + own) / 2 - mid) * n).sum(1))) < 1e-12
    d = gh - own
    assert np.max(np.abs(d[:, 0] * n[:, 1] - d[:, 1] * n[:, 0])) < 1e-12


def test_wall_distance_equals_reference_wdt(rae):
    """The 2-D wall distance (the reference's bisection over the wall edges)
    against the distances the reference itself wrote for this grid."""
    m, c, g = rae
    ref = m.read_wdt(WDT)
    wd = orc.wall_distance(c, g)
    rel = np.abs(wd - ref) / ref
    assert rel.max() < 1e-7
    assert np.mean(wd == ref) > 0.98  # bitwise in almost every cell (centroid ulps aside)
    # and an independent check: never below the exact point-to-edge distance,
    # never above the distance to the nearest wall node or edge midpoint (the
    # bisection's first candidates; it is not the exact distance: it keeps an
    # end point that beats the midpoint, off by up to ~1e-3 here)
    walls = np.flatnonzero(c["bc_type"] == 2)
    fp, fn = c["face_node_ptr"], c["face_nodes"]
    a = np.stack([c["x"][fn[fp[walls]]], c["y"][fn[fp[walls]]]], 1)
    b = np.stack([c["x"][fn[fp[walls] + 1]], c["y"][fn[fp[walls] + 1]]], 1)
    sel = np.random.default_rng(3).choice(m.n_cell, 400, replace=False)
    p = np.stack([g["xcc"][sel], g["ycc"][sel]], 1)[:, None, :]
    ab = (b - a)[None]
    t = np.clip(((p - a[None]) * ab).sum(2) / (ab * ab).sum(2), 0, 1)
    exact = np.linalg.norm(p - (a[None] + t[..., None] * ab), axis=2).min(1)
    tol = 1e-14 * (1.0 + np.linalg.norm(p[:, 0, :], axis=1))  # rounding of the projection itself
    assert np.all(wd[sel] >= exact - tol)
    cand = np.concatenate([a, b, (a + b) / 2])[None]
    upper = np.linalg.norm(p - cand, axis=2).min(1)
    assert np.all(wd[sel] <= upper)
    assert np.median(np.abs(wd[sel] - exact) / exact) < 1e-9


def test_reader_rejects_bad_files(tmp_path):
    bad = tmp_path / "x.fts"
    bad.write_bytes(b"not an hdf5 file" * 10)
    with pytest.raises(pa.PhlError, match="not an HDF5 file"):
        pa.Mesh.from_fts(str(bad))
    cut = tmp_path / "cut.fts"
    cut.write_bytes(open(FTS, "rb").read()[:200000])
    with pytest.raises(pa.PhlError):
        pa.Mesh.from_fts(str(cut))
    m = pa.Mesh.from_fts(FTS)
    with pytest.raises(pa.PhlError, match="expected 10"):
        wd = np.empty(10)
        pa._check(pa.load_library().phl_read_wdt(os.fsencode(WDT), 10, pa._dptr(wd)))
    assert m.n_cell == 36036


def test_2d_oracle_iterations_run(rae):
    """The oracle on the reference grid with the example's flow (M 0.2,
    alpha 2.79, Re 6.5e6, SA): a few iterations stay finite and drop the
    residual."""
    m, c, g = rae
    p = pa.default_params(mach=0.2, alpha_deg=2.79, reynolds=6.5e6, viscous=3, cfl_start=10.0, cfl_end=10.0)
    s = orc.OracleSolver(c, g, p)
    assert s.dim == 2
    s.set_wall_distance(orc.wall_distance(c, g))
    s.init_freestream()
    hist, nfix = s.iterate(1, 4)
    assert np.all(np.isfinite(hist)) and nfix == 0
