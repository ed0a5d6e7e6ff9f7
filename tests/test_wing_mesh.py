"""The swept-wing generator: closed cells, positive volumes, the boundary
made of the wing surface (area close to twice the planform), the root
symmetry plane and the farfield box, and the cell count formula."""
import numpy as np

import phenlei_amd as pa


def test_small_wing_is_a_valid_hybrid_mesh():
    nch, nsp, npr, nte = 10, 5, 4, 3
    m = pa.Mesh.swept_wing(nch, nsp, npr, 1e-3, nte)
    ni = nch + (3 * nch + 5) // 10 + (6 * nch + 5) // 10
    nk = nsp + max(2, (4 * nsp + 5) // 10)
    assert m.n_cell == ni * nk * 2 * (2 * npr + 6 * nte)
    c, met = m.connectivity(), m.metrics()
    assert np.all(met["vol"][:m.n_cell] > 0.0)
    nc = m.n_cell
    acc = np.zeros((nc, 3))
    n = np.stack([met["xfn"], met["yfn"], met["zfn"]], 1) * met["area"][:, None]
    np.add.at(acc, c["left"], n)
    inner = c["right"] < nc
    np.add.at(acc, c["right"][inner], -n[inner])
    assert np.abs(acc).max() <= 1e-10
    bc = c["bc_type"]
    area = met["area"][:m.n_bface]
    planform = 1.5 * 1.0 * (1.0 + 0.5) / 2.0
    wall = area[bc == 2].sum()
    assert 2.0 * planform < wall < 2.0 * planform * 1.02  # thin section: a little above twice the planform
    # the root plane: the whole box cross-section, 20 chords up and down
    yz = met["yfc"][:m.n_bface][bc == 3]
    assert np.allclose(met["zfc"][:m.n_bface][bc == 3], 0.0)
    assert yz.min() < -5.0 and yz.max() > 5.0
    assert np.isclose(c["y"].min(), -20.0) and np.isclose(c["y"].max(), 20.0)
    assert np.isclose(c["z"].max(), 20.0)
    assert set(np.unique(bc)) == {2, 3, 4}
    # wall faces lie on the planform, on both sides of the wing plane
    yw = met["yfc"][:m.n_bface][bc == 2]
    assert np.abs(yw).max() < 0.07 and (yw > 0).any() and (yw < 0).any()


def test_wing_cells_keep_their_columns_together():
    """Cells are numbered column by column, so the forward sweep's level
    count stays near (columns along i and k) + (cells in a column)."""
    m = pa.Mesh.swept_wing(8, 4, 3, 1e-3, 2)
    c = m.connectivity()
    nb = m.n_bface
    # every interior face joins cells at most one column apart in i or k:
    # a column holds 2 * (2 * 3 + 6 * 2) = 36 cells
    col = 36
    d = np.abs(c["left"][nb:] // col - c["right"][nb:] // col)
    ni = 8 + (3 * 8 + 5) // 10 + (6 * 8 + 5) // 10
    assert d.max() <= ni
