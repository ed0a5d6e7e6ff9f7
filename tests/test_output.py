"""Output and restart files (host side, no GPU): the restart block layout of
NSSolverUnstruct::DumpRestart / TurbSolverUnstr::DumpRestart read back
independently, and the Tecplot FEPolyhedron dump of
DumpToTecplotASCII::DumpFlowByFaceTopo parsed back, with node values that
are exact for a linear field wherever the inverse-distance stencil is
symmetric (CompNodeVar)."""
import numpy as np
import pytest

import phenlei_amd as pa


def test_restart_block_layout_and_round_trip(tmp_path):
    m = pa.Mesh.bump_hex(6, 4, 2, 0.05)
    rng = np.random.default_rng(3)
    q = rng.standard_normal((5, m.n_total))
    path = tmp_path / "flow.dat"
    pa.write_restart(path, q, outnstep=1234)
    raw = path.read_bytes()
    assert len(raw) == 4 + 8 * 5 * m.n_total
    assert np.frombuffer(raw[:4], dtype="<i4")[0] == 1234
    assert np.array_equal(np.frombuffer(raw[4:], dtype="<f8").reshape(5, m.n_total), q)
    step, back = pa.read_restart(path, 5, m.n_total)
    assert step == 1234 and np.array_equal(back, q)
    nut = rng.random(m.n_total)
    tpath = tmp_path / "turb.dat"
    pa.write_restart(tpath, nut, outnstep=7)
    step, nb = pa.read_restart(tpath, 1, m.n_total)
    assert step == 7 and np.array_equal(nb[0], nut)
    with pytest.raises(pa.PhlError) as e:  # a block of another zone size
        pa.read_restart(path, 5, m.n_total + 1)
    assert e.value.status == pa.PHL_ERR_ARG


def _parse_tecplot(path):
    lines = path.read_text().splitlines()
    head = {}
    i = 0
    while "=" in lines[i] and not lines[i].strip()[0].isdigit() and not lines[i].strip().startswith("-"):
        k, v = lines[i].split("=", 1)
        head[k.strip().lower()] = v.strip()
        i += 1
        if k.strip() == "TotalNumBoundaryConnections":
            break
    toks = " ".join(lines[i:]).split()
    return head, toks


def test_tecplot_block_zone_layout_and_node_values(tmp_path):
    # flat channel (no bump): a uniform box grid, so the inverse-distance
    # stencil of a node is symmetric unless a symmetry face drops its ghosts
    m = pa.Mesh.bump_hex(8, 4, 4, 0.0)
    conn, met = m.connectivity(), m.metrics()
    nt = m.n_total
    a = np.array([1.0, 0.3, -0.2, 0.5])
    cx, cy, cz = met["xcc"], met["ycc"], met["zcc"]
    rho = a[0] + a[1] * cx + a[2] * cy + a[3] * cz
    q = np.tile(np.array([1.0, 1.0, 0.0, 0.0, 1.0 / (1.4 * 0.25)])[:, None], (1, nt))
    q[0] = rho
    path = tmp_path / "flow.plt"
    m.write_tecplot(path, q, mach=0.5, title="box")
    head, toks = _parse_tecplot(path)
    assert head["zonetype"] == "FEPolyhedron"
    nn, nf, ne = int(head["nodes"]), int(head["faces"]), int(head["elements"])
    assert (nn, nf, ne) == (len(conn["x"]), m.n_face, m.n_cell)
    assert int(head["totalnumfacenodes"]) == len(conn["face_nodes"])
    vals = np.array(toks[: 11 * nn], dtype=float).reshape(11, nn)
    assert np.allclose(vals[0], conn["x"], rtol=0, atol=1e-11)
    assert np.allclose(vals[2], conn["z"], rtol=0, atol=1e-11)
    rest = toks[11 * nn:]
    counts = np.array(rest[:nf], dtype=int)
    assert np.array_equal(counts, np.diff(conn["face_node_ptr"]))
    fn = np.array(rest[nf:nf + counts.sum()], dtype=int) - 1
    assert np.array_equal(fn, conn["face_nodes"])
    lr = np.array(rest[nf + counts.sum():], dtype=int)
    left, right = lr[:nf], lr[nf:]
    assert np.array_equal(left, conn["left"] + 1)
    assert np.array_equal(right, np.where(conn["right"] < m.n_cell, conn["right"] + 1, 0))
    # node values: uniform variables stay uniform (Cp 0, u 1, Mach 0.5)
    assert np.allclose(vals[4], 1.0) and np.allclose(vals[10], 0.0, atol=1e-12)
    # the linear density is exact where the stencil is symmetric: nodes that
    # no symmetry face touches and that lie on at most one boundary plane
    # (a node on a plane takes the mirrored ghosts of that plane's faces)
    skip, planes = set(), {}
    ptr, nodes, bc = conn["face_node_ptr"], conn["face_nodes"], conn["bc_type"]
    normals = np.round(np.stack([met["xfn"], met["yfn"], met["zfn"]], axis=1), 6)
    for f in range(m.n_bface):
        for n in nodes[ptr[f]:ptr[f + 1]].tolist():
            if bc[f] == pa.BC["symmetry"]:
                skip.add(n)
            planes.setdefault(n, set()).add(tuple(normals[f]))
    skip.update(n for n, p in planes.items() if len(p) > 1)
    free = np.array(sorted(set(range(nn)) - skip))
    assert free.size > 0
    exact = a[0] + a[1] * conn["x"] + a[2] * conn["y"] + a[3] * conn["z"]
    assert np.max(np.abs(vals[3][free] - exact[free])) <= 1e-12
    assert np.max(np.abs(vals[3] - exact)) > 1e-6  # the symmetry-plane stencils are one-sided
