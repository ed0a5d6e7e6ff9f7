"""Pins of the oracle's node-based Green-Gauss gradient (gradientName
"ggnode", the reference default; ComputeNodeValue NSSolverUnstruct.cpp:
4734-5117, CompGradientGGNode_NEW Geo_UnstructGrid.cpp:4202-4315):

- node values of a linear field are exact wherever the inverse-distance
  stencil is symmetric (the interior nodes of a uniform box);
- the gradient of a linear field is exact in cells whose nodes are all
  exact, and zero for a uniform field everywhere;
- a per-node brute force of the priority rules (solid surface, then
  farfield, then other codes; symmetry nodes count as interior) on a small
  distorted mesh;
- the temperature node values lag one iteration (they are computed from
  the previous iteration's t, as the reference's PostSolve does)."""
import numpy as np

import phenlei_amd as pa
import oracle as orc

BC = pa.BC


def _pair(mesh, **pover):
    conn, met = mesh.connectivity(), mesh.metrics()
    p = pa.default_params(gradient=2, **pover)
    return conn, met, orc.OracleSolver(conn, met, p)


def _linear_state(met, cpu, a):
    x, y, z = met["xcc"], met["ycc"], met["zcc"]
    q = np.tile(cpu.freestream()[:, None], (1, len(x)))
    q[0] = a[0] + a[1] * x + a[2] * y + a[3] * z
    return q


def _interior_nodes(conn, nn, nb):
    """Nodes on no boundary face: on a uniform box their eight cells are
    equidistant, so the inverse-distance average of a linear field is exact
    (boundary nodes average face values whose ghosts the boundary conditions
    set, so they are not)."""
    ptr, nodes = conn["face_node_ptr"], conn["face_nodes"]
    bnodes = set(nodes[:ptr[nb]].tolist())
    return np.array(sorted(set(range(nn)) - bnodes))


def test_linear_field_node_values_and_gradient_are_exact_on_a_box():
    m = pa.Mesh.bump_hex(10, 6, 6, 0.0)
    conn, met, cpu = _pair(m)
    a = np.array([1.0, 0.2, -0.3, 0.4])
    q = _linear_state(met, cpu, a)
    cpu.set("q", q)
    cpu.residual()  # boundary conditions, node values, gradients
    nv = cpu.node_values()
    ok = _interior_nodes(conn, len(conn["x"]), m.n_bface)
    exact = a[0] + a[1] * conn["x"] + a[2] * conn["y"] + a[3] * conn["z"]
    assert ok.size > 0
    assert np.max(np.abs(nv[0][ok] - exact[ok])) <= 1e-13
    assert np.max(np.abs(nv[0] - exact)) > 1e-6  # one-sided stencils elsewhere
    # cells whose nodes are all exact: the gradient is the field's slope
    good = set(ok.tolist())
    cptr, cn = conn["cell_node_ptr"], conn["cell_nodes"]
    cells = [c for c in range(m.n_cell) if all(n in good for n in cn[cptr[c]:cptr[c + 1]])]
    assert len(cells) > 0
    g = np.stack([cpu.get("dqdx")[0], cpu.get("dqdy")[0], cpu.get("dqdz")[0]])
    assert np.max(np.abs(g[:, cells] - a[1:, None])) <= 1e-11
    # every ghost takes its owner's gradient
    left = conn["left"][:m.n_bface]
    gh = m.n_cell + np.arange(m.n_bface)
    assert np.array_equal(g[:, gh], g[:, left])


def test_uniform_field_has_zero_gradient_everywhere():
    """Freestream along a flat inviscid channel satisfies every boundary
    condition (slip wall and symmetry mirrors leave it unchanged), so every
    face value equals it and the gradients vanish (jittered tetrahedra too)."""
    m = pa.Mesh.bump_hybrid(8, 3, 4, 1e-3, 4, 0.1, 5, 0.0)
    conn, met, cpu = _pair(m, mach=0.4)
    cpu.init_freestream()
    cpu.residual()
    for name in ("dqdx", "dqdy", "dqdz", "dtdx", "dtdy", "dtdz"):
        assert np.max(np.abs(cpu.get(name))) <= 1e-9, name


def _brute_node_values(conn, met, q, t, n_cell):
    """Each node from the definition: its class by the priority of the
    boundary faces touching it, then the inverse-distance average over the
    contributors of that class (face centres carry the average of the cell
    and its ghost; interior nodes take the real cells around them)."""
    nn = len(conn["x"])
    X = np.stack([conn["x"], conn["y"], conn["z"]], axis=1)
    fptr, fn, bc = conn["face_node_ptr"], conn["face_nodes"], conn["bc_type"]
    cptr, cn = conn["cell_node_ptr"], conn["cell_nodes"]
    nb = len(bc)
    touching = [[] for _ in range(nn)]
    for f in range(nb):
        for n in fn[fptr[f]:fptr[f + 1]]:
            touching[n].append(f)
    cells_of = [[] for _ in range(nn)]
    for c in range(n_cell):
        for n in cn[cptr[c]:cptr[c + 1]]:
            cells_of[n].append(c)
    qt = np.vstack([q, t[None, :]])
    out = np.zeros((6, nn))
    fc = np.stack([met["xfc"], met["yfc"], met["zfc"]], axis=1)
    cc = np.stack([met["xcc"], met["ycc"], met["zcc"]], axis=1)
    for n in range(nn):
        codes = {bc[f] for f in touching[n]}
        if BC["solid_surface"] in codes:
            src = [f for f in touching[n] if bc[f] == BC["solid_surface"]]
        elif BC["farfield"] in codes:
            src = [f for f in touching[n] if bc[f] == BC["farfield"]]
        elif codes - {BC["symmetry"]}:
            src = [f for f in touching[n] if bc[f] != BC["symmetry"]]
        else:
            src = None
        if src is not None:
            w = np.array([1.0 / np.linalg.norm(X[n] - fc[f]) for f in src])
            v = np.array([0.5 * (qt[:, conn["left"][f]] + qt[:, n_cell + f]) for f in src])
        else:
            w = np.array([1.0 / np.linalg.norm(X[n] - cc[c]) for c in cells_of[n]])
            v = np.array([qt[:, c] for c in cells_of[n]])
        out[:, n] = (w[:, None] * v).sum(axis=0) / w.sum()
    return out


def test_node_values_equal_a_per_node_brute_force():
    m = pa.Mesh.bump_hybrid(6, 2, 3, 1e-2, 3, 0.15, 9)
    conn, met, cpu = _pair(m, mach=0.4, viscous=1, reynolds=1e4)
    rng = np.random.default_rng(4)
    q = np.tile(cpu.freestream()[:, None], (1, m.n_total))
    q[:, :m.n_cell] *= 1.0 + 0.05 * rng.uniform(-1, 1, (5, m.n_cell))
    cpu.set("q", q)
    cpu.residual()  # boundary, then node values from the state's own temperature
    qb = cpu.get("q")
    t = qb[4] / (cpu.get("t")[0] * 0 + qb[0]) / (1.0 / (1.4 * 0.4 ** 2))
    ref = _brute_node_values(conn, met, qb, t, m.n_cell)
    nv = cpu.node_values()
    assert np.max(np.abs(nv - ref) / np.maximum(np.abs(ref), 1e-12)) <= 1e-13


def test_temperature_node_values_lag_one_iteration():
    """Iteration n's gradient reads node temperatures of iteration n-1's t
    (ComputeNodeValue runs in PostSolve, before the next
    ComputeGamaAndTemperature); the node states are current."""
    m = pa.Mesh.bump_hybrid(6, 2, 3, 1e-2, 3, 0.15, 9)
    conn, met, cpu = _pair(m, mach=0.4, viscous=1, reynolds=1e4, cfl_start=2.0, cfl_end=2.0)
    cpu.init_freestream()
    cpu.iterate(1, 2)
    t_prev = cpu.get("t").copy()     # the temperature of step 2's UpdateIndependentField
    cpu.residual()                   # step 3's node values
    q3 = cpu.get("q")
    nv = cpu.node_values()
    ref_lag = _brute_node_values(conn, met, q3, t_prev, m.n_cell)
    t_now = q3[4] / (q3[0] * (1.0 / (1.4 * 0.4 ** 2)))
    ref_now = _brute_node_values(conn, met, q3, t_now, m.n_cell)
    def err(a, b):
        return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-12))
    assert err(nv[:5], ref_lag[:5]) <= 1e-13
    assert err(nv[5], ref_lag[5]) <= 1e-13
    assert err(nv[5], ref_now[5]) > 1e-9
