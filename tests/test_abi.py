"""C-ABI and host-side checks that need no GPU: the library loads, exports every
symbol include/phl.h declares, builds the synthetic meshes, and rejects
unsupported options with a status instead of exiting."""
import subprocess

import numpy as np
import pytest

import phenlei_amd as pa


def test_library_exports_every_declared_symbol():
    lib = pa.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = pa.exported_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if s not in exported]
    assert not missing, missing


def test_bump_channel_sizes_match_the_configuration():
    m = pa.Mesh.bump_hex(64, 32, 8, 0.05)
    assert (m.n_cell, m.n_face, m.n_bface) == (16384, 51968, 5632)
    conn = m.connectivity()
    counts = np.bincount(conn["bc_type"], minlength=5)
    assert counts[4] == 2 * 32 * 8        # inlet + outlet farfield
    assert counts[2] == 64 * 8            # lower wall
    assert counts[3] == 64 * 8 + 2 * 64 * 32  # top + sides symmetry
    # faces: boundary first, ghost index nCell+f, interior owner < neighbour
    nb = m.n_bface
    assert np.array_equal(conn["right"][:nb], m.n_cell + np.arange(nb))
    assert np.all(conn["left"][nb:] < conn["right"][nb:])


def test_face_normals_point_from_left_to_right():
    m = pa.Mesh.bump_hex(12, 6, 4, 0.05)
    conn, met = m.connectivity(), m.metrics()
    n = np.stack([met["xfn"], met["yfn"], met["zfn"]])
    d = np.stack([met["xcc"][conn["right"]] - met["xcc"][conn["left"]],
                  met["ycc"][conn["right"]] - met["ycc"][conn["left"]],
                  met["zcc"][conn["right"]] - met["zcc"][conn["left"]]])
    assert np.all((n * d).sum(axis=0) > 0)


def test_renumbering_is_a_permutation_of_the_same_grid():
    m = pa.Mesh.bump_hex(8, 4, 4, 0.05)
    r = m.renumbered("colour")
    assert r.n_colours == 2
    assert (r.n_cell, r.n_face, r.n_bface) == (m.n_cell, m.n_face, m.n_bface)
    v0 = np.sort(m.metrics()["vol"][:m.n_cell])
    v1 = np.sort(r.metrics()["vol"][:r.n_cell])
    assert np.allclose(v0, v1, rtol=1e-13)
    perm = np.arange(m.n_cell)[::-1].copy()
    rr = m.renumbered("perm", perm)
    assert np.allclose(rr.metrics()["vol"][:m.n_cell], m.metrics()["vol"][:m.n_cell][perm], rtol=1e-13)


def test_mesh_create_rejects_bad_connectivity():
    m = pa.Mesh.bump_hex(4, 2, 2, 0.05)
    c = m.connectivity()
    bad = c["right"].copy()
    bad[0] = 0  # a boundary face must point at its ghost cell
    with pytest.raises(pa.PhlError) as e:
        pa.Mesh.from_arrays(c["x"], c["y"], c["z"], c["face_node_ptr"], c["face_nodes"], c["left"], bad,
                            c["n_bface"], c["bc_type"], c["cell_node_ptr"], c["cell_nodes"])
    assert e.value.status == pa.PHL_ERR_ARG


@pytest.mark.parametrize("override", [dict(gradient=3),
                                      dict(entropy_fix=0), dict(entropy_fix=7),
                                      dict(viscous=2), dict(precision=7)])
def test_unsupported_options_return_a_status(override):
    m = pa.Mesh.bump_hex(4, 2, 2, 0.05)
    with pytest.raises(pa.PhlError) as e:
        pa.Solver(m, pa.default_params(**override))
    assert e.value.status == pa.PHL_ERR_UNSUPPORTED


@pytest.mark.parametrize("override", [dict(sweep_form=-1), dict(sweep_form=4)])
def test_bad_sweep_form_is_an_argument_error(override):
    m = pa.Mesh.bump_hex(4, 2, 2, 0.05)
    with pytest.raises(pa.PhlError) as e:
        pa.Solver(m, pa.default_params(**override))
    assert e.value.status == pa.PHL_ERR_ARG
    assert "sweep_form" in str(e.value)


def test_sweep_count_below_one_is_an_argument_error():
    m = pa.Mesh.bump_hex(4, 2, 2, 0.05)
    with pytest.raises(pa.PhlError) as e:
        pa.Solver(m, pa.default_params(n_lusgs_sweeps=0))
    assert e.value.status == pa.PHL_ERR_ARG


def test_lsq_on_a_2d_grid_is_rejected():
    import os
    m = pa.Mesh.from_fts(os.path.join(os.path.dirname(__file__), "golden", "rae2822_hybrid2d__4_0.fts"))
    with pytest.raises(pa.PhlError) as e:
        pa.Solver(m, pa.default_params(gradient=1))
    assert e.value.status == pa.PHL_ERR_UNSUPPORTED


def test_several_sweep_pairs_on_a_batch_are_rejected():
    a, b = pa.Mesh.bump_hex(4, 2, 2, 0.05), pa.Mesh.bump_hex(4, 2, 2, 0.03)
    u = pa.Mesh.concat([a, b])
    with pytest.raises(pa.PhlError) as e:
        pa.Solver(u, [pa.default_params(n_lusgs_sweeps=3), pa.default_params(n_lusgs_sweeps=3, mach=0.6)])
    assert e.value.status == pa.PHL_ERR_UNSUPPORTED


def test_turbulence_batch_is_rejected_before_any_device_call():
    a, b = pa.Mesh.bump_hex(4, 2, 2, 0.05), pa.Mesh.bump_hex(4, 2, 2, 0.03)
    u = pa.Mesh.concat([a, b])
    with pytest.raises(pa.PhlError) as e:
        pa.Solver(u, [pa.default_params(viscous=3), pa.default_params(viscous=3, mach=0.6)])
    assert e.value.status == pa.PHL_ERR_UNSUPPORTED


def test_concat_keeps_each_case_intact():
    a = pa.Mesh.bump_hex(6, 3, 2, 0.05)
    b = pa.Mesh.bump_hex(4, 5, 3, 0.02)
    u = pa.Mesh.concat([a, b])
    assert u.n_cases == 2
    assert (u.n_cell, u.n_face, u.n_bface) == (a.n_cell + b.n_cell, a.n_face + b.n_face, a.n_bface + b.n_bface)
    cu, mu = u.connectivity(), u.metrics()
    nb = u.n_bface
    assert np.array_equal(cu["right"][:nb], u.n_cell + np.arange(nb))
    for k, part in enumerate((a, b)):
        cs, gs = u.case_slices(k)
        mp = part.metrics()
        assert np.array_equal(mu["vol"][cs], mp["vol"][:part.n_cell])
        assert np.array_equal(mu["vol"][gs], mp["vol"][part.n_cell:])
        assert np.array_equal(mu["xcc"][gs], mp["xcc"][part.n_cell:])
    # faces of case 1 map onto its own faces, in order, with shifted cells
    cb = b.connectivity()
    b0 = u.case_bface_ptr[1]
    assert np.array_equal(cu["left"][b0:b0 + b.n_bface], cb["left"][:b.n_bface] + a.n_cell)
    assert np.array_equal(cu["bc_type"][b0:b0 + b.n_bface], cb["bc_type"])
    i0 = nb + (a.n_face - a.n_bface)
    assert np.array_equal(cu["left"][i0:], cb["left"][b.n_bface:] + a.n_cell)
    assert np.array_equal(cu["right"][i0:], cb["right"][b.n_bface:] + a.n_cell)
    with pytest.raises(pa.PhlError):
        u.renumbered("colour")
