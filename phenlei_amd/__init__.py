"""phenlei_amd: MI355X (gfx950) implementation of PHengLEI's unstructured
compressible residual + LU-SGS path, behind the C-ABI declared in include/phl.h.

This module is a thin ctypes binding over libphl.so.  It is the host-side
mirror used by tests and bench.py; a C++ host solver calls the same C-ABI
directly.  There is no CPU fallback: if libphl.so is
missing or does not load, importing the solver raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphl.so")

# Status codes (include/phl.h)
PHL_OK = 0
PHL_ERR_ARG = 1
PHL_ERR_HIP = 2
PHL_ERR_NUMERIC = 3
PHL_ERR_UNSUPPORTED = 4
PHL_ERR_ALLOC = 5
PHL_NOALLOC_MAX_CELLS = 16777216  # phl.h: cells of phl_residual_norm_noalloc's static scratch array

# Field identifiers (include/phl.h), shared with the oracle.
FIELD = dict(q=0, res=1, dqdx=2, dqdy=3, dqdz=4, limit=5, dq=6, spec_inv=7, spec_vis=8, dt=9,
             diag=10, t=11, visl=12, dtdx=13, dtdy=14, dtdz=15, delta_flux=16,
             vist=17, nut=18, res_t=19, spec_t=20, dq_t=21, wdist=22, dnudx=23, dnudy=24, dnudz=25,
             pressure_factor=26)
FIELD_ROWS = dict(q=5, res=5, dqdx=5, dqdy=5, dqdz=5, limit=1, dq=5, spec_inv=1, spec_vis=1, dt=1,
                  diag=1, t=1, visl=1, dtdx=1, dtdy=1, dtdz=1, delta_flux=5,
                  vist=1, nut=1, res_t=1, spec_t=1, dq_t=1, wdist=1, dnudx=1, dnudy=1, dnudz=1,
                  pressure_factor=1)
SETTABLE_ROWS_LEN = dict(q="total", dt="cell", vist="total", nut="total", wdist="cell")

BC = dict(extrapolation=1, solid_surface=2, symmetry=3, farfield=4, inflow=5, outflow=6)


class Params(ctypes.Structure):
    """phl_params (include/phl.h); identical layout to the oracle's orc_params."""
    _fields_ = [
        ("mach", ctypes.c_double), ("alpha_deg", ctypes.c_double), ("beta_deg", ctypes.c_double),
        ("gamma", ctypes.c_double), ("reynolds", ctypes.c_double), ("t_ref", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("prandtl_turb", ctypes.c_double), ("t_wall", ctypes.c_double),
        ("venkat_k", ctypes.c_double), ("entropy_scale", ctypes.c_double),
        ("cfl_start", ctypes.c_double), ("cfl_end", ctypes.c_double), ("ktmax", ctypes.c_double),
        ("lusgs_tol", ctypes.c_double), ("force_ref_length", ctypes.c_double),
        ("force_ref_area", ctypes.c_double), ("force_ref_x", ctypes.c_double),
        ("force_ref_y", ctypes.c_double), ("force_ref_z", ctypes.c_double),
        ("force_ref_span", ctypes.c_double), ("sa_nu_inf", ctypes.c_double), ("sa_vist_inf", ctypes.c_double),
        ("turb_cfl_scale", ctypes.c_double), ("turb_skew_angle", ctypes.c_double),
        ("viscous", ctypes.c_int), ("gradient", ctypes.c_int), ("limiter", ctypes.c_int),
        ("entropy_fix", ctypes.c_int), ("cfl_vary_step", ctypes.c_int),
        ("n_lusgs_sweeps", ctypes.c_int), ("precision", ctypes.c_int), ("storage_order", ctypes.c_int),
        ("face_loop", ctypes.c_int), ("sweep_form", ctypes.c_int), ("flow_init_step", ctypes.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def default_params(**overrides):
    """Reference defaults of cfd_para.hypara on this path, with overrides."""
    p = Params()
    p.mach, p.alpha_deg, p.beta_deg, p.gamma = 0.5, 0.0, 0.0, 1.4
    p.reynolds, p.t_ref, p.prandtl, p.prandtl_turb, p.t_wall = 6.5e6, 288.15, 0.72, 0.9, -1.0
    p.venkat_k, p.entropy_scale = 5.0, 1.0
    p.cfl_start, p.cfl_end, p.ktmax, p.lusgs_tol = 0.01, 10.0, 1.0e10, 0.01
    p.force_ref_length, p.force_ref_area, p.force_ref_span = 1.0, 1.0, 1.0
    p.sa_nu_inf, p.sa_vist_inf, p.turb_cfl_scale, p.turb_skew_angle = 3.0, 1.0e-3, 1.0, 0.0
    p.viscous, p.gradient, p.limiter, p.entropy_fix = 0, 0, 0, 3
    p.cfl_vary_step, p.n_lusgs_sweeps, p.precision = 500, 1, 0
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


# Order of phl_forces' output (PHL_FORCE_* in include/phl.h).
FORCE_NAMES = ("CA_f", "CA_p", "CN_f", "CN_p", "CZ_f", "CZ_p", "cpx", "cpy", "cpz",
               "Cmx_f", "Cmx_p", "Cmy_f", "Cmy_p", "Cmz_f", "Cmz_p",
               "CL", "CD", "side", "CD_pr", "CD_sf", "CD_CL2PA", "xcp")


class PhlError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"phl status {status}: {msg}")
        self.status = status


_lib = None

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_F = ctypes.POINTER(ctypes.c_float)


def _declare(lib):
    sig = {
        "phl_last_error": (ctypes.c_char_p, []),
        "phl_params_default": (None, [ctypes.POINTER(Params)]),
        "phl_mesh_create": (ctypes.c_int, [ctypes.c_int, _D, _D, _D, ctypes.c_int, _I, _I, _I, _I,
                                           ctypes.c_int, _I, ctypes.c_int, _I, _I, ctypes.POINTER(_P)]),
        "phl_mesh_bump_hex": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.POINTER(_P)]),
        "phl_mesh_sizes": (ctypes.c_int, [_P, _I, _I, _I, _I, _I, _I]),
        "phl_mesh_bump_hybrid": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                ctypes.c_int, ctypes.c_double, ctypes.c_uint, ctypes.c_double,
                                                ctypes.POINTER(_P)]),
        "phl_mesh_renumber": (ctypes.c_int, [_P, ctypes.c_int, _I, _I, ctypes.POINTER(_P)]),
        "phl_mesh_swept_wing": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_int, ctypes.POINTER(_P)]),
        "phl_mesh_connectivity": (ctypes.c_int, [_P, _D, _D, _D, _I, _I, _I, _I, _I, _I, _I]),
        "phl_mesh_metrics": (ctypes.c_int, [_P] + [_D] * 12),
        "phl_mesh_destroy": (None, [_P]),
        "phl_write_tecplot": (ctypes.c_int, [_P, _D, ctypes.c_double, ctypes.c_double, ctypes.c_char_p,
                                             ctypes.c_char_p]),
        "phl_write_restart": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]),
        "phl_read_restart": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _I, _D]),
        "phl_device_info": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _I,
                                           ctypes.POINTER(ctypes.c_double)]),
        "phl_create": (ctypes.c_int, [_P, ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(_P)]),
        "phl_destroy": (None, [_P]),
        "phl_set_state": (ctypes.c_int, [_P, _D]),
        "phl_get_state": (ctypes.c_int, [_P, _D]),
        "phl_get_residual": (ctypes.c_int, [_P, _D]),
        "phl_get_field": (ctypes.c_int, [_P, ctypes.c_int, _D]),
        "phl_set_field": (ctypes.c_int, [_P, ctypes.c_int, _D]),
        "phl_turb_boundary": (ctypes.c_int, [_P]),
        "phl_turb_viscosity": (ctypes.c_int, [_P]),
        "phl_turb_gradient": (ctypes.c_int, [_P]),
        "phl_turb_update_residuals": (ctypes.c_int, [_P]),
        "phl_turb_spectrum": (ctypes.c_int, [_P]),
        "phl_turb_lusgs": (ctypes.c_int, [_P]),
        "phl_turb_relax": (ctypes.c_int, [_P]),
        "phl_init_freestream": (ctypes.c_int, [_P]),
        "phl_boundary": (ctypes.c_int, [_P]),
        "phl_gamma_t": (ctypes.c_int, [_P]),
        "phl_pressure_factor": (ctypes.c_int, [_P]),
        "phl_gradient": (ctypes.c_int, [_P]),
        "phl_compute_node_value": (ctypes.c_int, [_P]),
        "phl_get_node_values": (ctypes.c_int, [_P, _D]),
        "phl_limiter": (ctypes.c_int, [_P]),
        "phl_update_residuals": (ctypes.c_int, [_P]),
        "phl_load_residuals": (ctypes.c_int, [_P]),
        "phl_inviscid_flux": (ctypes.c_int, [_P]),
        "phl_viscous_flux": (ctypes.c_int, [_P]),
        "phl_residual": (ctypes.c_int, [_P]),
        "phl_launch_counts": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
        "phl_set_sweep_wait": (ctypes.c_int, [_P, ctypes.c_int]),
        "phl_spectrum_radius": (ctypes.c_int, [_P]),
        "phl_time_step": (ctypes.c_int, [_P, ctypes.c_int]),
        "phl_lusgs": (ctypes.c_int, [_P, _I]),
        "phl_sweep_normals": (ctypes.c_int, [_P, ctypes.c_int, _I, _D]),
        "phl_set_init_phase": (ctypes.c_int, [_P, ctypes.c_int]),
        "phl_residual_norm": (ctypes.c_int, [_P, _D]),
        "phl_residual_norm_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                     ctypes.c_void_p]),
        "phl_forces": (ctypes.c_int, [_P, _D]),
        "phl_forces_case": (ctypes.c_int, [_P, ctypes.c_int, _D]),
        "phl_mesh_concat": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(_P)]),
        "phl_read_fts": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P), _I]),
        "phl_read_wdt": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _D]),
        "phl_mesh_dim": (ctypes.c_int, [_P]),
        "phl_mesh_cases": (ctypes.c_int, [_P, _I, _I, _I]),
        "phl_create_batch": (ctypes.c_int, [_P, ctypes.POINTER(Params), ctypes.c_int, ctypes.POINTER(_P)]),
        "phl_batch_size": (ctypes.c_int, [_P, _I]),
        "phl_set_frozen": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_ubyte)]),
        "phl_solve_batch": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, _I, _D]),
        "phl_iterate": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D]),
        "phl_solve": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, _I, _D]),
        "phl_synchronize": (ctypes.c_int, [_P]),
        "phl_set_launch_mode": (ctypes.c_int, [_P, ctypes.c_int]),
        "phl_sweep_levels": (ctypes.c_int, [_P, _I, _I]),
        "phl_set_precision": (ctypes.c_int, [_P, ctypes.c_int]),
        "phl_get_precision": (ctypes.c_int, [_P, _I]),
        "phl_set_precision_switch": (ctypes.c_int, [_P, ctypes.c_int]),
        "phl_autotune": (ctypes.c_int, [_P, _I, _I, _D]),
        "phl_set_kernel_timing": (ctypes.c_int, [_P, ctypes.c_int]),
        "phl_get_kernel_timing": (ctypes.c_int, [_P, _D, _I]),
        "phl_hbm_copy": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_double)]),
        "phl_mpnn_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.POINTER(_P)]),
        "phl_mpnn_destroy": (None, [_P]),
        "phl_mpnn_n_params": (ctypes.c_int, [_P, _I]),
        "phl_mpnn_set_params": (ctypes.c_int, [_P, _F]),
        "phl_mpnn_get_params": (ctypes.c_int, [_P, _F]),
        "phl_mpnn_set_graph": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _I]),
        "phl_mpnn_set_graph_mesh": (ctypes.c_int, [_P, _P]),
        "phl_mpnn_set_data": (ctypes.c_int, [_P, _F, _F]),
        "phl_mpnn_forward": (ctypes.c_int, [_P, _F]),
        "phl_mpnn_set_forward_precision": (ctypes.c_int, [_P, ctypes.c_int]),
        "phl_mpnn_gradient": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double), _F]),
        "phl_mpnn_train": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_float, _D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load_library(path=LIB_PATH):
    """Load libphl.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError(f"{path} not found: run __graft_entry__.build() or phenlei_amd/build.py")
        _lib = _declare(ctypes.CDLL(path))
    return _lib


def exported_symbols():
    """Names of the C-ABI functions declared in include/phl.h."""
    import re
    with open(os.path.join(os.path.dirname(_HERE), "include", "phl.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(phl_[a-z_0-9]+)\s*\(", text)))


def _check(status):
    if status != PHL_OK:
        msg = load_library().phl_last_error().decode(errors="replace")
        raise PhlError(status, msg)


def _dptr(a):
    return a.ctypes.data_as(_D)


def _iptr(a):
    return a.ctypes.data_as(_I)


class Mesh:
    """Host-side unstructured grid (phl_mesh)."""

    def __init__(self, handle):
        self._h = handle
        lib = load_library()
        v = [ctypes.c_int() for _ in range(6)]
        _check(lib.phl_mesh_sizes(self._h, *[ctypes.byref(x) for x in v]))
        (self.n_node, self.n_cell, self.n_face, self.n_bface, self.n_face_nodes,
         self.n_cell_nodes) = [x.value for x in v]
        self.n_total = self.n_cell + self.n_bface
        nc = ctypes.c_int()
        _check(lib.phl_mesh_cases(self._h, ctypes.byref(nc), None, None))
        self.n_cases = nc.value
        self.case_cell_ptr = np.empty(self.n_cases + 1, np.int32)
        self.case_bface_ptr = np.empty(self.n_cases + 1, np.int32)
        _check(lib.phl_mesh_cases(self._h, ctypes.byref(nc), _iptr(self.case_cell_ptr), _iptr(self.case_bface_ptr)))
        self.dim = lib.phl_mesh_dim(self._h)
        self.n_zones = 1

    @classmethod
    def from_fts(cls, path):
        """The reference's partitioned grid file (phl_read_fts): its zones
        joined into one; n_zones is the count in the file."""
        h = _P()
        nz = ctypes.c_int()
        _check(load_library().phl_read_fts(os.fsencode(path), ctypes.byref(h), ctypes.byref(nz)))
        m = cls(h)
        m.n_zones = nz.value
        return m

    def read_wdt(self, path):
        """Cell wall distances of the reference's .wdt file, joined cell order."""
        wd = np.empty(self.n_cell)
        _check(load_library().phl_read_wdt(os.fsencode(path), self.n_cell, _dptr(wd)))
        return wd

    @classmethod
    def bump_hex(cls, nx, ny, nz, bump=0.05):
        h = _P()
        _check(load_library().phl_mesh_bump_hex(nx, ny, nz, bump, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def swept_wing(cls, n_chord, n_span, n_prism=30, h1=1e-5, n_tet=29):
        """Semi-span swept wing (phl_mesh_swept_wing); (80, 40, 30, 1e-5, 29) is 3.98M cells."""
        h = _P()
        _check(load_library().phl_mesh_swept_wing(n_chord, n_span, n_prism, h1, n_tet, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def bump_hybrid(cls, nx, nz, n_prism=15, h1=1e-4, n_tet=20, jitter=0.1, seed=1, bump=0.05):
        h = _P()
        _check(load_library().phl_mesh_bump_hybrid(nx, nz, n_prism, h1, n_tet, jitter, seed, bump,
                                                   ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, x, y, z, face_node_ptr, face_nodes, left, right, n_bface, bc_type,
                    cell_node_ptr, cell_nodes):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z)]
        b = [np.ascontiguousarray(v, dtype=np.int32) for v in
             (face_node_ptr, face_nodes, left, right, bc_type, cell_node_ptr, cell_nodes)]
        h = _P()
        _check(load_library().phl_mesh_create(
            len(a[0]), _dptr(a[0]), _dptr(a[1]), _dptr(a[2]), len(b[2]), _iptr(b[0]), _iptr(b[1]),
            _iptr(b[2]), _iptr(b[3]), n_bface, _iptr(b[4]), len(b[5]) - 1, _iptr(b[5]), _iptr(b[6]),
            ctypes.byref(h)))
        return cls(h)

    @classmethod
    def concat(cls, parts):
        """Batched mesh: the disjoint union of `parts`, one case each (phl_mesh_concat)."""
        arr = (_P * len(parts))(*[p._h for p in parts])
        h = _P()
        _check(load_library().phl_mesh_concat(arr, len(parts), ctypes.byref(h)))
        out = cls(h)
        out._parts = list(parts)
        return out

    def case_slices(self, k):
        """(cell slice, ghost slice) of case k in [5][n_total] arrays."""
        c0, c1 = self.case_cell_ptr[k], self.case_cell_ptr[k + 1]
        b0, b1 = self.case_bface_ptr[k], self.case_bface_ptr[k + 1]
        return slice(c0, c1), slice(self.n_cell + b0, self.n_cell + b1)

    def renumbered(self, method="colour", perm=None):
        """New mesh with cells renumbered ('colour' or an explicit permutation new->old)."""
        h = _P()
        nc = ctypes.c_int()
        if method == "colour":
            _check(load_library().phl_mesh_renumber(self._h, 0, None, ctypes.byref(nc), ctypes.byref(h)))
        else:
            p = np.ascontiguousarray(perm, dtype=np.int32)
            _check(load_library().phl_mesh_renumber(self._h, 1, _iptr(p), ctypes.byref(nc), ctypes.byref(h)))
        out = Mesh(h)
        out.n_colours = nc.value
        return out

    def connectivity(self):
        n = self
        out = dict(
            x=np.empty(n.n_node), y=np.empty(n.n_node), z=np.empty(n.n_node),
            face_node_ptr=np.empty(n.n_face + 1, np.int32), face_nodes=np.empty(n.n_face_nodes, np.int32),
            left=np.empty(n.n_face, np.int32), right=np.empty(n.n_face, np.int32),
            bc_type=np.empty(max(n.n_bface, 1), np.int32)[:n.n_bface],
            cell_node_ptr=np.empty(n.n_cell + 1, np.int32), cell_nodes=np.empty(n.n_cell_nodes, np.int32))
        bc = np.empty(max(n.n_bface, 1), np.int32)
        _check(load_library().phl_mesh_connectivity(
            self._h, _dptr(out["x"]), _dptr(out["y"]), _dptr(out["z"]), _iptr(out["face_node_ptr"]),
            _iptr(out["face_nodes"]), _iptr(out["left"]), _iptr(out["right"]), _iptr(bc),
            _iptr(out["cell_node_ptr"]), _iptr(out["cell_nodes"])))
        out["bc_type"] = bc[:n.n_bface].copy()
        out["n_bface"] = n.n_bface
        return out

    def metrics(self):
        nf, nt = self.n_face, self.n_total
        names_f = ["xfc", "yfc", "zfc", "xfn", "yfn", "zfn", "area"]
        names_c = ["xcc", "ycc", "zcc", "vol"]
        out = {k: np.empty(nf) for k in names_f}
        out.update({k: np.empty(nt) for k in names_c})
        out["skew"] = np.empty(self.n_cell)
        _check(load_library().phl_mesh_metrics(self._h, *[_dptr(out[k]) for k in names_f + names_c + ["skew"]]))
        return out

    def write_tecplot(self, path, q, mach, gamma=1.4, title="phenlei_amd"):
        """Tecplot ASCII file of the primitive state q [5][nTotal]: one
        FEPolyhedron zone, node values of density, u, v, w, pressure,
        temperature, Mach number and Cp."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.shape != (5, self.n_total):
            raise ValueError("q must be [5][nTotal]")
        _check(load_library().phl_write_tecplot(self._h, _dptr(q), float(mach), float(gamma), title.encode(),
                                                os.fsencode(path)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phl_mesh_destroy(self._h)
            self._h = None


KERNEL_CLASSES = ["boundary", "grad_lim", "flux", "norm", "dt_diag", "sweep_fwd", "ghost_dq",
                  "sweep_bwd", "update", "turb"]


def hbm_copy_gbps(device=0, nbytes=2 * 1024**3, iters=10):
    """Best-of-iters streaming-copy bandwidth (read + write bytes / time)."""
    g = ctypes.c_double()
    _check(load_library().phl_hbm_copy(device, float(nbytes), iters, ctypes.byref(g)))
    return g.value


def write_restart(path, q, outnstep=0):
    """One zone's restart block: int outnstep, then q[n_eq][n_total] doubles
    (q [5][nTotal] primitives for the NS solver, [nTotal] nu~ for SA)."""
    q = np.ascontiguousarray(np.atleast_2d(q), dtype=np.float64)
    _check(load_library().phl_write_restart(os.fsencode(path), int(outnstep), q.shape[0], q.shape[1], _dptr(q)))


def read_restart(path, n_eq, n_total):
    """(outnstep, q[n_eq][n_total]) of a restart block."""
    q = np.empty((n_eq, n_total))
    step = ctypes.c_int()
    _check(load_library().phl_read_restart(os.fsencode(path), n_eq, n_total, ctypes.byref(step), _dptr(q)))
    return step.value, q


def residual_norm_noalloc(res_ptr, n_cell, ld, sum5_ptr, stream=None):
    """phl_residual_norm_noalloc on caller-owned device memory: res_ptr and
    sum5_ptr are device addresses (ints), stream a hipStream_t or None.  The
    work is only enqueued on the stream."""
    _check(load_library().phl_residual_norm_noalloc(ctypes.c_void_p(res_ptr), n_cell, ld, ctypes.c_void_p(sum5_ptr),
                                                    ctypes.c_void_p(stream)))


def device_info(device=0):
    name = ctypes.create_string_buffer(256)
    ncu = ctypes.c_int()
    hbm = ctypes.c_double()
    _check(load_library().phl_device_info(device, name, 256, ctypes.byref(ncu), ctypes.byref(hbm)))
    return dict(name=name.value.decode(), n_cu=ncu.value, hbm_bytes=hbm.value)


class Solver:
    """Device context (phl_ctx).  Method names follow the reference hooks."""

    def __init__(self, mesh, params, device=0):
        """params: one Params, or a list of Params (one per case) for a batched mesh."""
        self.mesh = mesh
        self.params = params
        h = _P()
        if isinstance(params, (list, tuple)):
            arr = (Params * len(params))(*params)
            _check(load_library().phl_create_batch(mesh._h, arr, device, ctypes.byref(h)))
            self.n_cases = len(params)
        else:
            _check(load_library().phl_create(mesh._h, ctypes.byref(params), device, ctypes.byref(h)))
            self.n_cases = 1
        self._h = h

    # state ---------------------------------------------------------------
    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(5, self.mesh.n_total)
        _check(_lib.phl_set_state(self._h, _dptr(q)))

    def get_state(self):
        out = np.empty((5, self.mesh.n_total))
        _check(_lib.phl_get_state(self._h, _dptr(out)))
        return out

    def get_residual(self):
        out = np.empty((5, self.mesh.n_total))
        _check(_lib.phl_get_residual(self._h, _dptr(out)))
        return out

    def get_field(self, name):
        rows = FIELD_ROWS[name]
        out = np.empty((rows, self.mesh.n_total))
        _check(_lib.phl_get_field(self._h, FIELD[name], _dptr(out)))
        return out if rows > 1 else out[0]

    def set_field(self, name, value):
        """Overwrite q, dt, vist, nut or wdist (phl_set_field)."""
        n = self.mesh.n_total if SETTABLE_ROWS_LEN[name] == "total" else self.mesh.n_cell
        v = np.ascontiguousarray(value, dtype=np.float64).reshape(FIELD_ROWS[name], -1)[:, :n]
        v = np.ascontiguousarray(v)
        _check(_lib.phl_set_field(self._h, FIELD[name], _dptr(v)))

    def init_freestream(self):
        _check(_lib.phl_init_freestream(self._h))

    # Spalart-Allmaras hooks (viscous = 3) -----------------------------------
    def turb_boundary(self):
        _check(_lib.phl_turb_boundary(self._h))

    def turb_viscosity(self):
        _check(_lib.phl_turb_viscosity(self._h))

    def turb_gradient(self):
        _check(_lib.phl_turb_gradient(self._h))

    def turb_update_residuals(self):
        _check(_lib.phl_turb_update_residuals(self._h))

    def turb_spectrum(self):
        _check(_lib.phl_turb_spectrum(self._h))

    def turb_lusgs(self):
        _check(_lib.phl_turb_lusgs(self._h))

    def turb_relax(self):
        _check(_lib.phl_turb_relax(self._h))

    # the path ------------------------------------------------------------
    def boundary(self):
        _check(_lib.phl_boundary(self._h))

    def gamma_t(self):
        _check(_lib.phl_gamma_t(self._h))

    def pressure_factor(self):
        _check(_lib.phl_pressure_factor(self._h))

    def gradient(self):
        _check(_lib.phl_gradient(self._h))

    def compute_node_value(self):
        """ComputeNodeValue + ModifyNodeValue (gradient 2, "ggnode")."""
        _check(_lib.phl_compute_node_value(self._h))

    def node_values(self):
        """Node values [6][nNode] (five primitives, then T) of the last node pass (ggnode)."""
        out = np.empty((6, self.mesh.n_node))
        _check(_lib.phl_get_node_values(self._h, _dptr(out)))
        return out

    def limiter(self):
        _check(_lib.phl_limiter(self._h))

    def load_residuals(self):
        _check(_lib.phl_load_residuals(self._h))

    def inviscid_flux(self):
        _check(_lib.phl_inviscid_flux(self._h))

    def viscous_flux(self):
        _check(_lib.phl_viscous_flux(self._h))

    def update_residuals(self):
        """NSSolver::UpdateResiduals: both face loops, added to the residual."""
        _check(_lib.phl_update_residuals(self._h))

    def residual(self):
        """LoadResiduals + Controller::UpdateResiduals fused: boundary, gradient, limiter, face loops."""
        _check(_lib.phl_residual(self._h))

    def launch_counts(self):
        """(graph replays, eagerly launched steps) of iterate since creation."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _check(_lib.phl_launch_counts(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_sweep_wait(self, ticks):
        """Wait bound of the dataflow sweeps in 100 MHz ticks (default 20,000,000)."""
        _check(_lib.phl_set_sweep_wait(self._h, int(ticks)))

    def spectrum_radius(self):
        _check(_lib.phl_spectrum_radius(self._h))

    def time_step(self, step):
        _check(_lib.phl_time_step(self._h, step))

    def lusgs(self):
        n = ctypes.c_int()
        _check(_lib.phl_lusgs(self._h, ctypes.byref(n)))
        return n.value

    def set_init_phase(self, on):
        """isInIniting for hook calls: first-order faces, Van Leer flux, zero limiter."""
        _check(_lib.phl_set_init_phase(self._h, int(bool(on))))

    def sweep_normals(self, turb=False):
        """Squared dq changes of the sweep pairs of the last LU-SGS call
        (n_lusgs_sweeps > 1; NS, or the turbulence model with turb), first 8."""
        n = ctypes.c_int()
        out = np.empty(8)
        _check(_lib.phl_sweep_normals(self._h, int(bool(turb)), ctypes.byref(n), _dptr(out)))
        return out[:min(n.value, 8)]

    def residual_norm(self):
        out = np.empty(5)
        _check(_lib.phl_residual_norm(self._h, _dptr(out)))
        return out

    def set_launch_mode(self, mode):
        """'graph' (default) replays captured hipGraphs of 8 steps; 'eager' launches each kernel
        in order on one stream; 'eager_overlap' launches eagerly with the turbulence sweeps on a
        second stream (as the graphs do)."""
        _check(_lib.phl_set_launch_mode(self._h, {"eager": 0, "graph": 1, "eager_overlap": 2}[mode]))

    def autotune(self):
        """Pick the flux / gradient workgroup tiles by timing (phl_autotune)."""
        ft, gt = ctypes.c_int(), ctypes.c_int()
        ms = np.empty(6)
        _check(_lib.phl_autotune(self._h, ctypes.byref(ft), ctypes.byref(gt), _dptr(ms)))
        return dict(flux_tile=ft.value, grad_tile=gt.value,
                    flux_ms=dict(zip((64, 128, 256), ms[:3].round(5).tolist())),
                    grad_ms=dict(zip((64, 128, 256), ms[3:].round(5).tolist())))

    def set_precision(self, precision):
        """'fp64' or 'mixed' (FP32 gradients and face fluxes, FP64 accumulation)."""
        _check(_lib.phl_set_precision(self._h, {"fp64": 0, "mixed": 1}[precision]))

    def precision(self):
        """'fp64' or 'mixed' (the current form of the context)."""
        v = ctypes.c_int()
        _check(_lib.phl_get_precision(self._h, ctypes.byref(v)))
        return "mixed" if v.value else "fp64"

    def set_precision_switch(self, stall_steps):
        """solve / solve_batch switch a mixed context to FP64 after stall_steps without halving."""
        _check(_lib.phl_set_precision_switch(self._h, stall_steps))

    def forces(self):
        """Wall force coefficients (phl_forces) as a dict keyed by FORCE_NAMES."""
        out = np.empty(len(FORCE_NAMES))
        _check(_lib.phl_forces(self._h, _dptr(out)))
        return dict(zip(FORCE_NAMES, out))

    def iterate(self, first_step, n_steps, history=True):
        """Density-residual RMS per step ([n_steps], or [n_steps, n_cases] for a batch)."""
        hist = np.empty((max(n_steps, 1), self.n_cases))
        _check(_lib.phl_iterate(self._h, first_step, n_steps, _dptr(hist) if history else None))
        if not history:
            return None
        return hist[:n_steps, 0] if self.n_cases == 1 else hist[:n_steps]

    def solve_batch(self, max_steps, drop):
        """Per-case steps to the drop (negative: not reached) and the [steps, n_cases] history."""
        steps = np.empty(self.n_cases, np.int32)
        hist = np.empty((max_steps, self.n_cases))
        _check(_lib.phl_solve_batch(self._h, max_steps, drop, _iptr(steps), _dptr(hist)))
        run = int(np.max(np.abs(steps)))
        return steps, hist[:run]

    def set_frozen(self, frozen):
        """Freeze cases of a batch (bool per case) from the next iterate on."""
        f = np.ascontiguousarray(np.asarray(frozen, dtype=np.uint8))
        _check(_lib.phl_set_frozen(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))))

    def forces_case(self, k):
        out = np.empty(len(FORCE_NAMES))
        _check(_lib.phl_forces_case(self._h, k, _dptr(out)))
        return dict(zip(FORCE_NAMES, out))

    def solve(self, max_steps, drop):
        hist = np.empty(max_steps)
        n = ctypes.c_int()
        _check(_lib.phl_solve(self._h, max_steps, drop, ctypes.byref(n), _dptr(hist)))
        return n.value, hist[:n.value]

    def synchronize(self):
        _check(_lib.phl_synchronize(self._h))

    def sweep_levels(self):
        """(forward, backward) dependency levels of the LU-SGS sweeps."""
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(_lib.phl_sweep_levels(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_kernel_timing(self, enable=True):
        _check(_lib.phl_set_kernel_timing(self._h, 1 if enable else 0))

    def kernel_timing(self):
        """{class: total ms}, number of timed steps (HIP events on the solver stream)."""
        ms = np.zeros(len(KERNEL_CLASSES))
        n = ctypes.c_int()
        _check(_lib.phl_get_kernel_timing(self._h, _dptr(ms), ctypes.byref(n)))
        return dict(zip(KERNEL_CLASSES, ms.tolist())), n.value

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.phl_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _fptr(a):
    return a.ctypes.data_as(_F)


class Mpnn:
    """Warm-start message-passing network on the device (phl_mpnn): hidden
    width 128, FP32 (BF16 inference optional), dense layers on MFMA, Adam
    training."""

    N_IN, N_OUT, HIDDEN = 8, 5, 128

    def __init__(self, layers=6, seed=0, device=0):
        self.layers = layers
        h = _P()
        _check(load_library().phl_mpnn_create(layers, seed, device, ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int()
        _check(_lib.phl_mpnn_n_params(self._h, ctypes.byref(n)))
        self.n_params = n.value
        self.n_nodes = 0

    def layout(self):
        """(name, shape) of the flat parameter vector, in order."""
        L, Hd = self.layers, self.HIDDEN
        return [("W0", (self.N_IN, Hd)), ("b0", (Hd,)), ("A", (L, Hd, Hd)), ("B", (L, Hd, Hd)), ("c", (L, Hd)),
                ("Wo", (Hd, self.N_OUT)), ("bo", (self.N_OUT,))]

    def get_params(self):
        flat = np.empty(self.n_params, np.float32)
        _check(_lib.phl_mpnn_get_params(self._h, _fptr(flat)))
        return self.unflatten(flat)

    def set_params(self, params):
        flat = np.concatenate([np.asarray(params[k], np.float32).reshape(-1) for k, _ in self.layout()])
        assert flat.size == self.n_params
        _check(_lib.phl_mpnn_set_params(self._h, _fptr(flat)))

    def unflatten(self, flat):
        out, o = {}, 0
        for k, shp in self.layout():
            n = int(np.prod(shp))
            out[k] = flat[o:o + n].reshape(shp).copy()
            o += n
        return out

    def set_graph(self, nbr=None, mesh=None):
        if mesh is not None:
            _check(_lib.phl_mpnn_set_graph_mesh(self._h, mesh._h))
            self.n_nodes = mesh.n_cell
        else:
            nbr = np.ascontiguousarray(nbr, np.int32)
            _check(_lib.phl_mpnn_set_graph(self._h, nbr.shape[0], nbr.shape[1], _iptr(nbr)))
            self.n_nodes = nbr.shape[0]

    def set_data(self, X, T=None):
        X = np.ascontiguousarray(X, np.float32).reshape(self.n_nodes, self.N_IN)
        Tp = None
        if T is not None:
            T = np.ascontiguousarray(T, np.float32).reshape(self.n_nodes, self.N_OUT)
            Tp = _fptr(T)
        _check(_lib.phl_mpnn_set_data(self._h, _fptr(X), Tp))

    def forward(self):
        y = np.empty((self.n_nodes, self.N_OUT), np.float32)
        _check(_lib.phl_mpnn_forward(self._h, _fptr(y)))
        return y

    def set_forward_precision(self, bf16):
        """forward()'s hidden-layer products: False FP32, True BF16 (FP32
        accumulation); training stays FP32."""
        _check(_lib.phl_mpnn_set_forward_precision(self._h, 1 if bf16 else 0))

    def gradient(self):
        loss = ctypes.c_double()
        g = np.empty(self.n_params, np.float32)
        _check(_lib.phl_mpnn_gradient(self._h, ctypes.byref(loss), _fptr(g)))
        return loss.value, self.unflatten(g)

    def train(self, steps, lr):
        hist = np.empty(max(steps, 1))
        _check(_lib.phl_mpnn_train(self._h, steps, lr, _dptr(hist)))
        return hist[:steps]

    def close(self):
        if getattr(self, "_h", None):
            _lib.phl_mpnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

