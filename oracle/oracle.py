"""ctypes wrapper of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py to check the HIP implementation.  The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
# the same restatement built with OpenMP: face loops split by cell owner, LU-SGS
# by dependency level; the same bits as liborc.so (the CPU baseline over all
# host cores)
LIB_PATH_OMP = os.path.join(HERE, "liborc_omp.so")

FIELD = dict(q=0, res=1, dqdx=2, dqdy=3, dqdz=4, limit=5, dq=6, spec_inv=7, spec_vis=8, dt=9,
             diag=10, t=11, visl=12, dtdx=13, dtdy=14, dtdz=15, delta_flux=16,
             vist=17, nut=18, res_t=19, spec_t=20, dq_t=21, wdist=22, dnudx=23, dnudy=24, dnudz=25,
             pressure_factor=26)
FIELD_ROWS = dict(q=5, res=5, dqdx=5, dqdy=5, dqdz=5, limit=1, dq=5, spec_inv=1, spec_vis=1, dt=1,
                  diag=1, t=1, visl=1, dtdx=1, dtdy=1, dtdz=1, delta_flux=5,
                  vist=1, nut=1, res_t=1, spec_t=1, dq_t=1, wdist=1, dnudx=1, dnudy=1, dnudz=1,
                  pressure_factor=1)

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_P = ctypes.c_void_p
_libs = {}


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def load(omp=False):
    path = LIB_PATH_OMP if omp else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            raise ImportError(f"{path} not built (run phenlei_amd/build.py)")
        lib = ctypes.CDLL(path)
        lib.orc_face_metrics.argtypes = [ctypes.c_int, _I, _I, _D, _D, _D] + [_D] * 7
        lib.orc_cell_metrics.argtypes = [ctypes.c_int] * 3 + [_I, _I, _I, _I, _I, _I, _D, _D, _D] + [_D] * 7 + [_D] * 4
        lib.orc_skewness.argtypes = [ctypes.c_int] * 3 + [_I, _I] + [_D] * 10 + [_D]
        lib.orc_face_metrics_2d.argtypes = [ctypes.c_int, _I, _I, _D, _D, _D] + [_D] * 7
        lib.orc_cell_metrics_2d.argtypes = [ctypes.c_int] * 3 + [_I, _I] + [_D] * 7 + [_D] * 4
        lib.orc_set_dim.argtypes = [_P, ctypes.c_int]
        lib.orc_create.restype = _P
        lib.orc_create.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I, _I, _I] + [_D] * 12
        lib.orc_destroy.argtypes = [_P]
        lib.orc_get.argtypes = [_P, ctypes.c_int, _D]
        lib.orc_set.argtypes = [_P, ctypes.c_int, _D]
        lib.orc_freestream.argtypes = [_P, _D]
        for name in ("orc_init_freestream", "orc_boundary", "orc_gamma_t", "orc_pressure_factor", "orc_viscosity",
                     "orc_gradient",
                     "orc_limiter", "orc_zero_residual", "orc_inviscid_flux", "orc_viscous_flux",
                     "orc_update_residuals", "orc_residual", "orc_spectrum", "orc_sa_boundary", "orc_sa_viscosity",
                     "orc_sa_gradient", "orc_sa_update_residuals", "orc_sa_spectrum", "orc_sa_lusgs",
                     "orc_sa_relax"):
            getattr(lib, name).argtypes = [_P]
        lib.orc_time_step.argtypes = [_P, ctypes.c_int]
        lib.orc_lusgs.argtypes = [_P, _D]
        lib.orc_sweep_normals.argtypes = [_P, ctypes.c_int, _D]
        lib.orc_set_init_phase.argtypes = [_P, ctypes.c_int]
        lib.orc_vanleer_flux.argtypes = [_D, _D] + [ctypes.c_double] * 4 + [_D]
        lib.orc_sweep_normals.restype = ctypes.c_int
        lib.orc_lusgs.restype = ctypes.c_int
        lib.orc_residual_norm.argtypes = [_P, _D]
        lib.orc_iterate.argtypes = [_P, ctypes.c_int, ctypes.c_int, _D]
        lib.orc_iterate.restype = ctypes.c_int
        lib.orc_forces.argtypes = [_P, _D]
        lib.orc_set_wall_distance.argtypes = [_P, _D]
        lib.orc_sa_face_matrices.argtypes = [_P, _D, _D]
        lib.orc_wall_distance.argtypes = [ctypes.c_int, _D, _D, _D, ctypes.c_int, _I, _I] + [_D] * 6 + [_D]
        lib.orc_wall_distance_2d.argtypes = [ctypes.c_int, _D, _D, _D, ctypes.c_int, _I, _I] + [_D] * 6 + [_D]
        lib.orc_force_coefficients.argtypes = [_D, ctypes.c_double, ctypes.c_double, _D]
        lib.orc_roe_flux.argtypes = [_D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_double, _D]
        lib.orc_roe_flux_pc.argtypes = [_D, _D] + [ctypes.c_double] * 4 + [ctypes.c_int] + [ctypes.c_double] * 4 + [_D]
        lib.orc_harten.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.orc_harten.restype = ctypes.c_double
        lib.orc_venfun.argtypes = [ctypes.c_double] * 3
        lib.orc_venfun.restype = ctypes.c_double
        lib.orc_mxdq.argtypes = [_D] + [ctypes.c_double] * 5 + [_D, ctypes.c_double, ctypes.c_int, _D]
        lib.orc_lusgs_rad.argtypes = [_D, _D] + [ctypes.c_double] * 5
        lib.orc_lusgs_rad.restype = ctypes.c_double
        lib.orc_entropy_coef.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.orc_entropy_coef.restype = ctypes.c_double
        lib.orc_sa_source_point.argtypes = [ctypes.c_double] * 6 + [_D]
        lib.orc_face_states.argtypes = [_P, ctypes.c_int, _D, _D]
        lib.orc_set_nodes.argtypes = [_P, ctypes.c_int, _D, _D, _D, _I, _I, _I, _I]
        lib.orc_node_values.argtypes = [_P]
        lib.orc_get_node_values.argtypes = [_P, _D]
        lib.orc_ptotal.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.orc_ptotal.restype = ctypes.c_double
        _libs[path] = lib
    return _libs[path]


def is_2d(conn):
    """A grid whose faces are all edges (two nodes) is 2-D."""
    return bool(np.all(np.diff(np.asarray(conn["face_node_ptr"])) == 2))


def face_metrics(conn):
    lib = load()
    nf = len(conn["left"])
    out = {k: np.empty(nf) for k in ("xfc", "yfc", "zfc", "xfn", "yfn", "zfn", "area")}
    fn = lib.orc_face_metrics_2d if is_2d(conn) else lib.orc_face_metrics
    fn(nf, _ip(conn["face_node_ptr"]), _ip(conn["face_nodes"]), _dp(conn["x"]), _dp(conn["y"]), _dp(conn["z"]),
       *[_dp(out[k]) for k in ("xfc", "yfc", "zfc", "xfn", "yfn", "zfn", "area")])
    return out


def cell_metrics(conn, fm):
    lib = load()
    nf = len(conn["left"])
    nb = conn["n_bface"]
    nc = len(conn["cell_node_ptr"]) - 1
    nt = nc + nb
    out = {k: np.empty(nt) for k in ("xcc", "ycc", "zcc", "vol")}
    if is_2d(conn):
        lib.orc_cell_metrics_2d(nc, nb, nf, _ip(conn["left"]), _ip(conn["right"]),
                                *[_dp(fm[k]) for k in ("xfc", "yfc", "zfc", "xfn", "yfn", "zfn", "area")],
                                *[_dp(out[k]) for k in ("xcc", "ycc", "zcc", "vol")])
    else:
        lib.orc_cell_metrics(nc, nb, nf, _ip(conn["left"]), _ip(conn["right"]), _ip(conn["face_node_ptr"]),
                             _ip(conn["face_nodes"]), _ip(conn["cell_node_ptr"]), _ip(conn["cell_nodes"]),
                             _dp(conn["x"]), _dp(conn["y"]), _dp(conn["z"]),
                             *[_dp(fm[k]) for k in ("xfc", "yfc", "zfc", "xfn", "yfn", "zfn", "area")],
                             *[_dp(out[k]) for k in ("xcc", "ycc", "zcc", "vol")])
    skew = np.empty(nc)
    lib.orc_skewness(nc, nb, nf, _ip(conn["left"]), _ip(conn["right"]),
                     *[_dp(fm[k]) for k in ("xfc", "yfc", "zfc", "xfn", "yfn", "zfn", "area")],
                     *[_dp(out[k]) for k in ("xcc", "ycc", "zcc")], _dp(skew))
    out["skew"] = skew
    return out


def vanleer_flux(qL, qR, n, gamma=1.4):
    qL = np.ascontiguousarray(qL, dtype=np.float64)
    qR = np.ascontiguousarray(qR, dtype=np.float64)
    out = np.empty(5)
    load().orc_vanleer_flux(_dp(qL), _dp(qR), float(n[0]), float(n[1]), float(n[2]), gamma, _dp(out))
    return out


def roe_flux(qL, qR, n, gamma=1.4, fix=3, eps=1e-4, convect=None, pc=None):
    """Roe flux per unit area; eps is the acoustic entropy-fix coefficient,
    convect the convective one (eps when None); pc the pressure factors of
    the face's two cells (methods 2 and 6)."""
    lib = load()
    qL = np.ascontiguousarray(qL, dtype=np.float64)
    qR = np.ascontiguousarray(qR, dtype=np.float64)
    out = np.empty(5)
    cv = eps if convect is None else convect
    if pc is None:
        lib.orc_roe_flux(_dp(qL), _dp(qR), float(n[0]), float(n[1]), float(n[2]), gamma, fix, eps, cv, _dp(out))
    else:
        lib.orc_roe_flux_pc(_dp(qL), _dp(qR), float(n[0]), float(n[1]), float(n[2]), gamma, fix, eps, cv,
                            float(pc[0]), float(pc[1]), _dp(out))
    return out


def mxdq(prim, n, area, gamma, dq, radius, ipn):
    lib = load()
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    dq = np.ascontiguousarray(dq, dtype=np.float64)
    out = np.empty(5)
    lib.orc_mxdq(_dp(prim), float(n[0]), float(n[1]), float(n[2]), area, gamma, _dp(dq), radius, ipn, _dp(out))
    return out


def lusgs_rad(ql, qr, n, area, gamma=1.4):
    ql = np.ascontiguousarray(ql, dtype=np.float64)
    qr = np.ascontiguousarray(qr, dtype=np.float64)
    return load().orc_lusgs_rad(_dp(ql), _dp(qr), float(n[0]), float(n[1]), float(n[2]), area, gamma)


class OracleSolver:
    """Sequential FP64 restatement of NSSolverUnstruct on one zone."""

    def __init__(self, conn, metrics, params, omp=False):
        """omp: the OpenMP build (same results, all host cores)."""
        self._L = lib = load(omp)
        self.n_cell = len(conn["cell_node_ptr"]) - 1
        self.n_bface = conn["n_bface"]
        self.n_total = self.n_cell + self.n_bface
        self._keep = [np.ascontiguousarray(conn[k], dtype=np.int32) for k in ("left", "right")]
        bc = np.ascontiguousarray(conn["bc_type"], dtype=np.int32)
        if bc.size == 0:
            bc = np.zeros(1, np.int32)
        self._keep.append(bc)
        m = [np.ascontiguousarray(metrics[k], dtype=np.float64) for k in
             ("xfc", "yfc", "zfc", "xfn", "yfn", "zfn", "area", "xcc", "ycc", "zcc", "vol", "skew")]
        self._keep += m
        self.params = params
        self._h = lib.orc_create(ctypes.byref(params), self.n_cell, self.n_bface, len(conn["left"]),
                                 _ip(self._keep[0]), _ip(self._keep[1]), _ip(self._keep[2]), *[_dp(a) for a in m])
        self.n_node = len(conn["x"])
        # node lists: the node-based gradient (gradient 2) needs them
        nodes = [np.ascontiguousarray(conn[k], dtype=np.float64) for k in ("x", "y", "z")]
        lists = [np.ascontiguousarray(conn[k], dtype=np.int32) for k in
                 ("face_node_ptr", "face_nodes", "cell_node_ptr", "cell_nodes")]
        lib.orc_set_nodes(self._h, self.n_node, *[_dp(a) for a in nodes], *[_ip(a) for a in lists])
        self.dim = 2 if is_2d(conn) else 3
        lib.orc_set_dim(self._h, self.dim)

    def compute_node_value(self):
        """ComputeNodeValue + ModifyNodeValue on the current state."""
        self._L.orc_node_values(self._h)

    def node_values(self):
        """Node values [6][nNode] (five primitives, then T) of the last node pass."""
        out = np.empty((6, self.n_node))
        self._L.orc_get_node_values(self._h, _dp(out))
        return out

    def get(self, name):
        rows = FIELD_ROWS[name]
        out = np.empty((rows, self.n_total))
        self._L.orc_get(self._h, FIELD[name], _dp(out))
        return out if rows > 1 else out[0]

    def set(self, name, value):
        v = np.ascontiguousarray(value, dtype=np.float64).reshape(-1)
        self._L.orc_set(self._h, FIELD[name], _dp(v))

    def freestream(self):
        out = np.empty(5)
        self._L.orc_freestream(self._h, _dp(out))
        return out

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        fn = getattr(self._L, "orc_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda *a: fn(self._h, *a)

    def set_wall_distance(self, wd):
        wd = np.ascontiguousarray(wd, dtype=np.float64)
        self._L.orc_set_wall_distance(self._h, _dp(wd))

    def sa_face_matrices(self, n_face):
        a, b = np.empty(n_face), np.empty(n_face)
        self._L.orc_sa_face_matrices(self._h, _dp(a), _dp(b))
        return a, b

    def face_states(self, f):
        """(qL, qR) of face f: MUSCL with the positivity fallback and BoundaryQlQrFix."""
        a, b = np.empty(5), np.empty(5)
        self._L.orc_face_states(self._h, int(f), _dp(a), _dp(b))
        return a, b

    def lusgs(self):
        sweeps = np.empty(8)
        return self._L.orc_lusgs(self._h, _dp(sweeps))

    def sweep_normals(self, turb=False):
        """Squared dq changes of each sweep pair of the last LU-SGS call (NS,
        or SA with turb), at most the first 8."""
        out = np.empty(8)
        n = self._L.orc_sweep_normals(self._h, int(bool(turb)), _dp(out))
        return out[:min(n, 8)]

    def residual_norm(self):
        out = np.empty(5)
        self._L.orc_residual_norm(self._h, _dp(out))
        return out

    def forces(self):
        """AirForceCoefParts + CompForce at the current state (22 values)."""
        out = np.empty(22)
        self._L.orc_forces(self._h, _dp(out))
        return out

    def iterate(self, first_step, n_steps):
        hist = np.empty(max(n_steps, 1))
        nfix = self._L.orc_iterate(self._h, first_step, n_steps, _dp(hist))
        return hist[:n_steps], nfix

    def close(self):
        if getattr(self, "_h", None):
            self._L.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def sa_source_point(nut, omega, wdist, rho, visl, reynolds):
    """SA source at one point: (source per volume, (P - D)/nu~, its
    linearisation, implicit diagonal term per volume)."""
    out = np.empty(4)
    load().orc_sa_source_point(nut, omega, wdist, rho, visl, reynolds, _dp(out))
    return out


def wall_distance(mesh_conn, metrics, bc_wall=2):
    """Cell-centre wall distance over the solid-surface faces (orc_wall_distance)."""
    n_cell = len(mesh_conn["cell_node_ptr"]) - 1
    nb = mesh_conn["n_bface"]
    walls = [f for f in range(nb) if mesh_conn["bc_type"][f] == bc_wall]
    fp = mesh_conn["face_node_ptr"]
    ptr = [0]
    nodes = []
    for f in walls:
        nodes.extend(mesh_conn["face_nodes"][fp[f]:fp[f + 1]])
        ptr.append(len(nodes))
    ptr = np.asarray(ptr, np.int32)
    nodes = np.asarray(nodes if nodes else [0], np.int32)
    w = np.asarray(walls if walls else [0], np.int64)
    wx = np.ascontiguousarray(metrics["xfc"][w])
    wy = np.ascontiguousarray(metrics["yfc"][w])
    wz = np.ascontiguousarray(metrics["zfc"][w])
    out = np.empty(n_cell)
    c = [np.ascontiguousarray(metrics[k][:n_cell]) for k in ("xcc", "ycc", "zcc")]
    xyz = [np.ascontiguousarray(mesh_conn[k], dtype=np.float64) for k in ("x", "y", "z")]
    fn = load().orc_wall_distance_2d if is_2d(mesh_conn) else load().orc_wall_distance
    fn(n_cell, _dp(c[0]), _dp(c[1]), _dp(c[2]), len(walls), _ip(ptr), _ip(nodes),
       _dp(wx), _dp(wy), _dp(wz), _dp(xyz[0]), _dp(xyz[1]), _dp(xyz[2]), _dp(out))
    return out
