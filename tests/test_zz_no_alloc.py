"""phl_residual_norm_noalloc (no allocation: block partials in a static
device array) against phl_residual_norm (partials in the context's buffer).
Both sum 256-cell blocks in storage order and then the block sums in a
fixed order, so on the same array in the same order they agree bitwise."""
import ctypes

import numpy as np
import pytest

import phenlei_amd as pa

pytestmark = pytest.mark.gpu


class _Hip:
    """The few HIP runtime calls the test needs for its own device buffers."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.lib.hipFree.argtypes = [ctypes.c_void_p]
        self.lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def _ok(self, e):
        assert e == 0, f"HIP error {e}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self._ok(self.lib.hipMalloc(ctypes.byref(p), a.nbytes))
        self._ok(self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))
        return p.value

    def zeros(self, n):
        return self.upload(np.zeros(n))

    def download(self, ptr, n):
        out = np.empty(n)
        self._ok(self.lib.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), out.nbytes, 2))
        return out

    def sync(self):
        self._ok(self.lib.hipDeviceSynchronize())

    def free(self, ptr):
        self._ok(self.lib.hipFree(ctypes.c_void_p(ptr)))


def _residual(mesh, steps):
    """A live residual of a context that stores cells in the mesh's order."""
    p = pa.default_params(mach=0.5, cfl_start=10.0, cfl_end=10.0, storage_order=1)
    s = pa.Solver(mesh, p)
    s.init_freestream()
    s.iterate(1, steps)
    s.residual()
    ref = s.residual_norm()
    res = s.get_residual()
    s.close()
    return res, ref


def test_noalloc_norm_equals_the_context_norm_on_two_calls():
    hip = _Hip()
    out = hip.zeros(5)
    try:
        for dims, steps in (((48, 24, 8), 7), ((16, 8, 4), 3)):  # the second call is smaller
            m = pa.Mesh.bump_hex(*dims, 0.05)
            res, ref = _residual(m, steps)
            nc = m.n_cell
            ld = res.shape[1]
            assert np.any(res[:, :nc] != 0.0)
            d = hip.upload(res)
            pa.residual_norm_noalloc(d, nc, ld, out)
            hip.sync()
            got = hip.download(out, 5)
            hip.free(d)
            assert np.array_equal(got, ref), (dims, got, ref)
    finally:
        hip.free(out)


def test_noalloc_norm_rejects_sizes_beyond_its_buffer():
    with pytest.raises(pa.PhlError) as e:
        pa.residual_norm_noalloc(1 << 20, pa.PHL_NOALLOC_MAX_CELLS + 1, pa.PHL_NOALLOC_MAX_CELLS + 1, 1 << 20)
    assert e.value.status == pa.PHL_ERR_ARG
    with pytest.raises(pa.PhlError):
        pa.residual_norm_noalloc(0, 10, 10, 1 << 20)
