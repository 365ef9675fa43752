"""cdc_max_signal_speed_noalloc, the entry point that allocates nothing,
through the C ABI against cdc_solver_max_signal_speed on the same states:
bitwise equal (a maximum of the same per-cell speeds), on two calls in a row
whose second has fewer cells, for both hydro kinds.  Device buffers come
from the HIP runtime the library itself loaded (no second runtime in the
process)."""
import ctypes

import numpy as np
import pytest

from cdc_amd import Solver, abi
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    lib = abi.load()
    assert lib.cdc_device_count() >= 1, "no HIP device: the HIP path must run on a GPU"
    rt = ctypes.CDLL("libamdhip64.so.7")      # the runtime libcdc_hip.so linked, already loaded
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
        getattr(rt, f).restype = ctypes.c_int
    return lib, rt


H2D, D2H = 1, 2


def noalloc_speed(lib, rt, kind, gamma, u):
    """Upload u ([..][v], float64) to device memory, call the entry point,
    read back the result."""
    a = np.ascontiguousarray(u, dtype=np.float64)
    nv = a.shape[-1]
    dev, out = ctypes.c_void_p(), ctypes.c_void_p()
    assert rt.hipMalloc(ctypes.byref(dev), a.nbytes) == 0
    assert rt.hipMalloc(ctypes.byref(out), 8) == 0
    try:
        assert rt.hipMemcpy(dev, a.ctypes.data, a.nbytes, H2D) == 0
        rc = lib.cdc_max_signal_speed_noalloc(kind, gamma, dev, a.size // nv, out, None)
        assert rc == 0, lib.cdc_last_error_message()
        assert rt.hipDeviceSynchronize() == 0
        r = np.zeros(1)
        assert rt.hipMemcpy(r.ctypes.data, out, 8, D2H) == 0
        return float(r[0])
    finally:
        rt.hipFree(dev)
        rt.hipFree(out)


@pytest.mark.parametrize("kind", ["iso", "euler"])
def test_noalloc_max_signal_speed_matches_solver(hip, kind):
    lib, rt = hip
    # a 4x4-block state of 64^2, then a single 16^2 block: fewer cells and
    # fewer partial maxima on the second call
    for nb, bs, seed in ((4, 64, 3), (1, 16, 4)):
        mesh, hydro, ph = H.small(kind, nb=nb, bs=bs, R=4.0)
        u = H.initial(mesh, hydro, ph, seed=seed)
        rng = np.random.default_rng(seed)
        u[..., 1:3] *= 1.0 + rng.uniform(-0.5, 0.5, size=u.shape[:-1] + (2,))   # distinct speeds
        g = Solver(mesh, hydro, ph, u)
        want = g.max_signal_speed()
        g.close()
        k = abi.CDC_EULER if hydro.nv == 4 else abi.CDC_ISOTHERMAL
        gm = hydro.gamma_law_index if hydro.nv == 4 else 1.0
        got = noalloc_speed(lib, rt, k, gm, u)
        assert got == want, (nb, bs, got, want)


def test_noalloc_rejects_bad_arguments(hip):
    lib, _ = hip
    fake = ctypes.c_void_p(0x1000)      # never dereferenced: every case fails its host check first
    assert lib.cdc_max_signal_speed_noalloc(abi.CDC_ISOTHERMAL, 1.0, fake, 0, fake, None) == abi.CDC_ERR_INVALID_ARGUMENT
    assert lib.cdc_max_signal_speed_noalloc(abi.CDC_ISOTHERMAL, 1.0, fake, (1 << 40) + 1, fake,
                                            None) == abi.CDC_ERR_INVALID_ARGUMENT
    assert lib.cdc_max_signal_speed_noalloc(7, 1.0, fake, 10, fake, None) == abi.CDC_ERR_INVALID_ARGUMENT
    assert lib.cdc_max_signal_speed_noalloc(abi.CDC_EULER, 1.0, fake, 10, fake, None) == abi.CDC_ERR_INVALID_ARGUMENT
    assert lib.cdc_max_signal_speed_noalloc(abi.CDC_ISOTHERMAL, 1.0, None, 10, fake, None) == abi.CDC_ERR_INVALID_ARGUMENT
