"""Error bounds of the stage kernel's FP64 1/x, 1/sqrt(x) and sqrt(x): the
hardware estimates refined by one Newton step, through cdc_math_probe,
against extended-precision references (x86 long double, 64-bit mantissa).
The bars sit just above the maxima measured over 524,288 log-uniform inputs
(10.8, 19 and 0.5 ulp)."""
import numpy as np
import pytest

from cdc_amd import abi

pytestmark = pytest.mark.gpu


def ulps(approx, ref):
    spacing = np.abs(np.spacing(ref.astype(np.float64))).astype(np.longdouble)
    return (np.abs(approx.astype(np.longdouble) - ref) / spacing).astype(np.float64)


def test_refined_estimates_error_bounds():
    lib = abi.load()
    assert lib.cdc_device_count() >= 1, "no HIP device: the HIP path must run on a GPU"
    rng = np.random.default_rng(11)
    n = 1 << 17
    x = (1.0 + rng.random(n)) * np.exp2((rng.random(n) - 0.5) * 120.0)   # about 1e-18 .. 1e18
    # exact powers of two and the values the solver meets most (O(1e-6 .. 1e3))
    x[:64] = np.exp2(np.arange(-32, 32))
    x[64:4096] = 10.0 ** rng.uniform(-6, 3, 4096 - 64)
    out = np.empty((n, 3))
    rc = lib.cdc_math_probe(x.ctypes.data, out.ctypes.data, n)
    assert rc == 0, lib.cdc_last_error_message().decode()
    xl = x.astype(np.longdouble)
    sq = np.sqrt(xl)
    assert ulps(out[:, 0], np.longdouble(1) / xl).max() <= 11.0
    assert ulps(out[:, 1], np.longdouble(1) / sq).max() <= 20.0
    assert ulps(out[:, 2], sq).max() <= 0.5 + 1e-9


def test_math_probe_rejects_bad_arguments():
    lib = abi.load()
    assert lib.cdc_math_probe(None, None, 4) == abi.CDC_ERR_INVALID_ARGUMENT
    x = np.ones(1)
    assert lib.cdc_math_probe(x.ctypes.data, x.ctypes.data, 0) == 0
