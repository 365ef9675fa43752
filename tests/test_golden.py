"""Golden fixtures (tests/golden/*.npz, written by tests/golden/gen_golden.py):
the oracle reproduces them bit for bit on CPU; the HIP path reproduces them
within the parity bar on GPU."""
import glob
import os

import numpy as np
import pytest

from cdc_amd import Solver
from tests import helpers as H
from tests.golden.gen_golden import CASES, STEPS, case_config

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(CASES)


def load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz")))


def test_every_case_has_a_fixture():
    assert sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz"))) == NAMES


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    g = load(name)
    mesh, hydro, ph = case_config(name)
    np.testing.assert_array_equal(H.initial(mesh, hydro, ph), g["initial"])
    o = H.oracle_for(mesh, hydro, ph, g["initial"])
    for k in range(1, max(STEPS) + 1):
        assert o.advance(1) == g["dt"][k - 1]
        if k in STEPS:
            s = o.get_state()
            np.testing.assert_array_equal(s["conserved"], g[f"conserved_{k}"])
            np.testing.assert_array_equal(s["ist"], g[f"ist_{k}"])
            np.testing.assert_array_equal(s["oec"], g[f"oec_{k}"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_golden(name):
    g = load(name)
    mesh, hydro, ph = case_config(name)
    s = Solver(mesh, hydro, ph, g["initial"])
    for k in range(1, max(STEPS) + 1):
        assert s.advance(1) == pytest.approx(g["dt"][k - 1], rel=1e-12)
        if k in STEPS:
            st = s.get_state()
            assert st["time"] == pytest.approx(float(g[f"time_{k}"]), rel=1e-12)
            assert H.field_err(st["conserved"], g[f"conserved_{k}"]) <= 1e-12
            keep = [0, 1, 2, 3, 4, 6]
            assert H.items_err(st["ist"][:, keep], g[f"ist_{k}"][:, keep]) <= 1e-10
            assert np.max(np.abs(st["oec"] - g[f"oec_{k}"])) <= 1e-12
