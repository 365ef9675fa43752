"""Device re-gridding between solver handles (cdc_solver_resample_from)
against the host resampler (bitwise: same pairwise-tree order), and the
fixed-dt advance against the CFL advance."""
import numpy as np
import pytest

from cdc_amd import Mesh, Solver, abi, resample_state
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    lib = abi.load()
    assert lib.cdc_device_count() >= 1, "no HIP device: the HIP path must run on a GPU"


@pytest.mark.parametrize("kind", ["iso", "euler"])
def test_device_resample_matches_host(kind):
    fm, hydro, ph = H.small(kind, nb=4, bs=16, R=4.0)
    u0 = H.initial(fm, hydro, ph)
    fine = Solver(fm, hydro, ph, u0)
    fine.advance(2)
    sf = fine.get_state()
    cm = Mesh(2, 8, 4.0)      # 64^2 -> 16^2, re-blocked
    cu0, _, _ = resample_state(fm, cm, u0)
    coarse = Solver(cm, hydro, ph, cu0)
    coarse.resample_from(fine)
    sc = coarse.get_state()
    want, wi, we = resample_state(fm, cm, sf["conserved"], sf["ist"], sf["oec"])
    np.testing.assert_array_equal(sc["conserved"], want)
    np.testing.assert_array_equal(sc["ist"], wi)
    np.testing.assert_array_equal(sc["oec"], we)
    assert sc["time"] == sf["time"] and sc["iteration"] == sf["iteration"]
    # and back up: injection, then the coarse state again
    up = Solver(fm, hydro, ph, u0)
    up.resample_from(coarse)
    back = Solver(cm, hydro, ph, cu0)
    back.resample_from(up)
    np.testing.assert_array_equal(back.get_state()["conserved"], want)


def test_advance_dt_is_advance_at_the_cfl_dt():
    mesh, hydro, ph = H.small("euler", nb=2, bs=16, R=4.0, rk_order="RK3")
    u0 = H.initial(mesh, hydro, ph)
    a, b = Solver(mesh, hydro, ph, u0), Solver(mesh, hydro, ph, u0)
    for _ in range(3):
        dt = a.advance(1)
        b.advance_dt(dt)
    sa, sb = a.get_state(), b.get_state()
    np.testing.assert_array_equal(sa["conserved"], sb["conserved"])
    np.testing.assert_array_equal(sa["ist"], sb["ist"])
    np.testing.assert_array_equal(sa["oec"], sb["oec"])
    with pytest.raises(Exception):
        b.advance_dt(-1.0)
