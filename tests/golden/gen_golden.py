"""Writes the golden fixtures tests/golden/*.npz from the CPU oracle.

The reference ships no tests, fixtures or golden vectors and cannot be built
here (Rust toolchain and its arithmetic crates are absent), so these vectors
come from the oracle once its building blocks are pinned
(tests/test_oracle_pins.py).  They freeze the oracle against regressions and
give the GPU path a fixed target.  Run from the repository root:

    python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "app-binary2d_amd")]

from tests import helpers as H  # noqa: E402

CASES = {
    "iso_rk2_1x16": dict(kind="iso", nb=1, bs=16, kw=dict(rk_order="RK2")),
    "iso_rk3_alpha_torquefree_2x16": dict(kind="iso", nb=2, bs=16,
                                          kw=dict(rk_order="RK3", viscosity={"alpha": 0.1},
                                                  sink_model="torque_free", binary_eccentricity=0.5,
                                                  binary_mass_ratio=0.5)),
    "euler_rk3_alpha_1x32": dict(kind="euler", nb=1, bs=32,
                                 kw=dict(rk_order="RK3", viscosity={"alpha": 0.1}, binary_eccentricity=0.5,
                                         binary_mass_ratio=0.5)),
    "euler_beta_rk2_2x16": dict(kind="euler_beta", nb=2, bs=16, kw=dict(rk_order="RK2")),
}
STEPS = (1, 10)


def generate(name, c):
    mesh, hydro, ph = H.small(c["kind"], nb=c["nb"], bs=c["bs"], R=4.0, **c["kw"])
    u0 = H.initial(mesh, hydro, ph)
    o = H.oracle_for(mesh, hydro, ph, u0)
    out = {"initial": u0}
    dts = []
    for k in range(1, max(STEPS) + 1):
        dts.append(o.advance(1))
        if k in STEPS:
            s = o.get_state()
            out[f"conserved_{k}"] = s["conserved"]
            out[f"ist_{k}"] = s["ist"]
            out[f"oec_{k}"] = s["oec"]
            out[f"time_{k}"] = np.array(s["time"])
    out["dt"] = np.array(dts)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


def case_config(name):
    c = CASES[name]
    return H.small(c["kind"], nb=c["nb"], bs=c["bs"], R=4.0, **c["kw"])


if __name__ == "__main__":
    for n, c in CASES.items():
        generate(n, c)
        print("wrote", n)
