"""Accretion and torque series from recorded time-series samples (the
reference's tools/time-series: extract_time_series flattens the samples of a
checkpoint into arrays, its plot shows the accreted mass -sink_k[0]).

A sample (State::time_series_sample, src/state.rs:135-146, as recorded by
cdc_amd.driver.Runner or read back from a checkpoint) holds the time, the
block sums of integrated_source_terms [7][nv] and orbital_elements_change
[7][4], and the summed torque impulses on the two bodies [2] (no reference
counterpart).  Rates are finite differences between consecutive samples,
placed at the interval midpoints.

    python -m cdc_amd.timeseries chkpt.0042.cbor --output series.npz
"""
import argparse

import numpy as np

from . import checkpoint as ckpt

ORBITAL_PERIOD = 2.0 * np.pi


def extract(samples):
    """List of samples -> dict of arrays: time [n], ist [n][7][nv],
    oec [n][7][4], torque [n][2] (zeros where a sample has none)."""
    if not samples:
        raise ValueError("time series data was empty")
    out = {"time": np.array([float(s["time"]) for s in samples]),
           "ist": np.array([np.asarray(s["ist"], dtype=np.float64) for s in samples]),
           "oec": np.array([np.asarray(s["oec"], dtype=np.float64) for s in samples])}
    out["torque"] = np.array([np.asarray(s.get("torque", np.zeros(2)), dtype=np.float64) for s in samples])
    return out


def accreted_mass(series):
    """Mass taken by sink 1 and sink 2 so far, [n][2] (-sink_k[0])."""
    ist = series["ist"]
    return -np.stack([ist[:, 0, 0], ist[:, 1, 0]], axis=1)


def rates(series):
    """Midpoint times [n-1], accretion rates Mdot_k [n-1][2] and torques on the
    bodies tau_k [n-1][2] between consecutive samples."""
    t = series["time"]
    dt = np.diff(t)
    if np.any(dt <= 0):
        raise ValueError("sample times must increase")
    tm = 0.5 * (t[1:] + t[:-1])
    mdot = np.diff(accreted_mass(series), axis=0) / dt[:, None]
    tau = np.diff(series["torque"], axis=0) / dt[:, None]
    return tm, mdot, tau


def from_checkpoint(path):
    return extract(ckpt.read(path)["time_series"])


def main(argv=None):
    ap = argparse.ArgumentParser(description="accretion and torque series from a checkpoint")
    ap.add_argument("filename")
    ap.add_argument("--output", default=None, help="npz file (default: time_series.NNNN.npz)")
    a = ap.parse_args(argv)
    s = from_checkpoint(a.filename)
    tm, mdot, tau = rates(s)
    out = a.output or a.filename.replace("chkpt", "time_series").replace(".cbor", ".npz")
    np.savez(out, time=s["time"], ist=s["ist"], oec=s["oec"], torque=s["torque"], rate_time=tm, mdot=mdot,
             torque_rate=tau)
    print(f"save {out}: {len(s['time'])} samples, orbits {s['time'][0] / ORBITAL_PERIOD:.3f} to "
          f"{s['time'][-1] / ORBITAL_PERIOD:.3f}")


if __name__ == "__main__":
    main()
