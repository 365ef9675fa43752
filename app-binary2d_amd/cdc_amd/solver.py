"""Host-side handle over the C ABI: BlockData + State + scheme::advance.

`Solver.advance(fold)` is scheme::advance (src/scheme.rs:323-349) on the GPU;
a hydro error raises `HydroError` (src/physics.rs:26-71) and leaves the
solver's state as it was before the call.  Arrays use the reference layout:
conserved [block][i][j][v], ist [block][7][nv], oec [block][7][4].
"""
import ctypes
from ctypes import byref, c_double, c_int64, c_void_p

import numpy as np

from . import abi


class CdcError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


class HydroError(CdcError):
    """HydroError: NegativeDensity / NegativePressure / OrbitalEvolutionError."""

    def __init__(self, err):
        self.kind = {abi.CDC_ERR_NEGATIVE_DENSITY: "NegativeDensity",
                     abi.CDC_ERR_NEGATIVE_PRESSURE: "NegativePressure",
                     abi.CDC_ERR_UNBOUND_ORBIT: "OrbitalEvolutionError"}.get(err.code, "Unknown")
        self.value = err.value
        self.position = (err.x, err.y)
        self.time = err.time
        self.binary = [(b.mass, b.position_x, b.position_y, b.velocity_x, b.velocity_y) for b in err.binary] \
            if err.has_binary else None
        super().__init__(err.code, err.message.decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(c_void_p)


def _check(lib, rc):
    if rc != abi.CDC_OK:
        raise CdcError(rc, lib.cdc_last_error_message().decode(errors="replace"))


def orbital_state_from_time(physics, t):
    lib = abi.load()
    out = (abi.CPointMass * 2)()
    c = physics.to_c()
    _check(lib, lib.cdc_orbital_state_from_time(byref(c), float(t), out))
    return np.array([[p.mass, p.position_x, p.position_y, p.velocity_x, p.velocity_y] for p in out])


def resample_state(src_mesh, dst_mesh, conserved, ist=None, oec=None):
    """Re-grid a boundary-layout state between meshes of one domain
    (cdc_resample_state).  Returns (conserved, ist, oec) on dst_mesh."""
    lib = abi.load()
    nv = conserved.shape[-1]
    u = np.ascontiguousarray(conserved, dtype=np.float64)
    nbd, bsd = dst_mesh.num_blocks, dst_mesh.block_size
    out = np.empty((nbd * nbd, bsd, bsd, nv))
    oi = np.empty((nbd * nbd, 7, nv))
    oe = np.empty((nbd * nbd, 7, 4))
    i = None if ist is None else np.ascontiguousarray(ist, dtype=np.float64)
    e = None if oec is None else np.ascontiguousarray(oec, dtype=np.float64)
    ms, md = src_mesh.to_c(), dst_mesh.to_c()
    _check(lib, lib.cdc_resample_state(byref(ms), byref(md), nv, _ptr(u), None if i is None else _ptr(i),
                                       None if e is None else _ptr(e), _ptr(out), _ptr(oi), _ptr(oe)))
    return out, oi, oe


def seeded_cavity(mesh, hydro, cavity_radius=2.5, cavity_exponent=12.0, density_floor=1e-4,
                  noise_amplitude=1e-3, softening_length=0.05, mach_number=10.0, seed=1,
                  taper_radius=0.0, taper_width=0.25):
    """The seeded cavity initial model (no reference counterpart), per cell."""
    lib = abi.load()
    m = abi.CCavityModel(cavity_radius, cavity_exponent, density_floor, noise_amplitude,
                         softening_length, mach_number, taper_radius, taper_width, seed)
    nb, bs = mesh.num_blocks, mesh.block_size
    out = np.zeros((nb * nb, bs, bs, hydro.nv), dtype=np.float64)
    cm, ch = mesh.to_c(), hydro.to_c()
    _check(lib, lib.cdc_model_seeded_cavity(byref(cm), byref(ch), byref(m), _ptr(out)))
    return out


class Solver:
    """One mesh, hydro system and physics on one GPU."""

    def __init__(self, mesh, hydro, physics, initial_conserved):
        self.lib = abi.load()
        abi.check_layout(self.lib)
        self.mesh, self.hydro, self.physics = mesh, hydro, physics
        self.nv = hydro.nv
        self.nblk = mesh.num_blocks ** 2
        shape = (self.nblk, mesh.block_size, mesh.block_size, self.nv)
        init = np.ascontiguousarray(initial_conserved, dtype=np.float64).reshape(shape)
        self._cm, self._ch, self._cp = mesh.to_c(), hydro.to_c(), physics.to_c()
        h = c_void_p()
        _check(self.lib, self.lib.cdc_solver_create(byref(self._cm), byref(self._ch), byref(self._cp),
                                                   _ptr(init), byref(h)))
        self.handle = h
        self.shape = shape

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cdc_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_physics(self, physics):
        self.physics = physics
        self._cp = physics.to_c()
        _check(self.lib, self.lib.cdc_solver_set_physics(self.handle, byref(self._cp)))

    def set_state(self, conserved, time=0.0, iteration=(0, 1), ist=None, oec=None, torque=None):
        """Upload a State; `torque` ([nblk][2] per-block torque impulses, as
        torque() returns them) restores the running torque totals, else they
        start from zero."""
        u = np.ascontiguousarray(conserved, dtype=np.float64).reshape(self.shape)
        i = None if ist is None else np.ascontiguousarray(ist, dtype=np.float64).reshape(self.nblk, 7, self.nv)
        e = None if oec is None else np.ascontiguousarray(oec, dtype=np.float64).reshape(self.nblk, 7, 4)
        _check(self.lib, self.lib.cdc_solver_set_state(self.handle, float(time), int(iteration[0]),
                                                      int(iteration[1]), _ptr(u),
                                                      None if i is None else _ptr(i),
                                                      None if e is None else _ptr(e)))
        if torque is not None:
            self.set_torque(torque)

    def set_torque(self, torque):
        """Restore the per-block torque impulses [nblk][2] of the current state."""
        t = np.ascontiguousarray(torque, dtype=np.float64).reshape(self.nblk, 2)
        _check(self.lib, self.lib.cdc_solver_set_torque(self.handle, _ptr(t)))

    def get_state(self):
        """Returns dict(time, iteration=(num, den), conserved, ist, oec)."""
        u = np.empty(self.shape, dtype=np.float64)
        ist = np.empty((self.nblk, 7, self.nv), dtype=np.float64)
        oec = np.empty((self.nblk, 7, 4), dtype=np.float64)
        t, n, d = c_double(), c_int64(), c_int64()
        _check(self.lib, self.lib.cdc_solver_get_state(self.handle, byref(t), byref(n), byref(d),
                                                      _ptr(u), _ptr(ist), _ptr(oec)))
        return {"time": t.value, "iteration": (n.value, d.value), "conserved": u, "ist": ist, "oec": oec}

    def time(self):
        """Committed simulation time (no state copy)."""
        t = c_double()
        _check(self.lib, self.lib.cdc_solver_get_state(self.handle, byref(t), None, None, None, None, None))
        return t.value

    def advance(self, fold=1):
        """scheme::advance: returns dt; raises HydroError (state unchanged)."""
        dt = c_double()
        err = abi.CError()
        rc = self.lib.cdc_solver_advance(self.handle, int(fold), byref(dt), byref(err))
        if rc in (abi.CDC_ERR_NEGATIVE_DENSITY, abi.CDC_ERR_NEGATIVE_PRESSURE, abi.CDC_ERR_UNBOUND_ORBIT):
            raise HydroError(err)
        _check(self.lib, rc)
        return dt.value

    def advance_dt(self, dt, fold=1):
        """`fold` RK steps at the given dt (no CFL); raises HydroError."""
        err = abi.CError()
        rc = self.lib.cdc_solver_advance_dt(self.handle, int(fold), float(dt), byref(err))
        if rc in (abi.CDC_ERR_NEGATIVE_DENSITY, abi.CDC_ERR_NEGATIVE_PRESSURE, abi.CDC_ERR_UNBOUND_ORBIT):
            raise HydroError(err)
        _check(self.lib, rc)

    def resample_from(self, src):
        """This solver's committed state := src's, re-gridded on the device."""
        _check(self.lib, self.lib.cdc_solver_resample_from(self.handle, src.handle))

    def save(self, snapshot):
        _check(self.lib, self.lib.cdc_solver_save(self.handle, snapshot.handle))

    def restore(self, snapshot):
        _check(self.lib, self.lib.cdc_solver_restore(self.handle, snapshot.handle))

    def torque(self):
        """Per-block torque impulses on body 1 and 2, [nblk][2]."""
        out = np.empty((self.nblk, 2))
        _check(self.lib, self.lib.cdc_solver_get_torque(self.handle, _ptr(out)))
        return out

    def time_series_sample(self):
        """State::time_series_sample: dict(time, ist [7][nv], oec [7][4],
        torque [2]) summed over blocks."""
        t = c_double()
        ist = np.empty((7, self.nv))
        oec = np.empty((7, 4))
        tq = np.empty(2)
        _check(self.lib, self.lib.cdc_solver_time_series_sample(self.handle, byref(t), _ptr(ist), _ptr(oec),
                                                               _ptr(tq)))
        return {"time": t.value, "ist": ist, "oec": oec, "torque": tq}

    def max_signal_speed(self):
        a = c_double()
        _check(self.lib, self.lib.cdc_solver_max_signal_speed(self.handle, byref(a)))
        return a.value

    def set_tile(self, tx=0, ty=0):
        _check(self.lib, self.lib.cdc_solver_set_tile(self.handle, int(tx), int(ty)))

    def get_tile(self):
        tx, ty = ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib, self.lib.cdc_solver_get_tile(self.handle, byref(tx), byref(ty)))
        return tx.value, ty.value

    def set_kernel(self, tx=0, ty=0, threads=0):
        _check(self.lib, self.lib.cdc_solver_set_kernel(self.handle, int(tx), int(ty), int(threads)))

    def get_kernel(self):
        """(tx, ty, threads) of the stage kernel in use."""
        tx, ty, nt = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib, self.lib.cdc_solver_get_kernel(self.handle, byref(tx), byref(ty), byref(nt)))
        return tx.value, ty.value, nt.value

    def autotune(self, reps=3):
        """Time every supported tile shape on the current state; keep the fastest.
        Returns (tx, ty, ms per stage launch)."""
        tx, ty, ms = ctypes.c_int32(), ctypes.c_int32(), c_double()
        _check(self.lib, self.lib.cdc_solver_autotune(self.handle, int(reps), byref(tx), byref(ty), byref(ms)))
        return tx.value, ty.value, ms.value

    def set_timing(self, enable=True):
        _check(self.lib, self.lib.cdc_solver_set_timing(self.handle, int(enable)))

    def get_timing(self):
        n, ms, cells = c_int64(), c_double(), c_double()
        _check(self.lib, self.lib.cdc_solver_get_timing(self.handle, byref(n), byref(ms), byref(cells)))
        return n.value, ms.value, cells.value


class Snapshot:
    """A device copy of a solver's committed state (cdc_snapshot_*)."""

    def __init__(self, solver):
        self.lib = abi.load()
        h = c_void_p()
        _check(self.lib, self.lib.cdc_snapshot_create(solver.handle, byref(h)))
        self.handle = h

    def time(self):
        t = c_double()
        _check(self.lib, self.lib.cdc_snapshot_time(self.handle, byref(t)))
        return t.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cdc_snapshot_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ensemble:
    """One RK step of many independent runs per call (cdc_ensemble_*).

    `members` are Solver objects of one hydro kind and rk_order; each active
    member advances one step at its own dt and ends bitwise where
    `member.advance(1)` would have left it.  The members stay owned by the
    caller and must outlive the ensemble.
    """

    def __init__(self, members, graph=True):
        self.lib = abi.load()
        self.members = list(members)
        n = len(self.members)
        arr = (c_void_p * n)(*[m.handle for m in self.members])
        h = c_void_p()
        _check(self.lib, self.lib.cdc_ensemble_create(arr, n, byref(h)))
        self.handle = h
        self.n = n
        if not graph:
            self.set_graph(False)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cdc_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_graph(self, enable=True):
        _check(self.lib, self.lib.cdc_ensemble_set_graph(self.handle, int(enable)))

    def get_tile(self):
        tx, ty = ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib, self.lib.cdc_ensemble_get_tile(self.handle, byref(tx), byref(ty)))
        return tx.value, ty.value

    def stats(self):
        n, z = c_int64(), c_double()
        _check(self.lib, self.lib.cdc_ensemble_get_stats(self.handle, byref(n), byref(z)))
        return n.value, z.value

    def advance(self, active=None):
        """One step of every active member.  Returns (dt, codes, errors): per
        member dt (nan when inactive), error code (0 = committed) and the
        HydroError of each failed member (None otherwise)."""
        n = self.n
        act = None if active is None else np.ascontiguousarray(active, dtype=np.int32).reshape(n)
        dt = np.full(n, np.nan)
        codes = np.zeros(n, dtype=np.int32)
        errs = (abi.CError * n)()
        rc = self.lib.cdc_ensemble_advance(self.handle, None if act is None else _ptr(act), _ptr(dt), _ptr(codes),
                                           ctypes.cast(errs, c_void_p))
        hydro = (abi.CDC_ERR_NEGATIVE_DENSITY, abi.CDC_ERR_NEGATIVE_PRESSURE, abi.CDC_ERR_UNBOUND_ORBIT)
        if rc != abi.CDC_OK and rc not in hydro:
            _check(self.lib, rc)
        out = [HydroError(errs[i]) if codes[i] in hydro else None for i in range(n)]
        return dt, codes, out
