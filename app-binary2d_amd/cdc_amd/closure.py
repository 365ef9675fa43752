"""Learned flux closure for coarse runs (cdc_closure_*), and the pipeline
that trains it on a down-sampled fine run and compares coarse runs with and
without it against the fine run.  No counterpart in the reference."""
import ctypes
import math
import time
from ctypes import byref, c_double, c_int32, c_void_p

import numpy as np

from . import abi
from .solver import CdcError, HydroError, _check, _ptr

NUM_PARAMS = None


def num_params():
    global NUM_PARAMS
    if NUM_PARAMS is None:
        NUM_PARAMS = int(abi.load().cdc_closure_num_params())
    return NUM_PARAMS


class Closure:
    """Net, Adam state and training samples on the GPU for one coarse grid."""

    def __init__(self, grid, domain_radius, max_samples=0, seed=1, feature_scale=(0.25, 0.5, 0.5),
                 output_scale=(1.0, 1.0, 1.0), loss_weight=(1.0, 1.0, 1.0), init_output_gain=1.0):
        self.lib = abi.load()
        c = abi.CClosureConfig()
        c.grid, c.max_samples, c.seed, c.domain_radius = int(grid), int(max_samples), int(seed), float(domain_radius)
        c.feature_scale[:] = list(feature_scale)
        c.output_scale[:] = list(output_scale)
        c.loss_weight[:] = list(loss_weight)
        c.init_output_gain = float(init_output_gain)
        h = c_void_p()
        _check(self.lib, self.lib.cdc_closure_create(byref(c), byref(h)))
        self.handle = h
        self.grid = int(grid)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.cdc_closure_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self):
        c = abi.CClosureConfig()
        _check(self.lib, self.lib.cdc_closure_get_config(self.handle, byref(c)))
        return c

    def params(self):
        out = np.empty(num_params(), dtype=np.float32)
        _check(self.lib, self.lib.cdc_closure_get_params(self.handle, _ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(num_params())
        _check(self.lib, self.lib.cdc_closure_set_params(self.handle, _ptr(p)))

    def set_scales(self, output_scale=None, loss_weight=None):
        o = None if output_scale is None else np.ascontiguousarray(output_scale, dtype=np.float64)
        w = None if loss_weight is None else np.ascontiguousarray(loss_weight, dtype=np.float64)
        _check(self.lib, self.lib.cdc_closure_set_scales(self.handle, None if o is None else _ptr(o),
                                                         None if w is None else _ptr(w)))

    def features(self, solver):
        out = np.empty((self.grid * self.grid, 4), dtype=np.float32)
        _check(self.lib, self.lib.cdc_closure_features(self.handle, solver.handle, _ptr(out)))
        return out.reshape(self.grid, self.grid, 4)

    def forward(self, solver):
        out = np.empty((self.grid * self.grid, 6), dtype=np.float32)
        _check(self.lib, self.lib.cdc_closure_forward(self.handle, solver.handle, _ptr(out)))
        return out.reshape(self.grid, self.grid, 6)

    def add_sample(self, base, plain, truth, dt):
        _check(self.lib, self.lib.cdc_closure_add_sample(self.handle, base.handle, plain.handle, truth.handle,
                                                         float(dt)))

    def samples(self):
        n = c_int32()
        ms = np.zeros(3)
        _check(self.lib, self.lib.cdc_closure_num_samples(self.handle, byref(n), _ptr(ms)))
        return n.value, ms

    def clear_samples(self):
        _check(self.lib, self.lib.cdc_closure_clear_samples(self.handle))

    def sample_grad(self, k):
        g = np.empty(num_params(), dtype=np.float32)
        l = c_double()
        _check(self.lib, self.lib.cdc_closure_sample_grad(self.handle, int(k), byref(l), _ptr(g)))
        return l.value, g

    def train(self, steps, lr=1e-3, seed=0):
        out = np.empty(max(1, steps), dtype=np.float32)
        _check(self.lib, self.lib.cdc_closure_train(self.handle, int(steps), float(lr), int(seed), _ptr(out)))
        return out[:steps]

    def advance(self, solver):
        """One corrected coarse step; returns dt, raises HydroError."""
        dt = c_double()
        err = abi.CError()
        rc = self.lib.cdc_closure_advance(self.handle, solver.handle, byref(dt), byref(err))
        if rc in (abi.CDC_ERR_NEGATIVE_DENSITY, abi.CDC_ERR_NEGATIVE_PRESSURE, abi.CDC_ERR_UNBOUND_ORBIT):
            raise HydroError(err)
        _check(self.lib, rc)
        return dt.value
