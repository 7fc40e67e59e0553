"""ctypes binding of the CPU oracle (liboracle.so / liboracle_fast.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / CPU baseline.  The product path
never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PARITY_LIB = os.path.join(HERE, "liboracle.so")
FAST_LIB = os.path.join(HERE, "liboracle_fast.so")

OR_POS, OR_VEL, OR_DENSITY, OR_PRESSURE, OR_ACCEL, OR_MASS = range(6)


class Params(C.Structure):
    _fields_ = [
        ("rest_density", C.c_float), ("gas_constant", C.c_float), ("viscosity", C.c_float),
        ("smoothing_length", C.c_float), ("particle_mass", C.c_float), ("timestep", C.c_float),
        ("gravity", C.c_float), ("damping", C.c_float), ("CFL_factor", C.c_float),
        ("xmin", C.c_float), ("xmax", C.c_float), ("ymin", C.c_float), ("ymax", C.c_float),
        ("zmin", C.c_float), ("zmax", C.c_float), ("neighbor_search_radius", C.c_float),
        ("formulation", C.c_int32),
    ]


class SurCfg(C.Structure):
    """Mirror of or_sur_cfg (surrogate network definition, see surrogate_oracle.c)."""
    _fields_ = [
        ("h1", C.c_int32), ("h2", C.c_int32), ("h3", C.c_int32), ("bf16", C.c_int32),
        ("inv_r", C.c_float), ("inv_v", C.c_float), ("inv_h", C.c_float),
        ("w_norm", C.c_float), ("rho0", C.c_float), ("rho_scale", C.c_float),
        ("msg_scale", C.c_float), ("acc_scale", C.c_float),
    ]


_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_U64 = C.POINTER(C.c_uint64)
_P = C.c_void_p
_SZ = C.c_size_t

_SIG = {
    "or_default_params": (None, [C.POINTER(Params)]),
    "or_create": (_P, [_SZ]),
    "or_destroy": (None, [_P]),
    "or_set_params": (None, [_P, C.POINTER(Params)]),
    "or_set_particles": (C.c_int, [_P, _SZ, _F, _F, _F]),
    "or_size": (_SZ, [_P]),
    "or_step": (C.c_float, [_P, C.c_float]),
    "or_update_neighbor_lists": (None, [_P]),
    "or_compute_densities": (None, [_P]),
    "or_compute_pressures": (None, [_P]),
    "or_compute_forces": (None, [_P]),
    "or_integrate_leapfrog": (None, [_P, C.c_float]),
    "or_apply_boundary_conditions": (None, [_P]),
    "or_compute_cfl_timestep": (C.c_float, [_P]),
    "or_get": (None, [_P, C.c_int, _F]),
    "or_neighbor_total": (_SZ, [_P]),
    "or_get_neighbors": (None, [_P, _I64, _I64]),
    "or_max_neighbors": (_SZ, [_P]),
    "or_total_mass": (C.c_float, [_P]),
    "or_total_energy": (C.c_float, [_P]),
    "or_num_cells": (_SZ, [_P]),
    "or_hash_position": (C.c_uint64, [C.c_int32, C.c_int32, C.c_int32]),
    "or_grid_coords": (None, [C.c_float, _F, _I32]),
    "or_cell_keys": (None, [_SZ, _F, C.c_float, _I32, _U64]),
    "or_W": (C.c_float, [_F, C.c_float, C.c_int32]),
    "or_gradW": (None, [_F, C.c_float, C.c_int32, _F]),
    "or_laplacianW": (C.c_float, [_F, C.c_float, C.c_int32]),
    "or_integrate_clamp": (None, [_SZ, _F, _F, _F, C.c_float, C.POINTER(Params)]),
    "or64_density_accel": (None, [_SZ, _F, _F, _F, _I64, _I64, C.POINTER(Params), _D, _D]),
    "or64_accel_term_scale": (None, [_SZ, _F, _F, _F, _F, _I64, _I64, C.POINTER(Params), _D]),
    "or_brute_neighbors": (_SZ, [_SZ, _F, C.c_float, _I64, _I64, _SZ]),
    "or_count_neighbors": (C.c_uint64, [_P, _U64]),
    "or_sample_step0": (C.c_int, [_P, _SZ, _I64, _I64, _F, _F, _D, _D, _D]),
    "or_sample_neighbors": (None, [_P, _SZ, _I64, _I64, _I64]),
    "or_num_threads": (C.c_int, []),
    "or_bf16_round": (C.c_float, [C.c_float]),
    "or_sur_nparams": (_SZ, [C.POINTER(SurCfg)]),
    "or_sur_forward": (None, [C.POINTER(SurCfg), _F, _SZ, _F, _F, _F, _I64, _I64, _D, _D,
                              _I64]),
    "or_sur_loss_grad": (C.c_double, [C.POINTER(SurCfg), _F, _SZ, _F, _F, _F, _I64, _I64,
                                      C.POINTER(C.c_uint8), _F, _D]),
    "or_sur_adam": (None, [_SZ, _F, _F, _F, _F, C.c_float, C.c_float, C.c_float, C.c_float,
                           C.c_int32]),
}

_libs = {}


def build() -> None:
    """Compiles liboracle.so and liboracle_fast.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(fast: bool = False) -> C.CDLL:
    path = FAST_LIB if fast else PARITY_LIB
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    for name, (res, args) in _SIG.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[path] = lib
    return lib


def _f(a):
    return a.ctypes.data_as(_F)


def default_params(**kw) -> Params:
    p = Params()
    load().or_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Oracle:
    """One oracle state; mirrors the reference SPHEngine step on the CPU."""

    def __init__(self, capacity: int, params: Params = None, fast: bool = False):
        self.lib = load(fast)
        self.h = self.lib.or_create(int(capacity))
        self.params = params or default_params()
        self.lib.or_set_params(self.h, C.byref(self.params))

    def __del__(self):
        try:
            self.lib.or_destroy(self.h)
        except Exception:
            pass

    def set_params(self, p: Params):
        self.params = p
        self.lib.or_set_params(self.h, C.byref(p))

    def set_particles(self, pos, vel=None, mass=None):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(mass, dtype=np.float32).reshape(n)
        self.lib.or_set_particles(self.h, n, _f(pos), None if vel is None else _f(vel),
                                  None if mass is None else _f(mass))

    @property
    def n(self) -> int:
        return int(self.lib.or_size(self.h))

    def step(self, dt: float = 0.0) -> float:
        return float(self.lib.or_step(self.h, float(dt)))

    def get(self, which: int) -> np.ndarray:
        n = self.n
        shape = (n, 3) if which in (OR_POS, OR_VEL, OR_ACCEL) else (n,)
        out = np.empty(shape, dtype=np.float32)
        self.lib.or_get(self.h, which, _f(out))
        return out

    def neighbors(self):
        n = self.n
        tot = self.lib.or_neighbor_total(self.h)
        offs = np.empty(n + 1, dtype=np.int64)
        idx = np.empty(max(tot, 1), dtype=np.int64)
        self.lib.or_get_neighbors(self.h, offs.ctypes.data_as(_I64), idx.ctypes.data_as(_I64))
        return offs, idx[:tot]

    def count_neighbors(self):
        """(sum of list lengths, longest list) of the current state, lists not stored."""
        mx = C.c_uint64(0)
        tot = self.lib.or_count_neighbors(self.h, C.byref(mx))
        return int(tot), int(mx.value)

    def sample_step0(self, sample):
        """Step-0 phases evaluated for the sampled particles only (and the densities of their
        neighbours, which their forces read): per sample the list length, FP32 density and
        acceleration, FP64 density and acceleration over the same lists, and the scale S_i."""
        sample = np.ascontiguousarray(sample, dtype=np.int64)
        k = sample.shape[0]
        out = dict(counts=np.empty(k, np.int64), rho=np.empty(k, np.float32),
                   acc=np.empty((k, 3), np.float32), rho64=np.empty(k, np.float64),
                   acc64=np.empty((k, 3), np.float64), scale=np.empty(k, np.float64))
        rc = self.lib.or_sample_step0(
            self.h, k, sample.ctypes.data_as(_I64), out["counts"].ctypes.data_as(_I64),
            _f(out["rho"]), _f(out["acc"]), out["rho64"].ctypes.data_as(_D),
            out["acc64"].ctypes.data_as(_D), out["scale"].ctypes.data_as(_D))
        if rc != 0:
            raise ValueError("sample index out of range")
        return out

    def sample_neighbors(self, sample):
        """Neighbour lists (CSR) of the sample after sample_step0."""
        sample = np.ascontiguousarray(sample, dtype=np.int64)
        k = sample.shape[0]
        offs = np.empty(k + 1, np.int64)
        self.lib.or_sample_neighbors(self.h, k, sample.ctypes.data_as(_I64),
                                     offs.ctypes.data_as(_I64), None)
        idx = np.empty(max(int(offs[-1]), 1), np.int64)
        self.lib.or_sample_neighbors(self.h, k, sample.ctypes.data_as(_I64),
                                     offs.ctypes.data_as(_I64), idx.ctypes.data_as(_I64))
        return offs, idx[: int(offs[-1])]


def cell_keys(pos: np.ndarray, cell: float):
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    n = pos.shape[0]
    coords = np.empty((n, 3), dtype=np.int32)
    keys = np.empty(n, dtype=np.uint64)
    load().or_cell_keys(n, _f(pos), float(cell), coords.ctypes.data_as(_I32),
                        keys.ctypes.data_as(_U64))
    return coords, keys


def W(r, h=0.02, formulation=0) -> float:
    r = np.ascontiguousarray(r, dtype=np.float32)
    return float(load().or_W(_f(r), float(h), formulation))


def gradW(r, h=0.02, formulation=0) -> np.ndarray:
    r = np.ascontiguousarray(r, dtype=np.float32)
    out = np.empty(3, dtype=np.float32)
    load().or_gradW(_f(r), float(h), formulation, _f(out))
    return out


def laplacianW(r, h=0.02, formulation=0) -> float:
    r = np.ascontiguousarray(r, dtype=np.float32)
    return float(load().or_laplacianW(_f(r), float(h), formulation))


def integrate_clamp(pos, vel, acc, dt, params: Params):
    pos = np.array(pos, dtype=np.float32).reshape(-1, 3)
    vel = np.array(vel, dtype=np.float32).reshape(-1, 3)
    acc = np.ascontiguousarray(acc, dtype=np.float32).reshape(-1, 3)
    load().or_integrate_clamp(pos.shape[0], _f(pos), _f(vel), _f(acc), float(dt), C.byref(params))
    return pos, vel


def density_accel_f64(pos, vel, mass, offs, idx, params: Params, with_acc=True):
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
    mass = np.ascontiguousarray(mass, dtype=np.float32)
    n = pos.shape[0]
    rho = np.empty(n, dtype=np.float64)
    acc = np.empty((n, 3), dtype=np.float64) if with_acc else None
    load().or64_density_accel(n, _f(pos), _f(vel), _f(mass),
                              np.ascontiguousarray(offs, np.int64).ctypes.data_as(_I64),
                              np.ascontiguousarray(idx, np.int64).ctypes.data_as(_I64),
                              C.byref(params), rho.ctypes.data_as(_D),
                              acc.ctypes.data_as(_D) if with_acc else None)
    return rho, acc


def accel_term_scale(pos, vel, mass, rho, offs, idx, params: Params):
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
    mass = np.ascontiguousarray(mass, dtype=np.float32)
    rho = np.ascontiguousarray(rho, dtype=np.float32)
    n = pos.shape[0]
    out = np.empty(n, dtype=np.float64)
    load().or64_accel_term_scale(n, _f(pos), _f(vel), _f(mass), _f(rho),
                                 np.ascontiguousarray(offs, np.int64).ctypes.data_as(_I64),
                                 np.ascontiguousarray(idx, np.int64).ctypes.data_as(_I64),
                                 C.byref(params), out.ctypes.data_as(_D))
    return out


def brute_neighbors(pos, r2):
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
    n = pos.shape[0]
    offs = np.empty(n + 1, dtype=np.int64)
    tot = load().or_brute_neighbors(n, _f(pos), float(r2), offs.ctypes.data_as(_I64), None, 0)
    idx = np.empty(max(tot, 1), dtype=np.int64)
    load().or_brute_neighbors(n, _f(pos), float(r2), offs.ctypes.data_as(_I64),
                              idx.ctypes.data_as(_I64), tot)
    return offs, idx[:tot]


# ---------------------------------------------------------------- surrogate network
def sur_cfg(h1=128, h2=128, h3=128, bf16=True, inv_r=25.0, inv_v=1.0, inv_h=50.0,
            w_norm=1.5, rho0=1000.0, rho_scale=0.01, msg_scale=1.0 / 64.0,
            acc_scale=10.0) -> SurCfg:
    return SurCfg(h1, h2, h3, int(bf16), inv_r, inv_v, inv_h, w_norm, rho0, rho_scale,
                  msg_scale, acc_scale)


def sur_nparams(cfg: SurCfg) -> int:
    return int(load().or_sur_nparams(C.byref(cfg)))


def bf16_round(x) -> np.ndarray:
    """Round-to-nearest-even to bfloat16, returned as float32 (vectorised restatement of
    or_bf16_round; the two are compared in the tests)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(nan, (u | 0x00400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32)


def _csr(offs, idx):
    return (np.ascontiguousarray(offs, np.int64), np.ascontiguousarray(idx, np.int64))


def sur_forward(cfg: SurCfg, params, pos, vel, rho, offs, idx, want_S=False):
    """Network output (n, 3) in normalised units; optionally S (n, h1) and edge counts."""
    params = np.ascontiguousarray(params, np.float32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    rho = np.ascontiguousarray(rho, np.float32)
    offs, idx = _csr(offs, idx)
    n = pos.shape[0]
    out = np.empty((n, 3), np.float64)
    S = np.empty((n, cfg.h1), np.float64) if want_S else None
    cnt = np.empty(n, np.int64) if want_S else None
    load().or_sur_forward(C.byref(cfg), _f(params), n, _f(pos), _f(vel), _f(rho),
                          offs.ctypes.data_as(_I64), idx.ctypes.data_as(_I64),
                          out.ctypes.data_as(_D), S.ctypes.data_as(_D) if want_S else None,
                          cnt.ctypes.data_as(_I64) if want_S else None)
    return (out, S, cnt) if want_S else out


def sur_loss_grad(cfg: SurCfg, params, pos, vel, rho, offs, idx, in_batch, target):
    """(loss, gradient) in FP64 over the particles with in_batch != 0."""
    params = np.ascontiguousarray(params, np.float32)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
    rho = np.ascontiguousarray(rho, np.float32)
    target = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
    mask = np.ascontiguousarray(in_batch, np.uint8)
    offs, idx = _csr(offs, idx)
    n = pos.shape[0]
    grad = np.empty(params.size, np.float64)
    loss = load().or_sur_loss_grad(C.byref(cfg), _f(params), n, _f(pos), _f(vel), _f(rho),
                                   offs.ctypes.data_as(_I64), idx.ctypes.data_as(_I64),
                                   mask.ctypes.data_as(C.POINTER(C.c_uint8)), _f(target),
                                   grad.ctypes.data_as(_D))
    return float(loss), grad


def sur_adam(w, m, v, g, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in place on float32 arrays w, m, v (FP64 arithmetic)."""
    g = np.ascontiguousarray(g, np.float32)
    load().or_sur_adam(w.size, _f(w), _f(m), _f(v), _f(g), lr, beta1, beta2, eps, int(t))
