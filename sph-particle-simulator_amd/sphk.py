"""ctypes binding of libsphk.so, the C ABI declared in include/sphk.h.

This is the thin host-side path that tests, the benchmark and Python users take
to the HIP kernels.  It never falls back to CPU code: if libsphk.so is missing
or no GPU is visible, constructing an Engine raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphk.so")

SPHK_POSITIONS = 0
SPHK_VELOCITIES = 1
SPHK_DENSITIES = 2
SPHK_PRESSURES = 3
SPHK_ACCELERATIONS = 4
SPHK_MASSES = 5

ERRORS = {-1: "invalid argument", -2: "HIP runtime error", -3: "capacity exceeded",
          -4: "domain beyond the dense cell grid", -5: "no GPU device"}


class SphkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sphk error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(C.Structure):
    """sphk_params: the fields of SPHParameters (reference src/sph_engine.h:14-34)."""

    _fields_ = [
        ("rest_density", C.c_float), ("gas_constant", C.c_float), ("viscosity", C.c_float),
        ("smoothing_length", C.c_float), ("particle_mass", C.c_float), ("timestep", C.c_float),
        ("gravity", C.c_float), ("damping", C.c_float), ("CFL_factor", C.c_float),
        ("xmin", C.c_float), ("xmax", C.c_float), ("ymin", C.c_float), ("ymax", C.c_float),
        ("zmin", C.c_float), ("zmax", C.c_float), ("neighbor_search_radius", C.c_float),
        ("formulation", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("total_time", C.c_double), ("neighbor_search_time", C.c_double),
        ("density_computation_time", C.c_double), ("force_computation_time", C.c_double),
        ("integration_time", C.c_double), ("max_neighbors", C.c_uint64),
        ("total_neighbor_queries", C.c_uint64), ("total_neighbors", C.c_uint64),
        ("steps", C.c_uint64), ("timed_steps", C.c_uint64),
    ]


class SurrogateConfig(C.Structure):
    """sphk_surrogate_config: macro-step length, judge period/tolerance, feature scales."""

    _fields_ = [
        ("substeps", C.c_int32), ("judge_every", C.c_int32), ("tolerance", C.c_float),
        ("inv_v", C.c_float), ("rho_scale", C.c_float), ("msg_scale", C.c_float),
        ("acc_scale", C.c_float),
    ]


class TrainConfig(C.Structure):
    """sphk_train_config: snapshots of the engine's own plain run used for training."""

    _fields_ = [
        ("snapshots", C.c_uint64), ("iters_per_snapshot", C.c_uint32),
        ("batch_chunks", C.c_uint32), ("dt", C.c_float), ("lr", C.c_float),
        ("seed", C.c_uint64),
    ]


class SurrogateReport(C.Structure):
    _fields_ = [
        ("macro_steps", C.c_uint64), ("plain_macro_steps", C.c_uint64), ("judges", C.c_uint64),
        ("last_error", C.c_float), ("max_error", C.c_float), ("max_abs_error", C.c_float),
        ("fell_back", C.c_int32), ("fallback_at", C.c_uint64), ("error_bound", C.c_float),
    ]


_P = C.c_void_p
_F = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_U32 = C.POINTER(C.c_uint32)
_U64 = C.POINTER(C.c_uint64)
_I64 = C.POINTER(C.c_int64)
_SZ = C.c_size_t

# name -> argtypes; every function returns int status unless noted
SIGNATURES = {
    "sphk_default_params": [C.POINTER(Params)],
    "sphk_create": [_SZ, C.POINTER(_P)],
    "sphk_destroy": [_P],
    "sphk_set_params": [_P, C.POINTER(Params)],
    "sphk_get_params": [_P, C.POINTER(Params)],
    "sphk_upload": [_P, _SZ, _SZ, _F, _F, _F],
    "sphk_size": [_P, C.POINTER(_SZ)],
    "sphk_clear": [_P],
    "sphk_step": [_P, C.c_float],
    "sphk_run_steps": [_P, _SZ, C.c_int],
    "sphk_download": [_P, C.c_int, _F, _SZ],
    "sphk_render_pack": [_P, C.c_void_p, C.c_void_p, _SZ],
    "sphk_total_energy_noalloc": [_P, _F],
    "sphk_get_time": [_P, _F],
    "sphk_get_step_count": [_P, _U64],
    "sphk_total_mass": [_P, _F],
    "sphk_total_energy": [_P, _F],
    "sphk_conservation_errors": [_P, _F, _F],
    "sphk_set_timing": [_P, C.c_int],
    "sphk_get_stats": [_P, C.POINTER(Stats)],
    "sphk_reset_stats": [_P],
    "sphk_snapshot_save": [_P],
    "sphk_snapshot_restore": [_P],
    "sphk_synchronize": [_P],
    "sphk_debug_cell_coords": [_P, _I32, _SZ],
    "sphk_debug_cell_keys": [_P, _U64, _SZ],
    "sphk_debug_sorted": [_P, _U32, _U32, _SZ],
    "sphk_debug_neighbors": [_P, _I64, _I64, _SZ, C.POINTER(_SZ)],
    "sphk_debug_grid": [_P, _I32, _I32, _U64],
    "sphk_debug_walk_words": [_P, _U32, _U32, _U32, _U32, _SZ, C.POINTER(_SZ), C.POINTER(_SZ)],
    "sphk_set_scenes": [_P, _SZ, _U64, _F],
    "sphk_scene_count": [_P, C.POINTER(_SZ)],
    "sphk_set_graph": [_P, C.c_int],
    "sphk_set_walk_shapes": [_P, C.c_int, C.c_int],
    "sphk_get_walk_shapes": [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "sphk_walk_shape_count": [],
    "sphk_walk_shape_name": [C.c_int],
    "sphk_surrogate_default_config": [C.POINTER(SurrogateConfig)],
    "sphk_surrogate_param_count": [C.POINTER(_SZ)],
    "sphk_surrogate_init": [_P, C.POINTER(SurrogateConfig), C.c_uint64],
    "sphk_surrogate_get_config": [_P, C.POINTER(SurrogateConfig)],
    "sphk_surrogate_get_params": [_P, _F, _SZ],
    "sphk_surrogate_set_params": [_P, _F, _SZ],
    "sphk_surrogate_predict": [_P, _F, _SZ],
    "sphk_surrogate_targets": [_P, C.c_float, _F, _SZ],
    "sphk_surrogate_loss_grad": [_P, _F, _SZ, C.c_uint64, C.c_uint32, _F, _F, _SZ,
                                 C.POINTER(C.c_uint8)],
    "sphk_surrogate_adam": [_P, _F, _SZ, C.c_float],
    "sphk_surrogate_train": [_P, C.POINTER(TrainConfig), _F, _SZ],
    "sphk_surrogate_macro_step": [_P, C.c_float],
    "sphk_surrogate_run": [_P, C.c_uint64, C.c_float, C.POINTER(SurrogateReport)],
    "sphk_surrogate_densities": [_P, _F, _SZ],
    "sphk_last_error": [C.c_char_p, _SZ],
    "sphk_version": [],
}

_STRING_RESULTS = {"sphk_version", "sphk_walk_shape_name"}

_lib: Optional[C.CDLL] = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads libsphk.so and declares every exported function."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name in _STRING_RESULTS else C.c_int
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def last_error() -> str:
    lib = load()
    buf = C.create_string_buffer(1024)
    lib.sphk_last_error(buf, 1024)
    return buf.value.decode()


def check(rc: int) -> None:
    if rc != 0:
        raise SphkError(rc, last_error())


def default_params(**overrides) -> Params:
    p = Params()
    check(load().sphk_default_params(C.byref(p)))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


class Engine:
    """One libsphk engine (one HIP stream on the current device)."""

    def __init__(self, capacity: int = 1_000_000, params: Optional[Params] = None):
        self.lib = load()
        self.h = _P()
        check(self.lib.sphk_create(int(capacity), C.byref(self.h)))
        self.capacity = int(capacity)
        if params is not None:
            self.set_params(params)

    def close(self):
        if self.h:
            self.lib.sphk_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: Params):
        check(self.lib.sphk_set_params(self.h, C.byref(p)))

    def params(self) -> Params:
        p = Params()
        check(self.lib.sphk_get_params(self.h, C.byref(p)))
        return p

    def upload(self, pos, vel=None, mass=None, offset: int = 0):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        mass = None if mass is None else np.ascontiguousarray(mass, dtype=np.float32).reshape(n)
        check(self.lib.sphk_upload(self.h, offset, n, _ptr(pos, C.c_float), _ptr(vel, C.c_float),
                                   _ptr(mass, C.c_float)))

    def size(self) -> int:
        n = _SZ()
        check(self.lib.sphk_size(self.h, C.byref(n)))
        return n.value

    def clear(self):
        check(self.lib.sphk_clear(self.h))

    def step(self, dt: float = 0.0):
        check(self.lib.sphk_step(self.h, float(dt)))

    def run_steps(self, k: int, adaptive: bool = True):
        check(self.lib.sphk_run_steps(self.h, int(k), int(bool(adaptive))))

    def synchronize(self):
        check(self.lib.sphk_synchronize(self.h))

    def _download(self, which: int, shape) -> np.ndarray:
        out = np.empty(shape, dtype=np.float32)
        check(self.lib.sphk_download(self.h, which, _ptr(out, C.c_float), out.size))
        return out

    def render_pack(self, out_ptr: int, count: int, colors_ptr: Optional[int] = None):
        """Writes the renderer's 9-float instance rows (position, velocity, colour) of every
        particle into the device buffer at out_ptr (count floats, at least 9 n)."""
        check(self.lib.sphk_render_pack(self.h, colors_ptr, out_ptr, count))

    def positions(self) -> np.ndarray:
        return self._download(SPHK_POSITIONS, (self.size(), 3))

    def velocities(self) -> np.ndarray:
        return self._download(SPHK_VELOCITIES, (self.size(), 3))

    def accelerations(self) -> np.ndarray:
        return self._download(SPHK_ACCELERATIONS, (self.size(), 3))

    def masses(self) -> np.ndarray:
        return self._download(SPHK_MASSES, (self.size(),))

    def densities(self) -> np.ndarray:
        """Capacity-length, like SPHEngine::get_densities."""
        return self._download(SPHK_DENSITIES, (self.capacity,))

    def pressures(self) -> np.ndarray:
        return self._download(SPHK_PRESSURES, (self.capacity,))

    def time(self) -> float:
        t = C.c_float()
        check(self.lib.sphk_get_time(self.h, C.byref(t)))
        return t.value

    def step_count(self) -> int:
        v = C.c_uint64()
        check(self.lib.sphk_get_step_count(self.h, C.byref(v)))
        return v.value

    def total_mass(self) -> float:
        v = C.c_float()
        check(self.lib.sphk_total_mass(self.h, C.byref(v)))
        return v.value

    def total_energy_noalloc(self) -> float:
        out = C.c_float()
        check(self.lib.sphk_total_energy_noalloc(self.h, C.byref(out)))
        return out.value

    def total_energy(self) -> float:
        v = C.c_float()
        check(self.lib.sphk_total_energy(self.h, C.byref(v)))
        return v.value

    def conservation_errors(self):
        a, b = C.c_float(), C.c_float()
        check(self.lib.sphk_conservation_errors(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_timing(self, on: bool):
        check(self.lib.sphk_set_timing(self.h, int(bool(on))))

    def stats(self) -> Stats:
        s = Stats()
        check(self.lib.sphk_get_stats(self.h, C.byref(s)))
        return s

    def reset_stats(self):
        check(self.lib.sphk_reset_stats(self.h))

    def snapshot_save(self):
        check(self.lib.sphk_snapshot_save(self.h))

    def snapshot_restore(self):
        check(self.lib.sphk_snapshot_restore(self.h))

    def set_scenes(self, offsets, bounds=None):
        """Independent scenes: particles [offsets[k], offsets[k+1]) with walls bounds[k]."""
        if offsets is None or len(offsets) == 0:
            check(self.lib.sphk_set_scenes(self.h, 0, None, None))
            return
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        b = None if bounds is None else np.ascontiguousarray(bounds, dtype=np.float32).reshape(-1, 6)
        check(self.lib.sphk_set_scenes(self.h, off.size - 1, _ptr(off, C.c_uint64),
                                       _ptr(b, C.c_float)))

    def scene_count(self) -> int:
        v = _SZ()
        check(self.lib.sphk_scene_count(self.h, C.byref(v)))
        return v.value

    def set_graph(self, on: bool):
        check(self.lib.sphk_set_graph(self.h, int(bool(on))))

    def set_walk_shapes(self, density: int, force: int):
        check(self.lib.sphk_set_walk_shapes(self.h, int(density), int(force)))

    def walk_shapes(self):
        a, b = C.c_int(), C.c_int()
        check(self.lib.sphk_get_walk_shapes(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- parity/debug ----
    def cell_coords(self) -> np.ndarray:
        n = self.size()
        out = np.empty((n, 3), dtype=np.int32)
        check(self.lib.sphk_debug_cell_coords(self.h, _ptr(out, C.c_int32), n))
        return out

    def cell_keys(self) -> np.ndarray:
        n = self.size()
        out = np.empty(n, dtype=np.uint64)
        check(self.lib.sphk_debug_cell_keys(self.h, _ptr(out, C.c_uint64), n))
        return out

    def sorted_order(self):
        n = self.size()
        order = np.empty(n, dtype=np.uint32)
        cells = np.empty(n, dtype=np.uint32)
        check(self.lib.sphk_debug_sorted(self.h, _ptr(order, C.c_uint32), _ptr(cells, C.c_uint32), n))
        return order, cells

    def neighbors(self):
        n = self.size()
        offs = np.empty(n + 1, dtype=np.int64)
        total = _SZ()
        check(self.lib.sphk_debug_neighbors(self.h, _ptr(offs, C.c_int64), None, 0, C.byref(total)))
        idx = np.empty(max(total.value, 1), dtype=np.int64)
        check(self.lib.sphk_debug_neighbors(self.h, _ptr(offs, C.c_int64), _ptr(idx, C.c_int64),
                                            idx.size, C.byref(total)))
        return offs, idx[: total.value]

    def walk_words(self):
        """The production density walk's stored candidate words for the current state:
        (items, nwords, wj0 [chunks, W], mask [chunks, W, 64]), W the words the engine stores
        per chunk (chunks with more words store their first W)."""
        nc = _SZ()
        stored = _SZ()
        check(self.lib.sphk_debug_walk_words(self.h, None, None, None, None, 0, C.byref(nc),
                                             C.byref(stored)))
        c = nc.value
        items = np.empty(max(c, 1), dtype=np.uint32)
        nw = np.empty(max(c, 1), dtype=np.uint32)
        wj0 = np.zeros((max(c, 1), 256), dtype=np.uint32)
        mask = np.zeros((max(c, 1), 256, 64), dtype=np.uint32)
        check(self.lib.sphk_debug_walk_words(self.h, _ptr(items, C.c_uint32), _ptr(nw, C.c_uint32),
                                             _ptr(wj0, C.c_uint32), _ptr(mask, C.c_uint32),
                                             c, C.byref(nc), None))
        w = stored.value
        return items[:c], nw[:c], wj0[:c, :w], mask[:c, :w]

    def grid(self):
        o = np.empty(3, dtype=np.int32)
        d = np.empty(3, dtype=np.int32)
        g = C.c_uint64()
        check(self.lib.sphk_debug_grid(self.h, _ptr(o, C.c_int32), _ptr(d, C.c_int32), C.byref(g)))
        return o, d, g.value


    # ---- learned macro-step surrogate ----
    def surrogate_init(self, seed: int = 0, config: Optional[SurrogateConfig] = None):
        check(self.lib.sphk_surrogate_init(self.h, C.byref(config) if config else None,
                                           C.c_uint64(seed)))

    def surrogate_config(self) -> SurrogateConfig:
        c = SurrogateConfig()
        check(self.lib.sphk_surrogate_get_config(self.h, C.byref(c)))
        return c

    def surrogate_params(self) -> np.ndarray:
        out = np.empty(surrogate_param_count(), np.float32)
        check(self.lib.sphk_surrogate_get_params(self.h, _ptr(out, C.c_float), out.size))
        return out

    def surrogate_set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        check(self.lib.sphk_surrogate_set_params(self.h, _ptr(p, C.c_float), p.size))

    def surrogate_predict(self) -> np.ndarray:
        n = self.size()
        out = np.empty((n, 3), np.float32)
        check(self.lib.sphk_surrogate_predict(self.h, _ptr(out, C.c_float), out.size))
        return out

    def surrogate_targets(self, dt: float) -> np.ndarray:
        n = self.size()
        out = np.empty((n, 3), np.float32)
        check(self.lib.sphk_surrogate_targets(self.h, float(dt), _ptr(out, C.c_float), out.size))
        return out

    def surrogate_loss_grad(self, target, batch_seed: int = 0, batch_chunks: int = 1 << 30):
        """(loss, gradient, in_batch mask) for targets in m/s^2 (caller order)."""
        n = self.size()
        t = np.ascontiguousarray(target, np.float32).reshape(n, 3)
        loss = C.c_float()
        grad = np.empty(surrogate_param_count(), np.float32)
        mask = np.empty(n, np.uint8)
        check(self.lib.sphk_surrogate_loss_grad(self.h, _ptr(t, C.c_float), t.size,
                                                C.c_uint64(batch_seed), int(batch_chunks),
                                                C.byref(loss), _ptr(grad, C.c_float), grad.size,
                                                _ptr(mask, C.c_uint8)))
        return loss.value, grad, mask

    def surrogate_adam(self, grad, lr: float = 1e-3):
        g = np.ascontiguousarray(grad, np.float32)
        check(self.lib.sphk_surrogate_adam(self.h, _ptr(g, C.c_float), g.size, float(lr)))

    def surrogate_train(self, snapshots: int, dt: float, batch_chunks: int = 1024,
                        iters_per_snapshot: int = 1, lr: float = 1e-3, seed: int = 0):
        """Trains on the engine's own trajectory; returns the per-snapshot mean loss."""
        tc = TrainConfig(int(snapshots), int(iters_per_snapshot), int(batch_chunks), float(dt),
                         float(lr), int(seed))
        curve = np.zeros(int(snapshots), np.float32)
        check(self.lib.sphk_surrogate_train(self.h, C.byref(tc), _ptr(curve, C.c_float),
                                            curve.size))
        return curve

    def surrogate_macro_step(self, dt: float):
        check(self.lib.sphk_surrogate_macro_step(self.h, float(dt)))

    def surrogate_run(self, macro_steps: int, dt: float) -> SurrogateReport:
        rep = SurrogateReport()
        check(self.lib.sphk_surrogate_run(self.h, C.c_uint64(macro_steps), float(dt),
                                          C.byref(rep)))
        return rep

    def surrogate_densities(self) -> np.ndarray:
        out = np.empty(self.size(), np.float32)
        check(self.lib.sphk_surrogate_densities(self.h, _ptr(out, C.c_float), out.size))
        return out


def surrogate_param_count() -> int:
    n = _SZ()
    check(load().sphk_surrogate_param_count(C.byref(n)))
    return n.value


def surrogate_default_config(**overrides) -> SurrogateConfig:
    c = SurrogateConfig()
    check(load().sphk_surrogate_default_config(C.byref(c)))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def version() -> str:
    return load().sphk_version().decode()


def walk_shape_names():
    lib = load()
    return [lib.sphk_walk_shape_name(i).decode() for i in range(lib.sphk_walk_shape_count())]
