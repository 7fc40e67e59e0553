"""Seeded synthetic scenes for the SPH step (dam break, drop into pool, batched ensemble).

Frames: the scene descriptions are z-up; the engine follows the reference, whose
gravity acts on y (equations::compute_external_forces, src/sph_engine.cpp:434-436).
Every scene is therefore built directly in the engine frame (x, y_up, z), i.e. the
description's (x, y, z) maps to (x, z, y).  Lattices are enumerated x-major
(i over x, then y, then z), the loop order of create_fluid_block
(src/particle.cpp:166-188).

Randomness is counter-based (splitmix64 of seed and a per-value counter), so the
same scene is produced bit-for-bit on any host regardless of thread count.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

SPACING = 0.01
GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, counter: np.ndarray) -> np.ndarray:
    """U[0,1) doubles from (seed, counter)."""
    with np.errstate(over="ignore"):
        base = splitmix64(np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
        u = splitmix64(np.asarray(counter, dtype=np.uint64) ^ base)
    return (u >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, counter: np.ndarray) -> np.ndarray:
    """Standard normals by Box-Muller over two counter streams."""
    c = np.asarray(counter, dtype=np.uint64)
    u1 = uniform01(seed, c * np.uint64(2))
    u2 = uniform01(seed, c * np.uint64(2) + np.uint64(1))
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclasses.dataclass
class Scene:
    name: str
    pos: np.ndarray            # (N,3) float32, engine frame
    vel: np.ndarray            # (N,3) float32
    mass: float                # per-particle mass (uniform)
    bounds: tuple              # (xmin, xmax, ymin, ymax, zmin, zmax), engine frame
    steps: int
    dt: float = 0.001
    smoothing_length: float = 0.02
    rest_density: float = 1000.0
    gas_constant: float = 2000.0
    viscosity: float = 0.001
    description: str = ""

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def params(self, sphk_mod, formulation: int = 0):
        p = sphk_mod.default_params()
        p.rest_density = self.rest_density
        p.gas_constant = self.gas_constant
        p.viscosity = self.viscosity
        p.smoothing_length = self.smoothing_length
        p.neighbor_search_radius = 2.0 * self.smoothing_length
        p.particle_mass = self.mass
        p.timestep = self.dt
        p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = self.bounds
        p.formulation = formulation
        return p

    def oracle_params(self, oracle_mod, formulation: int = 0):
        p = oracle_mod.default_params()
        p.rest_density = self.rest_density
        p.gas_constant = self.gas_constant
        p.viscosity = self.viscosity
        p.smoothing_length = self.smoothing_length
        p.neighbor_search_radius = float(np.float32(2.0) * np.float32(self.smoothing_length))
        p.particle_mass = self.mass
        p.timestep = self.dt
        p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = self.bounds
        p.formulation = formulation
        return p


def lattice(nx: int, ny: int, nz: int, spacing: float = SPACING, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """(nx*ny*nz, 3) float64 lattice, x-major (i over x outermost), engine frame."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    p = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64) * spacing
    return p + np.asarray(origin, dtype=np.float64)


def uniform_jitter(seed: int, n: int, amp: float, stream: int = 0) -> np.ndarray:
    c = np.arange(3 * n, dtype=np.uint64) + np.uint64(stream << 40)
    return ((2.0 * uniform01(seed, c) - 1.0) * amp).reshape(n, 3)


def dam_break_block(nx: int, ny_up: int, nz: int, box_xyz_up, seed: int = 0, steps: int = 100,
                    name: str = "dam_break") -> Scene:
    """Fluid block nx x ny_up(vertical) x nz at the box corner, uniform jitter +-0.05 spacing.

    box_xyz_up = (Lx, Ly, Lz_up) in the z-up description; returned in the engine frame."""
    base = lattice(nx, ny_up, nz)
    pos = base + uniform_jitter(seed, base.shape[0], 0.05 * SPACING)
    lx, ly, lz_up = box_xyz_up
    return Scene(name=name, pos=pos.astype(np.float32), vel=np.zeros_like(pos, dtype=np.float32),
                 mass=0.001, bounds=(0.0, lx, 0.0, lz_up, 0.0, ly), steps=steps)


def config0(seed: int = 0) -> Scene:
    """Dam break, 10,000 particles: 25 x 20 x 20 (z-up) lattice in a 1 x 0.6 x 0.4 box, 100 steps."""
    s = dam_break_block(25, 20, 20, (1.0, 0.6, 0.4), seed, 100, "dam_break_10k")
    s.description = "25x20x20 lattice, spacing 0.01, jitter +-5e-4, box 1x0.6x0.4 (z-up)"
    return s


def config1(seed: int = 0) -> Scene:
    """Dam break, 1,000,000 particles: 100^3 lattice in a 3 x 1.5 x 1.2 box, 1000 steps."""
    s = dam_break_block(100, 100, 100, (3.0, 1.5, 1.2), seed, 1000, "dam_break_1M")
    s.description = "100^3 lattice, spacing 0.01, jitter +-5e-4, box 3x1.5x1.2 (z-up)"
    return s


def small_block(nx: int, ny: int, nz: int, seed: int = 0, box=(1.0, 0.6, 0.4)) -> Scene:
    return dam_break_block(nx, ny, nz, box, seed, 10, f"block_{nx}x{ny}x{nz}")


def config2(seed: int = 0) -> Scene:
    """Drop into pool: 400 x 400 x 60 (z-up) pool + lattice sphere r=0.15 centred 0.4 m above the
    pool surface (z = 0.6 + 0.4), falling at 2 m/s; Gaussian jitter sigma = 0.1 spacing; box
    4.2 x 4.2 x 1.5; 500 steps.  The pool lattice starts at (0.1, 0.1, 0) so pool and sphere are
    centred on the box."""
    pool = lattice(400, 60, 400, origin=(0.1, 0.0, 0.1))  # engine frame: y is up
    r = 15
    a = np.arange(-r, r + 1)
    ii, jj, kk = np.meshgrid(a, a, a, indexing="ij")
    inside = (ii ** 2 + jj ** 2 + kk ** 2) <= r * r
    off = np.stack([ii[inside], jj[inside], kk[inside]], axis=1).astype(np.float64) * SPACING
    centre = np.array([2.1, 1.0, 2.1])  # engine frame (x, y_up, z)
    sphere = centre + off
    pos = np.concatenate([pool, sphere], axis=0)
    n = pos.shape[0]
    c = np.arange(3 * n, dtype=np.uint64) + np.uint64(1 << 41)
    pos = pos + (0.1 * SPACING) * normal(seed, c).reshape(n, 3)
    vel = np.zeros((n, 3), dtype=np.float32)
    vel[pool.shape[0]:, 1] = -2.0
    return Scene(name="drop_into_pool_9.6M", pos=pos.astype(np.float32), vel=vel, mass=0.001,
                 bounds=(0.0, 4.2, 0.0, 1.5, 0.0, 4.2), steps=500,
                 description="400x400x60 pool + r=0.15 sphere (14,147 pts) at z=1.0, v=-2 m/s")


def config3(seed: int = 0, n_scenes: int = 64, scale: float = 1.0) -> List[Scene]:
    """Batched ensemble: 64 dam-break scenes, N_s log-uniform in [10k, 500k] (times `scale`),
    block aspect ratios in [0.5, 2], box 2-4x the block extent per axis, 500 steps each."""
    scenes = []
    for s in range(n_scenes):
        c = np.uint64(1 << 42) + np.uint64(16 * s) + np.arange(8, dtype=np.uint64)
        u = uniform01(seed, c)
        ns = int(round(math.exp(math.log(1e4) + u[0] * (math.log(5e5) - math.log(1e4)))) * scale)
        ns = max(ns, 8)
        asp = 0.5 + 1.5 * u[1:4]
        base = (ns / float(np.prod(asp))) ** (1.0 / 3.0)
        dims = [max(1, int(round(base * a))) for a in asp]
        box_f = 2.0 + 2.0 * u[4:7]
        ext = [d * SPACING for d in dims]
        box = tuple(float(box_f[d] * ext[d]) for d in range(3))
        blk = lattice(dims[0], dims[1], dims[2])
        pos = blk + uniform_jitter(seed, blk.shape[0], 0.05 * SPACING, stream=64 + s)
        scenes.append(Scene(name=f"ensemble_{s}", pos=pos.astype(np.float32),
                            vel=np.zeros_like(pos, dtype=np.float32), mass=0.001,
                            bounds=(0.0, box[0], 0.0, box[1], 0.0, box[2]), steps=500,
                            description=f"dims {dims} box {box}"))
    return scenes


CONFIGS = {0: config0, 1: config1, 2: config2}
