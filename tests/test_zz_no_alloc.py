"""The allocation-free kinetic energy entry point (sphk_total_energy_noalloc) against
sphk_total_energy, through the C ABI: two calls in a row on one engine, the second with fewer
particles (the shared device scratch keeps the first call's partial sums), then after steps.

Bar: each term is the same FP32 (0.5 m)|v|^2 in both; the no-alloc sum is FP64 in a fixed
order, so it equals the FP64 sum of the FP32 terms to within one FP32 rounding, and differs
from sphk_total_energy's sequential FP32 sum by at most that sum's error bound, n 2^-24 of the
total (positive terms)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _terms_sum(vel, mass):
    v = vel.astype(np.float32)
    sp = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
    t = (np.float32(0.5) * mass.astype(np.float32)) * sp
    return float(np.sum(t.astype(np.float64)))


def _check(e, vel, mass, label):
    n = len(mass)
    got = e.total_energy_noalloc()
    again = e.total_energy_noalloc()
    ref = e.total_energy()
    exact = _terms_sum(vel, mass)
    print(f"{label}: n={n} noalloc {got!r} sequential {ref!r} fp64 {exact!r}", flush=True)
    assert got == again  # fixed order: the same bits every call
    assert abs(got - exact) <= np.spacing(np.float32(exact))
    assert abs(got - ref) <= n * 2.0 ** -24 * exact


def test_total_energy_noalloc_matches_total_energy(sphk):
    rng = np.random.default_rng(7)
    n1 = 300_000  # more particles than the first pass has threads: the strided loop runs
    pos = rng.uniform(-0.9, 0.9, (n1, 3)).astype(np.float32)
    vel = rng.normal(0, 0.7, (n1, 3)).astype(np.float32)
    mass = rng.uniform(0.5e-3, 2e-3, n1).astype(np.float32)
    e = sphk.Engine(n1, sphk.default_params())
    e.upload(pos, vel, mass)
    _check(e, vel, mass, "first call")
    n2 = 10_007  # second call, smaller input: fewer partial sums than the first call left
    e.upload(pos[:n2], vel[:n2], mass[:n2])
    assert e.size() == n2
    _check(e, vel[:n2], mass[:n2], "second call")
    # after steps the kernels own the velocities
    e.step(1e-4)
    e.step(1e-4)
    _check(e, e.velocities(), e.masses(), "after two steps")
