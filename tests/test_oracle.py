"""Pins the CPU oracle against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors for this path, and cannot be
compiled here (its sources need GLM, which is neither vendored nor installed), so each
oracle function is pinned by closed forms, brute force, lattice invariants or special
cases chosen so that a plausible slip (a dropped term, a wrong sign or index, a
swapped operand, a 27- vs 125-cell scan) fails at least one test.
"""
import math

import numpy as np
import pytest

from conftest import R2

H = np.float32(0.02)
SIGMA = np.float32(1.0) / (np.float32(math.pi) * H * H * H)


# ---------------------------------------------------------------- kernel closed forms
def test_W_closed_forms(oracle):
    # W(0) = sigma*2/3, W(q=1) = sigma/6, W(q=2) = 0, zero beyond the support
    assert oracle.W([0, 0, 0]) == pytest.approx(float(SIGMA) * 2 / 3, rel=1e-6)
    assert oracle.W([0.02, 0, 0]) == pytest.approx(float(SIGMA) / 6, rel=1e-6)
    assert oracle.W([0, 0.04, 0]) == 0.0
    assert oracle.W([0, 0, 0.05]) == 0.0
    # the reference sigma: 1/(pi h^3) = 39788.74 in FP32 (not normalised)
    assert float(SIGMA) == pytest.approx(39788.74, rel=1e-6)
    # the consistent formulation scales sigma by 3/2
    assert oracle.W([0, 0, 0], formulation=1) == pytest.approx(1.5 * oracle.W([0, 0, 0]), rel=1e-6)


def test_W_matches_piecewise_cubic_in_fp64(oracle):
    # sigma*(2/3 - q^2 + q^3/2) on [0,1], sigma*(2-q)^3/6 on (1,2]; radial direction irrelevant
    rng = np.random.default_rng(1)
    for _ in range(200):
        q = rng.uniform(0, 2.2)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = (q * 0.02 * d).astype(np.float32)
        qq = float(np.linalg.norm(r.astype(np.float64))) / 0.02
        s = 1.0 / (math.pi * 0.02 ** 3)
        exp = s * (2 / 3 - qq * qq + 0.5 * qq ** 3) if qq <= 1 else (s * (2 - qq) ** 3 / 6 if qq <= 2 else 0)
        assert oracle.W(r) == pytest.approx(exp, rel=2e-5, abs=1e-2)


def test_W_continuity_and_integral(oracle):
    # continuous value and slope at q = 1
    eps = 1e-4
    a = oracle.W([0.02 * (1 - eps), 0, 0])
    b = oracle.W([0.02 * (1 + eps), 0, 0])
    assert abs(a - b) < 1e-3 * oracle.W([0.02, 0, 0])
    # int W dV = 2/3 for the reference normalisation (1 for the consistent one)
    r = np.linspace(0, 0.04, 4001)
    w = np.array([oracle.W([x, 0, 0]) for x in r], dtype=np.float64)
    integral = np.trapezoid(4 * math.pi * r * r * w, r)
    assert integral == pytest.approx(2 / 3, rel=2e-4)
    w1 = np.array([oracle.W([x, 0, 0], formulation=1) for x in r], dtype=np.float64)
    assert np.trapezoid(4 * math.pi * r * r * w1, r) == pytest.approx(1.0, rel=2e-4)


def test_gradW_is_derivative_of_W_along_r(oracle):
    rng = np.random.default_rng(2)
    for _ in range(100):
        q = rng.uniform(0.05, 1.95)
        if abs(q - 1) < 0.02:
            continue
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = q * 0.02 * d
        g = oracle.gradW(r.astype(np.float32)).astype(np.float64)
        # finite difference of W along r-hat, in FP64 from the closed form
        s = 1.0 / (math.pi * 0.02 ** 3)

        def w(qq):
            return s * (2 / 3 - qq * qq + 0.5 * qq ** 3) if qq <= 1 else s * (2 - qq) ** 3 / 6

        dq = 1e-6
        dwdr = (w(q + dq) - w(q - dq)) / (2 * dq) / 0.02
        np.testing.assert_allclose(g, dwdr * d, rtol=2e-4, atol=1e-2 * abs(dwdr))
        # gradient points towards the centre (W decreasing): sign check
        assert np.dot(g, d) < 0


def test_gradW_zero_cases(oracle):
    assert np.all(oracle.gradW([0, 0, 0]) == 0)
    assert np.all(oracle.gradW([5e-7, 0, 0]) == 0)  # |r| < 1e-6 guard
    assert np.all(oracle.gradW([0.05, 0, 0]) == 0)  # outside support


def test_laplacian_closed_form(oracle):
    s = 1.0 / (math.pi * 0.02 ** 3)
    for q in (0.0, 0.3, 0.9, 1.2, 1.7, 2.0, 2.5):
        exp = s * (-2 + 3 * q) / 4e-4 if q <= 1 else (s * (2 - q) / 4e-4 if q <= 2 else 0.0)
        assert oracle.laplacianW([0, q * 0.02, 0]) == pytest.approx(exp, rel=1e-5, abs=1e-3)


# ---------------------------------------------------------------- hash keys
def test_hash_key_packing(oracle):
    lib = oracle.load()
    assert lib.or_hash_position(0, 0, 0) == 0
    assert lib.or_hash_position(1, 2, 3) == (1 << 42) | (2 << 21) | 3
    # negative coordinates are masked to 21 bits
    assert lib.or_hash_position(-1, 0, 0) == (0x1FFFFF << 42)
    assert lib.or_hash_position(0, -2, 5) == (0x1FFFFE << 21) | 5
    # aliasing at 2^21
    assert lib.or_hash_position(1 << 21, 0, 0) == 0


def test_cell_coords_rounding_and_specials(oracle):
    pos = np.array([
        [0.0, 0.03999999, 0.04],       # 0.04f*25 rounds to 1.0
        [-1e-9, -0.04, -0.0400001],    # negative coordinates
        [np.nan, 0.5, 0.5],            # NaN -> INT_MIN (x86 conversion)
        [1e30, 0.0, 0.0],              # out of int range -> INT_MIN
    ], dtype=np.float32)
    coords, keys = oracle.cell_keys(pos, 0.04)
    assert coords[0].tolist() == [0, 0, 1]
    assert coords[1].tolist() == [-1, -1, -2]
    assert coords[2, 0] == np.iinfo(np.int32).min and coords[2, 1] == 12
    assert coords[3, 0] == np.iinfo(np.int32).min
    # (p * 25.0f) in FP32, floored, for random values: exact integer maths
    rng = np.random.default_rng(3)
    p = rng.uniform(-3, 3, size=(1000, 3)).astype(np.float32)
    c, _ = oracle.cell_keys(p, 0.04)
    inv = np.float32(1.0) / np.float32(0.04)
    assert inv == np.float32(25.0)
    np.testing.assert_array_equal(c, np.floor(p * inv).astype(np.int32))


# ---------------------------------------------------------------- neighbour search
def _oracle_step0(oracle, pos, params=None, vel=None):
    o = oracle.Oracle(len(pos), params or oracle.default_params())
    o.set_particles(pos, vel)
    o.lib.or_update_neighbor_lists(o.h)
    return o


def test_neighbors_equal_brute_force(oracle):
    rng = np.random.default_rng(4)
    pos = rng.uniform(-0.1, 0.2, size=(1500, 3)).astype(np.float32)
    o = _oracle_step0(oracle, pos)
    offs, idx = o.neighbors()
    boffs, bidx = oracle.brute_neighbors(pos, R2)
    # same sets (brute force lists ascending j; the hash visits cells in (dx,dy,dz) order)
    for i in range(len(pos)):
        assert sorted(idx[offs[i]:offs[i + 1]]) == list(bidx[boffs[i]:boffs[i + 1]])
    # symmetric relation, self included
    pairs = set()
    for i in range(len(pos)):
        for j in idx[offs[i]:offs[i + 1]]:
            pairs.add((i, int(j)))
    assert all((j, i) in pairs for (i, j) in pairs)
    assert all((i, i) in pairs for i in range(len(pos)))


def test_sparse_scene_neighbors_equal_brute_force(oracle):
    # about as many occupied cells as particles (the cell table must grow past its first
    # size, set for dense scenes): same neighbour sets as brute force
    rng = np.random.default_rng(9)
    pos = rng.uniform(-2.0, 2.0, size=(4000, 3)).astype(np.float32)
    pos[: 400] = rng.uniform(0.0, 0.06, size=(400, 3)).astype(np.float32)  # a dense corner
    o = _oracle_step0(oracle, pos)
    offs, idx = o.neighbors()
    boffs, bidx = oracle.brute_neighbors(pos, R2)
    for i in range(len(pos)):
        assert sorted(idx[offs[i]:offs[i + 1]]) == list(bidx[boffs[i]:boffs[i + 1]])


def test_neighbor_order_is_cell_lexicographic(oracle):
    rng = np.random.default_rng(5)
    pos = rng.uniform(0, 0.16, size=(600, 3)).astype(np.float32)
    o = _oracle_step0(oracle, pos)
    offs, idx = o.neighbors()
    coords, _ = oracle.cell_keys(pos, 0.04)
    for i in range(0, len(pos), 7):
        lst = idx[offs[i]:offs[i + 1]]
        key = [tuple(coords[j] - coords[i]) + (int(j),) for j in lst]
        assert key == sorted(key)


def test_perfect_lattice_bulk(oracle):
    # unjittered 17^3 lattice at 0.01: the centre has 257 neighbours (251 strictly inside 4
    # spacings + the 6 axial points at d = 4 spacings) and bulk density 666.72 (FP64: 666.7195)
    n = 17
    p = np.stack(np.meshgrid(*(np.arange(n),) * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (p * 0.01).astype(np.float32)
    params = oracle.default_params()
    o = oracle.Oracle(len(pos), params)
    o.set_particles(pos)
    o.lib.or_update_neighbor_lists(o.h)
    o.lib.or_compute_densities(o.h)
    offs, idx = o.neighbors()
    c = (n // 2) * n * n + (n // 2) * n + n // 2
    cnt = offs[c + 1] - offs[c]
    lst = idx[offs[c]:offs[c + 1]]
    d2 = ((p[lst] - p[c]) ** 2).sum(1)
    assert (d2 < 16).sum() == 251
    assert cnt in (251, 257)  # the six ties at d = R pass or fail together by symmetry of rounding
    rho = o.get(oracle.OR_DENSITY)
    mass = np.full(len(pos), 0.001, dtype=np.float32)
    r64, _ = oracle.density_accel_f64(pos, np.zeros_like(pos), mass, offs, idx, params, with_acc=False)
    assert r64[c] == pytest.approx(666.7195, abs=2e-3)
    assert rho[c] == pytest.approx(r64[c], rel=1e-6)


def test_isolated_particle(oracle):
    params = oracle.default_params()
    o = oracle.Oracle(4, params)
    o.set_particles(np.array([[0.1, 0.2, 0.3]], np.float32))
    o.step(0.001)
    rho = o.get(oracle.OR_DENSITY)
    assert rho[0] == pytest.approx(0.001 * 39788.74 * 2 / 3, rel=1e-6)  # 26.52583
    acc = o.get(oracle.OR_ACCEL)
    np.testing.assert_array_equal(acc[0], np.array([0, -9.81, 0], np.float32))


def test_edge_pair_two_cells_apart_is_found(oracle):
    # A pair whose FP32 cell coordinates differ by 2 yet passes d^2 <= R^2.  0.04f*25 rounds up
    # to exactly 1.0, so b = 0.04f sits in cell 1, while a tiny negative a sits in cell -1; b - a
    # rounds back to 0.04f and d^2 == R^2.  A 27-cell scan misses this pair; the reference's
    # 125-cell scan (cell_radius 2) finds it.
    a = np.float32(-1e-10)
    b = np.float32(0.04)
    coords, _ = oracle.cell_keys(np.array([[a, 0.5, 0.5], [b, 0.5, 0.5]], np.float32), 0.04)
    assert coords[0, 0] == -1 and coords[1, 0] == 1
    dx = b - a
    assert dx * dx == R2
    pos = np.array([[a, 0.5, 0.5], [b, 0.5, 0.5]], dtype=np.float32)
    o = _oracle_step0(oracle, pos)
    offs, idx = o.neighbors()
    assert list(idx[offs[0]:offs[1]]) == [0, 1]
    assert list(idx[offs[1]:offs[2]]) == [0, 1]


def test_nan_particle_has_no_neighbors(oracle):
    pos = np.array([[0.1, 0.1, 0.1], [np.nan, 0.1, 0.1], [0.11, 0.1, 0.1]], dtype=np.float32)
    o = oracle.Oracle(3)
    o.set_particles(pos)
    o.step(0.001)
    offs, idx = o.neighbors()
    assert list(idx[offs[0]:offs[1]]) == [0, 2]
    assert offs[2] - offs[1] == 0
    rho = o.get(oracle.OR_DENSITY)
    assert rho[1] == pytest.approx(0.001 * 39788.74 * 2 / 3, rel=1e-6)


# ---------------------------------------------------------------- forces
def test_fp64_lattice_interior_acceleration_is_gravity(oracle):
    # Interior of a perfect lattice: neighbour densities and pressures are uniform, so the
    # pressure force sums m P/rho * gradW over a point-symmetric set -> 0; viscosity is 0 at rest.
    n = 17
    p = np.stack(np.meshgrid(*(np.arange(n),) * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (p * 0.01).astype(np.float32)
    params = oracle.default_params()
    o = _oracle_step0(oracle, pos, params)
    offs, idx = o.neighbors()
    mass = np.full(len(pos), 0.001, np.float32)
    vel = np.zeros_like(pos)
    rho, acc = oracle.density_accel_f64(pos, vel, mass, offs, idx, params)
    c = (n // 2) * n * n + (n // 2) * n + n // 2
    scale = oracle.accel_term_scale(pos, vel, mass, rho.astype(np.float32), offs, idx, params)
    assert abs(acc[c, 0]) < 1e-6 * scale[c]  # FP32-rounded lattice: ~4e-8
    assert abs(acc[c, 1] - (-9.81)) < 1e-6 * scale[c]
    assert abs(acc[c, 2]) < 1e-6 * scale[c]
    # near the +x face density, hence P = k(rho - rho0), falls towards the face; the SPH pressure
    # term approximates -grad(P)/rho, which points down the pressure gradient: +x.  A flipped
    # sign on the pressure term fails here.
    e = (n // 2 + 6) * n * n + (n // 2) * n + n // 2
    assert acc[e, 0] > 0.02 * scale[e]


def test_fp32_oracle_matches_fp64_on_jittered_block(oracle, scenes):
    s = scenes.small_block(12, 10, 9, seed=7)
    params = s.oracle_params(oracle)
    o = oracle.Oracle(s.n, params)
    o.set_particles(s.pos, s.vel)
    o.lib.or_update_neighbor_lists(o.h)
    o.lib.or_compute_densities(o.h)
    o.lib.or_compute_pressures(o.h)
    o.lib.or_compute_forces(o.h)
    offs, idx = o.neighbors()
    rho32 = o.get(oracle.OR_DENSITY)
    acc32 = o.get(oracle.OR_ACCEL)
    mass = np.full(s.n, 0.001, np.float32)
    rho64, acc64 = oracle.density_accel_f64(s.pos, s.vel, mass, offs, idx, params)
    np.testing.assert_allclose(rho32, rho64, rtol=2e-6)
    scale = oracle.accel_term_scale(s.pos, s.vel, mass, rho32, offs, idx, params)
    err = np.linalg.norm(acc32 - acc64, axis=1) / scale
    assert err.max() < 1e-5


def test_viscosity_term_sign(oracle):
    # two particles approaching each other: viscosity decelerates both (F_v ~ +(v_j - v_i))
    pos = np.array([[0.1, 0.5, 0.5], [0.12, 0.5, 0.5]], np.float32)
    vel = np.array([[1.0, 0, 0], [-1.0, 0, 0]], np.float32)
    params = oracle.default_params(gas_constant=0.0, gravity=0.0)
    o = oracle.Oracle(2, params)
    o.set_particles(pos, vel)
    o.lib.or_update_neighbor_lists(o.h)
    o.lib.or_compute_densities(o.h)
    o.lib.or_compute_pressures(o.h)
    o.lib.or_compute_forces(o.h)
    acc = o.get(oracle.OR_ACCEL)
    # q = 1: laplacian = sigma*(3-2)/h^2 > 0, so a_0 points along v_1 - v_0 = -x
    assert acc[0, 0] < 0 and acc[1, 0] > 0
    assert acc[0, 0] == pytest.approx(-acc[1, 0], rel=1e-6)


# ---------------------------------------------------------------- integrator, clamp, CFL
def test_leapfrog_closed_form(oracle):
    params = oracle.default_params(damping=0.5)
    x0 = np.array([[0.1, 0.2, 0.3]], np.float32)
    v0 = np.array([[1.0, -2.0, 0.5]], np.float32)
    a = np.array([[10.0, 20.0, -30.0]], np.float32)
    dt = 0.01
    x1, v1 = oracle.integrate_clamp(x0, v0, a, dt, params)
    # x' = x + (v + a dt/2) dt ;  v' = damping (v + a dt)
    np.testing.assert_allclose(x1, x0 + (v0 + 0.5 * a * dt) * dt, rtol=1e-6)
    np.testing.assert_allclose(v1, 0.5 * (v0 + a * dt), rtol=1e-6)


def test_wall_clamp(oracle):
    params = oracle.default_params(damping=1.0, xmin=0, xmax=1, ymin=0, ymax=1, zmin=0, zmax=1)
    x0 = np.array([[1.2, -0.3, 0.5], [0.5, 0.5, 0.5]], np.float32)
    v0 = np.array([[1.0, -2.0, 0.25], [0.0, 0.0, 0.0]], np.float32)
    x1, v1 = oracle.integrate_clamp(x0, v0, np.zeros_like(x0), 0.0, params)
    np.testing.assert_array_equal(x1[0], [1.0, 0.0, 0.5])
    np.testing.assert_array_equal(v1[0], np.array([-0.8, 1.6, 0.25], np.float32))
    np.testing.assert_array_equal(x1[1], x0[1])


def test_cfl_closed_form(oracle):
    params = oracle.default_params(timestep=1.0)
    o = oracle.Oracle(3, params)
    o.set_particles(np.zeros((3, 3), np.float32) + 0.1,
                    np.array([[3, 4, 0], [0, 0, 0], [np.nan, 0, 0]], np.float32))
    # no accelerations yet: dt_force = 0.4 sqrt(0.02/1e-6) = 56.6; dt_cfl = 0.4*0.02/5
    dt = o.lib.or_compute_cfl_timestep(o.h)
    assert dt == pytest.approx(0.4 * 0.02 / 5.000001, rel=1e-6)
    params.timestep = 1e-4
    o.set_params(params)
    assert o.lib.or_compute_cfl_timestep(o.h) == pytest.approx(1e-4)


def test_total_mass_definition(oracle):
    # get_total_mass is sum(rho) * h^3, not sum(m)
    o = oracle.Oracle(1)
    o.set_particles(np.array([[0.1, 0.2, 0.3]], np.float32))
    o.step(0.001)
    rho = o.get(oracle.OR_DENSITY)
    assert o.lib.or_total_mass(o.h) == pytest.approx(rho[0] * 0.02 ** 3, rel=1e-6)


def test_total_energy_closed_forms(oracle):
    # get_total_energy = sum 0.5 m |v|^2 (src/sph_engine.cpp:193-200).  Values chosen so that
    # every product and partial sum is exact in FP32, so the closed forms hold bit for bit.
    n = 64
    pos = (np.arange(3 * n, dtype=np.float32).reshape(n, 3) * np.float32(0.1))
    o = oracle.Oracle(n)
    # uniform velocity: N * 0.5 m |v|^2 = 64 * 0.5 * 0.5 * (0.25 + 0.0625) = 5.0
    vel = np.tile(np.array([0.5, 0.25, 0.0], np.float32), (n, 1))
    o.set_particles(pos, vel, np.full(n, 0.5, np.float32))
    assert o.lib.or_total_energy(o.h) == 5.0
    # two speeds, two masses: 32 * 0.5 * 2 * 1 + 32 * 0.5 * 0.25 * 4 = 32 + 16
    vel = np.zeros((n, 3), np.float32)
    vel[::2, 0] = 1.0
    vel[1::2, 1] = -2.0
    mass = np.where(np.arange(n) % 2 == 0, 2.0, 0.25).astype(np.float32)
    o.set_particles(pos, vel, mass)
    assert o.lib.or_total_energy(o.h) == 48.0
    # all three components enter: |(1, 2, 2)|^2 = 9
    vel = np.tile(np.array([1.0, 2.0, -2.0], np.float32), (n, 1))
    o.set_particles(pos, vel, np.full(n, 1.0, np.float32))
    assert o.lib.or_total_energy(o.h) == 0.5 * 9 * n
    # random state against FP64 (sequential FP32 sum: error below n * 2^-23 relative)
    rng = np.random.default_rng(4)
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    mass = rng.uniform(0.5, 2.0, n).astype(np.float32)
    o.set_particles(pos, vel, mass)
    ke64 = float(np.sum(0.5 * mass.astype(np.float64) * np.sum(vel.astype(np.float64) ** 2, 1)))
    assert o.lib.or_total_energy(o.h) == pytest.approx(ke64, rel=n * 2.0 ** -23)
    # the particles at rest add nothing
    o.set_particles(pos, np.zeros((n, 3), np.float32))
    assert o.lib.or_total_energy(o.h) == 0.0


def test_radius_below_support_drops_far_pairs(oracle):
    # h = 0.025 with the default radius 0.04 (examples/dam_break.cpp): a pair 0.045 apart has
    # q = 1.8 < 2 but lies beyond the radius, so it is not a neighbour and adds nothing; a
    # pair 0.035 apart is one.
    h = 0.025
    sig = np.float32(1.0) / (np.float32(math.pi) * np.float32(h) ** 3)
    params = oracle.default_params(smoothing_length=h, neighbor_search_radius=0.04,
                                   gravity=0.0)
    pos = np.array([[0.1, 0.5, 0.5], [0.145, 0.5, 0.5], [0.5, 0.5, 0.5], [0.535, 0.5, 0.5]],
                   np.float32)
    o = oracle.Oracle(4, params)
    o.set_particles(pos)
    o.lib.or_update_neighbor_lists(o.h)
    offs, idx = o.neighbors()
    assert idx[offs[0]:offs[1]].tolist() == [0]
    assert sorted(idx[offs[2]:offs[3]].tolist()) == [2, 3]
    o.lib.or_compute_densities(o.h)
    rho = o.get(oracle.OR_DENSITY)
    w0 = float(sig) * 2.0 / 3.0
    assert rho[0] == pytest.approx(0.001 * w0, rel=1e-6)
    q = 0.035 / h
    w = float(sig) * (2.0 - q) ** 3 / 6.0
    assert rho[2] == pytest.approx(0.001 * (w0 + w), rel=1e-5)


def test_sampled_step0_equals_full_phases(oracle, scenes):
    # The sampled evaluation (for scenes too large for per-particle lists) must give exactly
    # the full phases' values: same hash, same query, same per-particle loops.
    s = scenes.config0(seed=13)
    s.vel = np.random.default_rng(2).normal(0, 0.3, s.pos.shape).astype(np.float32)
    params = s.oracle_params(oracle)
    full = oracle.Oracle(s.n, params)
    full.set_particles(s.pos, s.vel)
    full.lib.or_update_neighbor_lists(full.h)
    full.lib.or_compute_densities(full.h)
    full.lib.or_compute_pressures(full.h)
    full.lib.or_compute_forces(full.h)
    offs, idx = full.neighbors()
    mass = np.full(s.n, s.mass, np.float32)
    rho64, acc64 = oracle.density_accel_f64(s.pos, s.vel, mass, offs, idx, params)
    scale = oracle.accel_term_scale(s.pos, s.vel, mass, full.get(2), offs, idx, params)
    sample = np.random.default_rng(1).choice(s.n, 300, replace=False)
    sample = np.concatenate([sample, sample[:5]])  # repeated entries are allowed
    o = oracle.Oracle(s.n, params)
    o.set_particles(s.pos, s.vel)
    r = o.sample_step0(sample)
    np.testing.assert_array_equal(r["counts"], np.diff(offs)[sample])
    np.testing.assert_array_equal(r["rho"], full.get(2)[sample])
    np.testing.assert_array_equal(r["acc"], full.get(4)[sample])
    np.testing.assert_array_equal(r["rho64"], rho64[sample])
    np.testing.assert_array_equal(r["acc64"], acc64[sample])
    np.testing.assert_array_equal(r["scale"], scale[sample])
    so, si = o.sample_neighbors(sample)
    for k, i in enumerate(sample):
        assert si[so[k]:so[k + 1]].tolist() == idx[offs[i]:offs[i + 1]].tolist()
    tot, mx = o.count_neighbors()
    assert tot == int(offs[-1]) and mx == int(np.diff(offs).max())
    with pytest.raises(ValueError):
        o.sample_step0([s.n])
