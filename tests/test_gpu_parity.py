"""Parity of the HIP step (through the C ABI of libsphk.so) with the CPU oracle.

Bars:
  * cell coordinates / 63-bit hash keys, per-cell membership and order, neighbour lists
    (content and order): bit-exact.
  * density: relative error <= 1e-5 against the FP32 oracle; the FP64 error is reported.
  * acceleration: |a_gpu - a_oracle| / S_i <= 1e-5 per particle, where S_i is the sum of the
    magnitudes of the pressure and viscosity terms of particle i divided by m_i (the sum
    cancels by up to 9 digits, so plain relative error on a is ill-conditioned).
  * leapfrog + wall clamp: bit-exact given the same accelerations.
"""
import numpy as np
import pytest

from conftest import R2

pytestmark = pytest.mark.gpu

RHO_TOL = 1e-5
ACC_TOL = 1e-5


def _engine(sphk, scene, formulation=0, capacity=None):
    e = sphk.Engine(capacity or scene.n, scene.params(sphk, formulation))
    e.upload(scene.pos, scene.vel)
    return e


def _oracle(oracle, scene, formulation=0):
    o = oracle.Oracle(scene.n, scene.oracle_params(oracle, formulation))
    o.set_particles(scene.pos, scene.vel)
    return o


def _oracle_phases(o):
    o.lib.or_update_neighbor_lists(o.h)
    o.lib.or_compute_densities(o.h)
    o.lib.or_compute_pressures(o.h)
    o.lib.or_compute_forces(o.h)


def _check_neighbors(e, o):
    offs_g, idx_g = e.neighbors()
    offs_o, idx_o = o.neighbors()
    np.testing.assert_array_equal(offs_g, offs_o)
    np.testing.assert_array_equal(idx_g, idx_o)
    return offs_o, idx_o


def _check_density(e, o, scene, params, label=""):
    n = scene.n
    rho_g = e.densities()[:n]
    rho_o = o.get(0 + 2)  # OR_DENSITY
    rel = np.abs(rho_g.astype(np.float64) - rho_o) / np.abs(rho_o)
    assert rel.max() <= RHO_TOL, f"{label} density rel err {rel.max():.3e}"
    return rho_g, rho_o, rel


def test_cell_keys_bitexact_config0(sphk, oracle, scenes):
    s = scenes.config0()
    e = _engine(sphk, s)
    coords_g = e.cell_coords()
    keys_g = e.cell_keys()
    coords_o, keys_o = oracle.cell_keys(s.pos, np.float32(0.04))
    np.testing.assert_array_equal(coords_g, coords_o)
    np.testing.assert_array_equal(keys_g, keys_o)
    # jitter puts some particles below 0: cell -1 must appear
    assert coords_o.min() == -1


def test_sorted_order_matches_hash_buckets(sphk, oracle, scenes):
    # SpatialHash::build appends ids in ascending order per key: the GPU's cell-sorted order,
    # grouped by cell, must list exactly the same ids in the same order.
    s = scenes.config0(seed=3)
    e = _engine(sphk, s)
    order, cells = e.sorted_order()
    _, keys_o = oracle.cell_keys(s.pos, np.float32(0.04))
    assert np.all(np.diff(cells.astype(np.int64)) >= 0)
    assert sorted(order.tolist()) == list(range(s.n))
    buckets = {}
    for i, k in enumerate(keys_o.tolist()):
        buckets.setdefault(k, []).append(i)
    bounds = np.flatnonzero(np.diff(cells.astype(np.int64))) + 1
    groups = np.split(order, bounds)
    assert len(groups) == len(buckets)
    for g in groups:
        k = int(keys_o[g[0]])
        assert g.tolist() == buckets[k]


def _scan_paths(n, bits):
    """Which scan each radix pass uses for n keys of `bits` bits (tiles of 2,048 keys): inside
    the downsweep (at most 8,192 histogram words), one look-back launch (at most 64 scan tiles
    of 4,096 words) or three launches."""
    passes = -(-bits // 11)
    per = -(-bits // passes)
    tiles = -(-n // 2048)
    out = []
    for p in range(passes):
        m = (1 << min(per, bits - p * per)) * tiles
        out.append("local" if m <= 8192 else "lookback" if m <= 64 * 4096 else "three")
    return out


@pytest.mark.parametrize("n", [20_000, 300_000, 1_200_000])
def test_sorted_order_every_scan_path(sphk, oracle, scenes, n):
    # the cell-sorted order is the stable order by (cell x, y, z, caller id) whichever scan the
    # radix passes take, and stays so over repeated builds (the look-back scan's state is
    # cleared by each launch for the next)
    s = scenes.config1()
    rng = np.random.default_rng(n)
    lo = np.array([0.0, 0.0, 0.0]) + 0.01
    hi = np.array([3.0, 1.2, 1.5]) - 0.01
    pos = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    e = sphk.Engine(n, s.params(sphk))
    coords, _ = oracle.cell_keys(pos, np.float32(0.04))
    want = np.lexsort((np.arange(n), coords[:, 2], coords[:, 1], coords[:, 0]))
    dims = coords.max(0) - coords.min(0) + 5
    paths = _scan_paths(n, int(np.prod(dims.astype(np.int64))).bit_length())
    print("scan paths", n, paths, flush=True)
    for rep in range(3):
        e.upload(pos)
        order, cells = e.sorted_order()
        np.testing.assert_array_equal(order, want.astype(np.uint32))
        assert np.all(np.diff(cells.astype(np.int64)) >= 0)


def test_graph_replay_matches_eager_mid_size(sphk, scenes):
    # 300,000 particles: the radix passes scan with the one-launch look-back, whose state the
    # replays must find cleared
    s = scenes.config1()
    rng = np.random.default_rng(4)
    pos = rng.uniform([0.05, 0.05, 0.05], [2.95, 1.15, 1.45], size=(300_000, 3)).astype(np.float32)
    outs = []
    for graph in (False, True):
        e = sphk.Engine(len(pos), s.params(sphk, 1))
        e.set_graph(graph)
        e.upload(pos)
        e.snapshot_save()
        res = []
        for k in range(3):
            e.snapshot_restore()
            e.step(1e-4)
            res.append((e.positions(), e.densities()))
        for k in range(3):
            e.step(1e-4)
        res.append((e.positions(), e.velocities()))
        outs.append(res)
    for a, b in zip(outs[0], outs[1]):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_neighbors_density_accel_step0_config0(sphk, oracle, scenes):
    s = scenes.config0()
    e = _engine(sphk, s)
    o = _oracle(oracle, s)
    _oracle_phases(o)
    offs, idx = _check_neighbors(e, o)
    e.step(s.dt)
    params = s.oracle_params(oracle)
    rho_g, rho_o, rel = _check_density(e, o, s, params, "config0")
    mass = np.full(s.n, s.mass, np.float32)
    rho64, acc64 = oracle.density_accel_f64(s.pos, s.vel, mass, offs, idx, params)
    rel64 = np.abs(rho_g - rho64) / rho64
    assert rel64.max() < 1e-5
    acc_g = e.accelerations()
    acc_o = o.get(4)
    scale = oracle.accel_term_scale(s.pos, s.vel, mass, rho_o, offs, idx, params)
    err = np.linalg.norm(acc_g.astype(np.float64) - acc_o, axis=1) / scale
    assert err.max() <= ACC_TOL, f"accel err {err.max():.3e}"
    err64 = np.linalg.norm(acc_g - acc64, axis=1) / scale
    assert err64.max() <= ACC_TOL
    # pressures capacity-length, P = k (rho - rho0)
    prs = e.pressures()
    np.testing.assert_allclose(prs[: s.n], np.float32(2000.0) * (rho_g - np.float32(1000.0)),
                               rtol=1e-6)


def test_integrate_clamp_bitexact(sphk, oracle, scenes):
    s = scenes.config0(seed=11)
    e = _engine(sphk, s)
    e.step(s.dt)
    acc = e.accelerations()
    x1, v1 = oracle.integrate_clamp(s.pos, s.vel, acc, s.dt, s.oracle_params(oracle))
    np.testing.assert_array_equal(e.positions(), x1)
    np.testing.assert_array_equal(e.velocities(), v1)


def test_neighbor_edge_pair_two_cells_apart(sphk, oracle):
    # (-1e-10, y, z) is in cell -1, (0.04f, y, z) in cell 1, and they are neighbours (d^2 == R^2).
    pos = np.array([[-1e-10, 0.5, 0.5], [0.04, 0.5, 0.5], [0.02, 0.5, 0.5],
                    [0.5, -1e-10, 0.3], [0.5, 0.04, 0.3], [0.3, 0.3, 0.04], [0.3, 0.3, -1e-10]],
                   np.float32)
    p = sphk.default_params(xmin=-1, xmax=1, ymin=-1, ymax=1, zmin=-1, zmax=1)
    e = sphk.Engine(len(pos), p)
    e.upload(pos)
    op = oracle.default_params()
    o = oracle.Oracle(len(pos), op)
    o.set_particles(pos)
    o.lib.or_update_neighbor_lists(o.h)
    offs, idx = _check_neighbors(e, o)
    assert 1 in idx[offs[0]:offs[1]].tolist()
    assert 4 in idx[offs[3]:offs[4]].tolist()
    assert 6 in idx[offs[5]:offs[6]].tolist()


def test_random_points_against_brute_force(sphk, oracle):
    rng = np.random.default_rng(12)
    pos = rng.uniform(-0.2, 0.3, size=(4000, 3)).astype(np.float32)
    e = sphk.Engine(len(pos), sphk.default_params())
    e.upload(pos)
    offs, idx = e.neighbors()
    boffs, bidx = oracle.brute_neighbors(pos, R2)
    for i in range(len(pos)):
        assert sorted(idx[offs[i]:offs[i + 1]].tolist()) == bidx[boffs[i]:boffs[i + 1]].tolist()


def _boundary_partner(b, u, r2):
    """FP32 point p = b + t u with the largest t (bisection on the FP32 test) such that
    (dx^2 + dy^2) + dz^2 <= r2 in FP32: a pair on the very edge of the radius."""
    def inside(t):
        p = (b.astype(np.float64) + t * u).astype(np.float32)
        d = (b - p).astype(np.float32)
        d2 = np.float32(np.float32(d[0] * d[0]) + np.float32(d[1] * d[1])) + np.float32(d[2] * d[2])
        return d2 <= r2, p
    lo, hi = 0.0, 0.05
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if inside(mid)[0]:
            lo = mid
        else:
            hi = mid
    return inside(lo)[1]


def test_edge_pairs_in_all_directions_against_brute_force(sphk, oracle):
    # Pairs exactly on the FP32 edge of the radius, in random directions and at coordinates
    # of several magnitudes: the candidate filter may pass extra pairs but must keep all of
    # these (the lists must equal the brute-force FP32 result).
    rng = np.random.default_rng(31)
    r2 = np.float32(R2)
    pts = []
    for centre in ((0.0, 0.0, 0.0), (3.7, -2.9, 1.3), (-5.1, 4.2, -3.3), (1e-3, 2.5, -4.9)):
        for _ in range(150):
            b = (np.array(centre) + rng.uniform(-0.3, 0.3, 3)).astype(np.float32)
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            pts += [b, _boundary_partner(b, u, r2)]
    pos = np.array(pts, np.float32)
    e = sphk.Engine(len(pos), sphk.default_params(xmin=-8, xmax=8, ymin=-8, ymax=8, zmin=-8,
                                                   zmax=8))
    e.upload(pos)
    offs, idx = e.neighbors()
    boffs, bidx = oracle.brute_neighbors(pos, r2)
    for i in range(len(pos)):
        assert sorted(idx[offs[i]:offs[i + 1]].tolist()) == bidx[boffs[i]:boffs[i + 1]].tolist()
    for k in range(0, len(pos), 2):
        assert k + 1 in idx[offs[k]:offs[k + 1]].tolist()


def test_matrix_core_filter_keeps_every_neighbour(sphk, oracle):
    # An unjittered lattice (spacing R/4, 64 particles per cell) puts thousands of pairs within
    # a few ulps of the radius, and its chunks are dense enough for the density walk's
    # matrix-core candidate filter.  The walk counts the neighbours that pass the reference's
    # exact test among the filter's candidates: the count must equal the brute-force count,
    # so the filter dropped none.
    g = np.arange(24, dtype=np.float64) * 0.01
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = (np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + [0.13, 0.21, 0.05]).astype(np.float32)
    p = sphk.default_params(xmin=0, xmax=1, ymin=0, ymax=1, zmin=0, zmax=1)
    e = sphk.Engine(len(pos), p)
    e.upload(pos)
    e.step(1e-5)
    st = e.stats()
    boffs, bidx = oracle.brute_neighbors(pos, np.float32(R2))
    counts = np.diff(boffs)
    assert st.total_neighbors == int(counts.sum())
    assert st.max_neighbors == int(counts.max())


def _stored_pairs(e, n):
    """(target, candidate) caller-id pairs of the words the production density walk stores,
    and the chunks whose words did not all fit the store."""
    items, nwords, wj0, mask = e.walk_words()
    order, cells = e.sorted_order()
    _, dims, _ = e.grid()
    nz = int(dims[2])
    k = np.arange(32, dtype=np.uint32)
    tgt, cand = [], []
    overflow = 0
    for c in range(len(items)):
        a = int(items[c])
        col = int(cells[a]) // nz
        lanes = np.arange(a, min(a + 64, n))
        lanes = lanes[(cells[lanes] // nz) == col]
        nw = int(nwords[c])
        if nw > wj0.shape[1]:  # more words than the engine stores: tested again
            overflow += 1
            continue
        bits = ((mask[c, :nw, :len(lanes), None] >> (31 - k)) & 1).astype(bool)  # [w, l, k]
        w_i, l_i, k_i = np.nonzero(bits)
        slots = wj0[c, w_i].astype(np.int64) + k_i
        assert slots.max(initial=0) < n, "candidate slot beyond the particles"
        tgt.append(order[lanes[l_i]].astype(np.int64))
        cand.append(order[slots].astype(np.int64))
    return np.concatenate(tgt), np.concatenate(cand), overflow


@pytest.mark.parametrize("scene_kind", ["lattice", "config0"])
@pytest.mark.parametrize("shape", ["seg32_waves4", "seg24_waves4", "seg32_waves4_split",
                                   "pp12_waves4"])
def test_stored_candidate_words_hold_every_pair(sphk, oracle, scenes, scene_kind, shape):
    # The words the force walk reads are the matrix-core filter's output in the production
    # density walk.  Pair by pair: every reference neighbour (brute force, d^2 <= r^2 with the
    # reference's arithmetic) and the particle itself are among its candidates, no candidate
    # appears twice (a duplicate would count a pair twice), and the candidates beyond the
    # radius are few (the filter filters).
    if scene_kind == "lattice":  # pairs within ulps of the radius, 64 particles per cell
        g = np.arange(24, dtype=np.float64) * 0.01
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        pos = (np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + [0.13, 0.21, 0.05]).astype(np.float32)
        p = sphk.default_params(xmin=0, xmax=1, ymin=0, ymax=1, zmin=0, zmax=1)
        e = sphk.Engine(len(pos), p)
        e.upload(pos)
    else:
        sc = scenes.config0()
        pos = sc.pos
        e = _engine(sphk, sc)
    names = sphk.walk_shape_names()
    e.set_walk_shapes(names.index(shape), names.index("seg32_waves4"))
    n = len(pos)
    tgt, cand, overflow = _stored_pairs(e, n)
    assert overflow == 0
    codes = tgt * n + cand
    assert len(np.unique(codes)) == len(codes), "a candidate stored twice"
    boffs, bidx = oracle.brute_neighbors(pos, np.float32(R2))
    bt = np.repeat(np.arange(n, dtype=np.int64), np.diff(boffs))
    assert np.isin(bt * n + bidx.astype(np.int64), codes).all(), "a neighbour was filtered out"
    ids = np.arange(n, dtype=np.int64)
    assert np.isin(ids * n + ids, codes).all(), "a particle is not its own candidate"
    slack = (len(codes) - len(bidx)) / len(bidx)  # brute-force lists hold the particle itself
    print(f"{scene_kind} {shape}: {len(bidx)} neighbour pairs, candidates beyond the radius "
          f"{slack:.3f} of them")
    assert 0.0 <= slack < 0.05  # measured 0.011 (lattice) and 0.006 (config 0)


def test_far_from_origin_against_brute_force(sphk, oracle):
    # Coordinates near 1000 (cell index ~25,000): the quantized filter keeps its coordinates
    # modulo 2^16 (x, y) and 2^32 (z), so neighbour differences must survive the wrap-around.
    rng = np.random.default_rng(41)
    r2 = np.float32(R2)
    base = np.array([1000.3, -777.7, 513.1])
    pts = [(base + rng.uniform(0, 0.25, 3)).astype(np.float32) for _ in range(2500)]
    for _ in range(200):
        b = (base + rng.uniform(0, 0.25, 3)).astype(np.float32)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pts += [b, _boundary_partner(b, u, r2)]
    pos = np.array(pts, np.float32)
    lo, hi = pos.min(0) - 0.5, pos.max(0) + 0.5
    bounds = dict(zip(("xmin", "ymin", "zmin"), map(float, lo)))
    bounds.update(zip(("xmax", "ymax", "zmax"), map(float, hi)))
    e = sphk.Engine(len(pos), sphk.default_params(**bounds))
    e.upload(pos)
    offs, idx = e.neighbors()
    boffs, bidx = oracle.brute_neighbors(pos, r2)
    for i in range(len(pos)):
        assert sorted(idx[offs[i]:offs[i + 1]].tolist()) == bidx[boffs[i]:boffs[i + 1]].tolist()


def test_nonfinite_particles(sphk, oracle):
    pos = np.array([[0.1, 0.1, 0.1], [np.nan, 0.1, 0.1], [0.11, 0.1, 0.1], [0.1, np.inf, 0.1],
                    [0.12, 0.1, 0.1]], np.float32)
    e = sphk.Engine(8, sphk.default_params())
    e.upload(pos)
    o = oracle.Oracle(8, oracle.default_params())
    o.set_particles(pos)
    o.lib.or_update_neighbor_lists(o.h)
    _check_neighbors(e, o)
    o.lib.or_compute_densities(o.h)
    o.lib.or_compute_pressures(o.h)
    o.lib.or_compute_forces(o.h)
    e.step(0.001)
    rho = e.densities()[:5]
    ro = o.get(2)
    np.testing.assert_allclose(rho, ro, rtol=1e-6)
    np.testing.assert_array_equal(e.accelerations()[1], np.array([0, -9.81, 0], np.float32))
    coords_g = e.cell_coords()
    coords_o, _ = oracle.cell_keys(e.positions(), np.float32(0.04))
    np.testing.assert_array_equal(coords_g, coords_o)


def test_wide_grid_chunk_path(sphk, oracle):
    # A grid of more than 8,192 (x, y) columns (here 105 x 105) lists its chunks in three
    # launches (counts, their scan, fill) and steps the non-finite bin's particles and the
    # clock in a launch of its own; the results match the oracle as on the one-launch path.
    rng = np.random.default_rng(21)
    pos = rng.uniform([0.05, 0.05, 0.05], [3.95, 3.95, 0.12], size=(60000, 3)).astype(np.float32)
    pos[7] = [np.nan, 0.1, 0.1]
    pos[11] = [0.2, np.inf, 0.3]
    bounds = dict(xmin=0, xmax=4, ymin=0, ymax=4, zmin=0, zmax=0.4)
    e = sphk.Engine(len(pos), sphk.default_params(**bounds))
    e.upload(pos)
    _, dims, _ = e.grid()
    assert int(dims[0]) * int(dims[1]) > 8192
    o = oracle.Oracle(len(pos), oracle.default_params(**bounds))
    o.set_particles(pos)
    _oracle_phases(o)
    offs, idx = _check_neighbors(e, o)
    e.step(0.001)
    assert e.time() == np.float32(0.001)
    rho_o = o.get(2)
    fin = np.isfinite(pos).all(axis=1)
    rel = np.abs(e.densities()[: len(pos)].astype(np.float64) - rho_o) / np.abs(rho_o)
    assert rel.max() <= RHO_TOL
    acc = e.accelerations()
    for i in (7, 11):
        np.testing.assert_array_equal(acc[i], np.array([0, -9.81, 0], np.float32))
    mass = np.full(len(pos), 0.001, np.float32)
    scale = oracle.accel_term_scale(pos, np.zeros_like(pos), mass, rho_o, offs, idx, o.params)
    # particles alone in their radius feel gravity only, exactly
    alone = fin & (scale == 0)
    np.testing.assert_array_equal(acc[alone].astype(np.float64), o.get(4)[alone])
    # Within 1e-5 of the sum of term magnitudes for particles with 8 or more neighbours (self
    # included).  With fewer, a lone neighbour at the edge of the kernel's support
    # (q ~ 1.9996) can be all there is: the sum is then ~1e6 times below the scene's median and
    # the term ill-conditioned (the FP32 oracle itself misses FP64 there by up to 3e-3 of the
    # sum), so those are held to 1e-2.
    nn = np.diff(offs)
    err = np.linalg.norm(acc.astype(np.float64) - o.get(4), axis=1) / np.where(scale > 0, scale, 1)
    dense = fin & (scale > 0) & (nn >= 8)
    sparse = fin & (scale > 0) & (nn < 8)
    assert dense.sum() > len(pos) // 2
    assert err[dense].max() <= ACC_TOL
    assert err[sparse].max(initial=0.0) <= 1e-2


def test_consistent_formulation_multistep(sphk, oracle, scenes):
    # the opt-in consistent formulation stays bounded at dt = 1e-4: 20 steps track the oracle
    s = scenes.small_block(10, 10, 10, seed=2)
    s.dt = 1e-4
    e = _engine(sphk, s, formulation=1)
    o = _oracle(oracle, s, formulation=1)
    for k in range(20):
        e.step(s.dt)
        o.step(s.dt)
    np.testing.assert_allclose(e.positions(), o.get(0), rtol=0, atol=2e-6)
    np.testing.assert_allclose(e.velocities(), o.get(1), rtol=0, atol=2e-3)
    rho = e.densities()[: s.n]
    np.testing.assert_allclose(rho, o.get(2), rtol=1e-4)
    assert e.step_count() == 20
    assert e.time() == pytest.approx(20 * 1e-4, rel=1e-5)


def test_reference_formulation_pileup_neighbors(sphk, oracle, scenes):
    # reference physics: the first step throws almost every particle into the box corners, so
    # later steps run on piles of exactly coincident positions.  Neighbour lists of the GPU
    # state must equal the oracle's hash query on that same state at every step.
    s = scenes.small_block(8, 8, 8, seed=4)
    e = _engine(sphk, s)
    o = oracle.Oracle(s.n, s.oracle_params(oracle))
    for k in range(4):
        o.set_particles(e.positions(), e.velocities())
        o.lib.or_update_neighbor_lists(o.h)
        _check_neighbors(e, o)
        e.step(s.dt)
    pos = e.positions()
    corners = np.isin(pos[:, 0], [0.0, 1.0]) & np.isin(pos[:, 1], [0.0, 0.4])
    assert corners.mean() > 0.3  # most of the block is thrown onto box edges in one step


def test_adaptive_timestep_matches_oracle(sphk, oracle, scenes):
    s = scenes.small_block(6, 6, 6, seed=5)
    s.vel = (np.random.default_rng(0).normal(size=s.vel.shape) * 0.3).astype(np.float32)
    e = _engine(sphk, s)
    o = _oracle(oracle, s)
    dt_o = o.lib.or_compute_cfl_timestep(o.h)
    t0 = e.time()
    e.step(0.0)
    assert e.time() - t0 == pytest.approx(dt_o, rel=0, abs=0)
    # second step uses |a[0]| of the first
    o.step(0.0)
    dt_o2 = o.lib.or_compute_cfl_timestep(o.h)
    t1 = np.float32(e.time())
    e.step(0.0)
    # the device adds dt to a float clock; dt (~1e-10, from |a_0| ~ 1e9) is below the clock's
    # ulp scale, so compare clocks to within one ulp
    expect = np.float32(t1 + np.float32(dt_o2))
    assert abs(np.float32(e.time()) - expect) <= np.spacing(expect)


def test_total_mass_and_energy(sphk, oracle, scenes):
    s = scenes.config0(seed=9)
    e = _engine(sphk, s)
    o = _oracle(oracle, s)
    e.step(s.dt)
    o.step(s.dt)
    assert e.total_mass() == pytest.approx(o.lib.or_total_mass(o.h), rel=1e-5)
    # exact invariant: the particle masses never change
    np.testing.assert_array_equal(e.masses(), np.full(s.n, s.mass, np.float32))
    me, ee = e.conservation_errors()
    assert ee == 0.0
    assert me == pytest.approx(abs(e.total_mass() - s.n * 0.001) / (s.n * 0.001), rel=1e-5)
    ke = e.total_energy()
    assert ke == pytest.approx(o.lib.or_total_energy(o.h), rel=1e-3)


def test_deterministic_and_snapshot_restore(sphk, scenes):
    s = scenes.config0(seed=1)
    e = _engine(sphk, s)
    e.snapshot_save()
    e.step(s.dt)
    a1 = e.positions()
    r1 = e.densities()
    e.snapshot_restore()
    np.testing.assert_array_equal(e.positions(), s.pos)
    e.step(s.dt)
    np.testing.assert_array_equal(e.positions(), a1)
    np.testing.assert_array_equal(e.densities(), r1)


def test_empty_and_capacity(sphk):
    e = sphk.Engine(4, sphk.default_params())
    e.step(0.001)  # no particles: no-op
    assert e.step_count() == 0
    pos = np.zeros((6, 3), np.float32) + 0.1
    with pytest.raises(sphk.SphkError) as ex:
        e.upload(pos)
    assert ex.value.code == -3
    assert e.size() == 4
    d = e.densities()
    assert d.shape == (4,)
    e.clear()
    assert e.size() == 0
    # append semantics (ParticleSystem::add_particles)
    e.upload(pos[:2])
    e.upload(pos[:1] + 0.01, offset=2)
    assert e.size() == 3


def test_capacity_length_densities(sphk, scenes):
    s = scenes.small_block(4, 4, 4)
    e = sphk.Engine(1000, s.params(sphk))
    e.upload(s.pos)
    e.step(s.dt)
    d = e.densities()
    assert d.shape == (1000,)
    assert np.all(d[s.n:] == 0)
    assert np.all(d[: s.n] > 0)


def test_stats(sphk, scenes):
    s = scenes.config0()
    e = _engine(sphk, s)
    e.set_timing(True)
    e.step(s.dt)
    st = e.stats()
    assert st.max_neighbors == 257
    assert st.total_neighbor_queries == s.n
    assert st.total_time > 0 and st.density_computation_time > 0
    e.reset_stats()
    assert e.stats().max_neighbors == 0


def test_config1_fullsize_step0(sphk, oracle, scenes):
    # the 1M-particle dam break at its real shape: density and acceleration of every particle
    s = scenes.config1()
    e = _engine(sphk, s)
    o = _oracle(oracle, s)
    _oracle_phases(o)
    e.set_timing(True)
    e.step(s.dt)
    st = e.stats()
    assert st.max_neighbors == o.lib.or_max_neighbors(o.h)
    assert st.total_neighbors == o.lib.or_neighbor_total(o.h)
    params = s.oracle_params(oracle)
    rho_g, rho_o, rel = _check_density(e, o, s, params, "config1")
    offs, idx = o.neighbors()
    mass = np.full(s.n, s.mass, np.float32)
    scale = oracle.accel_term_scale(s.pos, s.vel, mass, rho_o, offs, idx, params)
    err = np.linalg.norm(e.accelerations().astype(np.float64) - o.get(4), axis=1) / scale
    assert err.max() <= ACC_TOL
    # integration given the GPU's accelerations: bit-exact
    x1, v1 = oracle.integrate_clamp(s.pos, s.vel, e.accelerations(), s.dt, params)
    np.testing.assert_array_equal(e.positions(), x1)


def test_walk_shapes_bitexact(sphk, scenes):
    # every launch shape of the walks gives bit-identical results (same per-particle order)
    # (the split shapes walk a chunk with three wavefronts and sum the parts afterwards; they
    # mix freely with the single-wavefront ones, density and force chosen independently)
    # Both scenes leave candidate columns empty at the block's faces (an empty column must
    # not hide the start of the next column group from the segment and part breaks).
    # (the ping-pong shapes walk two segment buffers, lanes running ahead into the newer one)
    names = sphk.walk_shape_names()
    split = [v for v, nm in enumerate(names) if nm.endswith("_split")]
    pp = [v for v, nm in enumerate(names) if nm.startswith("pp")]
    assert split and pp, names
    pairs = [(v, v) for v in range(len(names))]
    pairs += [(split[0], 0), (0, split[0]), (split[-1], 2), (pp[0], pp[-1]), (pp[-1], split[0]),
              (split[-1], pp[0])]
    for s, dt in ((scenes.config0(seed=21), None), (scenes.small_block(10, 10, 10, seed=8), 1e-4)):
        dt = dt or s.dt
        ref = None
        for d, f in pairs:
            print(f"{s.name}: walk shapes density={names[d]} force={names[f]}", flush=True)
            e = _engine(sphk, s)
            e.set_walk_shapes(d, f)
            for _ in range(3):
                e.step(dt)
            out = (e.positions(), e.velocities(), e.densities(), e.accelerations())
            if ref is None:
                ref = out
            else:
                for a, b in zip(ref, out):
                    np.testing.assert_array_equal(a, b)
            offs, idx = e.neighbors()
            if d == 0 and f == 0:
                ref_n = (offs, idx)
            else:
                np.testing.assert_array_equal(offs, ref_n[0])
                np.testing.assert_array_equal(idx, ref_n[1])


@pytest.mark.parametrize("split", [False, True, "pp"])
def test_force_retest_path_matches_stored_masks(sphk, oracle, scenes, split):
    # a pile-up whose chunks have more candidate words than the engine stores (256 words of
    # 32); the force pass then re-tests them.  Both paths must agree with the oracle, in the
    # single-wavefront, the three-wavefront (split) and the ping-pong walks.
    rng = np.random.default_rng(3)
    pos = np.concatenate([rng.uniform(0.1, 0.13, size=(10000, 3)),
                          rng.uniform(0.3, 0.5, size=(500, 3))]).astype(np.float32)
    p = sphk.default_params(xmin=0, xmax=1, ymin=0, ymax=1, zmin=0, zmax=1)
    e = sphk.Engine(len(pos), p)
    names = sphk.walk_shape_names()
    if split == "pp":
        e.set_walk_shapes(names.index("pp12_waves4"), names.index("pp16_waves4"))
    elif split:
        v = names.index("seg16_waves4_split")
        e.set_walk_shapes(v, v)
    e.upload(pos)
    _, nwords, wj0, _ = e.walk_words()
    assert nwords.max() > wj0.shape[1], "no chunk takes the re-test path"
    o = oracle.Oracle(len(pos), oracle.default_params(xmin=0, xmax=1, ymin=0, ymax=1, zmin=0,
                                                      zmax=1))
    o.set_particles(pos)
    _oracle_phases(o)
    _check_neighbors(e, o)
    e.step(0.001)
    rho = e.densities()[: len(pos)]
    np.testing.assert_allclose(rho, o.get(2), rtol=1e-5)
    offs, idx = o.neighbors()
    mass = np.full(len(pos), 0.001, np.float32)
    scale = oracle.accel_term_scale(pos, np.zeros_like(pos), mass, o.get(2), offs, idx,
                                    o.params)
    err = np.linalg.norm(e.accelerations().astype(np.float64) - o.get(4), axis=1) / scale
    assert err.max() <= ACC_TOL


def test_word_store_fallback_bitexact(sphk):
    # An engine whose word buffers would pass 16 GiB (capacity 17M: 265,000 chunk slots) stores
    # 128 candidate words per chunk instead of 256; chunks of 129-256 words are then tested
    # again by the force walk instead of read back.  Both routes walk the same words in the
    # same order: the two engines give the same bits.
    rng = np.random.default_rng(5)
    pos = np.concatenate([rng.uniform(0.1, 0.13, size=(6000, 3)),
                          rng.uniform(0.3, 0.5, size=(1000, 3))]).astype(np.float32)
    p = sphk.default_params(xmin=0, xmax=1, ymin=0, ymax=1, zmin=0, zmax=1)
    outs = []
    for cap in (len(pos), 17_000_000):
        e = sphk.Engine(cap, p)
        e.upload(pos)
        _, nwords, wj0, _ = e.walk_words()
        stored = wj0.shape[1]
        assert stored == (256 if cap == len(pos) else 128)
        if cap == len(pos):
            # chunks that this engine reads back and the other tests again
            assert np.any((nwords > 128) & (nwords <= 256))
        for _ in range(2):
            e.step(1e-4)
        outs.append((e.positions()[: len(pos)], e.velocities()[: len(pos)],
                     e.densities()[: len(pos)]))
        e.close()
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(a, b)


def test_pileup_split_and_single_walks_bitexact(sphk):
    # Pile-up chunks have more candidate words than the engine stores: the single-wavefront
    # density walk steps its cursor through them, the split walk enumerates each part's words
    # in a word table, and the force walk re-tests them.  Every route walks the same words in
    # the same order, so all shapes give the same bits.
    rng = np.random.default_rng(11)
    pos = np.concatenate([rng.uniform(0.1, 0.13, size=(10000, 3)),
                          rng.uniform(0.3, 0.5, size=(2000, 3))]).astype(np.float32)
    names = sphk.walk_shape_names()
    outs = []
    e = sphk.Engine(len(pos), sphk.default_params(xmin=0, xmax=1, ymin=0, ymax=1, zmin=0, zmax=1))
    e.upload(pos)
    _, nwords, wj0, _ = e.walk_words()
    assert nwords.max() > wj0.shape[1], "no chunk takes the re-test path"
    for d, f in ((0, 0), (names.index("seg16_waves4_split"), names.index("seg32_waves4_split")),
                 (1, 0), (names.index("pp12_waves4"), names.index("pp16_waves4")),
                 (names.index("pp16_waves4"), names.index("pp12_waves4"))):
        p = sphk.default_params(xmin=0, xmax=1, ymin=0, ymax=1, zmin=0, zmax=1)
        e = sphk.Engine(len(pos), p)
        e.set_walk_shapes(d, f)
        e.upload(pos)
        for _ in range(2):
            e.step(1e-4)
        outs.append((e.positions(), e.velocities(), e.densities()))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            np.testing.assert_array_equal(a, b)


def test_graph_replay_matches_eager(sphk, scenes):
    # a captured step replayed many times equals the eager step bit for bit, with eager work
    # (downloads, restores) between the replays
    s = scenes.config0(seed=5)
    outs = []
    for graph in (False, True):
        e = _engine(sphk, s)
        e.set_graph(graph)
        e.snapshot_save()
        res = []
        for k in range(4):
            e.snapshot_restore()
            e.step(s.dt)
            res.append((e.positions(), e.velocities(), e.densities()))
        s2 = scenes.small_block(10, 10, 10, seed=8)
        e2 = _engine(sphk, s2, formulation=1)
        e2.set_graph(graph)
        for k in range(5):
            e2.step(1e-4)
            res.append((e2.positions(), e2.densities()))
        e2.step(0.0)  # adaptive dt: a new capture
        res.append((e2.positions(),))
        outs.append(res)
    for a, b in zip(outs[0], outs[1]):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def _ensemble(scenes_mod, n_scenes=6, scale=0.02, seed=3):
    return scenes_mod.config3(seed=seed, n_scenes=n_scenes, scale=scale)


def test_batched_scenes_equal_single_scene_runs(sphk, scenes):
    # one engine stepping several independent scenes gives each scene exactly the results of
    # an engine holding that scene alone
    ens = _ensemble(scenes)
    offs = np.cumsum([0] + [sc.n for sc in ens]).astype(np.uint64)
    bounds = np.array([sc.bounds for sc in ens], np.float32)
    pos = np.concatenate([sc.pos for sc in ens])
    vel = np.concatenate([sc.vel for sc in ens])
    eb = sphk.Engine(len(pos), ens[0].params(sphk))
    eb.upload(pos, vel)
    eb.set_scenes(offs, bounds)
    assert eb.scene_count() == len(ens)
    for k in range(3):
        eb.step(0.001)
    pb, vb, rb = eb.positions(), eb.velocities(), eb.densities()
    for i, sc in enumerate(ens):
        e = _engine(sphk, sc)
        for k in range(3):
            e.step(0.001)
        a, b = int(offs[i]), int(offs[i + 1])
        np.testing.assert_array_equal(pb[a:b], e.positions())
        np.testing.assert_array_equal(vb[a:b], e.velocities())
        np.testing.assert_array_equal(rb[a:b], e.densities()[: sc.n])


def test_batched_scenes_against_oracle(sphk, oracle, scenes):
    ens = _ensemble(scenes, n_scenes=4, scale=0.03, seed=9)
    offs = np.cumsum([0] + [sc.n for sc in ens]).astype(np.uint64)
    pos = np.concatenate([sc.pos for sc in ens])
    eb = sphk.Engine(len(pos), ens[0].params(sphk))
    eb.upload(pos)
    eb.set_scenes(offs, np.array([sc.bounds for sc in ens], np.float32))
    eb.set_graph(True)
    eb.step(0.001)
    rho = eb.densities()
    acc = eb.accelerations()
    for i, sc in enumerate(ens):
        o = _oracle(oracle, sc)
        _oracle_phases(o)
        a, b = int(offs[i]), int(offs[i + 1])
        np.testing.assert_allclose(rho[a:b], o.get(2), rtol=RHO_TOL)
        offs_o, idx_o = o.neighbors()
        mass = np.full(sc.n, sc.mass, np.float32)
        scale = oracle.accel_term_scale(sc.pos, sc.vel, mass, o.get(2), offs_o, idx_o, o.params)
        err = np.linalg.norm(acc[a:b].astype(np.float64) - o.get(4), axis=1) / scale
        assert err.max() <= ACC_TOL
    # scenes never see each other: neighbour lists stay inside each scene's id range
    offs_n, idx_n = eb.neighbors()
    for i in range(len(ens)):
        a, b = int(offs[i]), int(offs[i + 1])
        seg = idx_n[offs_n[a]:offs_n[b]]
        assert seg.min() >= a and seg.max() < b


# ---------------------------------------------------------------------------
# physics regimes the step-0 scenes above leave open
# ---------------------------------------------------------------------------
def _shape_index(sphk, name):
    return sphk.walk_shape_names().index(name)


def _params_pair(sphk, oracle, scene, formulation=0, **kw):
    p = scene.params(sphk, formulation)
    op = scene.oracle_params(oracle, formulation)
    for k, v in kw.items():
        setattr(p, k, v)
        setattr(op, k, v)
    return p, op


def _one_step_vs_oracle(sphk, oracle, scene, p, op, shape=None, dt=None):
    """One GPU step and the oracle's phases on the same state: neighbour lists, density and
    acceleration errors (FP32 oracle, FP64 over the same lists) and the conditioning scale."""
    e = sphk.Engine(scene.n, p)
    if shape is not None:
        v = _shape_index(sphk, shape)
        e.set_walk_shapes(v, v)
    e.upload(scene.pos, scene.vel)
    o = oracle.Oracle(scene.n, op)
    o.set_particles(scene.pos, scene.vel)
    _oracle_phases(o)
    offs, idx = _check_neighbors(e, o)
    e.step(dt or scene.dt)
    n = scene.n
    rho_g = e.densities()[:n]
    rho_o = o.get(2)
    mass = np.full(n, scene.mass, np.float32)
    rho64, acc64 = oracle.density_accel_f64(scene.pos, scene.vel, mass, offs, idx, op)
    scale = oracle.accel_term_scale(scene.pos, scene.vel, mass, rho_o, offs, idx, op)
    acc_g = e.accelerations().astype(np.float64)
    return dict(
        rho_rel=np.abs(rho_g - rho_o) / np.abs(rho_o),
        rho_rel64=np.abs(rho_g - rho64) / np.abs(rho64),
        acc_err=np.linalg.norm(acc_g - o.get(4), axis=1) / scale,
        acc_err64=np.linalg.norm(acc_g - acc64, axis=1) / scale,
        acc_g=acc_g, acc_o=o.get(4).astype(np.float64), scale=scale)


@pytest.mark.parametrize("h", [0.025, 0.015])
@pytest.mark.parametrize("shape", ["seg32_waves4", "seg32_waves4_split"])
def test_radius_differs_from_support(sphk, oracle, scenes, h, shape):
    # The reference keeps its 0.04 m search radius when h changes (examples/dam_break.cpp
    # sets h = 0.025 and leaves neighbor_search_radius at its default).  With R < 2h pairs
    # just beyond R would still carry kernel weight: they must not be counted.  R > 2h: pairs
    # between 2h and R are neighbours of zero weight.
    s = scenes.config0(seed=17)
    p, op = _params_pair(sphk, oracle, s, smoothing_length=h, neighbor_search_radius=0.04)
    r = _one_step_vs_oracle(sphk, oracle, s, p, op, shape)
    print(f"h={h} {shape}: rho rel {r['rho_rel'].max():.2e} (fp64 {r['rho_rel64'].max():.2e}), "
          f"accel {r['acc_err'].max():.2e} (fp64 {r['acc_err64'].max():.2e})")
    assert r["rho_rel"].max() <= RHO_TOL
    assert r["rho_rel64"].max() <= RHO_TOL
    assert r["acc_err"].max() <= ACC_TOL
    assert r["acc_err64"].max() <= ACC_TOL


def _moving(scene, seed, sigma=0.3):
    rng = np.random.default_rng(seed)
    scene.vel = (rng.normal(size=scene.vel.shape) * sigma).astype(np.float32)
    return scene


@pytest.mark.parametrize("formulation", [0, 1])
@pytest.mark.parametrize("shape", ["seg32_waves4", "seg32_waves4_split"])
def test_viscosity_term_isolated(sphk, oracle, scenes, formulation, shape):
    # gas_constant = 0 and gravity = 0: the acceleration is the viscosity sum alone
    # (compute_viscosity_force, src/sph_engine.cpp:414-432, with laplacianW_3d,
    # src/kernels.cpp:130-138); with seeded velocities N(0, 0.3) it is of order 1e4 m/s^2.
    # The scale S_i is then the sum of the viscosity terms' magnitudes alone.
    s = _moving(scenes.config0(seed=23), seed=5)
    p, op = _params_pair(sphk, oracle, s, formulation, gas_constant=0.0, gravity=0.0)
    r = _one_step_vs_oracle(sphk, oracle, s, p, op, shape)
    print(f"f={formulation} {shape}: viscosity accel err {r['acc_err'].max():.2e} "
          f"(fp64 {r['acc_err64'].max():.2e}), median |a| {np.median(np.linalg.norm(r['acc_o'], axis=1)):.3e}")
    assert r["acc_err"].max() <= ACC_TOL
    assert r["acc_err64"].max() <= ACC_TOL
    # the term is far from negligible: a sign or factor error would be of order S_i
    mag = np.linalg.norm(r["acc_o"], axis=1) / r["scale"]
    assert np.median(mag) > 1e-2


@pytest.mark.parametrize("shape", ["seg32_waves4", "seg32_waves4_split"])
def test_mixed_pressure_viscosity_moving_state(sphk, oracle, scenes, shape):
    # Reference formulation from a moving state, with a gas constant small enough that the
    # pressure and viscosity sums are of the same size: an error in either shows in S_i.
    s = _moving(scenes.config0(seed=29), seed=6)
    # |P| ~ 2e-4 * 333 Pa: about half of S_i is the viscosity sum (0.48 on this scene)
    p, op = _params_pair(sphk, oracle, s, 0, gas_constant=2e-4)
    r = _one_step_vs_oracle(sphk, oracle, s, p, op, shape)
    # the viscosity share of the scale, from the same run with k = 0
    p0, op0 = _params_pair(sphk, oracle, s, 0, gas_constant=0.0)
    r0 = _one_step_vs_oracle(sphk, oracle, s, p0, op0, shape)
    share = np.median((r0["scale"] - 9.81) / r["scale"])
    print(f"{shape}: accel err {r['acc_err'].max():.2e} (fp64 {r['acc_err64'].max():.2e}); "
          f"viscosity share of S_i {share:.2f}")
    assert 0.2 < share < 0.8
    assert r["acc_err"].max() <= ACC_TOL
    assert r["acc_err64"].max() <= ACC_TOL


def test_reference_step_from_moving_state(sphk, oracle, scenes):
    # the full reference parameters (k = 2000) from a moving state: pressure dominates S_i,
    # but the walks see nonzero neighbour velocities in every pair
    s = _moving(scenes.config0(seed=31), seed=7)
    p, op = _params_pair(sphk, oracle, s, 0)
    r = _one_step_vs_oracle(sphk, oracle, s, p, op, None)
    assert r["rho_rel"].max() <= RHO_TOL
    assert r["acc_err"].max() <= ACC_TOL
    assert r["acc_err64"].max() <= ACC_TOL


def _energies(pos, vel, mass, rho, p):
    m = mass.astype(np.float64)
    v = vel.astype(np.float64)
    ke = float(np.sum(0.5 * m * np.sum(v * v, 1)))
    pe = float(np.sum(m * -float(p.gravity) * (pos[:, 1].astype(np.float64) - float(p.ymin))))
    r = rho.astype(np.float64)
    k, r0 = float(p.gas_constant), float(p.rest_density)
    u = float(np.sum(m * k * (np.log(r / r0) + r0 / r - 1.0)))
    return ke, pe, u


@pytest.mark.parametrize("damping", [0.99, 1.0])
def test_energy_trajectory_matches_oracle(sphk, oracle, scenes, damping):
    # The energy record of the full run (kinetic, potential, internal energy of the equation
    # of state) along 100 consistent-formulation steps of a 1,000-particle block, GPU against
    # the oracle from the same state, with the config's damping and without.
    s = scenes.small_block(10, 10, 10, seed=14)
    s.dt = 1e-4
    p = s.params(sphk, formulation=1)
    p.damping = damping
    op = s.oracle_params(oracle, formulation=1)
    op.damping = damping
    e = sphk.Engine(s.n, p)
    e.upload(s.pos, s.vel)
    o = oracle.Oracle(s.n, op)
    o.set_particles(s.pos, s.vel)
    mass = np.full(s.n, s.mass, np.float32)
    for k in range(1, 101):
        e.step(s.dt)
        o.step(s.dt)
        if k % 25:
            continue
        eg = _energies(e.positions(), e.velocities(), mass, e.densities()[: s.n], p)
        eo = _energies(o.get(oracle.OR_POS), o.get(oracle.OR_VEL), mass,
                       o.get(oracle.OR_DENSITY), op)
        tot_g, tot_o = sum(eg), sum(eo)
        print(f"damping {damping} step {k}: E gpu {tot_g:.6e} oracle {tot_o:.6e}; "
              f"ke {eg[0]:.4e}/{eo[0]:.4e} u {eg[2]:.4e}/{eo[2]:.4e}")
        assert tot_g == pytest.approx(tot_o, rel=1e-4)
        assert eg[0] == pytest.approx(eo[0], rel=1e-3, abs=1e-3 * tot_o)
