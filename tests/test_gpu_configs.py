"""The workloads of the benchmark configurations at their full sizes, on the GPU, against the
oracle.

Scenes of millions of particles are too large for the oracle's per-particle neighbour vectors
(config 2 would need ~20 GB of lists), so the oracle is evaluated on a seeded sample
(Oracle.sample_step0: the hash over every particle, the query of the sampled particles and of
their neighbours, whose densities the forces read; pinned against the full phases in
tests/test_oracle.py), and the neighbour-count totals of every particle come from its
count-only query.  Bars as in test_gpu_parity.py: neighbour counts exact, density <= 1e-5
relative, acceleration <= 1e-5 of S_i, integration bit-exact given the GPU's accelerations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RHO_TOL = 1e-5
ACC_TOL = 1e-5
POOL = 400 * 400 * 60  # particles of the pool lattice in the drop-into-pool scene


def _check_sample(e, o, scene, sample, params, label):
    r = o.sample_step0(sample)
    rho_g = e.densities()[sample].astype(np.float64)
    acc_g = e.accelerations()[sample].astype(np.float64)
    rel = np.abs(rho_g - r["rho"]) / np.abs(r["rho"])
    rel64 = np.abs(rho_g - r["rho64"]) / r["rho64"]
    err = np.linalg.norm(acc_g - r["acc"], axis=1) / r["scale"]
    err64 = np.linalg.norm(acc_g - r["acc64"], axis=1) / r["scale"]
    print(f"{label}: {len(sample)} sampled particles, neighbours {r['counts'].min()}.."
          f"{r['counts'].max()}; rho rel {rel.max():.2e} (fp64 {rel64.max():.2e}); "
          f"accel {err.max():.2e} (fp64 {err64.max():.2e})", flush=True)
    assert rel.max() <= RHO_TOL, label
    assert rel64.max() <= RHO_TOL, label
    assert err.max() <= ACC_TOL, label
    assert err64.max() <= ACC_TOL, label
    return r


def _run_config2(sphk, oracle, s, sample):
    e = sphk.Engine(s.n, s.params(sphk))
    e.upload(s.pos, s.vel)
    e.set_timing(True)
    e.step(s.dt)
    st = e.stats()
    params = s.oracle_params(oracle)
    o = oracle.Oracle(s.n, params)
    o.set_particles(s.pos, s.vel)
    tot, mx = o.count_neighbors()
    print(f"{s.name}: {s.n} particles, neighbour total {tot} (mean {tot / s.n:.2f}), max {mx}",
          flush=True)
    assert st.total_neighbors == tot
    assert st.max_neighbors == mx
    assert st.total_neighbor_queries == s.n
    r = _check_sample(e, o, s, sample, params, s.name)
    # leapfrog + wall clamp of the sample, given the GPU's accelerations: bit-exact
    x1, v1 = oracle.integrate_clamp(s.pos[sample], s.vel[sample], e.accelerations()[sample],
                                    s.dt, params)
    np.testing.assert_array_equal(e.positions()[sample], x1)
    np.testing.assert_array_equal(e.velocities()[sample], v1)
    return r


def _pool_sample(s, rng, k_bulk=500, k_sphere=200):
    n_sphere = s.n - POOL
    y = s.pos[:POOL, 1]
    surface = np.flatnonzero(y > 0.585)  # the pool's top lattice layer
    wall = np.flatnonzero(s.pos[:POOL, 0] < 0.105)  # its first layer along x
    return np.unique(np.concatenate([
        rng.choice(POOL, k_bulk, replace=False),
        POOL + rng.choice(n_sphere, k_sphere, replace=False),
        rng.choice(surface, 100, replace=False),
        rng.choice(wall, 100, replace=False)]))


def test_config2_drop_into_pool_fullsize(sphk, oracle, scenes):
    # 9,614,147 particles: the 400 x 400 x 60 pool and the falling sphere, Gaussian jitter
    s = scenes.config2()
    assert s.n == 9_614_147
    sample = _pool_sample(s, np.random.default_rng(20))
    _run_config2(sphk, oracle, s, sample)


def test_config2_sphere_impact_fullsize(sphk, oracle, scenes):
    # The same scene at impact: the sphere lowered into the pool's surface, so the lattices
    # overlap and the neighbour counts around the splash about double (uneven chunks, long
    # candidate columns).  The sample concentrates on the overlap.
    s = scenes.config2()
    s.pos[POOL:, 1] -= np.float32(0.45)
    rng = np.random.default_rng(21)
    c = np.array([2.1, 0.55, 2.1], np.float32)
    near = np.flatnonzero(np.linalg.norm(s.pos - c, axis=1) < 0.17)
    sample = np.unique(np.concatenate([_pool_sample(s, rng, 300, 100),
                                       rng.choice(near, 400, replace=False)]))
    r = _run_config2(sphk, oracle, s, sample)
    assert r["counts"].max() > 400  # the overlap really is a pile-up


def test_config3_ensemble_fullsize(sphk, oracle, scenes):
    # 64 scenes of 10k..500k particles in one engine: per-scene results equal single-scene
    # engines bit for bit (smallest, largest and two seeded scenes), each against the oracle
    # on a sample, and the neighbour total of all 64 scenes equals the oracle's.
    ens = scenes.config3(seed=0)
    assert len(ens) == 64
    offs = np.cumsum([0] + [sc.n for sc in ens]).astype(np.uint64)
    pos = np.concatenate([sc.pos for sc in ens])
    n = int(offs[-1])
    eb = sphk.Engine(n, ens[0].params(sphk))
    eb.upload(pos)
    eb.set_scenes(offs, np.array([sc.bounds for sc in ens], np.float32))
    eb.set_graph(True)
    eb.set_timing(True)
    eb.step(0.001)
    st = eb.stats()
    rho0, acc0 = eb.densities(), eb.accelerations()
    tot = 0
    for sc in ens:
        o = oracle.Oracle(sc.n, sc.oracle_params(oracle))
        o.set_particles(sc.pos, sc.vel)
        tot += o.count_neighbors()[0]
        del o
    print(f"ensemble: {n} particles in 64 scenes, neighbour total {tot}", flush=True)
    assert st.total_neighbors == tot
    for _ in range(2):
        eb.step(0.001)
    pb, vb, rb = eb.positions(), eb.velocities(), eb.densities()
    sizes = np.array([sc.n for sc in ens])
    rng = np.random.default_rng(5)
    pick = [int(sizes.argmin()), int(sizes.argmax())]
    pick += [int(k) for k in rng.choice(np.setdiff1d(np.arange(64), pick), 2, replace=False)]
    for k in pick:
        sc = ens[k]
        a, b = int(offs[k]), int(offs[k + 1])
        e = sphk.Engine(sc.n, sc.params(sphk))
        e.upload(sc.pos, sc.vel)
        for _ in range(3):
            e.step(0.001)
        np.testing.assert_array_equal(pb[a:b], e.positions())
        np.testing.assert_array_equal(vb[a:b], e.velocities())
        np.testing.assert_array_equal(rb[a:b], e.densities()[: sc.n])
        del e
        # step 0 of the batched engine against the oracle on a sample of this scene
        params = sc.oracle_params(oracle)
        o = oracle.Oracle(sc.n, params)
        o.set_particles(sc.pos, sc.vel)
        sample = rng.choice(sc.n, min(300, sc.n), replace=False)
        r = o.sample_step0(sample)
        rel = np.abs(rho0[a:b][sample] - r["rho"]) / r["rho"]
        err = np.linalg.norm(acc0[a:b][sample].astype(np.float64) - r["acc"], axis=1) / r["scale"]
        err64 = np.linalg.norm(acc0[a:b][sample] - r["acc64"], axis=1) / r["scale"]
        print(f"scene {k} ({sc.n} particles): rho rel {rel.max():.2e}, accel {err.max():.2e} "
              f"(fp64 {err64.max():.2e})", flush=True)
        assert rel.max() <= RHO_TOL
        assert err.max() <= ACC_TOL and err64.max() <= ACC_TOL


def _sur_cfg(oracle, e, bf16=True):
    p = e.params()
    c = e.surrogate_config()
    f32 = np.float32
    return oracle.sur_cfg(bf16=bf16, inv_r=float(f32(1) / f32(p.neighbor_search_radius)),
                          inv_v=c.inv_v, inv_h=float(f32(1) / f32(p.smoothing_length)),
                          rho0=p.rest_density, rho_scale=c.rho_scale, msg_scale=c.msg_scale,
                          acc_scale=c.acc_scale)


def test_config4_surrogate_forward_fullsize(sphk, oracle, scenes):
    # The surrogate's network on the 1M dam break (consistent formulation, moving state):
    # its acceleration on a sample of particles against the oracle's forward pass over the
    # oracle's own neighbour lists of those particles (same bf16 rounding points).
    s = scenes.config1()
    s.vel = np.random.default_rng(8).normal(0, 0.3, s.pos.shape).astype(np.float32)
    e = sphk.Engine(s.n, s.params(sphk, 1))
    e.upload(s.pos, s.vel)
    e.surrogate_init(seed=4)
    params = e.surrogate_params()
    rng = np.random.default_rng(9)
    for a, b in ((1152, 1280), (17664, 17792), (34688, 34816), (35200, 35203)):  # biases
        params[a:b] = rng.normal(0, 0.3, b - a)
    e.surrogate_set_params(params)
    pred = e.surrogate_predict()
    rho = e.surrogate_densities()
    pos, vel = e.positions(), e.velocities()
    sample = rng.choice(s.n, 2000, replace=False)
    o = oracle.Oracle(s.n, s.oracle_params(oracle, 1))
    o.set_particles(pos, vel)
    o.sample_step0(sample)
    so, si = o.sample_neighbors(sample)
    # local arrays: the sample first, then the other particles its lists reach
    others = np.setdiff1d(np.unique(si), sample)
    glob = np.concatenate([sample, others])
    local = np.full(s.n, -1, np.int64)
    local[glob] = np.arange(len(glob))
    idx_l = local[si]
    offs_l = np.concatenate([so, np.full(len(others), so[-1], np.int64)])
    cfg = _sur_cfg(oracle, e)
    ref = oracle.sur_forward(cfg, params, pos[glob], vel[glob], rho[glob], offs_l,
                             idx_l)[: len(sample)] * cfg.acc_scale
    got = pred[sample]
    rms = np.sqrt(np.mean(ref ** 2))
    err = np.abs(got - ref)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"surrogate forward at 1M: {len(sample)} sampled, rel {rel:.2e}, "
          f"p99 {np.quantile(err, 0.99) / rms:.2e} of rms", flush=True)
    assert rel < 2e-3
    assert np.quantile(err, 0.99) < 1e-3 * rms
    assert err.max() < 2e-2 * rms


def test_config4_judge_and_fallback_fullsize(sphk, scenes):
    # The judge on the 1M scene.  A zero tolerance switches to plain SPH at the first judge
    # run (the state is then 12 plain steps, bit for bit); a loose one keeps the network
    # between judge runs (the state equals plain, network, plain macro-steps by hand).
    s = scenes.config1()
    dt = 1e-4
    e = sphk.Engine(s.n, s.params(sphk, 1))
    e.upload(s.pos, s.vel)
    e.surrogate_init(seed=9, config=sphk.surrogate_default_config(judge_every=2, tolerance=0.0))
    rep = e.surrogate_run(3, dt)
    assert (rep.macro_steps, rep.plain_macro_steps, rep.judges) == (0, 3, 1)
    assert rep.fell_back == 1 and rep.fallback_at == 1 and rep.last_error > 0
    e2 = sphk.Engine(s.n, s.params(sphk, 1))
    e2.upload(s.pos, s.vel)
    for _ in range(12):
        e2.step(dt)
    np.testing.assert_array_equal(e.positions(), e2.positions())
    np.testing.assert_array_equal(e.velocities(), e2.velocities())
    del e2
    e.upload(s.pos, s.vel)
    e.surrogate_init(seed=9, config=sphk.surrogate_default_config(judge_every=2,
                                                                 tolerance=1e30))
    rep = e.surrogate_run(3, dt)
    assert (rep.macro_steps, rep.plain_macro_steps, rep.judges, rep.fell_back) == (1, 2, 2, 0)
    assert rep.error_bound == pytest.approx(4 / 3 * rep.max_error * 2 ** 1.5, rel=1e-5)
    e3 = sphk.Engine(s.n, s.params(sphk, 1))
    e3.upload(s.pos, s.vel)
    e3.surrogate_init(seed=9)
    for _ in range(4):
        e3.step(dt)
    e3.surrogate_macro_step(dt)
    for _ in range(4):
        e3.step(dt)
    np.testing.assert_array_equal(e.positions(), e3.positions())
    np.testing.assert_array_equal(e.velocities(), e3.velocities())
