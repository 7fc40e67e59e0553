"""The learned macro-step surrogate on the GPU, through the C ABI: forward pass, gradients and
Adam against the CPU oracle (oracle/surrogate_oracle.c), training targets against the
oracle's plain steps, the training loop, the judge and its fallback."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32 = np.float32
BLOCKS = [("W1", 0, 1152), ("b1", 1152, 1280), ("W2", 1280, 17664), ("b2", 17664, 17792),
          ("Wn1", 17792, 34688), ("bn1", 34688, 34816), ("Wn2", 34816, 35200),
          ("bn2", 35200, 35203)]


def _engine(sphk, scene, formulation=1):
    e = sphk.Engine(scene.n, scene.params(sphk, formulation))
    e.upload(scene.pos, scene.vel)
    return e


def _cfg(oracle, e, bf16=True):
    p = e.params()
    c = e.surrogate_config()
    return oracle.sur_cfg(bf16=bf16, inv_r=float(F32(1) / F32(p.neighbor_search_radius)),
                          inv_v=c.inv_v, inv_h=float(F32(1) / F32(p.smoothing_length)),
                          rho0=p.rest_density, rho_scale=c.rho_scale, msg_scale=c.msg_scale,
                          acc_scale=c.acc_scale)


def _moving_block(scenes, seed=3):
    s = scenes.small_block(12, 10, 10, seed=seed)
    s.vel = np.random.default_rng(seed).normal(0, 0.3, s.pos.shape).astype(F32)
    return s


def _active_params(e, seed=0):
    # Xavier weights from the engine plus non-zero biases, so every layer is exercised
    rng = np.random.default_rng(seed)
    p = e.surrogate_params()
    for name, a, b in BLOCKS:
        if name.startswith("b"):
            p[a:b] = rng.normal(0, 0.3, b - a)
    e.surrogate_set_params(p)
    return p


def test_param_count_and_xavier_init(sphk, scenes):
    assert sphk.surrogate_param_count() == 35203
    e = _engine(sphk, scenes.small_block(4, 4, 4))
    e.surrogate_init(seed=11)
    p = e.surrogate_params()
    for name, a, b in BLOCKS:
        blk = p[a:b]
        if name.startswith("b"):
            assert np.all(blk == 0)
        else:
            fan_out, fan_in = {"W1": (128, 9), "W2": (128, 128), "Wn1": (128, 132),
                               "Wn2": (3, 128)}[name]
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            assert np.abs(blk).max() <= lim and np.abs(blk).max() > 0.9 * lim
            # uniform on (-lim, lim): variance lim^2 / 3
            assert np.var(blk) == pytest.approx(lim * lim / 3, rel=0.15)
    e2 = _engine(sphk, scenes.small_block(4, 4, 4))
    e2.surrogate_init(seed=11)
    np.testing.assert_array_equal(e2.surrogate_params(), p)


def test_forward_matches_oracle(sphk, oracle, scenes):
    s = _moving_block(scenes)
    e = _engine(sphk, s)
    e.surrogate_init(seed=7)
    params = _active_params(e)
    pred = e.surrogate_predict()
    np.testing.assert_array_equal(pred, e.surrogate_predict())  # deterministic
    rho = e.surrogate_densities()
    offs, idx = e.neighbors()
    pos, vel = e.positions(), e.velocities()
    cfg = _cfg(oracle, e)
    ref = oracle.sur_forward(cfg, params, pos, vel, rho, offs, idx) * cfg.acc_scale
    rms = np.sqrt(np.mean(ref**2))
    err = np.abs(pred - ref)
    # same bf16 rounding points: only FP32 accumulation order differs (and, rarely, the bf16
    # value an accumulated sum rounds to)
    assert np.linalg.norm(pred - ref) / np.linalg.norm(ref) < 2e-3
    assert np.quantile(err, 0.99) < 1e-3 * rms
    assert err.max() < 2e-2 * rms
    # against exact FP64 arithmetic the error is the bf16 rounding of operands (2^-9 relative)
    ref64 = oracle.sur_forward(_cfg(oracle, e, bf16=False), params, pos, vel, rho, offs,
                               idx) * cfg.acc_scale
    assert np.linalg.norm(pred - ref64) / np.linalg.norm(ref64) < 2e-2


@pytest.mark.parametrize("batch_chunks", [1 << 20, 5])
def test_gradient_matches_oracle(sphk, oracle, scenes, batch_chunks):
    s = _moving_block(scenes, seed=5)
    e = _engine(sphk, s)
    e.surrogate_init(seed=3)
    params = _active_params(e, seed=1)
    target = np.random.default_rng(2).normal(0, 10, (s.n, 3)).astype(F32)
    loss, grad, mask = e.surrogate_loss_grad(target, batch_seed=11, batch_chunks=batch_chunks)
    if batch_chunks > 1000:
        assert mask.all()
    else:
        assert 0 < mask.sum() <= 5 * 64
    rho = e.surrogate_densities()
    offs, idx = e.neighbors()
    cfg = _cfg(oracle, e)
    lo, go = oracle.sur_loss_grad(cfg, params, e.positions(), e.velocities(), rho, offs, idx,
                                  mask, target / cfg.acc_scale)
    assert loss == pytest.approx(lo, rel=1e-4)
    for name, a, b in BLOCKS:
        x, y = grad[a:b].astype(np.float64), go[a:b]
        rel = np.linalg.norm(x - y) / np.linalg.norm(y)
        cos = x @ y / (np.linalg.norm(x) * np.linalg.norm(y))
        assert rel < 5e-3 and cos > 0.9999, (name, rel, cos)


def test_adam_matches_oracle(sphk, oracle, scenes):
    e = _engine(sphk, scenes.small_block(4, 4, 4))
    e.surrogate_init(seed=2)
    w = e.surrogate_params()
    m, v = np.zeros_like(w), np.zeros_like(w)
    rng = np.random.default_rng(4)
    for t in range(1, 4):
        g = rng.normal(0, 10.0 ** -t, w.size).astype(F32)
        e.surrogate_adam(g, lr=1e-3)
        oracle.sur_adam(w, m, v, g, t, lr=1e-3)
    np.testing.assert_allclose(e.surrogate_params(), w, rtol=1e-6, atol=2e-7)


def test_targets_match_oracle_steps(sphk, oracle, scenes):
    s = scenes.small_block(10, 10, 10, seed=2)
    e = _engine(sphk, s)
    e.surrogate_init(seed=1)
    x0, v0 = e.positions(), e.velocities()
    dt = 1e-4
    got = e.surrogate_targets(dt)
    assert e.step_count() == 4 and e.time() == pytest.approx(4 * dt, rel=1e-6)
    o = oracle.Oracle(s.n, s.oracle_params(oracle, 1))
    o.set_particles(x0, v0)
    for _ in range(4):
        o.step(dt)
    d4 = F32(0.99) * F32(0.99) * F32(0.99) * F32(0.99)
    T = F32(dt) * F32(4)
    want = (o.get(1) - d4 * v0) * (F32(1) / (T * F32(10))) * F32(10)
    rms = np.sqrt(np.mean(want**2))
    # the plain steps agree with the oracle to FP32 rounding; T divides those differences
    np.testing.assert_allclose(got, want, rtol=0, atol=5e-5 * rms)
    np.testing.assert_allclose(e.positions(), o.get(0), rtol=0, atol=1e-6)


def test_training_reduces_loss(sphk, scenes):
    s = scenes.small_block(16, 12, 12, seed=4)
    e = _engine(sphk, s)
    e.surrogate_init(seed=5)
    curve = e.surrogate_train(snapshots=60, dt=1e-4, batch_chunks=16, iters_per_snapshot=4,
                              seed=1)
    assert np.all(np.isfinite(curve))
    assert curve[-10:].mean() < 0.1 * curve[:5].mean()
    assert e.step_count() == 60 * 4
    assert e.time() == pytest.approx(60 * 4 * 1e-4, rel=1e-4)


def test_judge_keeps_plain_result_and_falls_back(sphk, scenes):
    s = scenes.small_block(10, 8, 8, seed=6)
    dt = 1e-4
    # tolerance 0: the first judge run (macro-step 1) measures a nonzero divergence, so the
    # bound would pass the tolerance before the next judge run: plain SPH from then on
    e = _engine(sphk, s)
    e.surrogate_init(seed=9, config=sphk.surrogate_default_config(judge_every=2, tolerance=0.0))
    rep = e.surrogate_run(6, dt)
    assert (rep.macro_steps, rep.plain_macro_steps, rep.judges) == (0, 6, 1)
    assert rep.fell_back == 1 and rep.fallback_at == 1 and rep.last_error > 0
    assert rep.error_bound == 0.0  # no network step was kept
    assert e.step_count() == 24
    # every kept macro-step was plain: the state equals 24 plain steps
    e2 = _engine(sphk, s)
    for _ in range(24):
        e2.step(dt)
    np.testing.assert_array_equal(e.positions(), e2.positions())
    np.testing.assert_array_equal(e.velocities(), e2.velocities())
    assert e.time() == e2.time()
    # a loose tolerance never falls back: judge runs at macro-steps 1 and 5, the network
    # runs the others; the bound is (4/3) e_max (N - 1)^(3/2) with N = 8
    e3 = _engine(sphk, s)
    e3.surrogate_init(seed=9, config=sphk.surrogate_default_config(judge_every=4,
                                                                  tolerance=1e30))
    rep = e3.surrogate_run(8, dt)
    assert (rep.macro_steps, rep.plain_macro_steps, rep.judges, rep.fell_back) == (6, 2, 2, 0)
    assert rep.max_error >= rep.last_error > 0
    assert rep.error_bound == pytest.approx(4 / 3 * rep.max_error * 7 ** 1.5, rel=1e-5)
    assert e3.step_count() == 32
    # the same sequence by hand: plain, 3 network, plain, 3 network
    e4 = _engine(sphk, s)
    e4.surrogate_init(seed=9)
    for block in range(2):
        for _ in range(4):
            e4.step(dt)
        for _ in range(3):
            e4.surrogate_macro_step(dt)
    np.testing.assert_array_equal(e3.positions(), e4.positions())
    np.testing.assert_array_equal(e3.velocities(), e4.velocities())
    # a tolerance just below that bound: the run falls back at the judge run that measures
    # the largest divergence (1 or 5)
    e5 = _engine(sphk, s)
    e5.surrogate_init(seed=9, config=sphk.surrogate_default_config(
        judge_every=4, tolerance=float(0.999 * rep.error_bound)))
    rep5 = e5.surrogate_run(8, dt)
    assert rep5.fell_back == 1 and rep5.fallback_at in (1, 5)
    if rep5.fallback_at == 1:
        assert rep5.macro_steps == 0 and rep5.error_bound == 0.0
    else:
        assert rep5.macro_steps == 3
        assert rep5.error_bound == pytest.approx(rep.error_bound, rel=1e-5)


def test_errors_without_init_and_bad_arguments(sphk, scenes):
    e = _engine(sphk, scenes.small_block(4, 4, 4))
    with pytest.raises(sphk.SphkError):
        e.surrogate_predict()
    e.surrogate_init(seed=0)
    with pytest.raises(sphk.SphkError):
        e.surrogate_set_params(np.zeros(10, np.float32))
    with pytest.raises(sphk.SphkError):
        e.surrogate_macro_step(0.0)
    with pytest.raises(sphk.SphkError):
        e.surrogate_init(seed=0, config=sphk.surrogate_default_config(substeps=0))


def test_trained_network_beats_the_method_without_it(sphk, scenes):
    # Ablation: from a state reached by the engine's own trajectory, 8 macro-steps of the
    # trained network track 32 plain steps far better than the same macro-step integrator
    # with an untrained network or with gravity alone.  (Plain SPH at the macro-step dt, 2.2x
    # the acoustic CFL limit here, is measured alongside for the record.)
    s = scenes.small_block(20, 16, 16, seed=7)
    dt = 1e-4
    e = _engine(sphk, s)
    e.surrogate_init(seed=3)
    untrained = e.surrogate_params()
    curve = e.surrogate_train(snapshots=1000, dt=dt, batch_chunks=64, iters_per_snapshot=2,
                              seed=2)
    assert curve[-20:].mean() < 0.05 * curve[:5].mean()
    trained = e.surrogate_params()
    e.snapshot_save()
    for _ in range(32):
        e.step(dt)
    rho_ref = e.surrogate_densities()

    def err_after_macro(params):
        e.snapshot_restore()
        e.surrogate_set_params(params)
        for _ in range(8):
            e.surrogate_macro_step(dt)
        return float(np.abs(e.surrogate_densities() - rho_ref).mean())

    gravity = np.zeros_like(trained)
    gravity[BLOCKS[-1][1] + 1] = -9.81 / e.surrogate_config().acc_scale  # out_y = g / a_s
    e_trained = err_after_macro(trained)
    e_untrained = err_after_macro(untrained)
    e_gravity = err_after_macro(gravity)
    e.snapshot_restore()
    for _ in range(8):
        e.step(4 * dt)
    rho_big = e.surrogate_densities()
    e_big = float(np.nan_to_num(np.abs(rho_big - rho_ref), nan=1e30).mean())
    errs = dict(trained=e_trained, untrained=e_untrained, gravity=e_gravity, plain_4dt=e_big)
    print("mean |rho - rho_plain| after 8 macro-steps:", errs)
    assert e_trained < 0.5 * e_gravity, errs
    assert e_trained < 0.5 * e_untrained, errs
    assert e_trained < 10.0, errs  # 1% of the rest density
    # the main run's error bound holds: 8 macro-steps, judge runs at 1 and 5, against 32
    # plain steps from the same state (the bound assumes errors that add up like a random
    # walk; damping and the density's smoothing make them grow more slowly)
    e.snapshot_restore()
    e.surrogate_init(seed=3, config=sphk.surrogate_default_config(judge_every=4,
                                                                 tolerance=1e30))
    e.surrogate_set_params(trained)
    rep = e.surrogate_run(8, dt)
    err_run = float(np.abs(e.surrogate_densities() - rho_ref).mean())
    print(f"judged run: true error {err_run:.3f} kg/m^3, bound {rep.error_bound:.3f}")
    assert rep.judges == 2 and rep.macro_steps == 6
    assert err_run <= rep.error_bound
