"""Pins the CPU restatement of the surrogate network (oracle/surrogate_oracle.c): bf16
rounding, the hoisted message layer against a per-edge evaluation, the backward pass
against central finite differences, and Adam against its closed form."""
import numpy as np
import pytest


def _scene(oracle, n=40, seed=3, box=0.06):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    vel = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    rho = (1000 + rng.normal(0, 30, n)).astype(np.float32)
    offs, idx = oracle.brute_neighbors(pos, np.float32(0.04) * np.float32(0.04))
    return pos, vel, rho, offs, idx


def _params(oracle, cfg, seed=5, scale=0.6):
    rng = np.random.default_rng(seed)
    return (rng.normal(0, scale, oracle.sur_nparams(cfg))).astype(np.float32)


def test_bf16_round_matches_c_and_ties(oracle):
    lib = oracle.load()
    rng = np.random.default_rng(0)
    x = rng.normal(0, 10, 2000).astype(np.float32)
    x = np.concatenate([x, np.float32([0.0, -0.0, 1e-40, np.inf, -np.inf, 3.4e38])])
    got = np.array([lib.or_bf16_round(float(v)) for v in x], np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), oracle.bf16_round(x).view(np.uint32))
    # ties go to even: bf16 spacing at 1 is 2^-7
    f = np.float32
    assert oracle.bf16_round(f(1 + 2**-8)) == f(1.0)
    assert oracle.bf16_round(f(1 + 3 * 2**-8)) == f(1 + 2**-6)
    assert oracle.bf16_round(f(1 + 2**-8 + 2**-20)) == f(1 + 2**-7)
    assert np.isnan(lib.or_bf16_round(float("nan")))


def test_param_count(oracle):
    # W1 128x9 + b1 + W2 128x128 + b2 + Wn1 128x132 + bn1 + Wn2 3x128 + bn2
    assert oracle.sur_nparams(oracle.sur_cfg()) == 35203


def _per_edge_reference(oracle, cfg, params, pos, vel, rho, offs, idx):
    """Unhoisted evaluation: the second message layer applied to every edge, then summed."""
    h1, h2, h3 = cfg.h1, cfg.h2, cfg.h3
    p = params.astype(np.float64)
    k = 0

    def take(shape):
        nonlocal k
        size = int(np.prod(shape))
        out = p[k:k + size].reshape(shape)
        k += size
        return out

    W1, b1 = take((h1, 9)), take((h1,))
    W2, b2 = take((h2, h1)), take((h2,))
    Wn1, bn1 = take((h3, h2 + 4)), take((h3,))
    Wn2, bn2 = take((3, h3)), take((3,))
    n = pos.shape[0]
    out = np.zeros((n, 3))
    f32 = np.float32
    for i in range(n):
        m = np.zeros(h2)
        for t in range(offs[i], offs[i + 1]):
            j = idx[t]
            if j == i:
                continue
            r = pos[i] - pos[j]
            d2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]
            rl = np.sqrt(f32(d2))
            q = f32(rl * f32(cfg.inv_h))
            if q <= 1:
                w = (f32(2 / 3) - q * q) + f32(0.5) * q * q * q
            elif q <= 2:
                w = f32(1 / 6) * (f32(2) - q) ** 3
            else:
                w = f32(0)
            e = np.concatenate([r * f32(cfg.inv_r), (vel[i] - vel[j]) * f32(cfg.inv_v),
                                [rl * f32(cfg.inv_r), f32(w) * f32(cfg.w_norm),
                                 (rho[j] - f32(cfg.rho0)) * f32(cfg.rho_scale)]]).astype(np.float64)
            m += cfg.msg_scale * (W2 @ np.maximum(W1 @ e + b1, 0.0) + b2)
        x = np.concatenate([m, [(rho[i] - f32(cfg.rho0)) * f32(cfg.rho_scale)],
                            vel[i] * f32(cfg.inv_v)])
        z = np.maximum(Wn1 @ x + bn1, 0.0)
        out[i] = Wn2 @ z + bn2
    return out


def test_hoisted_message_layer_matches_per_edge(oracle):
    cfg = oracle.sur_cfg(h1=6, h2=5, h3=7, bf16=False)
    pos, vel, rho, offs, idx = _scene(oracle)
    params = _params(oracle, cfg)
    got = oracle.sur_forward(cfg, params, pos, vel, rho, offs, idx)
    want = _per_edge_reference(oracle, cfg, params, pos, vel, rho, offs, idx)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_forward_permutation_equivariant(oracle):
    cfg = oracle.sur_cfg(h1=8, h2=8, h3=8, bf16=True)
    pos, vel, rho, offs, idx = _scene(oracle, n=60, seed=9)
    params = _params(oracle, cfg, seed=2)
    out = oracle.sur_forward(cfg, params, pos, vel, rho, offs, idx)
    perm = np.random.default_rng(1).permutation(pos.shape[0])
    po, pi_ = oracle.brute_neighbors(pos[perm], np.float32(0.04) * np.float32(0.04))
    out_p = oracle.sur_forward(cfg, params, pos[perm], vel[perm], rho[perm], po, pi_)
    np.testing.assert_allclose(out_p, out[perm], rtol=1e-12, atol=1e-12)


def test_forward_closed_form_without_edge_weights(oracle):
    # W1 = 0: u = b1 for every edge, S = c * n * ReLU(b1)
    cfg = oracle.sur_cfg(h1=4, h2=3, h3=2, bf16=False)
    pos, vel, rho, offs, idx = _scene(oracle, n=30, seed=4)
    params = _params(oracle, cfg, seed=8)
    params[:4 * 9] = 0.0
    _, S, cnt = oracle.sur_forward(cfg, params, pos, vel, rho, offs, idx, want_S=True)
    b1 = params[36:40].astype(np.float64)
    n_edges = np.diff(offs) - 1  # self excluded
    np.testing.assert_array_equal(cnt, n_edges)
    np.testing.assert_allclose(S, cfg.msg_scale * n_edges[:, None] * np.maximum(b1, 0)[None, :],
                               rtol=1e-12)


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_central_differences(oracle, seed):
    cfg = oracle.sur_cfg(h1=5, h2=4, h3=6, bf16=False)
    pos, vel, rho, offs, idx = _scene(oracle, n=36, seed=seed + 20)
    params = _params(oracle, cfg, seed=seed + 30)
    rng = np.random.default_rng(seed)
    target = rng.normal(0, 1, (pos.shape[0], 3)).astype(np.float32)
    in_batch = (rng.random(pos.shape[0]) < 0.7).astype(np.uint8)
    loss, grad = oracle.sur_loss_grad(cfg, params, pos, vel, rho, offs, idx, in_batch, target)
    assert loss > 0
    picks = np.arange(params.size)  # every parameter of the small net
    fd = np.empty(picks.size)
    for t, k in enumerate(picks):
        eps = np.float32(2.0**-14) * max(np.float32(1.0), abs(params[k]))
        pp = params.copy()
        pm = params.copy()
        pp[k] = params[k] + eps
        pm[k] = params[k] - eps
        lp, _ = oracle.sur_loss_grad(cfg, pp, pos, vel, rho, offs, idx, in_batch, target)
        lm, _ = oracle.sur_loss_grad(cfg, pm, pos, vel, rho, offs, idx, in_batch, target)
        fd[t] = (lp - lm) / (float(pp[k]) - float(pm[k]))
    err = np.abs(fd - grad[picks])
    # central differences of a piecewise quadratic: exact up to the odd ReLU kinks crossed
    scale = np.abs(grad).max()
    assert np.median(err) <= 1e-7 * scale
    assert np.mean(err <= 1e-6 * scale + 1e-6 * np.abs(grad[picks])) >= 0.99


def test_loss_zero_at_own_output(oracle):
    cfg = oracle.sur_cfg(h1=6, h2=6, h3=6, bf16=False)
    pos, vel, rho, offs, idx = _scene(oracle, n=30, seed=11)
    params = _params(oracle, cfg, seed=12)
    out = oracle.sur_forward(cfg, params, pos, vel, rho, offs, idx)
    # a target equal to the float32 output leaves only the rounding of the target
    loss, grad = oracle.sur_loss_grad(cfg, params, pos, vel, rho, offs, idx,
                                      np.ones(pos.shape[0], np.uint8), out.astype(np.float32))
    assert loss < 1e-12 * max(1.0, float(np.mean(out**2)))


def test_adam_closed_form_under_constant_gradient(oracle):
    # with a constant gradient g the bias-corrected moments are exactly g and g^2, so
    # every step moves w by lr * g / (|g| + eps)
    rng = np.random.default_rng(3)
    n = 1000
    w0 = rng.normal(0, 1, n).astype(np.float32)
    g = rng.normal(0, 1e-2, n).astype(np.float32)
    w, m, v = w0.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)
    lr, eps, steps = 1e-3, 1e-8, 7
    for t in range(1, steps + 1):
        oracle.sur_adam(w, m, v, g, t, lr=lr, eps=eps)
    g64 = g.astype(np.float64)
    want = w0 - steps * lr * g64 / (np.abs(g64) + eps)
    np.testing.assert_allclose(w, want, rtol=0, atol=2e-6)
    # first step from zero moments: |update| ~ lr regardless of |g|
    w1, m1, v1 = w0.copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)
    oracle.sur_adam(w1, m1, v1, g * 100, 1, lr=lr, eps=eps)
    np.testing.assert_allclose(np.abs(w1 - w0), lr, rtol=1e-3)
