"""The branch-free spline forms the walks evaluate, against the reference's two-branch
cubic spline (CubicSplineKernel::W_3d, gradW_3d, laplacianW_3d: src/kernels.cpp:58-66,
96-108, 130-138), in FP32.

The density walk sums an eighth of 6W/sigma as u^3 + s^2 (-s/2) with u = clamp(1 - q/2),
s = clamp(1 - q) and scales the sum by 8 sigma / 6; the force walk uses (s - u)(s + u) for
dW/dq and u - 2 s for the Laplacian term, the constant factors riding on the per-particle
scales.  These tests check the rewrites on a dense sweep of r (FMA emulated by one FP64
operation rounded to FP32, the same rounding for both sides of an identity).
"""
import numpy as np

F = np.float32


def fma32(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F)


def clamp01(x):
    return np.minimum(np.maximum(x, F(0)), F(1))


H = F(0.02)
INV_H = F(1) / H
R = np.linspace(0.0, 2.3 * 0.02, 200_001, dtype=np.float64).astype(F)


def t_form(r):
    t = np.maximum(fma32(-INV_H * np.ones_like(r), r, F(2) * np.ones_like(r)), F(0))
    s = np.maximum(fma32(-INV_H * np.ones_like(r), r, np.ones_like(r)), F(0))
    return fma32(s * s, F(-4) * s, t * t * t)


def u_form(r):
    one = np.ones_like(r)
    u = clamp01(fma32(F(-0.5) * INV_H * one, r, one))
    s = clamp01(fma32(-INV_H * one, r, one))
    return fma32(s * s, F(-0.5) * s, u * u * u)


def test_density_u_form_is_the_t_form_over_8_bit_for_bit():
    t8 = t_form(R) / F(8)
    u = u_form(R)
    assert np.array_equal(t8.view(np.uint32), u.view(np.uint32))


def test_density_scale_is_eight_sixths_exactly():
    assert F(8.0) / F(6.0) == F(8) * (F(1) / F(6))


def test_t_form_is_the_reference_cubic_spline():
    q = (R.astype(np.float64) / float(H))
    ref = np.where(q <= 1, 1 - 1.5 * q**2 + 0.75 * q**3, np.where(q <= 2, 0.25 * (2 - q)**3, 0.0))
    # 6 W / sigma = 4 x the reference's bracket
    np.testing.assert_allclose(t_form(R).astype(np.float64), 4 * ref, rtol=0, atol=4e-6)


def test_force_forms_match_the_two_branch_expressions():
    one = np.ones_like(R)
    u = clamp01(fma32(F(-0.5) * INV_H * one, R, one))
    s = clamp01(fma32(-INV_H * one, R, one))
    dwq = (s - u) * (s + u)
    lap = fma32(F(-2) * one, s, u)
    q = R.astype(np.float64) / float(H)
    # the reference's brackets: gradW_3d sigma (-2q + 1.5q^2) and -sigma 0.5 (2 - q)^2 (times
    # r / (|r| h)); laplacianW_3d sigma (-2 + 3q) and sigma (2 - q) (over h^2).  The walk's
    # forms are half of each; grad_scale and lap_scale carry the 2.
    dw_ref = np.where(q <= 1, -2 * q + 1.5 * q * q, np.where(q <= 2, -0.5 * (2 - q)**2, 0.0)) / 2
    lap_ref = np.where(q <= 1, -2 + 3 * q, np.where(q <= 2, 2 - q, 0.0)) / 2
    # s and u carry one rounding each (2^-25 absolute near 1): the difference s - u loses
    # relative precision only as q -> 0, where dW/dq itself vanishes
    np.testing.assert_allclose(dwq.astype(np.float64), dw_ref, rtol=0, atol=2e-7)
    np.testing.assert_allclose(lap.astype(np.float64), lap_ref, rtol=0, atol=2e-7)
    # beyond the support both are exactly zero
    beyond = q >= 2.0000001
    assert np.all(dwq[beyond] == 0) and np.all(lap[beyond] == 0)
    # dwq is the derivative of the density walk's summand: d/dq (u^3 - s^3 / 2) =
    # 1.5 (s^2 - u^2) (central differences of the FP32 summand)
    # (every 200th point: steps of ~2e-3 in q keep the FP32 rounding noise below 1e-4)
    rs, qs = R[::200], q[::200]
    w8 = u_form(rs).astype(np.float64)
    dq = np.gradient(w8, qs)
    inner = (qs > 0.05) & (qs < 1.95) & (np.abs(qs - 1) > 0.01)
    np.testing.assert_allclose(dq[inner], 1.5 * dw_ref[::200][inner], rtol=0, atol=5e-4)
