"""The CPU oracle of the surrogate network, pinned by finite differences:
parameter gradients of the loss and position gradients of the outputs."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as MO  # noqa: E402


def _setup(seed=0, n=40):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.15, MO.NPARAM)
    freq = rng.normal(0, 1.5, (32, 3))
    centre = rng.normal(0, 1, 3)
    xyz = rng.uniform(-20, 20, (n, 3))
    e = rng.normal(0, 2, (n, 2))
    return p, freq, centre, 1 / 25.0, np.array([1.7, 0.6]), xyz, e


def test_loss_gradient_matches_finite_differences():
    p, freq, centre, s, osc, xyz, e = _setup()
    loss, g = MO.loss_grad(p, xyz, e, freq, centre, s, osc)
    rng = np.random.default_rng(3)
    # a sample of parameters from every block (weights and biases of each layer)
    blocks = [(MO.W1, MO.B1), (MO.B1, MO.W2), (MO.W2, MO.B2), (MO.B2, MO.W3), (MO.W3, MO.B3),
              (MO.B3, MO.W4), (MO.W4, MO.B4), (MO.B4, MO.NPARAM)]
    for lo, hi in blocks:
        for i in rng.integers(lo, hi, size=3):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (MO.loss_grad(pp, xyz, e, freq, centre, s, osc)[0]
                  - MO.loss_grad(pm, xyz, e, freq, centre, s, osc)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_position_gradient_matches_finite_differences():
    p, freq, centre, s, osc, xyz, _ = _setup(seed=1, n=8)
    y, dy = MO.eval_with_grad(p, xyz, freq, centre, s, osc)
    h = 1e-5
    for d in range(3):
        xp, xm = xyz.copy(), xyz.copy()
        xp[:, d] += h
        xm[:, d] -= h
        fd = (MO.eval_with_grad(p, xp, freq, centre, s, osc)[0]
              - MO.eval_with_grad(p, xm, freq, centre, s, osc)[0]) / (2 * h)
        assert np.allclose(fd, dy[:, :, d], rtol=1e-6, atol=1e-8)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], dtype=np.float32)
    r = MO.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie to even (mantissa ...0)
    assert r[2] == 1.0 + 2 ** -6                 # tie to even (mantissa ...1 -> up)
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def _jvp_setup(seed=4, n=30):
    p, freq, centre, s, osc, xyz, e = _setup(seed=seed, n=n)
    rng = np.random.default_rng(seed + 100)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    d = rng.normal(0, 0.3, (n, 2))
    return p, freq, centre, s, osc, xyz, e, dirs, d


def test_forward_mode_derivative_matches_finite_differences():
    """The tangent rows carry the outputs' derivative along each point's
    direction: equal to central differences of the outputs along it."""
    p, freq, centre, s, osc, xyz, _, dirs, _ = _jvp_setup()
    y, yd, _ = MO.jvp_forward(p, xyz, dirs, freq, centre, s)
    y0, _ = MO.forward(p, xyz, freq, centre, s)
    assert np.array_equal(y, y0)
    h = 1e-5
    fd = (MO.jvp_forward(p, xyz + h * dirs, dirs, freq, centre, s)[0]
          - MO.jvp_forward(p, xyz - h * dirs, dirs, freq, centre, s)[0]) / (2 * h)
    assert np.allclose(fd, yd, rtol=1e-6, atol=1e-9 * np.max(np.abs(yd)))


def test_forward_mode_loss_gradient_matches_finite_differences():
    """Gradient of the value + directional-derivative loss with respect to
    every parameter block, against central differences of the loss; with the
    derivative weight at zero it is the value-only loss's gradient."""
    p, freq, centre, s, osc, xyz, e, dirs, d = _jvp_setup()
    w = 3.0
    loss, g, _, _ = MO.jvp_loss_grad(p, xyz, dirs, e, d, w, freq, centre, s, osc)
    rng = np.random.default_rng(5)
    blocks = [(MO.W1, MO.B1), (MO.B1, MO.W2), (MO.W2, MO.B2), (MO.B2, MO.W3), (MO.W3, MO.B3),
              (MO.B3, MO.W4), (MO.W4, MO.B4), (MO.B4, MO.NPARAM)]
    for lo, hi in blocks:
        for i in rng.integers(lo, hi, size=3):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (MO.jvp_loss_grad(pp, xyz, dirs, e, d, w, freq, centre, s, osc)[0]
                  - MO.jvp_loss_grad(pm, xyz, dirs, e, d, w, freq, centre, s, osc)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    l0, g0, _, _ = MO.jvp_loss_grad(p, xyz, dirs, e, d, 0.0, freq, centre, s, osc)
    lv, gv = MO.loss_grad(p, xyz, e, freq, centre, s, osc)
    assert abs(l0 - lv) <= 1e-12 * lv
    assert np.allclose(g0, gv, rtol=1e-10, atol=1e-14)
