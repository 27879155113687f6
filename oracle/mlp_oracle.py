"""CPU oracle of the far-field surrogate network (TEST INFRASTRUCTURE ONLY).

A numpy restatement of the Fourier-feature MLP that the GPU kernels
(massccs_amd/csrc/mlp.hip) train and evaluate: features [sin(2 pi B r'),
cos(2 pi B r')] with r' = (r - centre) * in_scale, three 128-wide SiLU layers
and two linear outputs.  Arithmetic is FP64; with ``bf16=True`` the values
the GPU rounds to BF16 before a matrix product (features, weights,
activations, back-propagated pre-activation gradients) are rounded the same
way, so that the only remaining difference is FP32 accumulation order.  The
FP64 path is pinned by finite differences (tests/test_mlp_oracle.py).  Only
tests use this module; the network has no counterpart in the reference.
"""
import numpy as np

NF, NIN, NH, NOUT = 32, 64, 128, 2
W1 = 0
B1 = W1 + NIN * NH
W2 = B1 + NH
B2 = W2 + NH * NH
W3 = B2 + NH
B3 = W3 + NH * NH
W4 = B3 + NH
B4 = W4 + NH * NOUT
NPARAM = B4 + NOUT


def bf16(x):
    """Round to BF16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(NIN, NH), p[B1:W2], p[W2:B2].reshape(NH, NH), p[B2:W3],
            p[W3:B3].reshape(NH, NH), p[B3:W4], p[W4:B4].reshape(NH, NOUT), p[B4:])


def _r(x, on):
    return bf16(x) if on else x


def features(xyz, freq, centre, in_scale):
    rp = (np.asarray(xyz, dtype=np.float64) - np.asarray(centre)) * in_scale
    a = 2 * np.pi * rp @ np.asarray(freq, dtype=np.float64).T
    return np.concatenate([np.sin(a), np.cos(a)], axis=1), a


def silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s))


def forward(p, xyz, freq, centre, in_scale, bf=False):
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    x0, a = features(xyz, freq, centre, in_scale)
    x0 = _r(x0, bf)
    z1 = x0 @ _r(w1, bf) + b1
    h1, d1 = silu(z1)
    h1 = _r(h1, bf)
    z2 = h1 @ _r(w2, bf) + b2
    h2, d2 = silu(z2)
    h2 = _r(h2, bf)
    z3 = h2 @ _r(w3, bf) + b3
    h3, d3 = silu(z3)
    h3 = _r(h3, bf)
    y = h3 @ _r(w4, bf) + b4
    return y, dict(x0=x0, a=a, h1=h1, h2=h2, h3=h3, d1=d1, d2=d2, d3=d3)


def backward(p, c, dy, bf=False):
    """Gradients of sum(dy * y) w.r.t. the flat parameters and the features."""
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    g = np.zeros(NPARAM)
    g[W4:B4] = (c["h3"].T @ dy).ravel()
    g[B4:] = dy.sum(0)
    dz3 = (dy @ _r(w4, bf).T) * c["d3"]
    g[B3:W4] = dz3.sum(0)
    dz3b = _r(dz3, bf)
    g[W3:B3] = (c["h2"].T @ dz3b).ravel()
    dz2 = (dz3b @ _r(w3, bf).T) * c["d2"]
    g[B2:W3] = dz2.sum(0)
    dz2b = _r(dz2, bf)
    g[W2:B2] = (c["h1"].T @ dz2b).ravel()
    dz1 = (dz2b @ _r(w2, bf).T) * c["d1"]
    g[B1:W2] = dz1.sum(0)
    dz1b = _r(dz1, bf)
    g[W1:B1] = (c["x0"].T @ dz1b).ravel()
    dx0 = dz1b @ _r(w1, bf).T
    return g, dx0


def loss_grad(p, xyz, e, freq, centre, in_scale, out_scale, bf=False):
    """Mean over points and outputs of (y - e / out_scale)^2 and its gradient."""
    y, c = forward(p, xyz, freq, centre, in_scale, bf)
    t = np.asarray(e, dtype=np.float64) / np.asarray(out_scale)
    d = y - t
    n = len(y)
    loss = float((d * d).sum() / (2 * n))
    g, _ = backward(p, c, 2 * d / (2 * n), bf)
    return loss, g


def eval_with_grad(p, xyz, freq, centre, in_scale, out_scale, bf=False):
    """Outputs (n, 2) in kcal/mol and their position gradients (n, 2, 3)."""
    y, c = forward(p, xyz, freq, centre, in_scale, bf)
    freq = np.asarray(freq, dtype=np.float64)
    out = np.zeros((len(y), 2, 3))
    for o in range(2):
        dy = np.zeros_like(y)
        dy[:, o] = 1.0
        _, dx0 = backward(p, c, dy, bf)
        da = dx0[:, :NF] * np.cos(c["a"]) - dx0[:, NF:] * np.sin(c["a"])
        out[:, o, :] = (da @ freq) * 2 * np.pi * in_scale * out_scale[o]
    return y * np.asarray(out_scale), out


# ---------------------------------------------------------------- forward-mode derivative
# The network's derivative along a direction u per point is propagated
# forward with the values (a tangent row beside each value row):
#   feature tangents  d sin a = cos a da,  d cos a = -sin a da,  da = 2 pi s B u
#   layer             Z = H W + b,  dZ = dH W,  H' = SiLU(Z),  dH' = SiLU'(Z) dZ
#   output            y = H3 W4 + b4,  dy = dH3 W4
# Training fits y to the far field and dy to its directional derivative; the
# loss gradient back-propagates through both rows, the value row picking up
# SiLU''(Z) dZ times the tangent row's upstream gradient.  With bf16=True the
# GPU's roundings are mirrored: features and their tangents, activations and
# their tangents, and both rows' back-propagated pre-activation gradients go
# to BF16 before a product.

def silu2(z):
    """SiLU''(z)."""
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 - s) * (2.0 + z * (1.0 - 2.0 * s))


def jvp_forward(p, xyz, dirs, freq, centre, in_scale, bf=False):
    """Outputs y (n, 2) and their derivatives along the unit vectors dirs
    (n, 2; per A), in the network's scaled output units."""
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    freq = np.asarray(freq, dtype=np.float64)
    x0, a = features(xyz, freq, centre, in_scale)
    ad = 2 * np.pi * in_scale * (np.asarray(dirs, dtype=np.float64) @ freq.T)
    xd0 = np.concatenate([np.cos(a) * ad, -np.sin(a) * ad], axis=1)
    c = dict(x0=_r(x0, bf), xd0=_r(xd0, bf))
    h, hd = c["x0"], c["xd0"]
    for k, (w, b) in enumerate([(w1, b1), (w2, b2), (w3, b3)], 1):
        z = h @ _r(w, bf) + b
        zd = hd @ _r(w, bf)
        hv, s1 = silu(z)
        h, hd = _r(hv, bf), _r(s1 * zd, bf)
        c.update({"z%d" % k: z, "zd%d" % k: zd, "h%d" % k: h, "hd%d" % k: hd,
                  "s%d" % k: s1, "t%d" % k: silu2(z)})
    y = h @ _r(w4, bf) + b4
    yd = hd @ _r(w4, bf)
    return y, yd, c


def jvp_loss_grad(p, xyz, dirs, e, d, w, freq, centre, in_scale, out_scale, bf=False):
    """Loss sum over points and outputs of (y - e/out_scale)^2 +
    w (dy - d/out_scale)^2, divided by 2n, and its parameter gradient."""
    y, yd, c = jvp_forward(p, xyz, dirs, freq, centre, in_scale, bf)
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    osc = np.asarray(out_scale)
    re = y - np.asarray(e, dtype=np.float64) / osc
    rd = yd - np.asarray(d, dtype=np.float64) / osc
    n = len(y)
    loss = float(((re * re).sum() + w * (rd * rd).sum()) / (2 * n))
    dy, dyd = re / n, w * rd / n
    g = np.zeros(NPARAM)
    g[W4:B4] = (c["h3"].T @ dy + c["hd3"].T @ dyd).ravel()
    g[B4:] = dy.sum(0)
    dh, dhd = dy @ _r(w4, bf).T, dyd @ _r(w4, bf).T
    for k, (wl, lo, hi, blo, bhi) in [(3, (w3, W3, B3, B3, W4)), (2, (w2, W2, B2, B2, W3)),
                                      (1, (w1, W1, B1, B1, W2))]:
        dz = dh * c["s%d" % k] + dhd * c["t%d" % k] * c["zd%d" % k]
        dzd = dhd * c["s%d" % k]
        g[blo:bhi] = dz.sum(0)
        dzb, dzdb = _r(dz, bf), _r(dzd, bf)
        hin, hdin = (c["h%d" % (k - 1)], c["hd%d" % (k - 1)]) if k > 1 else (c["x0"], c["xd0"])
        g[lo:hi] = (hin.T @ dzb + hdin.T @ dzdb).ravel()
        dh, dhd = dzb @ _r(wl, bf).T, dzdb @ _r(wl, bf).T
    return loss, g, y, yd
