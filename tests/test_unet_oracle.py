"""Pins of the U-Net oracle (oracle/unet.py): convolution against a brute-force
loop and closed forms, every backward against finite differences, pooling /
upsampling adjointness, the whole network's gradient by finite differences."""
import numpy as np
import pytest

from oracle import unet as on


def _brute_conv(x, w, b):
    D, H, W, Cin = x.shape
    Cout = w.shape[0]
    y = np.zeros((D, H, W, Cout))
    for z in range(D):
        for yy in range(H):
            for xx in range(W):
                for co in range(Cout):
                    s = b[co]
                    for t, (oz, oy, ox) in enumerate(on.TAPS):
                        zz, y2, x2 = z + oz, yy + oy, xx + ox
                        if 0 <= zz < D and 0 <= y2 < H and 0 <= x2 < W:
                            s += w[co, t] @ x[zz, y2, x2]
                    y[z, yy, xx, co] = s
    return y


def test_conv_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((4, 3, 5, 2))
    w = rng.standard_normal((3, 27, 2))
    b = rng.standard_normal(3)
    np.testing.assert_allclose(on.conv3d(x, w, b), _brute_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv_delta_kernel_is_shift_and_constant_input():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5, 6, 7, 2))
    w = np.zeros((2, 27, 2))
    t = on.TAPS.index((1, 0, -1))          # y[z,y,x] = x[z+1, y, x-1]
    w[0, t, 0] = w[1, t, 1] = 1
    y = on.conv3d(x, w)
    np.testing.assert_array_equal(y[:-1, :, 1:], x[1:, :, :-1])
    assert np.all(y[-1] == 0) and np.all(y[:, :, 0] == 0)
    c = np.full((5, 5, 5, 3), 0.7)
    w = rng.standard_normal((4, 27, 3))
    b = rng.standard_normal(4)
    y = on.conv3d(c, w, b)
    np.testing.assert_allclose(y[2, 2, 2], b + 0.7 * w.sum(axis=(1, 2)), rtol=1e-12)


def _fd(f, x, k, h=1e-6):
    e = np.zeros_like(x)
    e.flat[k] = h
    return (f(x + e) - f(x - e)) / (2 * h)


def test_conv_gradients_finite_differences():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, 4, 3, 2))
    w = rng.standard_normal((3, 27, 2))
    R = rng.standard_normal((3, 4, 3, 3))
    dw, db = on.conv3d_dw(R, x)
    dx = on.conv3d_dx(R, w)
    for k in rng.choice(x.size, 6, replace=False):
        assert abs(_fd(lambda v: np.sum(on.conv3d(v, w) * R), x, k) - dx.flat[k]) < 1e-6
    for k in rng.choice(w.size, 6, replace=False):
        assert abs(_fd(lambda v: np.sum(on.conv3d(x, v) * R), w, k) - dw.flat[k]) < 1e-6
    b0 = np.zeros(3)
    for k in range(3):
        assert abs(_fd(lambda v: np.sum(on.conv3d(x, w, v) * R), b0, k) - db[k]) < 1e-6


def test_pool_upsample_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 6, 2, 3))
    y = rng.standard_normal((2, 3, 1, 3))
    assert np.isclose(np.sum(on.avgpool2(x) * y), np.sum(x * on.avgpool2_dx(y)))
    assert np.isclose(np.sum(on.upsample2(y) * x), np.sum(y * on.upsample2_dx(x)))
    np.testing.assert_allclose(on.avgpool2(on.upsample2(y)), y)


def test_network_gradient_finite_differences():
    rng = np.random.default_rng(4)
    p = on.init_params(rng, cin=2, c1=3, c2=4, c3=5)
    p = {k: (v[0], v[1] + 0.05 * rng.standard_normal(v[1].shape)) for k, v in p.items()}
    x = rng.standard_normal((4, 4, 4, 2))
    target = rng.standard_normal((4, 4, 4, 2))

    def loss(pp):
        return on.mse(on.forward(pp, x)[0], target)[0]

    out, a = on.forward(p, x)
    _, dout = on.mse(out, target)
    g = on.backward(p, x, a, dout)
    for name in ["e1a", "e2b", "bna", "d2a", "d1b", "head"]:
        for which in (0, 1):
            arr = p[name][which]
            for k in rng.choice(arr.size, 2, replace=False):
                def f(v, name=name, which=which):
                    q = dict(p)
                    q[name] = (v, p[name][1]) if which == 0 else (p[name][0], v)
                    return loss(q)
                num = _fd(f, arr, k, 1e-5)
                ana = g[name][which].flat[k]
                assert abs(num - ana) <= 1e-6 + 1e-4 * abs(ana), (name, which, num, ana)


def test_adam_first_step():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.1, 0.0])
    w1, m, v = on.adam(w, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # first bias-corrected step is lr * g / (|g| + eps)
    np.testing.assert_allclose(w1, w - 0.01 * g / (np.abs(g) + 1e-8))
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)
