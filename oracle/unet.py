"""TEST INFRASTRUCTURE -- FP64 NumPy oracle of the warm-start 3D U-Net.

The U-Net has no counterpart in the reference (it is an addition on top of
the ADMM path), so this module is the definition the HIP kernels are checked
against: direct (non-FFT) 3x3x3 convolution in channels-last layout, its
input and weight gradients, ReLU, 2x average pooling, 2x nearest
upsampling, channel concatenation, the 1x1x1 output head with a residual
connection, mean-squared error and Adam.  Everything is float64.

Pinning (tests/test_unet_oracle.py): each backward function against central
finite differences of its forward; the convolution against a brute-force
loop on a tiny volume and against the closed forms of a delta kernel (a
shift) and a constant input (sum of the weights); the whole network's
gradient against finite differences.

Layouts: activations [D, H, W, C]; conv weights [Cout, 27, Cin] with tap
t = 9*dz + 3*dy + dx for offsets (dz-1, dy-1, dx-1); bias [Cout].
"""
import numpy as np

TAPS = [(dz - 1, dy - 1, dx - 1) for dz in range(3) for dy in range(3) for dx in range(3)]


def _shifted(x, off):
    """x[z+oz, y+oy, x+ox] with zero padding outside."""
    oz, oy, ox = off
    D, H, W = x.shape[:3]
    out = np.zeros_like(x)
    zs = slice(max(0, -oz), min(D, D - oz))
    ys = slice(max(0, -oy), min(H, H - oy))
    xs = slice(max(0, -ox), min(W, W - ox))
    zd = slice(zs.start + oz, zs.stop + oz)
    yd = slice(ys.start + oy, ys.stop + oy)
    xd = slice(xs.start + ox, xs.stop + ox)
    out[zs, ys, xs] = x[zd, yd, xd]
    return out


def conv3d(x, w, b=None):
    """y[v, co] = b[co] + sum_{t, ci} w[co, t, ci] x[v + off_t, ci]."""
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, off in enumerate(TAPS):
        y += _shifted(x, off) @ w[:, t, :].T
    if b is not None:
        y += b
    return y


def conv3d_dx(dy, w):
    """gradient wrt x: dx[v, ci] = sum_{t, co} w[co, t, ci] dy[v - off_t, co]."""
    dx = np.zeros(dy.shape[:3] + (w.shape[2],))
    for t, off in enumerate(TAPS):
        dx += _shifted(dy, tuple(-o for o in off)) @ w[:, t, :]
    return dx


def conv3d_dw(dy, x):
    """gradients wrt w and b: dw[co, t, ci] = sum_v dy[v, co] x[v + off_t, ci]."""
    Cout, Cin = dy.shape[3], x.shape[3]
    dw = np.zeros((Cout, 27, Cin))
    dyf = dy.reshape(-1, Cout)
    for t, off in enumerate(TAPS):
        dw[:, t, :] = dyf.T @ _shifted(x, off).reshape(-1, Cin)
    return dw, dyf.sum(0)


def relu(x):
    return np.maximum(x, 0)


def relu_dx(dy, y):
    return dy * (y > 0)


def avgpool2(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def avgpool2_dx(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 0), 2, 1), 2, 2) / 8.0


def upsample2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def upsample2_dx(dy):
    D, H, W, C = dy.shape
    return dy.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def head(x, w, b, inp):
    """out[v, k] = inp[v, k] + b[k] + sum_c w[k, c] x[v, c] (residual 1x1x1)."""
    return inp + x @ w.T + b


def head_dx(dout, w, x):
    dx = dout @ w
    dw = dout.reshape(-1, dout.shape[-1]).T @ x.reshape(-1, x.shape[-1])
    return dx, dw, dout.reshape(-1, dout.shape[-1]).sum(0)


def mse(pred, target):
    """loss = mean((pred - target)^2), gradient wrt pred."""
    r = pred - target
    return float(np.mean(r * r)), 2.0 * r / r.size


def adam(w, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected), returns new (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


# ------------------------------------------------------------------ network
def init_params(rng, cin=2, c1=16, c2=32, c3=64):
    """He-initialised parameters of the 3-level U-Net (names as in the GPU
    model)."""
    def conv(co, ci):
        return rng.standard_normal((co, 27, ci)) * np.sqrt(2.0 / (27 * ci)), np.zeros(co)
    p = {}
    p["e1a"], p["e1b"] = conv(c1, cin), conv(c1, c1)
    p["e2a"], p["e2b"] = conv(c2, c1), conv(c2, c2)
    p["bna"], p["bnb"] = conv(c3, c2), conv(c3, c3)
    p["d2a"], p["d2b"] = conv(c2, c3 + c2), conv(c2, c2)
    p["d1a"], p["d1b"] = conv(c1, c2 + c1), conv(c1, c1)
    p["head"] = (rng.standard_normal((cin, c1)) * 0.01, np.zeros(cin))
    return p


CONVS = ["e1a", "e1b", "e2a", "e2b", "bna", "bnb", "d2a", "d2b", "d1a", "d1b"]


def forward(p, x):
    """U-Net: two conv+ReLU per level, average-pool down, nearest upsample
    up, skip concatenation [upsampled, skip], residual 1x1x1 head."""
    a = {}
    def cr(name, inp):
        w, b = p[name]
        a[name + "_in"] = inp
        y = relu(conv3d(inp, w, b))
        a[name] = y
        return y
    e1 = cr("e1b", cr("e1a", x))
    e2 = cr("e2b", cr("e2a", avgpool2(e1)))
    bn = cr("bnb", cr("bna", avgpool2(e2)))
    d2 = cr("d2b", cr("d2a", np.concatenate([upsample2(bn), e2], -1)))
    d1 = cr("d1b", cr("d1a", np.concatenate([upsample2(d2), e1], -1)))
    out = head(d1, *p["head"], x)
    a["d1"] = d1
    return out, a


def backward(p, x, a, dout):
    """Gradients of all parameters for the upstream gradient dout."""
    g = {}
    dd1, g_hw, g_hb = head_dx(dout, p["head"][0], a["d1"])
    g["head"] = (g_hw, g_hb)

    def cr_back(name, dy):
        w, _ = p[name]
        dpre = relu_dx(dy, a[name])
        dw, db = conv3d_dw(dpre, a[name + "_in"])
        g[name] = (dw, db)
        return conv3d_dx(dpre, w)

    c2 = p["d2b"][0].shape[0]
    c1 = p["d1b"][0].shape[0]
    dcat1 = cr_back("d1a", cr_back("d1b", dd1))
    dup1, de1 = dcat1[..., :c2], dcat1[..., c2:]
    dd2 = upsample2_dx(dup1)
    dcat2 = cr_back("d2a", cr_back("d2b", dd2))
    c3 = p["bnb"][0].shape[0]
    dup2, de2 = dcat2[..., :c3], dcat2[..., c3:]
    dbn = upsample2_dx(dup2)
    dp2 = cr_back("bna", cr_back("bnb", dbn))
    de2 = de2 + avgpool2_dx(dp2)
    dp1 = cr_back("e2a", cr_back("e2b", de2))
    de1 = de1 + avgpool2_dx(dp1)
    cr_back("e1a", cr_back("e1b", de1))
    del c1
    return g
