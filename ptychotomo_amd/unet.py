"""Warm-start 3D U-Net on the matrix cores (no counterpart in the reference).

A 3-level U-Net (channels 16-32-64, two 3x3x3 conv + ReLU per level, 2x
average pooling down, nearest 2x upsampling up, skip concatenation, a
residual 1x1x1 head) maps the 2-channel (delta, beta) volume after a few
ADMM iterations to the volume after many.  Activations and weights are bf16
in channels-last layout, accumulation FP32 (MFMA implicit-GEMM convolution,
``csrc/unet.hip``); master weights, gradients and the Adam state are FP32.

Forward, backward and the optimiser step run in the HIP library; PyTorch
holds the buffers.  Layer names and the layer order are those of the FP64
oracle the kernels are tested against.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._dev import device, ptr, stream

bf16 = torch.int16          # raw bf16 storage (the kernels reinterpret it)
f32 = torch.float32
f64 = torch.float64

CONVS = ["e1a", "e1b", "e2a", "e2b", "bna", "bnb", "d2a", "d2b", "d1a", "d1b"]


def _call(name, *args):
    return _lib.call(name, *args)


def _kp(cin):
    return (27 * cin + 15) // 16 * 16


class UNet3D:
    """U-Net for volumes [D, H, W, cin] with D, H, W divisible by 4."""

    def __init__(self, cin=2, c1=16, c2=32, c3=64, seed=0):
        self.cin, self.c = cin, (c1, c2, c3)
        shapes = {"e1a": (c1, cin), "e1b": (c1, c1), "e2a": (c2, c1), "e2b": (c2, c2),
                  "bna": (c3, c2), "bnb": (c3, c3), "d2a": (c2, c3 + c2), "d2b": (c2, c2),
                  "d1a": (c1, c2 + c1), "d1b": (c1, c1)}
        self.shapes = shapes
        rng = np.random.default_rng(seed)
        dev = device()
        self.w, self.b, self.wb, self.g, self.gb, self.m, self.v = {}, {}, {}, {}, {}, {}, {}
        for name in CONVS:
            co, ci = shapes[name]
            w = rng.standard_normal((co, 27, ci)) * np.sqrt(2.0 / (27 * ci))
            self.w[name] = torch.tensor(w, dtype=f32, device=dev).contiguous()
            self.b[name] = torch.zeros(co, dtype=f32, device=dev)
            self.wb[name] = torch.zeros(co * _kp(ci), dtype=bf16, device=dev)
            self.g[name] = torch.zeros_like(self.w[name])
            self.gb[name] = torch.zeros(co, dtype=f32, device=dev)
        hw = rng.standard_normal((cin, c1)) * 0.01
        self.w["head"] = torch.tensor(hw, dtype=f32, device=dev).contiguous()
        self.b["head"] = torch.zeros(cin, dtype=f32, device=dev)
        self.g["head"] = torch.zeros_like(self.w["head"])
        self.gb["head"] = torch.zeros(cin, dtype=f32, device=dev)
        for name in list(self.w):
            self.m[name] = (torch.zeros_like(self.w[name]), torch.zeros_like(self.b[name]))
            self.v[name] = (torch.zeros_like(self.w[name]), torch.zeros_like(self.b[name]))
        self.t = 0
        self._buf = {}
        self._wt = {}
        self.pack()

    # ------------------------------------------------------------------ buffers
    def pack(self):
        """bf16 copies of the conv weights from the FP32 masters."""
        for name in CONVS:
            co, ci = self.shapes[name]
            _call("pt_nn_pack_w", ptr(self.wb[name]), ptr(self.w[name]), co, ci, stream())

    def _t(self, key, n, dtype=bf16):
        t = self._buf.get(key)
        if t is None or t.numel() != n or t.dtype != dtype:
            t = torch.empty(n, dtype=dtype, device=device())
            self._buf[key] = t
        return t

    def _conv(self, name, x, dims, out_key, relu=True):
        D, H, W = dims
        co, ci = self.shapes[name]
        y = self._t(out_key, D * H * W * co)
        _call("pt_nn_conv3d", ptr(y), ptr(x), ptr(self.wb[name]), ptr(self.b[name]), D, H, W, ci,
              co, int(relu), stream())
        return y

    # ------------------------------------------------------------------ forward
    def forward(self, x):
        """x: float32 [D, H, W, cin] (device) -> float32 [D, H, W, cin]."""
        D, H, W, cin = x.shape
        assert cin == self.cin and D % 4 == 0 and H % 4 == 0 and W % 4 == 0
        c1, c2, c3 = self.c
        d1, d2, d3 = (D, H, W), (D // 2, H // 2, W // 2), (D // 4, H // 4, W // 4)
        V1, V2, V3 = D * H * W, d2[0] * d2[1] * d2[2], d3[0] * d3[1] * d3[2]
        st = stream()
        self.dims = (d1, d2, d3)
        xf = x.contiguous()
        self._x = xf
        xb = self._t("x", V1 * cin)
        _call("pt_nn_to_bf16", ptr(xb), ptr(xf), V1 * cin, st)
        e1a = self._conv("e1a", xb, d1, "e1a")
        e1b = self._conv("e1b", e1a, d1, "e1b")
        p1 = self._t("p1", V2 * c1)
        _call("pt_nn_pool", ptr(p1), ptr(e1b), *d1, c1, st)
        e2a = self._conv("e2a", p1, d2, "e2a")
        e2b = self._conv("e2b", e2a, d2, "e2b")
        p2 = self._t("p2", V3 * c2)
        _call("pt_nn_pool", ptr(p2), ptr(e2b), *d2, c2, st)
        bna = self._conv("bna", p2, d3, "bna")
        bnb = self._conv("bnb", bna, d3, "bnb")
        cat2 = self._t("cat2", V2 * (c3 + c2))
        _call("pt_nn_up", ptr(cat2), ptr(bnb), *d3, c3, c3 + c2, 0, st)
        _call("pt_nn_chan_copy", ptr(cat2), c3 + c2, c3, ptr(e2b), c2, 0, c2, V2, st)
        d2a = self._conv("d2a", cat2, d2, "d2a")
        d2b = self._conv("d2b", d2a, d2, "d2b")
        cat1 = self._t("cat1", V1 * (c2 + c1))
        _call("pt_nn_up", ptr(cat1), ptr(d2b), *d2, c2, c2 + c1, 0, st)
        _call("pt_nn_chan_copy", ptr(cat1), c2 + c1, c2, ptr(e1b), c1, 0, c1, V1, st)
        d1a = self._conv("d1a", cat1, d1, "d1a")
        d1b = self._conv("d1b", d1a, d1, "d1b")
        out = torch.empty((D, H, W, cin), dtype=f32, device=device())
        _call("pt_nn_head", ptr(out), ptr(d1b), ptr(self.w["head"]), ptr(self.b["head"]), ptr(xf),
              V1, c1, cin, st)
        self.acts = dict(x=xb, e1a=e1a, e1b=e1b, p1=p1, e2a=e2a, e2b=e2b, p2=p2, bna=bna,
                         bnb=bnb, cat2=cat2, d2a=d2a, d2b=d2b, cat1=cat1, d1a=d1a, d1b=d1b)
        return out

    # ------------------------------------------------------------------ backward
    def _flip(self, name):
        co, ci = self.shapes[name]
        wt = self._wt.get(name)
        if wt is None:
            wt = torch.empty(ci * _kp(co), dtype=bf16, device=device())
            self._wt[name] = wt
        _call("pt_nn_conv3d_wflip", ptr(wt), ptr(self.wb[name]), ci, co, stream())
        return wt

    def _conv_back(self, name, dy, dims, inp, split=None, need_dx=True):
        """Backward of conv+ReLU: dy (bf16, wrt the ReLU output) -> parameter
        gradients, and the input gradient (as one tensor, or as the channel
        ranges of `split` = [(c0, c1, key), ...])."""
        st = stream()
        D, H, W = dims
        V = D * H * W
        co, ci = self.shapes[name]
        y = self.acts[name]
        dpre = self._t("dpre_" + name, V * co)
        _call("pt_nn_relu_bwd", ptr(dpre), ptr(dy), ptr(y), V * co, st)
        part = self._t("wpart", 1 << 22, f32)
        bpart = self._t("bpart", 256 * 64, f64)
        _call("pt_nn_conv3d_wgrad", ptr(self.g[name]), ptr(self.gb[name]), ptr(part),
              part.numel(), ptr(bpart), ptr(dpre), ptr(inp), D, H, W, ci, co, st)
        if not need_dx:
            return None
        wt = self._flip(name)
        kpt = _kp(co)
        outs = []
        for c0, c1, key in (split or [(0, ci, "dx_" + name)]):
            dx = self._t(key, V * (c1 - c0))
            _call("pt_nn_conv3d", ptr(dx), ptr(dpre), ctypes.c_void_p(wt.data_ptr() + 2 * c0 * kpt),
                  None, D, H, W, co, c1 - c0, 0, st)
            outs.append(dx)
        return outs if split else outs[0]

    def backward(self, dout):
        """dout: float32 [D, H, W, cin] gradient of the loss wrt the output;
        fills self.g / self.gb."""
        st = stream()
        c1, c2, c3 = self.c
        d1, d2, d3 = self.dims
        V1 = d1[0] * d1[1] * d1[2]
        a = self.acts
        dd1 = self._t("dd1", V1 * c1)
        hp = self._t("hpart", 256 * 64, f64)
        dwb = torch.empty(self.cin * (c1 + 1), dtype=f32, device=device())
        _call("pt_nn_head_bwd", ptr(dd1), ptr(dwb), ptr(hp), hp.numel(), ptr(dout.contiguous()),
              ptr(a["d1b"]), ptr(self.w["head"]), V1, c1, self.cin, st)
        dwb = dwb.view(self.cin, c1 + 1)
        self.g["head"].copy_(dwb[:, :c1])
        self.gb["head"].copy_(dwb[:, c1])
        g = self._conv_back("d1b", dd1, d1, a["d1a"])
        dup1, de1 = self._conv_back("d1a", g, d1, a["cat1"],
                                    split=[(0, c2, "dup1"), (c2, c2 + c1, "de1s")])
        dd2 = self._t("dd2", d2[0] * d2[1] * d2[2] * c2)
        _call("pt_nn_up_bwd", ptr(dd2), ptr(dup1), *d2, c2, c2, 0, st)
        g = self._conv_back("d2b", dd2, d2, a["d2a"])
        dup2, de2 = self._conv_back("d2a", g, d2, a["cat2"],
                                    split=[(0, c3, "dup2"), (c3, c3 + c2, "de2s")])
        dbn = self._t("dbn", d3[0] * d3[1] * d3[2] * c3)
        _call("pt_nn_up_bwd", ptr(dbn), ptr(dup2), *d3, c3, c3, 0, st)
        g = self._conv_back("bnb", dbn, d3, a["bna"])
        dp2 = self._conv_back("bna", g, d3, a["p2"])
        de2t = self._t("de2t", d2[0] * d2[1] * d2[2] * c2)
        _call("pt_nn_pool_bwd", ptr(de2t), ptr(dp2), ptr(de2), *d2, c2, st)
        g = self._conv_back("e2b", de2t, d2, a["e2a"])
        dp1 = self._conv_back("e2a", g, d2, a["p1"])
        de1t = self._t("de1t", V1 * c1)
        _call("pt_nn_pool_bwd", ptr(de1t), ptr(dp1), ptr(de1), *d1, c1, st)
        g = self._conv_back("e1b", de1t, d1, a["e1a"])
        self._conv_back("e1a", g, d1, a["x"], need_dx=False)

    # ------------------------------------------------------------------ training
    def loss(self, out, target, grad=True):
        """mean squared error (FP64 sum); returns (loss tensor, gradient)."""
        n = out.numel()
        part = self._t("mse", 256, f64)
        gr = torch.empty_like(out) if grad else None
        _call("pt_nn_mse", ptr(part), 256, ptr(gr), ptr(out), ptr(target.contiguous()), n,
              stream())
        return part.sum() / n, gr

    def step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        """Adam on every parameter (FP32 masters), refreshing the bf16 copies."""
        self.t += 1
        st = stream()
        for name in self.w:
            conv = name in self.shapes
            K = 27 * self.shapes[name][1] if conv else 0
            Kp = _kp(self.shapes[name][1]) if conv else 0
            for which, (p, g) in enumerate(((self.w[name], self.g[name]),
                                            (self.b[name], self.gb[name]))):
                wb = self.wb[name] if (conv and which == 0) else None
                _call("pt_nn_adam", ptr(p), ptr(wb), ptr(g), ptr(self.m[name][which]),
                      ptr(self.v[name][which]), p.numel(), K if wb is not None else 0,
                      Kp if wb is not None else 0, lr, b1, b2, eps, self.t, st)

    def train_step(self, x, target, lr=1e-3):
        out = self.forward(x)
        loss, gr = self.loss(out, target)
        self.backward(gr)
        self.step(lr)
        return loss

    # ------------------------------------------------------------------ persistence
    def state_dict(self):
        d = {}
        for name in self.w:
            d[name + ".w"] = self.w[name].cpu().numpy()
            d[name + ".b"] = self.b[name].cpu().numpy()
        return d

    def load_state_dict(self, d):
        for name in self.w:
            self.w[name].copy_(torch.from_numpy(np.asarray(d[name + ".w"], np.float32)))
            self.b[name].copy_(torch.from_numpy(np.asarray(d[name + ".b"], np.float32)))
        self.pack()
