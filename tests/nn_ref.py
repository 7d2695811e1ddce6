"""CPU reference of the learned-guess U-Net (numpy): forward and backward
of the same layers, parameter layout and loss as the GPU network, optionally
rounding to BF16 exactly where the GPU stores BF16 (inputs, weights,
activations, and the gradient entering each convolution), so that GPU and
CPU differ only by FP32 accumulation order.  Test infrastructure only.
"""
from __future__ import annotations

import numpy as np

NIN = 3  # input channels: residual, extrapolated guess, previous correction

# (buffer, channels, level) and the operation list, as in nn.hpp
BUFS = {"X0": (16, 0), "E1A": (16, 0), "CAT1": (48, 0), "P1": (16, 1), "E2A": (32, 1), "CAT2": (128, 1),
        "P2": (32, 2), "E3A": (96, 2), "CAT3": (288, 2), "P3": (96, 3), "B1": (192, 3), "B2": (192, 3),
        "D3A": (96, 2), "D3B": (96, 2), "D2A": (32, 1), "D2B": (32, 1), "D1A": (16, 0), "D1B": (16, 0),
        "Y": (16, 0)}

OPS = [
    ("conv", ("X0", 0, 16), ("E1A", 0, 16), 1),
    ("conv", ("E1A", 0, 16), ("CAT1", 32, 16), 1),
    ("pool", ("CAT1", 32, 16), ("P1", 0, 16), 0),
    ("conv", ("P1", 0, 16), ("E2A", 0, 32), 1),
    ("conv", ("E2A", 0, 32), ("CAT2", 96, 32), 1),
    ("pool", ("CAT2", 96, 32), ("P2", 0, 32), 0),
    ("conv", ("P2", 0, 32), ("E3A", 0, 96), 1),
    ("conv", ("E3A", 0, 96), ("CAT3", 192, 96), 1),
    ("pool", ("CAT3", 192, 96), ("P3", 0, 96), 0),
    ("conv", ("P3", 0, 96), ("B1", 0, 192), 1),
    ("conv", ("B1", 0, 192), ("B2", 0, 192), 1),
    ("up", ("B2", 0, 192), ("CAT3", 0, 192), 0),
    ("conv", ("CAT3", 0, 288), ("D3A", 0, 96), 1),
    ("conv", ("D3A", 0, 96), ("D3B", 0, 96), 1),
    ("up", ("D3B", 0, 96), ("CAT2", 0, 96), 0),
    ("conv", ("CAT2", 0, 128), ("D2A", 0, 32), 1),
    ("conv", ("D2A", 0, 32), ("D2B", 0, 32), 1),
    ("up", ("D2B", 0, 32), ("CAT1", 0, 32), 0),
    ("conv", ("CAT1", 0, 48), ("D1A", 0, 16), 1),
    ("conv", ("D1A", 0, 16), ("D1B", 0, 16), 1),
    ("conv", ("D1B", 0, 16), ("Y", 0, 16), 0),
]


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def param_layout():
    """[(op index, weight offset, Cout, Cin)] and the total count."""
    out, off = [], 0
    for k, (kind, a, b, act) in enumerate(OPS):
        if kind == "conv":
            cin, cout = a[2], b[2]
            out.append((k, off, cout, cin))
            off += cout * 9 * cin + cout
    return out, off


def _im2col(x):
    """x: (H, W, C) -> (H*W, 9*C) with taps ordered (dy, dx), zero padding."""
    H, W, C = x.shape
    xp = np.zeros((H + 2, W + 2, C), x.dtype)
    xp[1:-1, 1:-1] = x
    cols = [xp[dy:dy + H, dx:dx + W] for dy in range(3) for dx in range(3)]
    return np.stack(cols, axis=2).reshape(H * W, 9 * C)


def _col2im(g, H, W, C):
    """Adjoint of _im2col."""
    g = g.reshape(H, W, 9, C)
    xp = np.zeros((H + 2, W + 2, C), g.dtype)
    t = 0
    for dy in range(3):
        for dx in range(3):
            xp[dy:dy + H, dx:dx + W] += g[:, :, t]
            t += 1
    return xp[1:-1, 1:-1]


class Net:
    def __init__(self, H, W, params, emulate_bf16=True, dtype=np.float64):
        self.H, self.W = H, W
        self.p = np.asarray(params, dtype)
        self.em = emulate_bf16
        self.dtype = dtype
        self.layout, self.n = param_layout()
        assert self.p.size == self.n

    def _r(self, a):
        return bf16(a).astype(self.dtype) if self.em else a

    def weights(self, k):
        for kk, off, cout, cin in self.layout:
            if kk == k:
                w = self.p[off:off + cout * 9 * cin].reshape(cout, 9, cin)
                b = self.p[off + cout * 9 * cin: off + cout * 9 * cin + cout]
                return off, w, b
        raise KeyError(k)

    def forward(self, x2):
        """x2: (NIN, H*W) planar inputs -> y (H*W,), keeps activations."""
        H, W = self.H, self.W
        A = {name: np.zeros((H >> lv, W >> lv, c), self.dtype) for name, (c, lv) in BUFS.items()}
        for c in range(NIN):
            A["X0"][:, :, c] = self._r(x2[c].reshape(H, W))
        for k, (kind, a, b, act) in enumerate(OPS):
            xin = A[a[0]][:, :, a[1]:a[1] + a[2]]
            if kind == "conv":
                _, w, bias = self.weights(k)
                wq = self._r(w)
                h, wd, _ = xin.shape
                pre = _im2col(xin) @ wq.reshape(wq.shape[0], -1).T + bias
                out = np.where(pre > 0, pre, 0.1 * pre) if act else pre
                out = out.reshape(h, wd, -1)
                A[b[0]][:, :, b[1]:b[1] + b[2]] = out if b[0] == "Y" else self._r(out)
            elif kind == "pool":
                h, wd, c = xin.shape
                out = xin.reshape(h // 2, 2, wd // 2, 2, c).mean(axis=(1, 3))
                A[b[0]][:, :, b[1]:b[1] + b[2]] = self._r(out)
            else:
                out = np.repeat(np.repeat(xin, 2, axis=0), 2, axis=1)
                A[b[0]][:, :, b[1]:b[1] + b[2]] = out
        self.A = A
        return A["Y"][:, :, 0].reshape(-1).copy()

    def loss_and_grad(self, x2, t):
        y = self.forward(x2)
        H, W = self.H, self.W
        P = H * W
        e = y - t
        loss = float(np.mean(e * e))
        A = self.A
        G = {name: np.zeros_like(v) for name, v in A.items()}
        G["Y"][:, :, 0] = (2.0 * e / P).reshape(H, W)
        grad = np.zeros(self.n, self.dtype)
        for k in range(len(OPS) - 1, -1, -1):
            kind, a, b, act = OPS[k]
            if kind == "conv":
                off, w, bias = self.weights(k)
                cout, _, cin = w.shape
                gout = G[b[0]][:, :, b[1]:b[1] + b[2]]
                yout = A[b[0]][:, :, b[1]:b[1] + b[2]]
                d = gout * np.where(yout > 0, 1.0, 0.1) if act else gout
                d = self._r(d)
                xin = A[a[0]][:, :, a[1]:a[1] + a[2]]
                h, wd, _ = xin.shape
                d2 = d.reshape(h * wd, cout)
                cols = _im2col(xin)
                grad[off:off + cout * 9 * cin] += (d2.T @ cols).reshape(-1)
                grad[off + cout * 9 * cin: off + cout * 9 * cin + cout] += d2.sum(axis=0)
                if a[0] != "X0":
                    wq = self._r(w).reshape(cout, 9 * cin)
                    G[a[0]][:, :, a[1]:a[1] + a[2]] += _col2im(d2 @ wq, h, wd, cin)
            elif kind == "pool":
                gout = G[b[0]][:, :, b[1]:b[1] + b[2]]
                G[a[0]][:, :, a[1]:a[1] + a[2]] += 0.25 * np.repeat(np.repeat(gout, 2, axis=0), 2, axis=1)
            else:
                gout = G[b[0]][:, :, b[1]:b[1] + b[2]]
                h, wd, c = gout.shape
                G[a[0]][:, :, a[1]:a[1] + a[2]] += gout.reshape(h // 2, 2, wd // 2, 2, c).sum(axis=(1, 3))
        return loss, grad
