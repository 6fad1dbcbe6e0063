"""CPU reference of the learned initial guess's network (test
infrastructure): the same four-level encoder-decoder as
projectracer_amd/csrc/pt_unet.hip, in numpy, with the same flat parameter
layout, forward pass, masked mean-square loss and backward pass.

With bf16=True every value the GPU stores in BF16 (inputs, conv weights,
activations after each convolution or pooling, the ReLU-masked output
gradients fed to the MFMA kernels) is rounded to BF16 here at the same
point, and sums are formed in float64; the GPU's FP32 accumulation order is
the remaining difference.  With bf16=False everything is float64: that
version is the one the finite-difference checks pin.
"""
from __future__ import annotations

import numpy as np

IN_CH = 8
NUSED_IN = 5
CONV_DIMS = [(IN_CH, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 256), (256, 256),
             (256, 128), (256, 128), (128, 128), (128, 64), (128, 64), (64, 64), (64, 32), (64, 32), (32, 32)]


def layout():
    """Offsets of each conv's weights (9, Cout, Cin) and bias, then the head."""
    off, convs = 0, []
    for cin, cout in CONV_DIMS:
        w = off
        off += 9 * cin * cout
        b = off
        off += cout
        convs.append((w, b, cin, cout))
    head_w = off
    off += 32
    head_b = off
    off += 1
    return convs, head_w, head_b, off


def bf16(a):
    """Round to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


class UNetRef:
    def __init__(self, params, bf16_mode=True):
        self.p = np.asarray(params, dtype=np.float64)
        self.bf = bf16_mode
        self.convs, self.head_w, self.head_b, self.n = layout()
        assert self.p.size == self.n

    def _r(self, a):
        return bf16(a).astype(np.float64) if self.bf else np.asarray(a, dtype=np.float64)

    def W(self, i):
        w, b, cin, cout = self.convs[i]
        return self._r(self.p[w:w + 9 * cin * cout].reshape(9, cout, cin)), self.p[b:b + cout]

    @staticmethod
    def _shifted(x):
        """(N,H,W,C) -> (N,H,W,9,C): x at (y+dy, x+dx), tap = 3 (dy+1) + (dx+1), zero outside."""
        N, H, W, C = x.shape
        xp = np.zeros((N, H + 2, W + 2, C))
        xp[:, 1:-1, 1:-1] = x
        return np.stack([xp[:, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1)],
                        axis=3)

    def conv(self, i, x, relu=True):
        w, b = self.W(i)
        y = np.einsum("nhwtc,toc->nhwo", self._shifted(x), w, optimize=True) + b
        if relu:
            y = np.maximum(y, 0.0)
        return self._r(y)

    def conv_bwd(self, i, x, y, dy):
        """Gradients of y = relu(conv(x)): (dW (9,Cout,Cin), db, dx)."""
        w, _ = self.W(i)
        g = self._r(dy * (y > 0))
        dW = np.einsum("nhwo,nhwtc->toc", g, self._shifted(x), optimize=True)
        db = g.sum(axis=(0, 1, 2))
        # data gradient: the convolution of g with the flipped, transposed weights
        wf = w[::-1].transpose(0, 2, 1)   # (tap', Cin, Cout)
        dx = np.einsum("nhwto,tco->nhwc", self._shifted(g), wf, optimize=True)
        return dW, db, dx

    @staticmethod
    def pool(x):
        N, H, W, C = x.shape
        return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))

    @staticmethod
    def pool_bwd(dy):
        return 0.25 * np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2)

    @staticmethod
    def up(x):
        return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)

    @staticmethod
    def up_bwd(dy):
        N, H, W, C = dy.shape
        return dy.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))

    def forward(self, x5):
        """x5: (N, H, W, 5) -> prediction (N, H, W); keeps the activations."""
        N, H, W, _ = x5.shape
        x = np.zeros((N, H, W, IN_CH))
        x[..., :NUSED_IN] = x5
        a = {"x0": self._r(x)}
        a["e0a"] = self.conv(0, a["x0"])
        a["e0b"] = self.conv(1, a["e0a"])
        a["p1"] = self._r(self.pool(a["e0b"]))
        a["e1a"] = self.conv(2, a["p1"])
        a["e1b"] = self.conv(3, a["e1a"])
        a["p2"] = self._r(self.pool(a["e1b"]))
        a["e2a"] = self.conv(4, a["p2"])
        a["e2b"] = self.conv(5, a["e2a"])
        a["p3"] = self._r(self.pool(a["e2b"]))
        a["e3a"] = self.conv(6, a["p3"])
        a["e3b"] = self.conv(7, a["e3a"])
        a["u3"] = self.up(a["e3b"])
        a["c3"] = self.conv(8, a["u3"])
        a["cat2"] = np.concatenate([a["c3"], a["e2b"]], axis=3)
        a["d2a"] = self.conv(9, a["cat2"])
        a["d2b"] = self.conv(10, a["d2a"])
        a["u2"] = self.up(a["d2b"])
        a["c2"] = self.conv(11, a["u2"])
        a["cat1"] = np.concatenate([a["c2"], a["e1b"]], axis=3)
        a["d1a"] = self.conv(12, a["cat1"])
        a["d1b"] = self.conv(13, a["d1a"])
        a["u1"] = self.up(a["d1b"])
        a["c1"] = self.conv(14, a["u1"])
        a["cat0"] = np.concatenate([a["c1"], a["e0b"]], axis=3)
        a["d0a"] = self.conv(15, a["cat0"])
        a["d0b"] = self.conv(16, a["d0a"])
        hw = self.p[self.head_w:self.head_w + 32]
        y = a["d0b"] @ hw + self.p[self.head_b]
        self.a = a
        return y

    @staticmethod
    def lap(f):
        """5-point Laplacian on each (H, W) image, zero outside it."""
        out = -4.0 * f
        out[:, :, 1:] += f[:, :, :-1]
        out[:, :, :-1] += f[:, :, 1:]
        out[:, 1:, :] += f[:, :-1, :]
        out[:, :-1, :] += f[:, 1:, :]
        return out

    def loss_grad(self, x5, t, m, lam=0.0):
        """Masked loss (sum m e^2 + lam sum (m lap e)^2) / sum m, e = y - t,
        and the flat gradient."""
        y = self.forward(x5)
        a = self.a
        msum = max(m.sum(), 1.0)
        e = y - t
        le = m * self.lap(e)
        loss = float(((m * e ** 2).sum() + lam * (le ** 2).sum()) / msum)
        dy = (2.0 * m * e + 2.0 * lam * self.lap(le)) / msum
        g = np.zeros(self.n)

        def put(i, dW, db):
            w, b, cin, cout = self.convs[i]
            g[w:w + 9 * cin * cout] = dW.ravel()
            g[b:b + cout] = db

        hw = self.p[self.head_w:self.head_w + 32]
        g[self.head_w:self.head_w + 32] = np.einsum("nhw,nhwc->c", dy, a["d0b"])
        g[self.head_b] = dy.sum()
        d = dy[..., None] * hw
        dW, db, d = self.conv_bwd(16, a["d0a"], a["d0b"], d); put(16, dW, db)
        dW, db, dcat0 = self.conv_bwd(15, a["cat0"], a["d0a"], d); put(15, dW, db)
        dW, db, d = self.conv_bwd(14, a["u1"], a["c1"], dcat0[..., :32]); put(14, dW, db)
        d = self.up_bwd(d)
        dW, db, d = self.conv_bwd(13, a["d1a"], a["d1b"], d); put(13, dW, db)
        dW, db, dcat1 = self.conv_bwd(12, a["cat1"], a["d1a"], d); put(12, dW, db)
        dW, db, d = self.conv_bwd(11, a["u2"], a["c2"], dcat1[..., :64]); put(11, dW, db)
        d = self.up_bwd(d)
        dW, db, d = self.conv_bwd(10, a["d2a"], a["d2b"], d); put(10, dW, db)
        dW, db, dcat2 = self.conv_bwd(9, a["cat2"], a["d2a"], d); put(9, dW, db)
        dW, db, d = self.conv_bwd(8, a["u3"], a["c3"], dcat2[..., :128]); put(8, dW, db)
        d = self.up_bwd(d)
        dW, db, d = self.conv_bwd(7, a["e3a"], a["e3b"], d); put(7, dW, db)
        dW, db, d = self.conv_bwd(6, a["p3"], a["e3a"], d); put(6, dW, db)
        de2b = dcat2[..., 128:] + self.pool_bwd(d)
        dW, db, d = self.conv_bwd(5, a["e2a"], a["e2b"], de2b); put(5, dW, db)
        dW, db, d = self.conv_bwd(4, a["p2"], a["e2a"], d); put(4, dW, db)
        de1b = dcat1[..., 64:] + self.pool_bwd(d)
        dW, db, d = self.conv_bwd(3, a["e1a"], a["e1b"], de1b); put(3, dW, db)
        dW, db, d = self.conv_bwd(2, a["p1"], a["e1a"], d); put(2, dW, db)
        de0b = dcat0[..., 32:] + self.pool_bwd(d)
        dW, db, d = self.conv_bwd(1, a["e0a"], a["e0b"], de0b); put(1, dW, db)
        dW, db, _ = self.conv_bwd(0, a["x0"], a["e0a"], d); put(0, dW, db)
        return loss, g
