"""The CPU reference of the learned guess's network (tests/unet_ref.py),
pinned before it serves as the GPU kernels' check: its analytic gradient
against central finite differences in float64 on a small image, for
parameters of every layer and the head; its BF16 rounding against the
known rounding of a few values."""
import numpy as np

from unet_ref import UNetRef, layout, bf16


def _params(seed=3):
    convs, hw, hb, n = layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(n)
    for (w, b, cin, cout) in convs:
        p[w:w + 9 * cin * cout] = rng.normal(0, np.sqrt(2.0 / (9 * cin)), 9 * cin * cout)
        p[b:b + cout] = rng.normal(0, 0.05, cout)
    p[hw:hw + 32] = rng.normal(0, 0.3, 32)
    p[hb] = 0.1
    return p


def test_bf16_rounding():
    vals = np.array([1.0, 1.00390625, 1.005859375, -2.75, 3.0e-5, 1.0 + 2 ** -8 + 2 ** -9], dtype=np.float32)
    got = bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1.005859375 = 1 + 3*2^-9 rounds up
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0078125)
    assert got[3] == -2.75
    assert got[5] == np.float32(1.0078125)


def test_gradient_matches_finite_differences():
    p = _params()
    rng = np.random.default_rng(5)
    N, H, W = 1, 8, 16
    x = rng.normal(0, 1, (N, H, W, 5))
    t = rng.normal(0, 1, (N, H, W))
    m = (rng.random((N, H, W)) > 0.2).astype(np.float64)
    net = UNetRef(p, bf16_mode=False)
    loss, g = net.loss_grad(x, t, m)
    convs, hw, hb, n = layout()
    picks = [hw + 3, hb]
    for (w, b, cin, cout) in convs:
        picks += list(rng.integers(w, w + 9 * cin * cout, 3)) + [b + int(rng.integers(0, cout))]
    h = 1e-6
    checked = 0
    for k in picks:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (UNetRef(pp, False).loss_grad(x, t, m)[0] - UNetRef(pm, False).loss_grad(x, t, m)[0]) / (2 * h)
        if abs(fd) < 1e-7 and abs(g[k]) < 1e-7:
            continue
        assert abs(fd - g[k]) <= 1e-4 * max(abs(fd), abs(g[k])) + 1e-9, (k, fd, g[k])
        checked += 1
    assert checked >= 40
