"""The learned guess's network on the GPU (hand-written BF16-MFMA kernels,
projectracer_amd/csrc/pt_unet.hip) against its CPU reference
(tests/unet_ref.py, pinned by finite differences in test_unet_ref.py) on the
same parameters and inputs: the forward pass, the masked mean-square loss,
the gradient of every layer, and an AdamW step.  The reference rounds to
BF16 where the GPU stores BF16, so the remaining difference is the order of
FP32 accumulation (and a BF16 rounding that order can flip)."""
import numpy as np
import pytest

from projectracer_amd.learned import GuessNet
from unet_ref import UNetRef, layout
from test_unet_ref import _params

pytestmark = pytest.mark.gpu


def _inputs(N=2, H=16, W=32, seed=11):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (N, H, W, 5)).astype(np.float32)
    t = rng.normal(0, 1, (N, H, W)).astype(np.float32)
    m = (rng.random((N, H, W)) > 0.1).astype(np.float32)
    return x, t, m


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_and_gradients_match_cpu_reference():
    print("[unet] forward / loss / gradients, N=2 16x32, vs the CPU reference", flush=True)
    p = _params().astype(np.float32)
    net = GuessNet(seed=1)
    net.set_params(p)
    x, t, m = _inputs()
    ref = UNetRef(p.astype(np.float64), bf16_mode=True)
    y_ref = ref.forward(x.astype(np.float64))
    y = net.forward(x)
    ey = _rel(y, y_ref)
    loss, g = net.loss_grad(x, t, m)
    loss_ref, g_ref = ref.loss_grad(x.astype(np.float64), t.astype(np.float64), m.astype(np.float64))
    print(f"[unet] forward rel err {ey:.2e}; loss {loss:.6e} vs {loss_ref:.6e}", flush=True)
    assert ey <= 1e-2
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    convs, hw, hb, n = layout()
    worst = 0.0
    for i, (w, b, cin, cout) in enumerate(convs):
        ew = _rel(g[w:w + 9 * cin * cout], g_ref[w:w + 9 * cin * cout])
        eb = _rel(g[b:b + cout], g_ref[b:b + cout])
        print(f"[unet] conv {i:2d} ({cin:3d}->{cout:3d}): dW rel err {ew:.2e}, db rel err {eb:.2e}", flush=True)
        worst = max(worst, ew, eb)
    eh = _rel(g[hw:hb + 1], g_ref[hw:hb + 1])
    print(f"[unet] head: rel err {eh:.2e}", flush=True)
    assert worst <= 3e-2 and eh <= 1e-2


def test_adamw_step_matches_formula():
    p = _params(seed=4).astype(np.float32)
    net = GuessNet(seed=2)
    net.set_params(p)
    x, t, m = _inputs(N=1, H=8, W=16, seed=3)
    _, g = net.loss_grad(x, t, m)
    lr, wd = 1e-3, 1e-2
    net.adamw_step(lr, wd)
    got = net.params()
    g64, p64 = g.astype(np.float64), p.astype(np.float64)
    mh = (0.1 * g64) / 0.1
    vh = (0.001 * g64 * g64) / 0.001
    want = p64 - lr * (mh / (np.sqrt(vh) + 1e-8) + wd * p64)
    err = np.max(np.abs(got - want))
    print(f"[unet] AdamW first step: max abs err {err:.2e}", flush=True)
    assert err <= 1e-6


def test_training_steps_reduce_the_loss():
    net = GuessNet(seed=7)
    rng = np.random.default_rng(0)
    N, H, W = 4, 32, 64
    x = rng.normal(0, 1, (N, H, W, 5)).astype(np.float32)
    # a smooth target the network can fit: a blurred combination of two inputs
    t = (0.5 * x[..., 0] - 0.3 * x[..., 4]).astype(np.float32)
    m = np.ones((N, H, W), dtype=np.float32)
    losses = []
    for _ in range(60):
        loss, _ = net.loss_grad(x, t, m)
        losses.append(loss)
        net.adamw_step(2e-3, 0.0)
    print(f"[unet] loss {losses[0]:.4f} -> {losses[-1]:.4f} in 60 AdamW steps", flush=True)
    assert losses[-1] < 0.2 * losses[0]
