"""Host side of the learned initial guess: the network object of the C ABI
(pt_net_*), its parameters, forward pass, loss gradient and AdamW step.

No reference counterpart: the reference warm-starts each stream-function
solve from the previous psi (PS/iterative_solver.f90:134).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .simulation import PtError


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class GuessNet:
    """The encoder-decoder on the device (BF16 weights on MFMA, FP32 master
    copy and AdamW state)."""

    def __init__(self, seed: int = 0, lib=None):
        self.lib = lib or _abi.load_library()
        self.ptr = C.c_void_p()
        self._check(self.lib.pt_net_create(int(seed), C.byref(self.ptr)), "pt_net_create")
        n = C.c_int64()
        self._check(self.lib.pt_net_num_params(self.ptr, C.byref(n)), "pt_net_num_params")
        self.nparam = n.value

    def _check(self, rc, where):
        if rc != _abi.PT_OK:
            detail = self.lib.pt_last_error()
            raise PtError(rc, where, detail.decode() if detail else "")

    def close(self):
        if self.ptr:
            self.lib.pt_net_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        self._check(self.lib.pt_net_get_params(self.ptr, _fp(out)), "pt_net_get_params")
        return out

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.nparam:
            raise ValueError(f"{a.size} parameters given, the network has {self.nparam}")
        self._check(self.lib.pt_net_set_params(self.ptr, _fp(a)), "pt_net_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        N, H, W, c = x.shape
        if c != 5:
            raise ValueError("inputs are (N, H, W, 5)")
        y = np.empty((N, H, W), dtype=np.float32)
        self._check(self.lib.pt_net_forward(self.ptr, N, H, W, _fp(x), _fp(y)), "pt_net_forward")
        return y

    def loss_grad(self, x, t, mask):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        N, H, W, _ = x.shape
        g = np.empty(self.nparam, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.pt_net_loss_grad(self.ptr, N, H, W, _fp(x), _fp(t), _fp(m), C.byref(loss), _fp(g)),
                    "pt_net_loss_grad")
        return loss.value, g

    def adamw_step(self, lr: float, weight_decay: float = 0.0):
        self._check(self.lib.pt_net_adamw_step(self.ptr, float(lr), float(weight_decay)), "pt_net_adamw_step")
