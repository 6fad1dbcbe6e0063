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

    # -- training set and training on the device ------------------------------
    def dataset(self, H: int, W: int, capacity: int):
        """Room for `capacity` (inputs, correction) pairs of H x W (= nz x nx) solves."""
        self._check(self.lib.pt_net_dataset(self.ptr, int(H), int(W), int(capacity)), "pt_net_dataset")

    def dataset_size(self) -> int:
        n = C.c_int32()
        self._check(self.lib.pt_net_dataset_size(self.ptr, C.byref(n)), "pt_net_dataset_size")
        return n.value

    def dataset_get(self, slot: int, H: int, W: int):
        """Stored pair `slot`: inputs (H, W, 5) and target (H, W)."""
        x = np.empty((H, W, 5), dtype=np.float32)
        t = np.empty((H, W), dtype=np.float32)
        self._check(self.lib.pt_net_dataset_get(self.ptr, int(slot), _fp(x), _fp(t)), "pt_net_dataset_get")
        return x, t

    def train(self, iters: int, batch: int, lr: float = 1e-3, weight_decay: float = 1e-4, seed: int = 0) -> float:
        """AdamW on random batches of the stored pairs; the last batch's loss."""
        loss = C.c_double()
        self._check(self.lib.pt_net_train(self.ptr, int(iters), int(batch), float(lr), float(weight_decay),
                                          int(seed), C.byref(loss)), "pt_net_train")
        return loss.value

    def set_loss(self, lap_weight: float):
        """Weight of the error's Laplacian term in the training loss."""
        self._check(self.lib.pt_net_set_loss(self.ptr, float(lap_weight)), "pt_net_set_loss")

    def out_scale(self) -> float:
        s = C.c_double()
        self._check(self.lib.pt_net_out_scale(self.ptr, C.byref(s)), "pt_net_out_scale")
        return s.value


def attach(sim, net, guess: bool = True, capture: bool = False):
    """Attach `net` to a Simulation: its solves propose the network's guess
    (guess=True) and/or add their (inputs, correction) pairs to the
    network's training set (capture=True).  net=None detaches."""
    mode = _abi.PT_GUESS_LEARNED if guess else _abi.PT_GUESS_WARM
    ptr = net.ptr if net is not None else None
    rc = sim.lib.pt_set_guess_net(sim.ctx, ptr, mode, 1 if capture else 0)
    sim._check(rc, "pt_set_guess_net")


def guess_counts(sim):
    """(solves that evaluated the network, solves that kept its proposal)."""
    t, u = C.c_int64(), C.c_int64()
    sim._check(sim.lib.pt_guess_counts(sim.ctx, C.byref(t), C.byref(u)), "pt_guess_counts")
    return t.value, u.value


def guess_last(sim):
    """The last proposal: (warm start's L1 residual sum, proposal's, step alpha)."""
    a, b, al = C.c_double(), C.c_double(), C.c_double()
    sim._check(sim.lib.pt_guess_last(sim.ctx, C.byref(a), C.byref(b), C.byref(al)), "pt_guess_last")
    return a.value, b.value, al.value
