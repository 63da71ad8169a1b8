"""ctypes binding of the cost-model MLP (include/vsgm_mlp.h) in
libvsgm_hip.so: forward, backward and Adam run in the library's HIP
kernels; numpy arrays in and out."""
import ctypes as C
from typing import Optional

import numpy as np

from .hip import DeviceContext, hiplib
from .api import VsgmError

HIDDEN = 64
MAX_IN = 128

_declared = False


def _lib():
    global _declared
    L = hiplib()
    if not _declared:
        P, u64, i = C.c_void_p, C.c_uint64, C.c_int
        sig = {
            "vsgm_mlp_create": ([P, i, u64, C.POINTER(P)], i),
            "vsgm_mlp_free": ([P], i),
            "vsgm_mlp_param_count": ([P, C.POINTER(u64)], i),
            "vsgm_mlp_get_params": ([P, P, u64], i),
            "vsgm_mlp_set_params": ([P, P, u64], i),
            "vsgm_mlp_forward": ([P, P, u64, P], i),
            "vsgm_mlp_backward": ([P, P, P, u64, C.POINTER(C.c_float), P, u64], i),
            "vsgm_mlp_adam_step": ([P, C.c_float, C.c_float, C.c_float, C.c_float], i),
            "vsgm_mlp_set_mfma": ([P, i], i),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        _declared = True
    return L


class CostModel:
    """3-layer MLP (in -> 64 -> 64 -> 1, ReLU), BF16 weights, FP32 sums."""

    def __init__(self, in_dim: int, seed: int = 0, ctx: Optional[DeviceContext] = None):
        self.ctx = ctx or DeviceContext(0)
        self.in_dim = in_dim
        h = C.c_void_p()
        self.ctx.check(_lib().vsgm_mlp_create(self.ctx.h, in_dim, seed, C.byref(h)), "mlp create")
        self.h = h.value
        n = C.c_uint64()
        self.ctx.check(_lib().vsgm_mlp_param_count(self.h, C.byref(n)), "mlp params")
        self.n_params = n.value

    def close(self):
        if self.h:
            _lib().vsgm_mlp_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def use_mfma(self, on: bool = True):
        """Matrix-core kernels (default) or the FP32-activation kernels."""
        self.ctx.check(_lib().vsgm_mlp_set_mfma(self.h, int(on)), "mlp mfma")

    # parameters --------------------------------------------------------
    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self.ctx.check(_lib().vsgm_mlp_get_params(self.h, p.ctypes.data, p.size), "mlp get")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise VsgmError(f"expected {self.n_params} parameters, got {p.size}")
        self.ctx.check(_lib().vsgm_mlp_set_params(self.h, p.ctypes.data, p.size), "mlp set")

    # compute ------------------------------------------------------------
    def _x(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.ndim != 2 or x.shape[1] != self.in_dim:
            raise VsgmError(f"inputs must be [n, {self.in_dim}]")
        return x

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = self._x(x)
        if len(x) == 0:
            return np.zeros(0, np.float32)
        dx = self.ctx.array(x.reshape(-1))
        dy = self.ctx.empty(len(x), np.float32)
        self.ctx.check(_lib().vsgm_mlp_forward(self.h, dx.ptr, len(x), dy.ptr), "mlp forward")
        return dy.numpy()

    def backward(self, x: np.ndarray, t: np.ndarray, want_grad: bool = True):
        """Loss and gradient (host copy when want_grad) of one batch."""
        x = self._x(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        dx, dt = self.ctx.array(x.reshape(-1)), self.ctx.array(t)
        loss = C.c_float()
        g = np.zeros(self.n_params, np.float32) if want_grad else None
        self.ctx.check(_lib().vsgm_mlp_backward(self.h, dx.ptr, dt.ptr, len(x), C.byref(loss),
                                                 g.ctypes.data if want_grad else None, self.n_params if want_grad else 0),
                       "mlp backward")
        return float(loss.value), g

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int = 30, batch: int = 256, lr: float = 3e-3, seed: int = 0,
              beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        """Minibatch Adam on device-resident data; returns the mean loss of
        every epoch (measured on the minibatches as they were stepped)."""
        x = self._x(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        rng = np.random.default_rng(seed)
        losses = []
        for _ in range(epochs):
            perm = rng.permutation(len(x))
            dx = self.ctx.array(x[perm].reshape(-1))
            dt = self.ctx.array(t[perm])
            tot, nb = 0.0, 0
            for s in range(0, len(x), batch):
                m = min(batch, len(x) - s)
                loss = C.c_float()
                self.ctx.check(_lib().vsgm_mlp_backward(self.h, dx.ptr + 4 * s * self.in_dim, dt.ptr + 4 * s, m,
                                                         C.byref(loss), None, 0), "mlp backward")
                self.ctx.check(_lib().vsgm_mlp_adam_step(self.h, lr, beta1, beta2, eps), "mlp adam")
                tot += loss.value
                nb += 1
            losses.append(tot / max(nb, 1))
        return losses
