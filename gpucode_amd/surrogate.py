"""Python side of the sweep surrogate (include/dem_surrogate.h): an MLP
8 -> 128 -> 128 -> 128 -> 2 (SiLU) trained and evaluated by the HIP kernels
of libgpucode_amd.so, plus the normalisation of sweep points and outcomes.
"""
import ctypes as C

import numpy as np

from . import _capi as A
from ._capi import check, fptr


class MLP:
    IN, HIDDEN, OUT = 8, 128, 2

    def __init__(self, seed=0):
        self._m = C.c_void_p()
        check(A.lib().sur_create(int(seed), C.byref(self._m)))

    @staticmethod
    def nparams():
        return A.lib().sur_nparams()

    def close(self):
        if self._m:
            A.lib().sur_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.nparams(), np.float32)
        check(A.lib().sur_get_params(self._m, fptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.nparams(),)
        check(A.lib().sur_set_params(self._m, fptr(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), self.OUT), np.float32)
        check(A.lib().sur_forward(self._m, len(x), fptr(x), fptr(y)))
        return y

    def forward_noalloc(self, n, x_dev, y_dev):
        """sur_forward_noalloc: x_dev (n x 8 FP32) and y_dev (n x 2 FP32) are
        device addresses (ints) on the network's GPU."""
        check(A.lib().sur_forward_noalloc(self._m, int(n), C.c_void_p(x_dev), C.c_void_p(y_dev)))

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.nparams(), np.float32)
        loss = C.c_float()
        check(A.lib().sur_loss_grad(self._m, len(x), fptr(x), fptr(t), C.byref(loss), fptr(g)))
        return loss.value, g

    def train(self, x, t, epochs, batch=256, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        check(A.lib().sur_train(self._m, len(x), fptr(x), fptr(t), int(epochs), int(batch),
                                float(lr), int(seed), C.byref(loss)))
        return loss.value


def unpack(p):
    """Split the flat parameter vector into (W1, b1, W2, b2, W3, b3, W4, b4)."""
    H, I, O = MLP.HIDDEN, MLP.IN, MLP.OUT
    shapes = [(H, I), (H,), (H, H), (H,), (H, H), (H,), (O, H), (O,)]
    out, k = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(np.asarray(p[k:k + n]).reshape(s))
        k += n
    return out


class Normalizer:
    """Affine maps of the sweep points to [-1, 1] (from their ranges) and of
    the outcomes to zero mean and unit variance (from the training set)."""

    def __init__(self, lo, hi, y):
        self.lo, self.hi = np.asarray(lo, float), np.asarray(hi, float)
        y = np.asarray(y, float)
        self.ym, self.ys = y.mean(0), y.std(0) + 1e-12

    def x(self, x):
        return (2.0 * (np.asarray(x, float) - self.lo) / (self.hi - self.lo) - 1.0).astype(np.float32)

    def y(self, y):
        return ((np.asarray(y, float) - self.ym) / self.ys).astype(np.float32)

    def y_inv(self, yn):
        return np.asarray(yn, float) * self.ys + self.ym
