"""The learned surrogate of an expensive body's signed distance (``cc_mlp_*`` in
include/cutcell_hip.h): a Fourier-feature MLP trained and evaluated by the library's own BF16-MFMA
kernels.  No counterpart in the reference, whose bodies are SDFLibrary closures evaluated exactly."""
import ctypes

import numpy as np

from ._lib import check, lib, make_geometry



class Surrogate:
    """3 -> Fourier features (``nfreq`` frequencies, N(0, sigma^2)) -> ``depth`` x ``width`` SiLU
    -> 1.  Parameters are FP32 on the device; the matrix products run in BF16 with FP32
    accumulation.  Training maps its box [lo, hi]^3 onto the unit box the features see."""

    def __init__(self, nfreq=64, width=256, depth=4, sigma=8.0, seed=0):
        L = lib()
        h = ctypes.c_void_p()
        check(L.cc_mlp_create(nfreq, width, depth, float(sigma), int(seed), ctypes.byref(h)))
        self._h = h
        self.nfreq, self.width, self.depth, self.sigma = nfreq, width, depth, sigma
        n, f3 = ctypes.c_int64(), ctypes.c_int64()
        check(L.cc_mlp_num_params(h, ctypes.byref(n), ctypes.byref(f3)))
        self.nparam = n.value

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().cc_mlp_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.empty(self.nparam, np.float32)
        B = np.empty((self.nfreq, 3), np.float32)
        check(lib().cc_mlp_get_params(self._h, p.ctypes.data, B.ctypes.data))
        return p, B

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        check(lib().cc_mlp_set_params(self._h, p.ctypes.data))

    def eval(self, x):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1, 3))
        y = np.empty(x.shape[0], np.float32)
        check(lib().cc_mlp_eval(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1, 3))
        y = np.ascontiguousarray(np.asarray(y, dtype=np.float32))
        g = np.empty(self.nparam, np.float32)
        loss = ctypes.c_float()
        check(lib().cc_mlp_loss_grad(self._h, x.ctypes.data, y.ctypes.data, x.shape[0],
                                       ctypes.byref(loss), g.ctypes.data))
        return float(loss.value), g

    def train(self, geo, nsamples=1 << 22, lo=-1.0, hi=1.0, band=0.05, epochs=24, batch=16384,
              lr=5e-3, seed=1):
        """Fit to exact samples of ``geo`` (half uniform in [lo, hi]^3, half within ``band`` of the
        surface); returns the per-epoch mean training loss.  Afterwards the features see
        (x - lo) / (hi - lo)."""
        g = make_geometry(geo)
        hist = np.zeros(epochs, np.float32)
        check(lib().cc_mlp_train(self._h, ctypes.byref(g), int(nsamples), float(lo), float(hi),
                                   float(band), int(epochs), int(batch), float(lr), int(seed),
                                   hist.ctypes.data))
        return hist
