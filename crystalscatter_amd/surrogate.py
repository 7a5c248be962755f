"""Surrogate network of the fits over the C ABI (csx_mlp_*, include/csx.h).

The network maps a normalised parameter set and a pixel's (qx, qy) to the normalised
log10-intensity of that pixel.  Forward, backward and Adam run in the library's BF16 MFMA
kernels; torch provides device memory and the stream.  There is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import CsxError


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class MLP:
    """One network on one device; buffers for up to max_batch rows per call."""

    NPAR, NFREQ = _lib.MLP_NPAR, _lib.MLP_NFREQ

    def __init__(self, device=0, max_batch=65536, seed=1):
        import torch
        self._L = _lib.lib()
        h = ctypes.c_void_p()
        rc = self._L.csx_mlp_create(int(device), int(max_batch), int(seed), ctypes.byref(h))
        if rc != 0:
            raise CsxError("csx_mlp_create failed (%d)" % rc)
        self._h = h
        self.device = torch.device("cuda", device)
        self.max_batch = max_batch
        self.nparams = int(self._L.csx_mlp_nparams())
        self._stream = torch.cuda.Stream(device=self.device)
        self._chk(self._L.csx_mlp_set_stream(self._h, ctypes.c_void_p(self._stream.cuda_stream)), "set_stream")

    def close(self):
        if getattr(self, "_h", None):
            self._L.csx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            b = ctypes.create_string_buffer(512)
            self._L.csx_mlp_last_error(self._h, b, 512)
            raise CsxError("csx_mlp_%s failed (%d): %s" % (what, rc, b.value.decode(errors="replace")))

    @property
    def stream(self):
        return self._stream

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        fb = np.empty(2 * self.NFREQ, dtype=np.float32)
        self._chk(self._L.csx_mlp_get_params(self._h, p.ctypes.data_as(_lib._PF), fb.ctypes.data_as(_lib._PF)),
                  "get_params")
        return p, fb

    def set_params(self, p=None, fb=None):
        pa = None if p is None else np.ascontiguousarray(p, dtype=np.float32)
        fa = None if fb is None else np.ascontiguousarray(fb, dtype=np.float32)
        self._chk(self._L.csx_mlp_set_params(self._h, None if pa is None else pa.ctypes.data_as(_lib._PF),
                                             None if fa is None else fa.ctypes.data_as(_lib._PF)), "set_params")

    def get_grads(self):
        g = np.empty(self.nparams, dtype=np.float32)
        self._chk(self._L.csx_mlp_get_grads(self._h, g.ctypes.data_as(_lib._PF)), "get_grads")
        return g

    def _rows(self, par, sets, q):
        n = int(q.shape[0])
        if not (1 <= n <= self.max_batch):
            raise ValueError("1..max_batch rows per call")
        for t, dt in ((par, "float32"), (sets, "int32"), (q, "float32")):
            if t.device != self.device or str(t.dtype) != "torch." + dt or not t.is_contiguous():
                raise ValueError("inputs must be contiguous %s tensors on %s" % (dt, self.device))
        if par.dim() != 2 or par.shape[1] != self.NPAR or q.shape[-1] != 2 or sets.shape[0] != n:
            raise ValueError("par is (nsets, 6), sets (n,), q (n, 2)")
        return n, int(par.shape[0])

    def forward(self, par, sets, q, out=None):
        """y[i] = network(par[sets[i]], q[i]); device tensors; returns a float32 tensor."""
        import torch
        n, ns = self._rows(par, sets, q)
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=self.device)
        self._stream.wait_stream(torch.cuda.current_stream(self.device))
        self._chk(self._L.csx_mlp_forward(self._h, n, ns, _p(par), _p(sets), _p(q), _p(out)), "forward")
        torch.cuda.current_stream(self.device).wait_stream(self._stream)
        return out

    def _target(self, target, n):
        if (target.device != self.device or str(target.dtype) != "torch.float32" or not target.is_contiguous()
                or target.numel() != n):
            raise ValueError("target must be a contiguous float32 tensor of n values on %s" % self.device)

    def gradient(self, par, sets, q, target):
        """Loss 1/n sum (y - target)^2 (synchronous) and the gradient in the model."""
        import torch
        n, ns = self._rows(par, sets, q)
        self._target(target, n)
        loss = ctypes.c_float()
        self._stream.wait_stream(torch.cuda.current_stream(self.device))
        self._chk(self._L.csx_mlp_gradient(self._h, n, ns, _p(par), _p(sets), _p(q), _p(target), ctypes.byref(loss)),
                  "gradient")
        return loss.value

    def train_step(self, par, sets, q, target, lr, want_loss=False):
        import torch
        n, ns = self._rows(par, sets, q)
        self._target(target, n)
        loss = ctypes.c_float()
        self._stream.wait_stream(torch.cuda.current_stream(self.device))
        self._chk(self._L.csx_mlp_train_step(self._h, n, ns, _p(par), _p(sets), _p(q), _p(target), ctypes.c_float(lr),
                                             ctypes.byref(loss) if want_loss else None), "train_step")
        return loss.value if want_loss else None

    def synchronize(self):
        self._chk(self._L.csx_mlp_synchronize(self._h), "synchronize")
