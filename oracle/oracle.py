"""ctypes wrapper of the CPU oracle (liboracle.so) -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  The product never does.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run: make -C oracle)")
        L = ctypes.CDLL(path)
        vp, cp, d, i = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_int
        L.csxo_new.restype = vp
        pd = ctypes.POINTER(d)
        L.csxo_scale_intensity.argtypes = [pd, ctypes.c_long, i]
        L.csxo_scale_clip.argtypes = [pd, ctypes.c_long, i, i, i, i]
        L.csxo_rphi.argtypes = [pd, i, i, i, i, i, i, i, pd, i, i, i]
        L.csxo_ifft.argtypes = [pd, i, i, i, i, i, i, i, i, i, pd]
        L.csxo_ifft.restype = ctypes.c_long
        L.csxo_free.argtypes = [vp]
        L.csxo_set.argtypes = [vp, cp, d]
        L.csxo_set.restype = i
        L.csxo_get.argtypes = [vp, cp]
        L.csxo_get.restype = d
        L.csxo_prepare.argtypes = [vp]
        L.csxo_prepare.restype = i
        L.csxo_combo_key.argtypes = [vp]
        L.csxo_combo_key.restype = ctypes.c_uint
        L.csxo_bounds.argtypes = [vp, ctypes.POINTER(i)]
        L.csxo_idx2q.argtypes = [vp, i, i, ctypes.POINTER(d)]
        L.csxo_pixel.argtypes = [vp, d, d, d, ctypes.POINTER(d)]
        L.csxo_pixel.restype = i
        L.csxo_image.argtypes = [vp, ctypes.POINTER(d), i]
        L.csxo_image.restype = i
        L.csxo_image_points.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.c_long,
                                        ctypes.POINTER(d), i]
        L.csxo_image_points.restype = i
        L.csxo_fit_fqs.argtypes = [vp, ctypes.POINTER(d), i, i, i, i, ctypes.POINTER(i),
                                   ctypes.POINTER(d), ctypes.POINTER(ctypes.c_long),
                                   ctypes.POINTER(ctypes.c_long), i]
        L.csxo_fit_fqs.restype = d
        L.csxo_table.argtypes = [vp, cp, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(i)]
        L.csxo_table.restype = vp
        L.csxo_scalar.argtypes = [vp, cp]
        L.csxo_scalar.restype = d
        L.csxo_gamma.argtypes = [d]
        L.csxo_gamma.restype = d
        L.csxo_gaussnorm3.argtypes = [d]
        L.csxo_gaussnorm3.restype = d
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class OracleEngine:
    """FP64 CPU restatement of SasCalc_GENERIC_calculation for the hot path."""

    def __init__(self, params=None):
        self._L = lib()
        self._h = self._L.csxo_new()
        if params:
            self.set(params)

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.csxo_free(self._h)
            self._h = None

    def set(self, params):
        for k, v in params.items():
            if self._L.csxo_set(self._h, k.encode(), float(v)) != 0:
                raise KeyError(k)

    def prepare(self):
        rc = self._L.csxo_prepare(self._h)
        if rc != 0:
            raise RuntimeError("oracle prepare failed (%d)" % rc)

    def combo_key(self):
        return self._L.csxo_combo_key(self._h)

    def bounds(self):
        b = (ctypes.c_int * 4)()
        self._L.csxo_bounds(self._h, b)
        return tuple(b)

    def idx2q(self, ihex, i):
        q = (ctypes.c_double * 3)()
        self._L.csxo_idx2q(self._h, int(ihex), int(i), q)
        return tuple(q)

    def pixel(self, qx, qy, qz):
        v = ctypes.c_double()
        if self._L.csxo_pixel(self._h, qx, qy, qz, ctypes.byref(v)) != 0:
            raise RuntimeError("combination not restated")
        return v.value

    def image(self, nthreads=1):
        x0, x1, y0, y1 = self.bounds()
        out = np.empty((y1 - y0) * (x1 - x0), dtype=np.float64)
        if self._L.csxo_image(self._h, _dp(out), int(nthreads)) != 0:
            raise RuntimeError("oracle image failed")
        return out

    def image_points(self, ihex, ii, nthreads=1):
        ihex = np.ascontiguousarray(ihex, dtype=np.int32)
        ii = np.ascontiguousarray(ii, dtype=np.int32)
        out = np.empty(ihex.size, dtype=np.float64)
        if self._L.csxo_image_points(self._h, _ip(ihex), _ip(ii), ihex.size, _dp(out), int(nthreads)) != 0:
            raise RuntimeError("oracle points failed")
        return out

    def fit_fqs(self, data, w, h, bstop=0, border=0, rects=None, nthreads=1):
        data = np.ascontiguousarray(data, dtype=np.float64)
        fqs = np.zeros(w * h, dtype=np.float64)
        r = np.full(16, -1, dtype=np.int32)
        if rects:
            for k, rc in enumerate(rects[:4]):
                r[4 * k:4 * k + 4] = rc
        cnt, nan = ctypes.c_long(), ctypes.c_long()
        s = self._L.csxo_fit_fqs(self._h, _dp(data), w, h, bstop, border, _ip(r), _dp(fqs),
                                 ctypes.byref(cnt), ctypes.byref(nan), int(nthreads))
        return s, cnt.value, nan.value, fqs

    def table(self, name):
        n, kind = ctypes.c_long(), ctypes.c_int()
        p = self._L.csxo_table(self._h, name.encode(), ctypes.byref(n), ctypes.byref(kind))
        if not p:
            raise KeyError(name)
        ct = {0: ctypes.c_double, 1: ctypes.c_int, 2: ctypes.c_float}[kind.value]
        arr = ctypes.cast(p, ctypes.POINTER(ct * n.value)).contents
        return np.ctypeslib.as_array(arr).copy()

    def scalar(self, name):
        return self._L.csxo_scalar(self._h, name.encode())


def gamma(z):
    return lib().csxo_gamma(float(z))


def gaussnorm3(a):
    return lib().csxo_gaussnorm3(float(a))


_OBJ = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double))


def simplex(f, x0, numfit, minv, maxv, stp=0.1, maxit=200, ftol=1e-6):
    """Nelder-Mead as doSimplexFit2D runs it (returns x, fbest, iterations, evaluations)."""
    L = lib()
    L.csxo_simplex.restype = ctypes.c_int
    L.csxo_simplex.argtypes = [_OBJ, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                               ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_int),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
    m = len(x0)
    x = np.array(x0, dtype=np.float64)
    lo = np.array(minv, dtype=np.float64)
    hi = np.array(maxv, dtype=np.float64)
    cb = _OBJ(lambda user, p: float(f(np.ctypeslib.as_array(p, shape=(m,)).copy())))
    it, fb, nev = ctypes.c_int(), ctypes.c_double(), ctypes.c_long()
    rc = L.csxo_simplex(cb, None, numfit, m, _dp(x), _dp(lo), _dp(hi), stp, maxit, ftol, ctypes.byref(it),
                        ctypes.byref(fb), ctypes.byref(nev))
    if rc != 0:
        raise RuntimeError("simplex failed")
    return x, fb.value, it.value, nev.value


def _pd(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def scale_intensity(img, linlog, extra=True):
    """scaleIntensity of an image (a copy); extra: the engine also scans one zero past it."""
    a = np.ascontiguousarray(np.append(img, 0.0) if extra else img, dtype=np.float64).copy()
    lib().csxo_scale_intensity(_pd(a), a.size, int(linlog))
    return a[:len(img)]


def scale_clip(data, scale=False, clip=False, clip40=False, genlog=False):
    a = np.ascontiguousarray(data, dtype=np.float64).copy()
    lib().csxo_scale_clip(_pd(a), a.size, int(scale), int(clip), int(clip40), int(genlog))
    return a


def rphi(img, bounds, xc, yc, s, scale=False, clip=False, clip40=False):
    """generateRphi of img (IDX order over bounds x0,x1,y0,y1); cells never written stay 0."""
    a = np.ascontiguousarray(img, dtype=np.float64)
    out = np.zeros(s * s)
    x0, x1, y0, y1 = (int(b) for b in bounds)
    lib().csxo_rphi(_pd(a), x0, x1, y0, y1, int(xc), int(yc), int(s), _pd(out), int(scale), int(clip), int(clip40))
    return out


def ifft(data, s, forward=False, out_type=0, scaled=False, clip=False, clip40=False, swap=False, cutx=0, cuty=0):
    """calculateIFFT of an s x s input (see csx_oracle.h); None where the reference returns none."""
    a = np.ascontiguousarray(data, dtype=np.float64)
    out = np.zeros(s * s)
    n = lib().csxo_ifft(_pd(a), int(s), int(forward), int(out_type), int(scaled), int(clip), int(clip40), int(swap),
                        int(cutx), int(cuty), _pd(out))
    return out[:n] if n else None
