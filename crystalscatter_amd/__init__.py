"""MI355X-native CrystalScatter per-pixel 2D scattering path.

Python mirror of the reference engine interface (SasCalc_GENERIC_calculation,
Source/sc_calc_generic_gpu.h:77-1000) over the C ABI of libcsx.so
(include/csx.h).  Every pixel is evaluated by the HIP kernels; there is no
CPU fallback.

    from crystalscatter_amd import SasCalc
    calc = SasCalc(device=0, params={"LType": 5, "EditRadius": 8, ...})
    img = calc.doCalculation()            # IDX order, like data() of the reference
    calc.image2d(img)                     # rows = i (qx), columns = ihex (qy)
"""
import ctypes

import numpy as np

from ._lib import (CSX_PREC_MIXED, CSX_PREC_REF64, CsxError, Grid, Mask, NCOEF, NIM2, Scalars,
                   lib)

__all__ = ["SasCalc", "Context", "CsxError", "CSX_PREC_MIXED", "CSX_PREC_REF64", "PRECISIONS", "simplex",
           "nm_batch", "nm_eval", "Random", "render_batch", "tpv_images"]

PRECISIONS = {"ref64": CSX_PREC_REF64, "mixed": CSX_PREC_MIXED, CSX_PREC_REF64: CSX_PREC_REF64,
              CSX_PREC_MIXED: CSX_PREC_MIXED}


def _prec(p):
    if p not in PRECISIONS:
        raise ValueError("precision must be 'ref64' or 'mixed'")
    return PRECISIONS[p]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class SasCalc:
    """Engine with the reference's key=value parameters (Source/sc_calc_generic.cpp:36-138).

    device < 0 gives a host-only engine (preparation and tables, no GPU work).
    """

    def __init__(self, device=0, params=None):
        self._L = lib()
        self._h = self._L.csx_engine_new(int(device))
        if not self._h:
            raise CsxError("could not create the engine on device %d (no HIP device?)" % device)
        self.device = device
        if params:
            self.set_params(params)

    def close(self):
        if getattr(self, "_h", None):
            self._L.csx_engine_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _err(self):
        b = ctypes.create_string_buffer(512)
        self._L.csx_engine_last_error(self._h, b, 512)
        return b.value.decode(errors="replace")

    def set(self, key, value):
        if self._L.csx_engine_set(self._h, key.encode(), float(value)) != 0:
            raise KeyError("unknown parameter key %r" % key)

    def set_params(self, params):
        for k, v in params.items():
            self.set(k, v)

    def get(self, key):
        return self._L.csx_engine_get(self._h, key.encode())

    def load_ini(self, path):
        """Set the parameters of a reference .ini file (SC_CalcCons::loadParameter,
        Source/sc_calcCons.cpp:107-399).  Returns (fitflags, unused keys)."""
        from .formats import read_ini
        params, fitflags, unused = read_ini(path)
        for k, v in params.items():
            try:
                self.set(k, v)
            except KeyError:
                unused.append(k)
        return fitflags, unused

    def set_keyvalues(self, text, center_midpoint=False):
        """Set the parameters of a key=value list ("radius=5;cbparticle=cube;...", the
        console's --mcpval, Source/sc_mainCons.cpp:858-893).  Returns the pairs not applied:
        names the list does not define and console-only settings."""
        from .formats import read_keyvalues
        params, ignored = read_keyvalues(text, center_midpoint)
        for k, v in params.items():
            try:
                self.set(k, v)
            except KeyError:
                ignored.append("%s=%r" % (k, v))
        return ignored

    def prepare(self):
        rc = self._L.csx_engine_prepare(self._h)
        if rc != 0:
            raise CsxError("prepare failed (%d): %s" % (rc, self._err()))

    def scalars(self):
        return self._L.csx_engine_scalars(self._h).contents

    def bounds(self):
        """Bounds of the image array (minX, maxX, minY, maxY of the reference): the computed
        block, or the whole image when ExpandImage fills it from a quadrant or half."""
        s = self.scalars()
        return s.axmin, s.axmax, s.aymin, s.aymax

    def block(self):
        """Bounds of the computed block (the fit array's and the batched images' layout)."""
        s = self.scalars()
        return s.xmin, s.xmax, s.ymin, s.ymax

    def doCalculation(self, precision="mixed"):
        """Image in the reference IDX order ((ihex-xmin) + W*(i-ymin)), like data()."""
        self.prepare()
        x0, x1, y0, y1 = self.bounds()
        out = np.empty((x1 - x0) * (y1 - y0), dtype=np.float64)
        b = (ctypes.c_int32 * 4)()
        rc = self._L.csx_engine_calc(self._h, _prec(precision), _dp(out), b)
        if rc != 0:
            raise CsxError("doCalculation failed (%d): %s" % (rc, self._err()))
        return out

    def doCalculationDevice(self, d_out_ptr, precision="mixed"):
        """Image into caller-owned device memory (e.g. torch tensor .data_ptr())."""
        b = (ctypes.c_int32 * 4)()
        rc = self._L.csx_engine_calc_device(self._h, _prec(precision), ctypes.c_void_p(d_out_ptr), b)
        if rc != 0:
            raise CsxError("doCalculation failed (%d): %s" % (rc, self._err()))
        return tuple(b)

    def image2d(self, flat):
        x0, x1, y0, y1 = self.bounds()
        return np.asarray(flat).reshape(y1 - y0, x1 - x0)

    def setFitData(self, w, h, data):
        data = np.ascontiguousarray(data, dtype=np.float64).ravel()
        if data.size != w * h:
            raise ValueError("fit data must hold w*h values")
        rc = self._L.csx_engine_set_fit_data(self._h, int(w), int(h), _dp(data))
        if rc != 0:
            raise CsxError("setFitData failed (%d): %s" % (rc, self._err()))

    def setNoFitRect(self, idx, x0, y0, x1, y1):
        if self._L.csx_engine_set_nofit_rect(self._h, idx, x0, y0, x1, y1) != 0:
            raise ValueError("rectangle index must be 0..3")

    def doFitCalculation(self, bstop=0, border=0, precision="mixed"):
        """Returns (sum, cnt, nancnt) like doFitCalculation(numThreads=GPU, bstop, border, cnt, nancnt)."""
        s = ctypes.c_double()
        c, n = ctypes.c_int64(), ctypes.c_int64()
        rc = self._L.csx_engine_fit(self._h, _prec(precision), int(bstop), int(border), ctypes.byref(s),
                                    ctypes.byref(c), ctypes.byref(n))
        if rc != 0:
            raise CsxError("doFitCalculation failed (%d): %s" % (rc, self._err()))
        return s.value, c.value, n.value

    def table(self, name):
        n, kind = ctypes.c_int64(), ctypes.c_int()
        p = self._L.csx_engine_table(self._h, name.encode(), ctypes.byref(n), ctypes.byref(kind))
        if not p:
            raise KeyError(name)
        ct = {0: ctypes.c_double, 1: ctypes.c_int32, 2: ctypes.c_float}[kind.value]
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct * n.value)).contents).copy()

    def last_kernel_ms(self):
        return self._L.csx_last_kernel_ms(self._L.csx_engine_ctx(self._h))

    def tpvPerformRandom(self, ids, rng=None):
        """Random variation of the listed variables (tpvPerformRandom,
        Source/sc_calc_generic_cpu.cpp:1028-1143): ids is a list of "key=factor" strings; an
        empty list clears.  Returns the reference's report lines.  The varied values apply to
        the next calculation; get() still returns the parameters as set."""
        text = "\n".join(ids)
        if text and rng is None:
            raise ValueError("a Random generator is needed")
        rep = ctypes.create_string_buffer(4096)
        rc = self._L.csx_engine_tpv(self._h, text.encode(), rng._h if rng is not None else None, rep, 4096)
        if rc != 0:
            raise CsxError("tpvPerformRandom: " + self._err())
        return rep.value.decode()

    def tpv_value(self, key):
        """A parameter as the last tpvPerformRandom varied it."""
        return self._L.csx_engine_tpv_value(self._h, key.encode())

    def stage_image(self, ctx, d_out_ptr, cap):
        """Stage this engine's image into ctx's pending batch (see render_batch); d_out_ptr
        has room for cap doubles."""
        if self._L.csx_engine_stage_image(self._h, ctx._h, ctypes.c_void_p(d_out_ptr), int(cap)) != 0:
            raise CsxError("stage_image: " + self._err())

    def stage_points(self, ctx, keys, X, d_out_ptr, stride, threads=16):
        """Stage len(X) images into ctx's pending batch (csx_engine_stage_points): image k has
        this engine's parameters with keys[j] = X[k, j] ("a,b" sets both), output at
        d_out_ptr + k*stride doubles.  Follow with csx_render_batch_run."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        n, npar = X.shape
        if npar != len(keys):
            raise ValueError("X must have one column per key")
        ckeys = (ctypes.c_char_p * npar)(*[k.encode() for k in keys])
        if self._L.csx_engine_stage_points(self._h, ctx._h, n, npar, ckeys, _dp(X), ctypes.c_void_p(d_out_ptr),
                                           int(stride), int(threads)) != 0:
            raise CsxError("stage_points: " + self._err())


class Random:
    """The C library's rand() sequence after srand(seed), in a private state (csx_rand)."""

    def __init__(self, seed=1):
        self._L = lib()
        self._h = self._L.csx_rand_new(int(seed) & 0xFFFFFFFF)
        if not self._h:
            raise CsxError("csx_rand_new failed")

    def close(self):
        if getattr(self, "_h", None):
            self._L.csx_rand_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def next(self):
        return int(self._L.csx_rand_next(self._h))

    def zuf(self, n):
        """n values of the reference's zuf() = (rand()/RAND_MAX*200 - 100)/100."""
        out = np.empty(int(n), dtype=np.float64)
        if self._L.csx_rand_zuf(self._h, _dp(out), out.size) != 0:
            raise CsxError("csx_rand_zuf failed")
        return out

    def skip(self, n):
        """Advance by n rand() values without drawing them."""
        if self._L.csx_rand_skip(self._h, int(n)) != 0:
            raise CsxError("csx_rand_skip failed")

    def zuf_device(self, ctx, out):
        """The next out.numel() zuf() values into a contiguous float64 CUDA tensor, drawn on
        the device (csx_rand_zuf_device); equal to zuf(out.numel())."""
        if out.dtype != torch_float64() or not out.is_cuda or not out.is_contiguous():
            raise CsxError("zuf_device needs a contiguous float64 CUDA tensor")
        import torch
        torch.cuda.current_stream().synchronize()
        ctx._chk(self._L.csx_rand_zuf_device(ctx._h, self._h, out.numel(), ctypes.c_void_p(out.data_ptr())),
                 "csx_rand_zuf_device")
        ctx.synchronize()
        return out


class Context:
    """Device context of the C ABI (csx_create): upload / render / batched objective."""

    def __init__(self, device=0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.csx_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise CsxError("csx_create failed (%d)" % rc)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.csx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _err(self):
        b = ctypes.create_string_buffer(512)
        self._L.csx_last_error(self._h, b, 512)
        return b.value.decode(errors="replace")

    def _chk(self, rc, what):
        if rc != 0:
            raise CsxError("%s failed (%d): %s" % (what, rc, self._err()))

    def set_stream(self, stream_ptr):
        self._chk(self._L.csx_set_stream(self._h, ctypes.c_void_p(stream_ptr)), "csx_set_stream")

    def upload(self, scalars, carr, carr11pm=None, latpar1=None, latpar2=None, latpar3=None):
        carr = np.ascontiguousarray(carr, dtype=np.float64)
        args = [self._h, ctypes.byref(scalars), _dp(carr)]
        keep = [carr]
        if carr11pm is not None and len(carr11pm):
            c11 = np.ascontiguousarray(carr11pm, dtype=np.float64)
            keep.append(c11)
            args.append(_dp(c11))
        else:
            args.append(None)
        for a, ct in ((latpar1, ctypes.c_int32), (latpar2, ctypes.c_int32), (latpar3, ctypes.c_float)):
            if a is None:
                args.append(None)
            else:
                arr = np.ascontiguousarray(a, dtype=np.int32 if ct is ctypes.c_int32 else np.float32)
                keep.append(arr)
                args.append(arr.ctypes.data_as(ctypes.POINTER(ct)))
        self._chk(self._L.csx_upload(*args), "csx_upload")

    def render(self, grid, d_out_ptr, precision="mixed"):
        g = Grid(*grid)
        self._chk(self._L.csx_render(self._h, ctypes.byref(g), _prec(precision), ctypes.c_void_p(d_out_ptr)),
                  "csx_render")

    def last_kernel_ms(self):
        return self._L.csx_last_kernel_ms(self._h)

    def synchronize(self):
        self._chk(self._L.csx_synchronize(self._h), "csx_synchronize")

    def set_device_prep(self, on):
        """Batched fits build lattice tables on the device (True, the default) or on the host."""
        self._chk(self._L.csx_set_device_prep(self._h, 1 if on else 0), "csx_set_device_prep")

    def device_prep_counts(self):
        """(sets whose tables the device built, entries whose rounding was checked against the
        C library's exp)."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(self._L.csx_device_prep_counts(self._h, ctypes.byref(a), ctypes.byref(b)), "csx_device_prep_counts")
        return a.value, b.value

    _TABLE_DTYPES = {0: np.float32, 1: np.uint8, 2: np.uint8, 3: np.int32, 4: np.float32, 5: np.float64}

    def fit_batch_table(self, set_index, which):
        """Table `which` of set `set_index` of the last batch run (csx_fit_batch_table): 0
        latpar3, 1 pair records (bytes), 2 FP64 pair records (bytes), 3 bin table, 4 isotropic
        list, 5 (pair-record count, bin scale, |q| min, |q| max, weight bound)."""
        need = ctypes.c_size_t()
        self._chk(self._L.csx_fit_batch_table(self._h, int(set_index), int(which), None, 0, ctypes.byref(need)),
                  "csx_fit_batch_table")
        out = np.empty(need.value, dtype=np.uint8)
        self._chk(self._L.csx_fit_batch_table(self._h, int(set_index), int(which), out.ctypes.data_as(ctypes.c_void_p),
                                              need.value, ctypes.byref(need)), "csx_fit_batch_table")
        return out.view(self._TABLE_DTYPES[which])


def _check_out(out, n):
    if out.dtype != torch_float64() or not out.is_cuda or out.dim() != 2 or out.shape[0] < n or not out.is_contiguous():
        raise ValueError("out must be a contiguous float64 CUDA tensor of shape (>= %d, npx)" % n)


def torch_float64():
    import torch
    return torch.float64


def render_batch(ctx, engines, precision="mixed", out=None):
    """Images of many engines (same grid size and combination) in one launch
    (csx_render_batch_run).  Returns a float64 CUDA tensor (len(engines), npx), IDX order."""
    import torch
    if not engines:
        raise ValueError("no engines")
    engines[0].prepare()
    x0, x1, y0, y1 = engines[0].block()
    npx = (x1 - x0) * (y1 - y0)
    if out is None:
        out = torch.empty((len(engines), npx), dtype=torch.float64, device="cuda")
    _check_out(out, len(engines))
    L = ctx._L
    ctx._chk(L.csx_fit_batch_begin(ctx._h), "csx_fit_batch_begin")
    for k, e in enumerate(engines):
        e.stage_image(ctx, out[k].data_ptr(), out.shape[1])
    ctx._chk(L.csx_render_batch_run(ctx._h, _prec(precision)), "csx_render_batch_run")
    return out


def tpv_images(ctx, engine, ids, n, rng, precision="mixed", out=None):
    """n images of engine's parameters, each after one tpvPerformRandom(ids, rng) call (the
    AI-data loop's random variation, Source/sc_mainCons.cpp:1236-1246), in one launch.
    Returns (images as a (n, npx) CUDA tensor, the report of every call)."""
    import torch
    engine.tpvPerformRandom([], None)
    engine.prepare()
    x0, x1, y0, y1 = engine.block()
    npx = (x1 - x0) * (y1 - y0)
    if out is None:
        out = torch.empty((n, npx), dtype=torch.float64, device="cuda")
    _check_out(out, n)
    L = ctx._L
    ctx._chk(L.csx_fit_batch_begin(ctx._h), "csx_fit_batch_begin")
    reports = []
    for k in range(n):
        reports.append(engine.tpvPerformRandom(ids, rng))
        engine.stage_image(ctx, out[k].data_ptr(), out.shape[1])
    ctx._chk(L.csx_render_batch_run(ctx._h, _prec(precision)), "csx_render_batch_run")
    return out, reports


def _nm_mode(mode):
    modes = {"reference": _lib.NM_REFERENCE, "corrected": _lib.NM_CORRECTED}
    if mode not in modes:
        raise ValueError("mode must be 'reference' or 'corrected'")
    return modes[mode]


def simplex(f, x0, lo, hi, stp=0.1, maxit=200, ftol=1e-6, nfree=None, mode="reference"):
    """Nelder-Mead on a Python objective f(x) -> float with the simplex of doSimplexFit2D
    (Source/sc_simplexfit2d.cpp:254-611).  x0, lo, hi have npar entries (the reference's
    parameter hash); the first nfree (default all) are fitted.  mode "reference" follows the
    reference exactly (see csx_nm_run_cb), "corrected" is the textbook simplex over the fitted
    entries.  Returns (x, f, iters, evals), x with all npar entries."""
    L = _lib.lib()
    n = len(x0)
    if n > _lib.NM_MAXPAR:
        raise ValueError("at most %d parameters" % _lib.NM_MAXPAR)
    nfree = n if nfree is None else nfree
    cb = _lib.OBJECTIVE(lambda user, p: float(f(np.ctypeslib.as_array(p, shape=(n,)).copy())))
    arr = lambda v: np.ascontiguousarray(v, dtype=np.float64)
    x0, lo, hi = arr(x0), arr(lo), arr(hi)
    r = _lib.NMResult()
    rc = L.csx_nm_run_cb(n, nfree, _nm_mode(mode), cb, None, _dp(x0), _dp(lo), _dp(hi), stp, maxit, ftol,
                         ctypes.byref(r))
    if rc != 0:
        raise CsxError("csx_nm_run_cb failed (%d)" % rc)
    return np.array(r.x[:n]), r.f, r.iters, r.evals


def _nm_fits(fits, n):
    arr = (_lib.NMFit * len(fits))()
    for i, fd in enumerate(fits):
        a = arr[i]
        a.base = fd["engine"]._h
        a.d_data = fd["d_data"]
        a.w, a.h = fd["w"], fd["h"]
        for k in range(16):
            a.mask.rects[k] = -1
        for r, rect in enumerate(fd.get("rects", [])[:4]):
            for k in range(4):
                a.mask.rects[4 * r + k] = rect[k]
        a.mask.bstop, a.mask.border = fd.get("bstop", 0), fd.get("border", 0)
        for j in range(n):
            a.x0[j], a.lo[j], a.hi[j] = fd.get("x0", fd["lo"])[j], fd["lo"][j], fd["hi"][j]
    return arr


def nm_eval(ctx, fits, keys, X, precision="mixed", threads=0):
    """The batched objective (csx_nm_eval) at points X[i] (nv x npar) of every fit i, in one
    launch; returns an (nfits, nv) array (1e20 outside the bounds)."""
    L = _lib.lib()
    n = len(keys)
    X = np.ascontiguousarray(X, dtype=np.float64)
    nf, nv = X.shape[0], X.shape[1]
    if X.shape != (len(fits), nv, n):
        raise ValueError("X must be (nfits, nv, npar)")
    arr = _nm_fits(fits, n)
    ckeys = (ctypes.c_char_p * n)(*[k.encode() for k in keys])
    out = np.empty((nf, nv))
    rc = L.csx_nm_eval(ctx._h if isinstance(ctx, Context) else ctx, nf, arr, n, ckeys, nv, _dp(X), _prec(precision),
                       threads, _dp(out))
    if rc != 0:
        raise CsxError("csx_nm_eval failed (%d)" % rc)
    return out


def nm_batch(ctx, fits, keys, stp=0.1, maxit=200, ftol=1e-6, precision="mixed", threads=0, nfree=None,
             mode="reference", screen=None, margin=0.5):
    """Batched Nelder-Mead (csx_nm_batch).  fits: list of dicts with 'engine' (SasCalc, may be
    host-only), 'd_data' (device pointer), 'w', 'h', 'x0', 'lo', 'hi', optional 'bstop',
    'border', 'rects'.  keys: the npar parameter names, the first nfree (default all) fitted.
    mode: "reference" (doSimplexFit2D exactly) or "corrected" (see simplex).  screen: optional
    callable screen(fit_ids [npts], X [npts, npar]) -> scores [npts] (lower is better): trial
    points scored above (1 + margin) times their simplex's worst vertex score are not evaluated
    (csx_nm_batch_screened).  Returns (list of (x, f, iters, evals), stats dict: rounds, evals,
    prep_ms, stage_ms, run_ms, pruned, screen_ms)."""
    L = _lib.lib()
    n = len(keys)
    nfree = n if nfree is None else nfree
    arr = _nm_fits(fits, n)
    ckeys = (ctypes.c_char_p * n)(*[k.encode() for k in keys])
    res = (_lib.NMResult * len(fits))()
    st = _lib.NMStats()
    opts = None
    if screen is not None:
        errors = []

        def _cb(user, npts, fit_p, x_p, out_p):
            try:
                fit = np.ctypeslib.as_array(fit_p, shape=(npts,)).copy()
                X = np.ctypeslib.as_array(x_p, shape=(npts, n)).copy()
                sc = np.asarray(screen(fit, X), dtype=np.float64).reshape(npts)
                np.ctypeslib.as_array(out_p, shape=(npts,))[:] = sc
                return 0
            except Exception as exc:  # reported after the call; the driver stops on a non-zero return
                errors.append(exc)
                return 1

        cb = _lib.SCREEN(_cb)
        opts = _lib.NMScreenOpts(cb, None, float(margin))
    rc = L.csx_nm_batch_screened(ctx._h if isinstance(ctx, Context) else ctx, len(fits), arr, n, nfree,
                                 _nm_mode(mode), ckeys, stp, maxit, ftol, _prec(precision), threads,
                                 ctypes.byref(opts) if opts is not None else None, res, ctypes.byref(st))
    if screen is not None and errors:
        raise errors[0]
    if rc != 0:
        raise CsxError("csx_nm_batch failed (%d)" % rc)
    stats = {k: getattr(st, k) for k, _ in _lib.NMStats._fields_}
    return [(np.array(r.x[:n]), r.f, r.iters, r.evals) for r in res], stats
