"""Training images on the GPU: the console's AI-data loop (generateAIfiles,
Source/sc_mainCons.cpp:916-1440) for one parameter line, batched.

Per image the reference runs: tpvPerformRandom, the calculation, then optionally
scaleIntensity (linear), noise d + z sqrt(d), a randomly placed beam stop and random lines of
zeroed pixels, and the r,phi resampling of the (log-scaled) image with its log rescaling.  All
random numbers come from one rand() sequence in that order, image after image.  Here one
native call (csx_engine_tpv_batch) draws them in the same order -- random variation, noise
values, beam-stop and line values -- and prepares and stages the images on host threads; the
device then renders all images in one launch and post-processes them batch-wide
(csx_render_batch_run, csx_post_*), so the images equal a reference run that starts from the
same rand() state.

Options (names after the reference's command-line settings, Source/sc_mainCons.cpp:1000-1120):
scale_scat (tpvScaleScat), noise (tpvAddNoise), bs = (x0, y0, xs, ys) factors (tpvBSx0..ys),
lines_x / lines_x_width and lines_y / lines_y_width (tpvLinesX, ...), rphi_size (fftRphiSize)
with rphi_scale / rphi_clip / rphi_clip40, fft_input_lin (fftInputLin), and the FFT image of
the r,phi image: fft_size (fftOutSize, equal to rphi_size: the reference's FFT without FFTW only
transforms inputs of the output's size), fft_type (fftOutFormat), fft_scale / fft_clip /
fft_clip40, fft_swap (fftOutputSwap), fft_cut (fftOutCutX/Y).  File writing is not part of
this path.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import CsxError, _lib, _prec


@dataclass
class AIOptions:
    scale_scat: bool = False
    noise: bool = False
    bs: tuple = (0.0, 0.0, 0.0, 0.0)
    lines_x: int = 0
    lines_x_width: int = 1
    lines_y: int = 0
    lines_y_width: int = 1
    rphi_size: int = 0
    rphi_scale: bool = False
    rphi_clip: bool = False
    rphi_clip40: bool = False
    fft_input_lin: bool = False
    fft_size: int = 0          # fftOutSize: the FFT image of the r,phi image (same size)
    fft_type: int = 0          # fftOutFormat: 0 real, 1 imaginary, 2 modulus, 3 re^2 - im^2
    fft_scale: bool = False
    fft_clip: bool = False
    fft_clip40: bool = False
    fft_swap: bool = False
    fft_cut: tuple = (0, 0)    # fftOutCutX, fftOutCutY


def _n_extra(opts):
    """zuf() values the beam stop and the lines of one image draw."""
    return (4 if opts.bs[0] > 0 else 0) + max(0, opts.lines_x) + max(0, opts.lines_y)


def _zero_rects(extra, opts, bounds, beam):
    """The beam stop and the lines of every image as rectangles (xa, xb, ya, yb; inclusive) for
    csx_post_zero_rects, and the beam-stop centres, from the zuf() values each image drew
    (extra: one row per image, in the reference's order; Source/sc_mainCons.cpp:1298-1348).
    The arithmetic is the reference's, elementwise: int() of the same double expressions."""
    x0, x1, y0, y1 = bounds
    n = extra.shape[0]
    full = lambda v: np.full(n, v, dtype=np.int64)
    bsx0, bsy0 = full(int(beam[0])), full(int(beam[1]))
    rects, col = [], 0
    if opts.bs[0] > 0:
        bsx0 = np.trunc(bsx0 * (1.0 + opts.bs[0] * extra[:, 0])).astype(np.int64)
        bsy0 = np.trunc(bsy0 * (1.0 + opts.bs[1] * extra[:, 1])).astype(np.int64)
        bsdx = np.trunc(2 * (1.0 + opts.bs[2] * extra[:, 2])).astype(np.int64)
        bsdy = np.trunc(2 * (1.0 + opts.bs[3] * extra[:, 3])).astype(np.int64)
        # the reference walks the image's x and y and tests the box
        rects.append(np.stack([np.maximum(bsx0 - bsdx, x0), np.minimum(bsx0 + bsdx, x1 - 1),
                               np.maximum(bsy0 - bsdy, y0), np.minimum(bsy0 + bsdy, y1 - 1)], 1))
        col = 4
    if opts.lines_x > 0:
        xc = (x1 - x0 + 1) // opts.lines_x
        for i in range(opts.lines_x):
            xp = np.trunc((x0 + i * xc) + np.abs(extra[:, col]) * xc).astype(np.int64)
            col += 1
            rects.append(np.stack([xp, xp + opts.lines_x_width - 1, full(y0), full(y1 - 1)], 1))
    if opts.lines_y > 0:
        yc = (y1 - y0 + 1) // opts.lines_y
        for i in range(opts.lines_y):
            yp = np.trunc((y0 + i * yc) + np.abs(extra[:, col]) * yc).astype(np.int64)
            col += 1
            rects.append(np.stack([full(x0), full(x1 - 1), yp, yp + opts.lines_y_width - 1], 1))
    big = np.iinfo(np.int32)
    r = np.ascontiguousarray(np.clip(np.stack(rects, 1), big.min, big.max).astype(np.int32)) if rects else None
    return r, list(zip(bsx0.tolist(), bsy0.tolist()))


def generate(ctx, engine, ids, n, rng, opts=None, precision="mixed", threads=16):
    """n training images of engine's parameters with random variation `ids` (see
    SasCalc.tpvPerformRandom), post-processed per `opts`.  Returns a dict with "scat"
    ((n, npx) float64 CUDA tensor, IDX order), "rphi" ((n, s*s), when rphi_size > 0), "fft"
    (when fft_size > 0), "bounds" and "centres" (the beam-stop centre per image).  The
    parameter values of each image are those tpvPerformRandom gives (see SasCalc.tpv_value
    for one call at a time)."""
    import torch
    opts = opts or AIOptions()
    if not 1 <= n <= 65535:
        raise ValueError("1..65535 images per call (the post-processing batch limit)")
    if opts.fft_size and opts.fft_size != opts.rphi_size:
        raise ValueError("the FFT image needs fft_size == rphi_size (the r,phi image is its input)")
    L = _lib.lib()
    engine.tpvPerformRandom([], None)
    engine.prepare()
    bounds = tuple(engine.block())
    x0, x1, y0, y1 = bounds
    npx = (x1 - x0) * (y1 - y0)
    beam = (engine.get("BeamPosX"), engine.get("BeamPosY"))
    scat = torch.empty((n, npx), dtype=torch.float64, device="cuda")
    zt = torch.empty((n, npx), dtype=torch.float64, device="cuda") if opts.noise else None
    _torch_done()
    nx = _n_extra(opts)
    extra = np.empty((n, max(nx, 1)))
    ctx._chk(L.csx_fit_batch_begin(ctx._h), "csx_fit_batch_begin")
    # random variation, noise (drawn on the device) and beam-stop/line values in the loop's
    # order, preparation and staging on host threads (csx_engine_tpv_batch)
    pd = ctypes.POINTER(ctypes.c_double)
    rc = L.csx_engine_tpv_batch(engine._h, ctx._h, "\n".join(ids).encode(), rng._h, n,
                                ctypes.c_void_p(scat.data_ptr()), npx,
                                ctypes.c_void_p(zt.data_ptr()) if zt is not None else None, nx,
                                extra.ctypes.data_as(pd) if nx else None, int(threads))
    if rc != 0:
        raise CsxError("csx_engine_tpv_batch: " + engine._err())
    rects, centres = _zero_rects(extra[:, :nx], opts, bounds, beam)
    ctx._chk(L.csx_render_batch_run(ctx._h, _prec(precision)), "csx_render_batch_run")
    d = ctypes.c_void_p(scat.data_ptr())
    if opts.scale_scat:
        ctx._chk(L.csx_post_scale_intensity(ctx._h, d, npx, n, 1), "csx_post_scale_intensity")
    keep = [zt]
    if opts.noise:
        ctx._chk(L.csx_post_noise(ctx._h, d, ctypes.c_void_p(zt.data_ptr()), n * npx, int(opts.scale_scat)),
                 "csx_post_noise")
    if rects is not None:
        ctx._chk(L.csx_post_zero_rects(ctx._h, d, npx, n, rects.ctypes.data_as(ctypes.c_void_p), rects.shape[1],
                                       x0, x1, y0, y1), "csx_post_zero_rects")
    out = {"scat": scat, "bounds": bounds, "centres": centres}
    if opts.rphi_size > 0:
        s = int(opts.rphi_size)
        src = scat
        if not opts.fft_input_lin:
            ctx._chk(L.csx_synchronize(ctx._h), "csx_synchronize")
            src = scat.clone()
            _torch_done()
            ctx._chk(L.csx_post_log_scale(ctx._h, ctypes.c_void_p(src.data_ptr()), npx, n, x1 - x0, x0, y0, 1),
                     "csx_post_log_scale")
        rphi = torch.empty((n, s * s), dtype=torch.float64, device="cuda")
        cen = np.ascontiguousarray(np.array(centres, dtype=np.int64).reshape(n, 2).astype(np.int32))
        ctx._chk(L.csx_post_rphi_many(ctx._h, ctypes.c_void_p(src.data_ptr()), x0, x1, y0, y1,
                                      cen.ctypes.data_as(ctypes.c_void_p), s, n, ctypes.c_void_p(rphi.data_ptr()),
                                      int(opts.rphi_scale), int(opts.rphi_clip), int(opts.rphi_clip40)),
                 "csx_post_rphi_many")
        ctx._chk(L.csx_post_scale_clip(ctx._h, ctypes.c_void_p(rphi.data_ptr()), s * s, n, 0, 0, 0, 1),
                 "csx_post_scale_clip")
        out["rphi"] = rphi
        keep.append(src)
        if opts.fft_size:
            cx, cy = (int(v) for v in opts.fft_cut)
            cut = 0 < cx < s and 0 < cy < s
            fft = torch.empty((n, cx * cy if cut else s * s), dtype=torch.float64, device="cuda")
            ctx._chk(L.csx_post_ifft(ctx._h, ctypes.c_void_p(rphi.data_ptr()), s, n, 0, int(opts.fft_type),
                                     int(opts.fft_scale), int(opts.fft_clip), int(opts.fft_clip40),
                                     int(opts.fft_swap), cx, cy, ctypes.c_void_p(fft.data_ptr())), "csx_post_ifft")
            out["fft"] = fft
    ctx._chk(L.csx_synchronize(ctx._h), "csx_synchronize")
    return out


def _torch_done():
    """Work queued on torch's stream (copies, clones) is finished before the library's stream
    reads it."""
    import torch
    torch.cuda.current_stream().synchronize()


def post_scale_intensity(ctx, imgs, linlog):
    """scaleIntensity of every row of imgs (float64 CUDA tensor (nimg, npx)), in place."""
    _chk_rows(imgs)
    _torch_done()
    ctx._chk(_lib.lib().csx_post_scale_intensity(ctx._h, ctypes.c_void_p(imgs.data_ptr()), imgs.shape[1],
                                                 imgs.shape[0], int(linlog)), "csx_post_scale_intensity")
    ctx.synchronize()


def post_scale_clip(ctx, data, scale=False, clip=False, clip40=False, genlog=False):
    """scaleAndClipData of every row of data, in place."""
    _chk_rows(data)
    _torch_done()
    ctx._chk(_lib.lib().csx_post_scale_clip(ctx._h, ctypes.c_void_p(data.data_ptr()), data.shape[1], data.shape[0],
                                            int(scale), int(clip), int(clip40), int(genlog)), "csx_post_scale_clip")
    ctx.synchronize()


def post_rphi(ctx, imgs, bounds, centre, s, scale=False, clip=False, clip40=False):
    """generateRphi of every row of imgs around one centre; returns (nimg, s*s)."""
    import torch
    _chk_rows(imgs)
    x0, x1, y0, y1 = (int(b) for b in bounds)
    if (x1 - x0) * (y1 - y0) != imgs.shape[1]:
        raise ValueError("bounds do not match the image size")
    out = torch.empty((imgs.shape[0], s * s), dtype=torch.float64, device="cuda")
    _torch_done()
    ctx._chk(_lib.lib().csx_post_rphi(ctx._h, ctypes.c_void_p(imgs.data_ptr()), x0, x1, y0, y1, int(centre[0]),
                                      int(centre[1]), int(s), imgs.shape[0], ctypes.c_void_p(out.data_ptr()),
                                      int(scale), int(clip), int(clip40)), "csx_post_rphi")
    ctx.synchronize()
    return out


def post_ifft(ctx, data, s, forward=False, out_type=0, scale=False, clip=False, clip40=False, swap=False, cut=(0, 0)):
    """calculateIFFT of every row of data (s x s each); returns (nimg, s*s) or the cut-out."""
    import torch
    _chk_rows(data)
    if data.shape[1] != s * s:
        raise ValueError("rows must hold s*s values")
    cx, cy = (int(v) for v in cut)
    n = cx * cy if (0 < cx < s and 0 < cy < s) else s * s
    out = torch.empty((data.shape[0], n), dtype=torch.float64, device="cuda")
    _torch_done()
    ctx._chk(_lib.lib().csx_post_ifft(ctx._h, ctypes.c_void_p(data.data_ptr()), int(s), data.shape[0], int(forward),
                                      int(out_type), int(scale), int(clip), int(clip40), int(swap), cx, cy,
                                      ctypes.c_void_p(out.data_ptr())), "csx_post_ifft")
    ctx.synchronize()
    return out


def _chk_rows(t):
    import torch
    if t.dtype != torch.float64 or not t.is_cuda or t.dim() != 2 or not t.is_contiguous():
        raise CsxError("expected a contiguous float64 CUDA tensor (rows, values)")
