"""CPU oracle of the AI-data post-processing (crystalscatter_amd.aidata) -- test infrastructure
only: the checker for the batched GPU pipeline.  Only tests/ import it.

Restates, per image and in the reference's order, the steps of the console's AI-data loop
after each calculation (Source/sc_mainCons.cpp:1254-1400): optional scaleIntensity (linear),
noise d + z sqrt(d) with one zuf() per pixel, the beam stop (four zuf() draws), the lines
(one zuf() per line), the log scaling of the image for the r,phi input (axes skipped), the
r,phi resampling with its log rescaling and the FFT image of the r,phi image (calculateIFFT).  scaleIntensity, generateRphi and scaleAndClipData
are the C restatements of liboracle.so (pinned bit for bit against the reference's own
routines, tests/golden/post.npz); the random variation comes from the product engine's
tpvPerformRandom, pinned against the reference's run (tests/golden/tpv_fcc.npz).
"""
import math

import numpy as np

import oracle as O


def _zuf(rng):
    return float(rng.zuf(1)[0])


def post_image(img, bounds, beam, rng, opts):
    """One image's post-processing, drawing from rng exactly where the reference does.
    Returns (image, centre, rphi or None)."""
    x0, x1, y0, y1 = (int(b) for b in bounds)
    sx, sy = x1 - x0, y1 - y0
    n = sx * sy
    src = np.array(img, dtype=np.float64)
    if opts.scale_scat:
        src = O.scale_intensity(src, 1)
    if opts.noise:
        z = rng.zuf(n)
        with np.errstate(invalid="ignore"):  # d = d + z sqrt(d), z/100 for scaled images (:1279-1287)
            src = src + ((z / 100.0) if opts.scale_scat else z) * np.sqrt(src)
    bsx0, bsy0 = int(beam[0]), int(beam[1])
    bsdx = bsdy = 2
    if opts.bs[0] > 0:
        bsx0 = int(bsx0 * (1.0 + opts.bs[0] * _zuf(rng)))
        bsy0 = int(bsy0 * (1.0 + opts.bs[1] * _zuf(rng)))
        bsdx = int(bsdx * (1.0 + opts.bs[2] * _zuf(rng)))
        bsdy = int(bsdy * (1.0 + opts.bs[3] * _zuf(rng)))
        X, Y = np.meshgrid(np.arange(x0, x1), np.arange(y0, y1))
        box = (X >= bsx0 - bsdx) & (X <= bsx0 + bsdx) & (Y >= bsy0 - bsdy) & (Y <= bsy0 + bsdy)
        src[((X - x0) + sx * (Y - y0))[box]] = 0
    for count, width, horizontal in ((opts.lines_x, opts.lines_x_width, True),
                                     (opts.lines_y, opts.lines_y_width, False)):
        if count <= 0:
            continue
        span = ((x1 - x0 + 1) if horizontal else (y1 - y0 + 1)) // count
        start = x0 if horizontal else y0
        for i in range(count):
            p = int(start + i * span + math.fabs(_zuf(rng)) * span)
            for a in range(p, p + width):
                if horizontal:
                    ii = (a - x0) + sx * (np.arange(y0, y1) - y0)
                else:
                    ii = (np.arange(x0, x1) - x0) + sx * (a - y0)
                ii = ii[(ii >= 0) & (ii < n)]
                src[ii] = 0
    rphi = None
    if opts.rphi_size > 0:
        data = src
        if not opts.fft_input_lin:
            # log scaling of the image for the r,phi input, pixels on the axes left out of the
            # range when the image contains the origin (:1366-1384)
            lo, hi = 1e26, 0.0
            for i in range(n):
                x, y = i % sx + x0, i // sx + y0
                if (x0 < 0 and x != 0 and y != 0) or x0 == 0:
                    if src[i] > 0:
                        v = math.log10(src[i])
                        if v < lo:
                            lo = v
                        elif v > hi:
                            hi = v
            data = np.where(src > 0, (np.log10(np.where(src > 0, src, 1.0)) - lo) / (hi - lo), 0.0)
        rphi = O.rphi(data, (x0, x1, y0, y1), bsx0, bsy0, opts.rphi_size, opts.rphi_scale, opts.rphi_clip,
                      opts.rphi_clip40)
        rphi = O.scale_clip(rphi, genlog=True)
        if opts.fft_size:
            fft = O.ifft(rphi, opts.rphi_size, False, opts.fft_type, opts.fft_scale, opts.fft_clip, opts.fft_clip40,
                         opts.fft_swap, opts.fft_cut[0], opts.fft_cut[1])
            return src, (bsx0, bsy0), rphi, fft
    return src, (bsx0, bsy0), rphi
