"""Host side of the training-image pipeline: the beam-stop and line rectangles built for all
images at once (crystalscatter_amd.aidata._zero_rects) zero exactly the pixels the per-image
CPU restatement of the reference's loops zeroes (oracle/aidata_oracle.py,
Source/sc_mainCons.cpp:1298-1348), including lines past the image edge."""
import os
import sys

import numpy as np
import pytest

from crystalscatter_amd import Random, aidata

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import aidata_oracle as AO  # noqa: E402


def _apply(rects, bounds, npx):
    """csx_post_zero_rects's rule on the CPU: cell (x, y) zeroes (x - x0) + sx (y - y0) when in range."""
    x0, x1, y0, y1 = bounds
    img = np.ones(npx)
    for xa, xb, ya, yb in rects:
        for y in range(ya, yb + 1):
            for x in range(xa, xb + 1):
                ii = (x - x0) + (x1 - x0) * (y - y0)
                if 0 <= ii < npx:
                    img[ii] = 0
    return img


@pytest.mark.parametrize("bounds,beam,kw", [
    ((-16, 16, -16, 16), (3, -4), dict(bs=(0.2, 0.2, 0.5, 0.5), lines_x=2, lines_y=1)),
    ((0, 24, 0, 20), (20, 2), dict(bs=(0.9, 0.9, 2.0, 2.0), lines_x=3, lines_x_width=4, lines_y=2, lines_y_width=3)),
    ((-10, 10, -12, 12), (0, 0), dict(lines_y=5, lines_y_width=2)),
    ((-8, 8, -8, 8), (7, 7), dict(bs=(0.5, 0.5, 1.0, 1.0))),
])
def test_rectangles_zero_what_the_reference_zeroes(bounds, beam, kw):
    opts = aidata.AIOptions(**kw)
    x0, x1, y0, y1 = bounds
    npx = (x1 - x0) * (y1 - y0)
    nimg, nx = 40, aidata._n_extra(opts)
    rng = Random(5)
    extra = rng.zuf(nimg * nx).reshape(nimg, nx)
    rects, centres = aidata._zero_rects(extra, opts, bounds, beam)
    # the oracle draws the same values from a fresh generator, image after image
    rng2 = Random(5)
    for k in range(nimg):
        ref, centre, _ = AO.post_image(np.ones(npx), bounds, beam, rng2, opts)
        assert centre == centres[k]
        assert np.array_equal(_apply(rects[k], bounds, npx), ref), k
