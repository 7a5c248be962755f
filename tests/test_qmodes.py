"""The 1D q mode and the ExpandImage layouts, against the reference's own images
(tests/golden/qmodes.npz, make_golden.py qmodes): image bounds of the host preparation, the
mirror fill of ExpandImage (Source/sc_calc_generic_cpu.cpp:1246-1269) restated as the
gather the device kernel k_expand evaluates, and the 1D q grid (convert_idx2q,
Source/sc_calc_generic_gpu.cu:252-257) through the oracle's pixel values (CPU only)."""
import numpy as np
import pytest

from helpers import fixture, manifest, rel_err

from crystalscatter_amd import SasCalc

CASES = manifest()["fixtures"]["qmodes.npz"]["cases"]
FIX = fixture("qmodes.npz")


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_bounds(name):
    e = SasCalc(device=-1, params=CASES[name])
    e.set("ignore_cbs", 1)
    e.prepare()
    assert tuple(e.bounds()) == tuple(int(v) for v in FIX[name + "_bounds"])


def expand_gather(block, z, quad):
    """The whole image from the computed block, pixel by pixel, as k_expand gathers it."""
    rw = z
    rymin = 0 if quad == 1 else -z
    at = lambda x, y: block[x + rw * (y - rymin)]
    out = np.zeros(4 * z * z)
    for y in range(-z, z):
        for x in range(-z, z):
            v = 0.0
            if x != -z:
                if quad == 1:
                    v = at(abs(x), y if y >= 0 else -y - 1)
                elif x > 0:
                    v = at(x, y)
                elif x == 0:
                    v = at(0, -y) if y > 0 else at(0, y)
                elif y != -z:
                    v = at(-x, -y)
            out[(x + z) + 2 * z * (y + z)] = v
    return out


@pytest.mark.parametrize("quad", [1, 2])
def test_expand_fill_matches_reference(quad):
    """The reference's ExpandImage output equals the gather applied to its own computed
    block, bit for bit, including the pixels its copy loops never write (0 in a fresh
    engine)."""
    z = CASES["q%d_block" % quad]["GridPoints"]
    full = expand_gather(FIX["q%d_block" % quad], z, quad)
    np.testing.assert_array_equal(full, FIX["q%d_expand" % quad])


def test_expand_is_not_plain_symmetry():
    """The half's column x = 0 keeps its own lower half but takes the mirrored lower half
    above the axis: the in-place order of the reference's loop."""
    z = CASES["q2_block"]["GridPoints"]
    b = FIX["q2_block"]
    col = lambda img, w, y0: np.array([img[0 + w * (y - y0)] for y in range(-z, z)])
    before = col(b, z, -z)
    after = np.array([FIX["q2_expand"][(0 + z) + 2 * z * (y + z)] for y in range(-z, z)])
    np.testing.assert_array_equal(after[: z + 1], before[: z + 1])
    np.testing.assert_array_equal(after[z + 1:], before[z - 1:0:-1])


@pytest.mark.parametrize("name", ["d1_sphere", "d1_fcc"])
def test_1d_grid_through_the_oracle(oracle_lib, name):
    p = CASES[name]
    o = oracle_lib.OracleEngine({k: v for k, v in p.items() if k not in ("use1d", "EditQsteps", "EditQmin")})
    o.prepare()
    n = p["EditQsteps"]
    qx = (p["EditQmax"] - p["EditQmin"]) / n * np.arange(n)
    got = np.array([o.pixel(q, 1e-20, 1e-20) for q in qx])
    assert rel_err(got, FIX[name], floor=1e-300).max() < 1e-14
