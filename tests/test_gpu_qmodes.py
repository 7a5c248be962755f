"""The 1D q mode and the ExpandImage layouts on the GPU (csx_engine_calc), against the
reference's own images (tests/golden/qmodes.npz): the expanded image array, bit for bit
where the reference mode's pixels are, and the 1D row."""
import numpy as np
import pytest

from helpers import fixture, manifest, rel_err

from crystalscatter_amd import SasCalc

pytestmark = pytest.mark.gpu

CASES = manifest()["fixtures"]["qmodes.npz"]["cases"]
FIX = fixture("qmodes.npz")


@pytest.mark.parametrize("prec,tol", [("ref64", 1e-13), ("mixed", 1e-4)])
@pytest.mark.parametrize("name", sorted(CASES))
def test_qmode_images(name, prec, tol):
    e = SasCalc(device=0, params=CASES[name])
    e.set("ignore_cbs", 1)
    img = e.doCalculation(prec)
    ref = FIX[name]
    assert img.shape == ref.shape
    assert tuple(e.bounds()) == tuple(int(v) for v in FIX[name + "_bounds"])
    r = rel_err(img, ref)
    print("%s %s: max rel %.3e" % (name, prec, r.max()), flush=True)
    assert r.max() < tol
    assert np.array_equal(img == 0, ref == 0)  # pixels the fill never writes


def test_expand_equals_block_mirror():
    """The expanded image's computed quadrant equals the block render exactly, and its other
    quadrants are copies of it (same precision, same kernel)."""
    blk = SasCalc(device=0, params=CASES["q1_block"])
    b = blk.doCalculation("mixed")
    full = SasCalc(device=0, params=CASES["q1_expand"]).doCalculation("mixed")
    z = CASES["q1_block"]["GridPoints"]
    F = full.reshape(2 * z, 2 * z)  # [y + z, x + z]
    B = b.reshape(z, z)             # [y, x]
    np.testing.assert_array_equal(F[z:, z:], B)
    np.testing.assert_array_equal(F[z:, 1:z][:, ::-1], B[:, 1:])
    np.testing.assert_array_equal(F[:z, z:][::-1, :], B)
