"""GPU parity of the isotropic cube (key 0x04000C07, Source/sc_gpu_pCub_lNon_oIso.h) through the
C ABI: against the reference's fixtures, against the CPU oracle at larger sizes and in detector
geometry, and through the fit objective.  The cube's P(q) is FP64 series/Porod arithmetic in
both precisions, so the bar is FP64 rounding everywhere."""
import numpy as np
import pytest

from helpers import fixture, params, rel_err

from crystalscatter_amd import SasCalc

pytestmark = pytest.mark.gpu

CUBE = dict(LType=12, ComboBoxParticle=4, Ordis=7, ComboBoxInterior=0, EditRadius=10, EditSigma=0.1,
            EditQmax=2, I0=1, Base=0, iso=0)


def _calc(p, prec):
    e = SasCalc(device=0, params=p)
    e.set("ignore_cbs", 1)
    return e, e.doCalculation(prec)


def _report(tag, r):
    print("%s: max rel %.3e  rms rel %.3e" % (tag, np.nanmax(r), np.sqrt(np.nanmean(r * r))), flush=True)


@pytest.mark.parametrize("prec", ["ref64", "mixed"])
@pytest.mark.parametrize("name", ["cube_a.npz", "cube_b.npz"])
def test_cube_fixtures(name, prec):
    f = fixture(name)
    e, img = _calc(params(name), prec)
    assert tuple(e.bounds()) == tuple(int(b) for b in f["bounds"])
    r = rel_err(img, f["image"])
    _report("%s %s" % (name, prec), r)
    assert r.max() < 1e-13


@pytest.mark.parametrize("prec", ["ref64", "mixed"])
def test_cube_512_against_oracle(oracle_lib, prec):
    """512x512 (radial table path): every pixel against the oracle, and I(-q) = I(q)."""
    p = dict(CUBE, GridPoints=256)
    e, img = _calc(p, prec)
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    ref = o.image(nthreads=16)
    r = rel_err(img, ref)
    _report("cube 512^2 " + prec, r)
    assert r.max() < 1e-13
    a = img.reshape(512, 512)[1:, 1:]
    assert np.array_equal(a, a[::-1, ::-1])


def test_cube_detector_geometry_against_oracle(oracle_lib):
    """Beam-centre geometry (q from pixel size, distance and wavelength; no radial table)."""
    p = dict(CUBE, EditRadius=6, EditSigma=0.15, GridPoints=96, CenterBeam=1, EditPixelX=2.5,
             EditPixelY=2.5, EditPixelNoX=192, EditPixelNoY=192, EditDet=10, EditWavelength=0.154,
             iso=0.1, ifluc=0.5, rfluc=3)
    e, img = _calc(p, "ref64")
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    r = rel_err(img, o.image(nthreads=16))
    _report("cube detector geometry", r)
    # q goes through atan2/sin/cos here (device libm within an ulp or two of glibc), and near
    # its convergence radius the alternating series sums terms 1e8 times larger than P(q): in
    # the oracle itself a 1-ulp change of qx moves such pixels by up to 1.3e-6
    assert np.median(r) == 0 and r.max() < 1e-5


@pytest.mark.parametrize("prec", ["ref64", "mixed"])
def test_cube_fit_objective_against_oracle(oracle_lib, prec):
    p = dict(CUBE, EditRadius=4, EditSigma=0.25, EditQmax=3, GridPoints=16, BeamPosX=16, BeamPosY=16)
    data = np.random.default_rng(11).uniform(0.01, 10.0, 32 * 32)
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    e = SasCalc(device=0, params=p)
    e.set("ignore_cbs", 1)
    e.setFitData(32, 32, data)
    for bstop, border, rect in [(0, 0, None), (4, 0, None), (0, 2, (3, 4, 9, 12))]:
        for i in range(4):
            e.setNoFitRect(i, -1, -1, -1, -1)
        if rect:
            e.setNoFitRect(0, *rect)
        s, c, n = e.doFitCalculation(bstop, border, prec)
        so, co, no, _ = o.fit_fqs(data, 32, 32, bstop, border, [rect] if rect else None)
        assert abs(s / so - 1) < 1e-12 and c == co and n == no, (bstop, border, rect, s, so)


def test_cube_other_interior_is_flat():
    """Interiors other than homogeneous: P(q) = 1, the image is base + I0 (1 + iso) + ifluc term."""
    p = dict(CUBE, ComboBoxInterior=1, GridPoints=16, I0=2, iso=0.5, Base=0.25)
    e, img = _calc(p, "mixed")
    assert np.all(img == 0.25 + 2 * 1.5)
