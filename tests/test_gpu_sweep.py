"""Sweep over every combination key the device path serves (the reference's all-combinations
regression, Source/sc_maingui.cpp:9049, restricted to the served keys): each renders a small
image in both precisions and equals the oracle's FP64 image, pixel by pixel.  A key that the
library does not serve must be refused at preparation, never rendered by another formula."""
import numpy as np
import pytest

from helpers import CFG1, CFG2, CFG3, rel_err

from crystalscatter_amd import CsxError, SasCalc

pytestmark = pytest.mark.gpu

G = 12  # 24x24 images

NO_LATTICE = [
    ("sphere iso", dict(CFG1)),
    ("sphere gauss", dict(CFG1, Ordis=0)),
    ("sphere zdir", dict(CFG1, Ordis=6)),
    ("cube iso", dict(CFG1, ComboBoxParticle=4, EditRadius=10, EditQmax=2)),
    ("cube zdir", dict(CFG1, ComboBoxParticle=4, Ordis=6, EditRadius=10, EditQmax=2)),
    ("cylinder iso", dict(CFG2, Ordis=7, EditQmax=2)),
    ("cylinder zdir", dict(CFG2, Ordis=6, theta=0, phi=0)),
    ("cylinder gauss", dict(CFG2)),
]
LATTICE = [("sphere lt%d ord%d shp%d" % (lt, od, pk + 1),
            dict(CFG3, LType=lt, Ordis=od, ComboBoxPeak=pk, uca=28, ucb=28, ucc=28 if lt != 8 else 40,
                 EditPeakPar={4: 30, 5: 1.5}.get(pk, 0), EditAzi=7))
           for lt in (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 17, 18) for od in (0, 7) for pk in range(6)]
# spheres under the other orientation distributions (which enter neither the sphere's factors
# nor the peak sums)
LATTICE += [("sphere lt5 ord%d shp2" % od, dict(CFG3, Ordis=od, uca=28, ucb=28, ucc=28, EditAzi=7))
            for od in (1, 2, 3, 4, 5, 6, 8, 9, 10, 11, 12, 13)]
BCT_CYL = [("cylinder bct zdir", dict(LType=8, ComboBoxParticle=1, Ordis=6, ComboBoxPeak=1, theta=0, phi=0,
                                       EditSigma=0.06, SigmaL=0.06, RadButDebyeScherrer=1, EditQmax=1, I0=1,
                                       Base=0, HKLmax=5, EditRadius=4, Length=30, uca=25, ucb=25, ucc=40,
                                       EditDomainSize=250, EditAzi=5))]
BCT_CYL.append(("cylinder bct zdir hkl0", dict(BCT_CYL[0][1], HKLmax=0)))
# perfectly oriented cylinders on the other lattices (calc_GENERIC's cylinder z-dir branch)
BCT_CYL += [("cylinder lt%d zdir" % lt, dict(BCT_CYL[0][1], LType=lt, ucc=25, ucgamma=120 if lt == 1 else 90))
            for lt in (1, 2, 3, 4, 5, 7, 9, 10, 11, 17)]
# isotropic cylinders (L/R >= 2) on lattices
BCT_CYL += [("cylinder lt%d iso" % lt, dict(BCT_CYL[0][1], LType=lt, Ordis=7, ucc=25, ucgamma=120 if lt == 1 else 90))
            for lt in (1, 3, 4, 5, 7, 8, 17)]
# isotropic homogeneous cubes on lattices
BCT_CYL += [("cube lt%d iso" % lt, dict(CFG3, LType=lt, ComboBoxParticle=4, Ordis=7, EditRadius=10, EditSigma=0.1,
                                        uca=28, ucb=28, ucc=28, EditQmax=1.5, ucgamma=120 if lt == 1 else 90))
            for lt in (1, 4, 5, 7, 9, 17)]
# the other peak shapes with cylinders (both orientations) and cubes
BCT_CYL += [("%s shp%d" % (name, pk + 1), dict(p, ComboBoxPeak=pk, EditPeakPar={4: 30, 5: 1.5}.get(pk, 0)))
            for name, p in (BCT_CYL[0], BCT_CYL[-1], [c for c in BCT_CYL if c[0] == "cylinder lt5 iso"][0])
            for pk in (0, 2, 3, 4, 5)]
CASES = NO_LATTICE + LATTICE + BCT_CYL


@pytest.mark.parametrize("name,p", CASES, ids=[c[0].replace(" ", "-") for c in CASES])
def test_served_key_matches_oracle(oracle_lib, name, p):
    p = dict(p, GridPoints=G)
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    ref = o.image(nthreads=8)
    assert np.isfinite(ref).all() and np.abs(ref).max() > 0
    for prec, tol in (("ref64", 1e-11), ("mixed", 1e-4)):
        e = SasCalc(device=0, params=p)
        e.set("ignore_cbs", 1)  # the fit workload's combination fails checkAllParameters' particle rule
        img = e.doCalculation(prec)
        r = rel_err(img, ref, floor=1e-300)
        ok = np.abs(ref) > 1e-300
        assert r[ok].max() < tol, (name, prec, float(r[ok].max()))
        assert np.all(img[~ok] == ref[~ok])


@pytest.mark.parametrize("p", [dict(CFG3, ComboBoxPeak=6), dict(CFG3, ComboBoxPeak=7),
                               dict(CFG3, LType=0), dict(CFG3, ComboBoxParticle=2),
                               dict(CFG3, CheckBoxTwinned=1),
                               dict(CFG3, ComboBoxParticle=4, Ordis=7, ComboBoxInterior=1)],
                         ids=["gamma-peak", "anisotropic-gaussian-peak", "lamellar-lattice", "disk", "twinned",
                              "core-shell-cube-lattice"])
def test_unserved_combinations_are_refused(p):
    e = SasCalc(device=0, params=dict(p, GridPoints=G))
    with pytest.raises(CsxError):
        e.doCalculation("mixed")
