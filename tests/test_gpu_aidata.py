"""Training-data generation on the GPU: many parameter sets rendered in one launch
(csx_render_batch_run) against one render per set, and random variation (tpvPerformRandom)
feeding the batch against the reference's own run (tests/golden/tpv_fcc.npz)."""
import numpy as np
import pytest

from helpers import CFG3, fixture, manifest, rel_err

from crystalscatter_amd import Context, Random, SasCalc, render_batch, tpv_images

pytestmark = pytest.mark.gpu

CUBE = dict(LType=12, ComboBoxParticle=4, Ordis=7, ComboBoxInterior=0, EditRadius=10, EditSigma=0.1,
            EditQmax=2, I0=1, Base=0, iso=0, GridPoints=40)


def _variants():
    fcc = [dict(CFG3, GridPoints=40, EditRadius=r, uca=a, ucb=a, ucc=a) for r, a in ((8, 30), (6, 24), (9, 33), (7, 28))]
    bct = [dict(p, GridPoints=24) for p in manifest()["fixtures"]["cfg4_bct.npz"]["params"][:3]]
    cube = [dict(CUBE, EditRadius=r) for r in (5, 8, 12)]
    return [("fcc", fcc), ("bct", bct), ("cube", cube)]


@pytest.mark.parametrize("prec", ["ref64", "mixed"])
def test_batch_equals_one_render_per_set(prec):
    ctx = Context(0)
    for name, sets in _variants():
        engines = []
        for p in sets:
            e = SasCalc(device=-1, params=p)
            e.set("ignore_cbs", 1)
            engines.append(e)
        out = render_batch(ctx, engines, prec).cpu().numpy()
        for k, p in enumerate(sets):
            one = SasCalc(device=0, params=p)
            one.set("ignore_cbs", 1)
            img = one.doCalculation(prec)
            r = rel_err(out[k], img)
            print("%s %s set %d: max rel %.3e" % (name, prec, k, r.max()), flush=True)
            # same pixel code; the single render takes |q|-only terms from its radial table,
            # which holds the same values
            assert r.max() <= (1e-15 if prec == "ref64" else 1e-6), (name, k)


@pytest.mark.parametrize("prec,tol", [("ref64", 1e-12), ("mixed", 1e-4)])
def test_tpv_batch_matches_the_reference_run(prec, tol):
    f = fixture("tpv_fcc.npz")
    m = manifest()["fixtures"]["tpv_fcc.npz"]
    e = SasCalc(device=-1, params=m["params"])
    e.set("ignore_cbs", 1)
    imgs, reports = tpv_images(Context(0), e, m["tpv"], m["n"], Random(m["seed"]), prec)
    imgs = imgs.cpu().numpy()
    assert "#".join(reports) + "#" == bytes(f["report"]).decode()
    for k in range(m["n"]):
        r = rel_err(imgs[k], f["images"][k])
        print("tpv %s image %d: max rel %.3e" % (prec, k, np.nanmax(r)), flush=True)
        assert np.nanmax(r) < tol
