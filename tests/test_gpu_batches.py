"""Batched evaluation on the GPU over uneven detectors, and the reference's own fits re-run by
the batched simplex.

- Fit objective: BCT parameter sets on 32x32, 48x48, 64x64 and 80x80 detectors staged into
  ONE csx_fit_batch_run; each sum equals the set's own csx_fit_fqs (doFitCalculation), and
  the 32x32 set equals the reference's own objective (tests/golden/cfg4_bct.npz).
- Images: the same four sizes staged into ONE csx_render_batch_run; each image equals the
  set's own render.
- Fits: the four fits of tests/golden/simplex_fits.npz (the reference's doSimplexFit2D runs)
  through csx_nm_batch in reference mode on the FP64 objective: same iterations, same number
  of evaluations, same final vertex.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from helpers import GOLDEN, fixture, manifest, rel_err

from crystalscatter_amd import Context, SasCalc, _lib, nm_batch

pytestmark = pytest.mark.gpu

SIZES = [16, 24, 32, 40]  # GridPoints: 32x32 .. 80x80


def _tables(h):
    carr = np.concatenate([h.table(n) for n in ("carr1p", "carr2p", "carr3p", "carr4p", "carr5p", "carr6p", "carr7p",
                                                "carr8p", "carr9p", "carr1f", "carr2f", "carr3f", "carr4f", "carr5f",
                                                "carr6f", "carr7f", "carr8f", "carr9f")])
    return carr, h.table("latpar1"), h.table("latpar2"), h.table("latpar3")


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


@pytest.mark.parametrize("prec", ["ref64", "mixed"])
def test_uneven_detectors_in_one_fit_batch(prec):
    import torch
    m = manifest()["fixtures"]["cfg4_bct.npz"]
    f = fixture("cfg4_bct.npz")
    rng = np.random.default_rng(11)
    ctx = Context(0)
    L = _lib.lib()
    mask = _lib.Mask()
    for i in range(16):
        mask.rects[i] = -1
    mask.bstop, mask.border = 4, 0
    seq, keep, datas = [], [], []
    assert L.csx_fit_batch_begin(ctx._h) == 0
    for k, n in enumerate(SIZES):
        # set 0 (32x32) is the reference's own fit problem (draw 0, beam stop at 16, 16)
        p = dict(m["fit_params"]) if k == 0 else dict(m["params"][k], GridPoints=n)
        w = 2 * n
        data = f["fitdata"] if n == 16 else rng.uniform(0.01, 10.0, w * w)
        e = SasCalc(device=0, params=p)
        e.setFitData(w, w, data)
        seq.append(e.doFitCalculation(4, 0, prec))
        h = SasCalc(device=-1, params=p)
        h.set("ignore_cbs", 1)
        h.prepare()
        carr, l1, l2, l3 = _tables(h)
        d = torch.tensor(data, dtype=torch.float64, device="cuda")
        keep += [carr, l1, l2, l3, h]
        datas.append(d)
        rc = L.csx_fit_batch_stage(ctx._h, ctypes.byref(h.scalars()), _ptr(carr, ctypes.c_double),
                                   _ptr(l1, ctypes.c_int32), _ptr(l2, ctypes.c_int32), _ptr(l3, ctypes.c_float),
                                   w, w, ctypes.byref(mask), ctypes.c_void_p(d.data_ptr()))
        assert rc == 0
    sums = np.zeros(len(SIZES))
    cnts = np.zeros(len(SIZES), dtype=np.int64)
    assert L.csx_fit_batch_run(ctx._h, 1 if prec == "mixed" else 0, _ptr(sums, ctypes.c_double),
                               _ptr(cnts, ctypes.c_int64)) == 0
    np.testing.assert_array_equal(sums, np.array([s for s, _, _ in seq]))
    np.testing.assert_array_equal(cnts, np.array([c for _, c, _ in seq]))
    ref = [x for x in m["fits"] if x["bstop"] == 4 and x["border"] == 0 and x["rect"] is None][0]
    assert m["fit_params"]["GridPoints"] == 16
    assert cnts[0] == ref["cnt"] and abs(sums[0] / ref["fqs"] - 1) < (1e-12 if prec == "ref64" else 1e-6)


@pytest.mark.parametrize("prec", ["ref64", "mixed"])
def test_uneven_detectors_in_one_render_batch(prec):
    import torch
    m = manifest()["fixtures"]["cfg4_bct.npz"]
    ctx = Context(0)
    L = _lib.lib()
    assert L.csx_fit_batch_begin(ctx._h) == 0
    outs, engines = [], []
    for k, n in enumerate(SIZES):
        h = SasCalc(device=-1, params=dict(m["params"][k], GridPoints=n))
        h.set("ignore_cbs", 1)
        out = torch.empty((2 * n) ** 2, dtype=torch.float64, device="cuda")
        h.stage_image(ctx, out.data_ptr(), out.numel())
        outs.append(out)
        engines.append(h)
    assert L.csx_render_batch_run(ctx._h, 1 if prec == "mixed" else 0) == 0
    for k, n in enumerate(SIZES):
        one = SasCalc(device=0, params=dict(m["params"][k], GridPoints=n))
        one.set("ignore_cbs", 1)
        img = one.doCalculation(prec)
        r = rel_err(outs[k].cpu().numpy(), img)
        print("bct %dx%d %s batch vs single: max rel %.3e" % (2 * n, 2 * n, prec, r.max()), flush=True)
        assert r.max() <= (1e-15 if prec == "ref64" else 1e-6)
    # 32x32 is fixture draw 0: the reference's own image
    if prec == "ref64":
        assert rel_err(outs[0].cpu().numpy(), fixture("cfg4_bct.npz")["image_0"]).max() < 1e-11


def test_reference_fits_on_the_gpu():
    """The reference's own simplex runs (oracle/ref_fit.cpp) re-run by csx_nm_batch, reference
    mode, FP64 objective on the GPU (within ~1e-11 of the reference's, so every comparison of
    the simplex falls the same way): the same iterations, evaluations and final vertex."""
    import torch
    fx = np.load(os.path.join(GOLDEN, "simplex_fits.npz"))
    with open(os.path.join(GOLDEN, "manifest.json")) as fh:
        cases = json.load(fh)["fixtures"]["simplex_fits.npz"]["cases"]
    ctx = Context(0)
    for name, c in sorted(cases.items()):
        keys = c["used"] + c["fixed"]
        base = SasCalc(device=-1, params=c["start"])
        base.set("ignore_cbs", 1)
        d = torch.tensor(fx[name + "_target"], dtype=torch.float64, device="cuda")
        fit = dict(engine=base, d_data=d.data_ptr(), w=c["w"], h=c["w"], x0=fx[name + "_fitstart"],
                   lo=[c["bounds"][k][0] for k in keys], hi=[c["bounds"][k][1] for k in keys],
                   bstop=c["bstop"], border=c["border"])
        res, _ = nm_batch(ctx, [fit], keys, stp=c["stp"], maxit=c["maxit"], ftol=c["ftol"], precision="ref64",
                          nfree=len(c["used"]), mode="reference")
        x, fb, it, ev = res[0]
        print("%s: iters %d/%d evals %d/%d" % (name, it, c["repetitions"], ev, c["numImgCalc"]), flush=True)
        assert it == c["repetitions"] and ev == c["numImgCalc"], name
        np.testing.assert_array_equal(x, fx[name + "_fitres"])


def _stage_sizes(ctx, sizes, seed, keep):
    """Stage BCT sets of the given GridPoints (uneven detectors) with random data."""
    import torch
    L = _lib.lib()
    m = manifest()["fixtures"]["cfg4_bct.npz"]
    rng = np.random.default_rng(seed)
    mask = _lib.Mask()
    for i in range(16):
        mask.rects[i] = -1
    mask.bstop, mask.border = 3, 0
    assert L.csx_fit_batch_begin(ctx._h) == 0
    for k, n in enumerate(sizes):
        h = SasCalc(device=-1, params=dict(m["params"][k % 6], GridPoints=n))
        h.set("ignore_cbs", 1)
        h.prepare()
        carr, l1, l2, l3 = _tables(h)
        d = torch.tensor(rng.uniform(0.01, 10.0, 4 * n * n), dtype=torch.float64, device="cuda")
        keep += [carr, l1, l2, l3, h, d]
        assert L.csx_fit_batch_stage(ctx._h, ctypes.byref(h.scalars()), _ptr(carr, ctypes.c_double),
                                     _ptr(l1, ctypes.c_int32), _ptr(l2, ctypes.c_int32), _ptr(l3, ctypes.c_float),
                                     2 * n, 2 * n, ctypes.byref(mask), ctypes.c_void_p(d.data_ptr())) == 0


def _run(ctx, nsets, prec=1):
    L = _lib.lib()
    sums = np.zeros(nsets)
    cnts = np.zeros(nsets, dtype=np.int64)
    assert L.csx_fit_batch_run(ctx._h, prec, _ptr(sums, ctypes.c_double), _ptr(cnts, ctypes.c_int64)) == 0
    return sums, cnts


def test_batch_graph_replays_equal_eager_rounds():
    """Rounds of different shapes launched as graph replays (the graph built once, updated in
    place) against the same rounds launched operation by operation, with eager work and
    synchronisations between the replays: every replay equals its eager round bit for bit."""
    import torch
    L = _lib.lib()
    ctx = Context(0)
    rounds = [[16, 24], [32, 16, 40, 8], [24], [40, 40, 16], [16, 24]]
    n0 = L.csx_batch_graph_launches(ctx._h)
    for r, sizes in enumerate(rounds):
        for prec in (1, 0):
            keep = []
            assert L.csx_set_batch_graph(ctx._h, 0) == 0
            _stage_sizes(ctx, sizes, r, keep)
            eager = _run(ctx, len(sizes), prec)
            x = torch.randn(4096, device="cuda").sum()  # eager work between replays
            torch.cuda.synchronize()
            assert L.csx_set_batch_graph(ctx._h, 1) == 0
            _stage_sizes(ctx, sizes, r, keep)
            graph = _run(ctx, len(sizes), prec)
            np.testing.assert_array_equal(graph[0], eager[0])
            np.testing.assert_array_equal(graph[1], eager[1])
            del x
    assert L.csx_batch_graph_launches(ctx._h) - n0 == 2 * len(rounds)


@pytest.mark.parametrize("key_params", ["bct", "fcc"])
def test_tuning_variants_give_the_same_sums(key_params):
    """Every kernel variant of the batched objective (different register budgets) gives the
    same sums and counts bit for bit."""
    from crystalscatter_amd import autotune
    import torch
    L = _lib.lib()
    key = 0x01020806 if key_params == "bct" else 0x00020500
    rng = np.random.default_rng(3)
    params = autotune._sets(key, 64, 12, rng)
    data = torch.tensor(rng.uniform(0.01, 10.0, 64 * 64), dtype=torch.float64, device="cuda")
    ctx = Context(0)
    ref = None
    for v in range(_lib.TUNE_VARIANTS):
        assert L.csx_tune_force(ctx._h, v) == 0
        keep = autotune._stage(ctx, params, 64, data)
        got = _run(ctx, len(params))
        del keep
        if ref is None:
            ref = got
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])
    assert L.csx_tune_force(ctx._h, -1) == 0


BATCH_KINDS = {
    "sphere-iso": dict(LType=12, ComboBoxParticle=0, Ordis=7, EditRadius=5, EditSigma=0.1, EditQmax=2, I0=1, Base=0),
    "sphere-zdir": dict(LType=12, ComboBoxParticle=0, Ordis=6, EditRadius=5, EditSigma=0.1, EditQmax=2, I0=1, Base=0),
    "cube-iso": dict(LType=12, ComboBoxParticle=4, Ordis=7, EditRadius=10, EditSigma=0.1, EditQmax=2, I0=1, Base=0),
    "cylinder-iso": dict(LType=12, ComboBoxParticle=1, Ordis=7, EditRadius=3, Length=30, EditSigma=0.06, SigmaL=0.06,
                         EditQmax=1, I0=1, Base=0),
    "cylinder-zdir": dict(LType=12, ComboBoxParticle=1, Ordis=6, EditRadius=3, Length=30, EditSigma=0.06,
                          SigmaL=0.06, theta=30, phi=45, EditQmax=1, I0=1, Base=0),
    "cylinder-iso-fcc": dict(LType=5, ComboBoxParticle=1, Ordis=7, ComboBoxPeak=1, EditRadius=3, Length=30,
                             EditSigma=0.06, SigmaL=0.06, uca=30, ucb=30, ucc=30, EditDomainSize=250, EditAzi=5,
                             RadButDebyeScherrer=1, HKLmax=5, EditQmax=1, I0=1, Base=0),
    "cube-iso-sc": dict(LType=7, ComboBoxParticle=4, Ordis=7, ComboBoxPeak=1, EditRadius=10, EditSigma=0.1, uca=25,
                        ucb=25, ucc=25, EditDomainSize=200, EditAzi=5, RadButDebyeScherrer=1, HKLmax=5, EditQmax=1,
                        I0=1, Base=0),
}
# the same lattice kinds with the q grid from the detector geometry (no radial table: the
# particle's factors are evaluated per pixel)
_BEAM = dict(CenterBeam=1, EditPixelX=2.5, EditPixelY=2.5, EditPixelNoX=64, EditPixelNoY=64, EditDet=2.0,
             EditWavelength=0.154)
BATCH_KINDS["cylinder-iso-fcc-beam"] = dict(BATCH_KINDS["cylinder-iso-fcc"], **_BEAM)
BATCH_KINDS["cube-iso-sc-beam"] = dict(BATCH_KINDS["cube-iso-sc"], **_BEAM)
BATCH_KINDS["cylinder-iso-beam"] = dict(BATCH_KINDS["cylinder-iso"], **_BEAM)
BATCH_KINDS["sphere-fcc-beam"] = dict(LType=5, ComboBoxParticle=0, Ordis=0, ComboBoxPeak=1, EditRadius=8,
                                      EditSigma=0.08, uca=30, ucb=30, ucc=30, EditDomainSize=200,
                                      RadButDebyeScherrer=1, EditDebyeWaller=1.5, EditAzi=5, HKLmax=5, EditQmax=1,
                                      I0=1, Base=0, **_BEAM)
BATCH_KINDS["cylinder-zdir-bct-beam"] = dict(LType=8, ComboBoxParticle=1, Ordis=6, ComboBoxPeak=1, theta=0, phi=0,
                                             EditRadius=4, Length=30, EditSigma=0.06, SigmaL=0.06, uca=25, ucb=25,
                                             ucc=40, EditDomainSize=250, EditAzi=5, RadButDebyeScherrer=1, HKLmax=5,
                                             EditQmax=1, I0=1, Base=0, **_BEAM)


@pytest.mark.parametrize("prec", ["ref64", "mixed"])
@pytest.mark.parametrize("kind", sorted(BATCH_KINDS))
def test_batched_images_equal_single_renders(kind, prec):
    """Every kind's images staged into one csx_render_batch_run (three parameter sets, radius
    and cell varied) equal the sets' own renders -- the batch reads each set's constants from
    the sets array, the single render from its kernel arguments: bit for bit, except the mixed
    lattice sums, which the batch evaluates in one pass (within 1e-6, as for the BCT sets)."""
    import torch
    from crystalscatter_amd import render_batch
    base = dict(BATCH_KINDS[kind], GridPoints=16)
    sets = [dict(base, EditRadius=base["EditRadius"] * f, uca=30 * f, ucb=30 * f, ucc=30 * f) for f in (1.0, 1.1, 0.9)]
    engines, singles = [], []
    for p in sets:
        e = SasCalc(device=0, params=p)
        e.set("ignore_cbs", 1)
        singles.append(torch.tensor(e.doCalculation(prec).ravel()))
        b = SasCalc(device=-1, params=p)
        b.set("ignore_cbs", 1)
        engines.append(b)
    ctx = Context(0)
    out = torch.full((len(sets), singles[0].numel()), -7.0, dtype=torch.float64, device="cuda")
    render_batch(ctx, engines, precision=prec, out=out)
    torch.cuda.synchronize()
    got = out.cpu()
    ctx.close()
    for k, ref in enumerate(singles):
        assert torch.isfinite(ref).all() and (ref != 0).any(), (kind, k)
        if prec == "mixed" and base["LType"] != 12:
            assert rel_err(got[k].numpy(), ref.numpy()).max() <= 1e-6, (kind, k)
        else:
            assert torch.equal(got[k], ref), (kind, prec, k, int((got[k] != ref).sum()))
