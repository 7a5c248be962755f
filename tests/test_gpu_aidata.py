"""Training-data generation on the GPU: many parameter sets rendered in one launch
(csx_render_batch_run) against one render per set, and random variation (tpvPerformRandom)
feeding the batch against the reference's own run (tests/golden/tpv_fcc.npz)."""
import numpy as np
import pytest

from helpers import CFG3, fixture, manifest, params, rel_err

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


def _post_case(name):
    f = fixture("post.npz")
    c = manifest()["fixtures"]["post.npz"]["cases"][name]
    return f, c["post"], f[name + "_image"], f[name + "_bounds"]


@pytest.mark.parametrize("name", ["fcc_rphi", "fcc_rphi_off", "fcc_rphi_clip"])
def test_rphi_matches_the_reference(oracle_lib, name):
    import torch
    from crystalscatter_amd import aidata
    f, post, img, b = _post_case(name)
    ctx = Context(0)
    imgs = torch.from_numpy(np.stack([img, img * 2.0])).cuda()
    r = aidata.post_rphi(ctx, imgs, b, (post["bsx"], post["bsy"]), post["rphi"], post.get("rphiscale", 0),
                         post.get("rphiclip", 0), post.get("rphiclip40", 0))
    got = r[0].cpu().numpy()
    err = np.abs(got - f[name + "_rphi"]).max() / np.abs(f[name + "_rphi"]).max()
    print("%s rphi: max err %.3e (of the maximum)" % (name, err), flush=True)
    # device cos/sin within an ulp of glibc move the bilinear sample points by ~1e-14 pixel
    assert err < 1e-12
    aidata.post_scale_clip(ctx, r, genlog=True)
    lg = r[0].cpu().numpy()
    assert np.abs(lg - f[name + "_rphi_log"]).max() < 1e-12
    # the second image of the batch against the oracle (the log scaling's start value 0 for
    # the maximum makes it depend on the overall scale)
    o = oracle_lib.rphi(img * 2.0, b, post["bsx"], post["bsy"], post["rphi"], post.get("rphiscale", 0),
                        post.get("rphiclip", 0), post.get("rphiclip40", 0))
    assert np.abs(r[1].cpu().numpy() - oracle_lib.scale_clip(o, genlog=True)).max() < 1e-12


def test_rphi_many_centres_against_the_oracle(oracle_lib):
    """One launch with a centre per image (cell maps shared by rounded maximum radius) gives
    each image's r,phi as the oracle does for that centre alone."""
    import ctypes
    import torch
    from crystalscatter_amd import _lib
    f, post, img, b = _post_case("fcc_rphi")
    centres = [(0, 0), (3, -4), (-7, 5), (3, -4), (12, 9), (-1, 0)]
    n, s = len(centres), 32
    imgs = torch.from_numpy(np.stack([img * (1.0 + 0.5 * k) for k in range(n)])).cuda()
    out = torch.empty((n, s * s), dtype=torch.float64, device="cuda")
    cen = np.array(centres, dtype=np.int32)
    ctx = Context(0)
    torch.cuda.synchronize()
    x0, x1, y0, y1 = (int(v) for v in b)
    ctx._chk(_lib.lib().csx_post_rphi_many(ctx._h, ctypes.c_void_p(imgs.data_ptr()), x0, x1, y0, y1,
                                           cen.ctypes.data_as(ctypes.c_void_p), s, n,
                                           ctypes.c_void_p(out.data_ptr()), 0, 0, 0), "csx_post_rphi_many")
    ctx.synchronize()
    got = out.cpu().numpy()
    for k, (xc, yc) in enumerate(centres):
        ref = oracle_lib.rphi(img * (1.0 + 0.5 * k), b, xc, yc, s, 0, 0, 0)
        assert np.abs(got[k] - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), k


def test_rphi_many_distinct_radii_in_chunks():
    """Many centres with distinct maximum radii (one cell map each): at s = 2048 the maps
    (16 MiB each) no longer fit one launch's budget, so the images go in several launches;
    every image equals its own one-centre call."""
    import ctypes
    import torch
    from crystalscatter_amd import _lib
    f, post, img, b = _post_case("fcc_rphi")
    x0, x1, y0, y1 = (int(v) for v in b)
    n, s = 20, 2048
    centres = np.array([(2 * k - 19, (k * 7) % 13 - 6) for k in range(n)], dtype=np.int32)
    imgs = torch.from_numpy(np.stack([img * (1.0 + 0.25 * k) for k in range(n)])).cuda()
    out = torch.empty((n, s * s), dtype=torch.float64, device="cuda")
    one = torch.empty((1, s * s), dtype=torch.float64, device="cuda")
    ctx = Context(0)
    L = _lib.lib()
    torch.cuda.synchronize()
    ctx._chk(L.csx_post_rphi_many(ctx._h, ctypes.c_void_p(imgs.data_ptr()), x0, x1, y0, y1,
                                  centres.ctypes.data_as(ctypes.c_void_p), s, n, ctypes.c_void_p(out.data_ptr()),
                                  0, 0, 0), "csx_post_rphi_many")
    ctx.synchronize()
    for k in range(n):
        ctx._chk(L.csx_post_rphi(ctx._h, ctypes.c_void_p(imgs[k].data_ptr()), x0, x1, y0, y1, int(centres[k, 0]),
                                 int(centres[k, 1]), s, 1, ctypes.c_void_p(one.data_ptr()), 0, 0, 0), "csx_post_rphi")
        ctx.synchronize()
        assert torch.equal(out[k], one[0]), k


def test_tpv_batch_failure_leaves_the_generator_unchanged():
    """A batch whose random variation produces an invalid set fails before any launch and puts
    the generator (and the variation state) back, so a retry draws the same sequence."""
    from crystalscatter_amd import CsxError
    from crystalscatter_amd import aidata
    p = dict(CFG3, GridPoints=16)
    e = SasCalc(device=-1, params=p)
    e.set("ignore_cbs", 1)
    a, b = Random(4242), Random(4242)
    a.zuf(5)
    b.zuf(5)
    # uca varied by up to +-5000 %: most draws are below the 0.01 the parameter check requires
    with pytest.raises(CsxError):
        aidata.generate(Context(0), e, ["uca=50"], 64, a, aidata.AIOptions(noise=False), "mixed")
    assert [a.next() for _ in range(40)] == [b.next() for _ in range(40)]
    with pytest.raises(ValueError):
        aidata.generate(Context(0), e, ["uca=0.05"], 70000, a, aidata.AIOptions(noise=False), "mixed")


@pytest.mark.parametrize("name", ["fcc_rphi", "fcc_rphi_off", "fcc_rphi_clip"])
def test_fft_matches_the_reference(name):
    """The FFT image of the reference's r,phi image (calculateIFFT without FFTW)."""
    import torch
    from crystalscatter_amd import aidata
    f, post, img, b = _post_case(name)
    x = torch.from_numpy(np.stack([f[name + "_rphi_log"]] * 3)).cuda()
    out = aidata.post_ifft(Context(0), x, post["rphi"], post.get("fftforward", 0), post["fft"],
                           post.get("fftscale", 0), post.get("fftclip", 0), post.get("fftclip40", 0),
                           post.get("fftswap", 0), (post.get("fftcutx", 0), post.get("fftcuty", 0))).cpu().numpy()
    ref = f[name + "_fft"]
    for row in out:
        err = np.abs(row - ref).max() / max(np.abs(ref).max(), 1e-300)
        print("%s fft: max err %.3e (of the maximum)" % (name, err), flush=True)
        # same butterflies and twiddles; only sqrt of the modulus and the log clip go
        # through the device libm
        assert err < 1e-13


@pytest.mark.parametrize("name,lin", [("fcc_rphi", 0), ("sphere_lin", 1)])
def test_scale_intensity_matches_the_reference(name, lin):
    import torch
    from crystalscatter_amd import aidata
    f, post, img, b = _post_case(name)
    t = torch.from_numpy(np.stack([img, img])).cuda()
    aidata.post_scale_intensity(Context(0), t, lin)
    for row in t.cpu().numpy():
        assert np.abs(row - f[name + "_scaled"]).max() < 1e-14


@pytest.mark.parametrize("opts_kw", [dict(noise=True, bs=(0.2, 0.2, 0.5, 0.5), lines_x=2, lines_y=1, rphi_size=128),
                                     dict(scale_scat=True, noise=True, rphi_size=128, fft_input_lin=True,
                                          rphi_scale=True, fft_size=128, fft_type=2, fft_swap=True,
                                          fft_scale=True, fft_cut=(64, 32))])
def test_ai_pipeline_matches_the_oracle(oracle_lib, opts_kw):
    """The batched pipeline against the per-image CPU restatement, from the same rand() state."""
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import aidata_oracle as AO
    from crystalscatter_amd import aidata
    p = dict(CFG3, GridPoints=32, BeamPosX=3, BeamPosY=-4)
    ids = ["EditRadius=0.1", "uca=0.05", "ucb=1", "ucc=1", "I0=0.3", "EditDomainSize=0.4"]
    opts = aidata.AIOptions(**opts_kw)
    e = SasCalc(device=-1, params=p)
    e.set("ignore_cbs", 1)
    out = aidata.generate(Context(0), e, ids, 5, Random(11), opts, "ref64")
    scat = out["scat"].cpu().numpy()
    rphi = out["rphi"].cpu().numpy()
    # CPU: the same sequence image by image
    e2 = SasCalc(device=-1, params=p)
    e2.set("ignore_cbs", 1)
    rng = Random(11)
    for k in range(5):
        e2.tpvPerformRandom(ids, rng)
        q = dict(p)
        q.update({key: e2.tpv_value(key) for key in ["EditRadius", "uca", "ucb", "ucc", "I0", "width_zuf"]})
        o = oracle_lib.OracleEngine(q)
        o.prepare()
        res = AO.post_image(o.image(nthreads=8), o.bounds(), (3, -4), rng, opts)
        img, centre, rp = res[:3]
        if opts.fft_size:
            ft = out["fft"][k].cpu().numpy()
            assert np.abs(ft - res[3]).max() < 1e-10 * max(1.0, np.abs(res[3]).max()), k
        assert tuple(out["centres"][k]) == centre
        ok = np.isfinite(img)
        assert np.array_equal(np.isfinite(scat[k]), ok) and np.array_equal(scat[k] == 0, img == 0)
        assert rel_err(scat[k][ok], img[ok]).max() < 1e-12, k
        assert np.abs(rphi[k] - rp).max() < 1e-10, k


@pytest.mark.parametrize("n", [0, 1, 30, 31, 495, 496, 497, 31743, 31744, 31745, 100003])
def test_device_zuf_equals_the_host_sequence(n):
    """rand() continued on the device (jumped start states, 31 terms in registers) gives the
    host's values bit for bit and leaves the generator where n draws leave it."""
    import torch
    ctx = Context(0)
    a, b = Random(31337), Random(31337)
    a.zuf(17)
    b.zuf(17)
    out = torch.full((max(n, 1),), 7.0, dtype=torch.float64, device="cuda")[:n]
    a.zuf_device(ctx, out)
    ref = b.zuf(n)
    assert np.array_equal(out.cpu().numpy(), ref)
    assert [a.next() for _ in range(50)] == [b.next() for _ in range(50)]


def test_batch_rejects_sets_that_need_carr11pm():
    """Gaussian-oriented cylinders and oriented cubes read the carr11pm table, which a batch
    does not carry: staging them is an error on the host (nothing is launched)."""
    import torch
    from crystalscatter_amd import CsxError
    ctx = Context(0)
    out = torch.empty(64 * 64, dtype=torch.float64, device="cuda")
    for p in (dict(params("cfg2_cylgauss.npz"), GridPoints=32), dict(CUBE, Ordis=6, GridPoints=32)):
        e = SasCalc(device=-1, params=p)
        e.set("ignore_cbs", 1)
        e.prepare()
        with pytest.raises(CsxError, match="carr11pm"):
            e.stage_image(ctx, out.data_ptr(), out.numel())
