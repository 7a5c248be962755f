"""GPU parity: the HIP kernels (through the C ABI) against the reference's golden
fixtures and the CPU oracle.  REF64 must agree to FP64 rounding (device libm differs from
glibc by <= 1 ulp); MIXED to 1e-4 relative (the benchmark tolerance)."""
import ctypes

import numpy as np
import pytest

from helpers import CFG1, CFG3, fixture, manifest, params, rel_err

from crystalscatter_amd import CsxError, SasCalc, _lib

pytestmark = pytest.mark.gpu


def _calc(p, prec, ignore_cbs=True):
    e = SasCalc(device=0, params=p)
    e.set("ignore_cbs", 1 if ignore_cbs else 0)
    img = e.doCalculation(prec)
    return e, img


def _report(tag, r):
    print("%s: max rel %.3e  rms rel %.3e" % (tag, np.nanmax(r), np.sqrt(np.nanmean(r * r))), flush=True)


@pytest.mark.parametrize("centerbeam", [0, 1])
def test_qgrid_bit_exact(oracle_lib, centerbeam):
    import torch
    p = manifest()["fixtures"]["vitess_sphere.npz"]["params"] if centerbeam else dict(CFG1)
    e = SasCalc(device=0, params=p)
    e.prepare()
    x0, x1, y0, y1 = e.bounds()
    n = (x1 - x0) * (y1 - y0)
    d = torch.empty(3 * n, dtype=torch.float64, device="cuda")
    g = _lib.Grid(x0, x1, y0, y1)
    L = _lib.lib()
    assert L.csx_qgrid(L.csx_engine_ctx(e._h), ctypes.byref(g), ctypes.c_void_p(d.data_ptr())) == 0
    q = d.cpu().numpy().reshape(-1, 3)
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    W = x1 - x0
    k = np.arange(n)
    ref = np.array([o.idx2q(j % W + x0, j // W + y0) for j in k])
    if centerbeam:
        # detector geometry goes through atan2/sin/cos (device libm within an ulp or two of
        # glibc) and 1-cos(theta) for qz: compare against |q| of the pixel
        qn = np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
        assert (np.abs(q - ref).max(axis=1) / qn).max() < 1e-14
    else:
        np.testing.assert_array_equal(q, ref)


@pytest.mark.parametrize("prec", ["ref64", "mixed"])
def test_cfg1_sphere_fp64(prec):
    f = fixture("cfg1_sphere.npz")
    e, img = _calc(params("cfg1_sphere.npz"), prec)
    assert tuple(e.bounds()) == tuple(int(b) for b in f["bounds"])
    r = rel_err(img, f["image"])
    _report("cfg1 " + prec, r)
    assert r.max() < 1e-13


def test_vitess_golden_geometry():
    f = fixture("vitess_sphere.npz")
    p = manifest()["fixtures"]["vitess_sphere.npz"]["params"]
    e, img = _calc(p, "ref64")
    r = rel_err(img, f["image"])
    _report("vitess golden", r)
    assert np.median(r) < 1e-14 and r.max() < 5e-3 and (r > 1e-12).sum() <= 11499


@pytest.mark.parametrize("prec", ["ref64", "mixed"])
def test_cfg2_reference_mode_exact(prec):
    f = fixture("cfg2_cylgauss.npz")
    e, img = _calc(params("cfg2_cylgauss.npz"), prec)
    np.testing.assert_array_equal(img, f["image"])


def test_cfg2_physics_mode_ref64_matches_oracle(oracle_lib):
    """Physics mode (fix_orcase=1): the GPU's FP64 nested Romberg against the oracle's."""
    from helpers import CFG2
    p = dict(CFG2, GridPoints=16, fix_orcase=1)
    e, img = _calc(p, "ref64")
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    r = rel_err(img, o.image(nthreads=16))
    _report("cfg2 physics ref64 vs oracle", r)
    assert r.max() < 1e-12


def test_cfg2_physics_mode_mixed_vs_tight_quadrature():
    """Physics mode, FP32 quadrature kernel at full size (512x512): against an independent
    nested adaptive quadrature of the same integral (scipy, epsrel 1e-10/1e-11) at 96 pixels
    across the detector (tests/golden/cfg2_tight.npz, tests/golden/make_cfg2_tight.py).  The
    reference's own Romberg (eps 1e-4) is off by up to a few percent at large q, so it is not
    the bar here."""
    from helpers import CFG2
    p = dict(CFG2, fix_orcase=1)
    e, img = _calc(p, "mixed")
    x0, x1, y0, y1 = e.bounds()
    assert (x1 - x0, y1 - y0) == (512, 512)
    t = fixture("cfg2_tight.npz")
    got = img[(t["ihex"] - x0) + (x1 - x0) * (t["i"] - y0)]
    r = rel_err(got, t["value"])
    _report("cfg2 physics mixed vs tight quadrature (96 px)", r)
    assert r.size == 96 and r.max() < 2e-6


def test_cfg2_physics_small_q_series_same_in_both_precisions():
    """Below 0.4*limq1 both precisions take the FP64 double series."""
    from helpers import CFG2
    p = dict(CFG2, GridPoints=16, fix_orcase=1)
    e, a = _calc(p, "mixed")
    _, b = _calc(p, "ref64")
    s = e.scalars()
    x0, x1, y0, y1 = e.bounds()
    ih, ii = np.meshgrid(np.arange(x0, x1), np.arange(y0, y1))
    q = np.hypot(ih, ii).ravel() * s.qmax / s.zmax
    small = q < 0.4 * s.limq[0] * 0.999
    assert small.sum() > 0
    np.testing.assert_array_equal(a[small], b[small])


@pytest.mark.parametrize("theta,phi", [(60, 20), (10, 80)])
def test_cylinders_reference_mode_other_orientations(oracle_lib, theta, phi):
    from helpers import CFG2
    p = dict(CFG2, GridPoints=16, theta=theta, phi=phi)
    e, img = _calc(p, "ref64")
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    r = rel_err(img, o.image(nthreads=16))
    assert r.max() < 1e-13


@pytest.mark.parametrize("prec,tol", [("ref64", 1e-13), ("mixed", 1e-4)])
def test_cfg3_fcc_small_grid(prec, tol):
    f = fixture("cfg3_fcc_g32.npz")
    e, img = _calc(params("cfg3_fcc_g32.npz"), prec)
    r = rel_err(img, f["image"])
    _report("cfg3 64x64 " + prec, r)
    assert r.max() < tol


@pytest.mark.parametrize("prec,tol", [("ref64", 1e-12), ("mixed", 1e-4)])
def test_cfg3_fcc_full_size(prec, tol):
    f = fixture("cfg3_fcc_1024_samples.npz")
    e, img = _calc(CFG3, prec)
    assert img.size == 1024 * 1024
    r = rel_err(img[f["idx"]], f["values"])
    _report("cfg3 1024^2 sampled " + prec, r)
    assert r.max() < tol
    # size-independent invariant: I(-q) = I(q).  FP64: Friedel partners are summed in the
    # same order up to rounding.  Mixed: mirrored pixels sit in different waves, whose pair
    # windows drop different terms below 2^-32 of a pair weight, so the two sides differ by
    # at most the dropped mass (about 1e-8 relative here), far below the 1e-4 bound.
    a = img.reshape(1024, 1024)[1:, 1:]
    assert rel_err(a, a[::-1, ::-1]).max() < (1e-13 if prec == "ref64" else 1e-6)


def test_cfg3_mixed_every_pixel_against_ref64():
    """All 1,048,576 pixels of the headline pattern: the mixed kernel against the GPU's FP64
    reference mode, which is itself pinned to the reference's own pixels (above, <= 1e-12)."""
    e, mixed = _calc(CFG3, "mixed")
    _, ref = _calc(CFG3, "ref64")
    assert mixed.size == ref.size == 1024 * 1024
    assert np.isfinite(ref).all()
    r = rel_err(mixed, ref)
    _report("cfg3 1024^2 every pixel mixed vs ref64", r)
    assert r.max() < 1e-4 and np.sqrt(np.mean(r * r)) < 1e-6


@pytest.mark.parametrize("n", [128, 256])
def test_bct_mixed_fit_sizes_against_ref64(n):
    """The BCT mixed kernel (precise pair loop, conditioning guard) at fit detector sizes
    256x256 and 512x512 against the GPU's FP64 reference mode, on the six fixture draws."""
    worst, rms = 0.0, 0.0
    for p in manifest()["fixtures"]["cfg4_bct.npz"]["params"]:
        q = dict(p, GridPoints=n)
        _, mixed = _calc(q, "mixed")
        _, ref = _calc(q, "ref64")
        assert mixed.size == (2 * n) ** 2
        ok = np.abs(ref) > 0
        r = rel_err(mixed[ok], ref[ok])
        worst = max(worst, float(r.max()))
        rms = max(rms, float(np.sqrt(np.mean(r * r))))
    print("bct %dx%d mixed vs ref64: max rel %.3e  worst rms %.3e" % (2 * n, 2 * n, worst, rms), flush=True)
    assert worst < 1e-4


def test_cfg3_full_size_against_oracle_random_pixels(oracle_lib):
    e, img = _calc(CFG3, "mixed")
    o = oracle_lib.OracleEngine(CFG3)
    o.prepare()
    rng = np.random.default_rng(7)
    idx = rng.choice(img.size, 6000, replace=False)
    ref = o.image_points(idx % 1024 - 512, idx // 1024 - 512, nthreads=16)
    r = rel_err(img[idx], ref)
    _report("cfg3 1024^2 vs oracle", r)
    assert r.max() < 1e-4


@pytest.mark.parametrize("prec,tol", [("ref64", 1e-11), ("mixed", 1e-4)])
def test_cfg4_bct_sets(prec, tol):
    f = fixture("cfg4_bct.npz")
    worst = 0.0
    for k, p in enumerate(manifest()["fixtures"]["cfg4_bct.npz"]["params"]):
        e, img = _calc(p, prec)
        r = rel_err(img, f["image_%d" % k])
        worst = max(worst, r.max())
    print("cfg4 %s worst rel %.3e" % (prec, worst))
    assert worst < tol


@pytest.mark.parametrize("prec,tol", [("ref64", 1e-12), ("mixed", 1e-6)])
def test_cfg4_fit_objective_and_masks(prec, tol):
    import torch
    f = fixture("cfg4_bct.npz")
    m = manifest()["fixtures"]["cfg4_bct.npz"]
    e = SasCalc(device=0, params=m["fit_params"])
    e.setFitData(32, 32, f["fitdata"])
    for fit in m["fits"]:
        for i in range(4):
            e.setNoFitRect(i, -1, -1, -1, -1)
        if fit["rect"]:
            e.setNoFitRect(0, *fit["rect"])
        s, c, n = e.doFitCalculation(fit["bstop"], fit["border"], prec)
        assert abs(s / fit["fqs"] - 1) < tol, (fit, s)
        assert c == fit["cnt"] and n == fit["nancnt"]


def test_batched_objective_equals_sequential():
    import torch
    from crystalscatter_amd import Context
    m = manifest()["fixtures"]["cfg4_bct.npz"]
    f = fixture("cfg4_bct.npz")
    seq = []
    sets = []
    data = torch.tensor(f["fitdata"], dtype=torch.float64, device="cuda")
    for p in m["params"]:
        e = SasCalc(device=0, params=p)
        e.setFitData(32, 32, f["fitdata"])
        s, c, n = e.doFitCalculation(4, 0, "mixed")
        seq.append(s)
        h = SasCalc(device=-1, params=p)
        h.set("ignore_cbs", 1)
        h.prepare()
        sets.append(h)
    ctx = Context(0)
    L = _lib.lib()
    assert L.csx_fit_batch_begin(ctx._h) == 0
    mask = _lib.Mask()
    for i in range(16):
        mask.rects[i] = -1
    mask.bstop, mask.border = 4, 0
    keep = []
    for h in sets:
        carr = np.concatenate([h.table(n) for n in ("carr1p", "carr2p", "carr3p", "carr4p", "carr5p", "carr6p", "carr7p",
                                                    "carr8p", "carr9p", "carr1f", "carr2f", "carr3f", "carr4f", "carr5f",
                                                    "carr6f", "carr7f", "carr8f", "carr9f")])
        l1, l2, l3 = h.table("latpar1"), h.table("latpar2"), h.table("latpar3")
        keep += [carr, l1, l2, l3]
        rc = L.csx_fit_batch_stage(ctx._h, ctypes.byref(h.scalars()), carr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   l1.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   l2.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                   l3.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 32, 32, ctypes.byref(mask),
                                   ctypes.c_void_p(data.data_ptr()))
        assert rc == 0
    sums = np.zeros(len(sets))
    cnts = np.zeros(len(sets), dtype=np.int64)
    assert L.csx_fit_batch_run(ctx._h, 1, sums.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               cnts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))) == 0
    np.testing.assert_array_equal(sums, np.array(seq))


def test_error_paths():
    from crystalscatter_amd import Context
    ctx = Context(0)
    L = _lib.lib()
    g = _lib.Grid(0, 8, 0, 8)
    assert L.csx_render(ctx._h, ctypes.byref(g), 1, ctypes.c_void_p(1)) == -4   # before upload
    e = SasCalc(device=0, params=dict(CFG1))
    with pytest.raises(ValueError):
        e.doCalculation("fp16")
    with pytest.raises(CsxError):
        SasCalc(device=0, params=dict(CFG1, uca=0.001)).doCalculation()


def test_back_to_back_renders_keep_their_own_radial_tables():
    """Renders queued without synchronisation alternate the radial tables (the pre-pass of one
    render overlaps the previous image kernel): every image equals its own isolated render."""
    import torch
    CFG3_SMALL = dict(CFG3, GridPoints=64)
    variants = [dict(CFG3_SMALL, EditRadius=r, uca=a, ucb=a, ucc=a) for r, a in ((8, 30), (6, 24), (9, 33), (7, 28))]
    engines = [SasCalc(device=0, params=v) for v in variants]
    n = 128 * 128
    alone = []
    for e in engines:
        out = torch.empty(n, dtype=torch.float64, device="cuda")
        e.doCalculationDevice(out.data_ptr(), "mixed")
        torch.cuda.synchronize()
        alone.append(out.clone())
    # one context, parameters switched between queued renders
    e = SasCalc(device=0, params=variants[0])
    outs = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(8)]
    for k in range(8):
        e.set_params(variants[k % 4])
        e.doCalculationDevice(outs[k].data_ptr(), "mixed")
    torch.cuda.synchronize()
    for k in range(8):
        assert torch.equal(outs[k], alone[k % 4]), k
    # same parameters, no upload in between: the renders overlap pre-pass and image kernel
    e.set_params(variants[1])
    for k in range(8):
        e.doCalculationDevice(outs[k].data_ptr(), "mixed")
    torch.cuda.synchronize()
    for k in range(8):
        assert torch.equal(outs[k], alone[1]), k


@pytest.mark.parametrize("prec,tol", [("ref64", 1e-13), ("mixed", 1e-4)])
@pytest.mark.parametrize("name", ["bcc_sphere.npz", "sc_sphere_iso.npz", "fcc_sphere_iso.npz", "bct_sphere.npz",
                                  "sphere_zdir.npz", "cyl_iso.npz", "cyl_iso_b.npz", "cube_zdir.npz",
                                  "cube_zdir_b.npz", "fcc_sphere_lorentz.npz", "bcc_sphere_mod1lor.npz",
                                  "sc_sphere_mod2lor_iso.npz", "fcc_sphere_pvoigt.npz", "bct_sphere_pearson.npz",
                                  "fcc_sphere_pearson_iso.npz", "ia3d_sphere.npz", "pn3m_sphere_iso.npz",
                                  "im3m_sphere_lorentz.npz", "fd3m_sphere.npz", "fd3m_sphere_pvoigt_iso.npz",
                                  "fcc_sphere_hkl0.npz", "hcp_sphere.npz", "ortho_sphere_iso.npz",
                                  "fcc_sphere_ordis1.npz", "bcc_sphere_zdir_lorentz.npz", "hex_sphere.npz"])
def test_generic_kernel_keys(name, prec, tol):
    """Combinations of the reference's generic kernel served here (spheres on the BCC, SC, BCT,
    Ia3d, Pn3m, Im3m and Fd3m lattices, lattices without reflections, other orientation
    distributions, isotropic cylinders, the other peak shapes) against its own output.  The peak shapes other than the Gaussian run
    the FP64 sums in either precision."""
    f = fixture(name)
    e, img = _calc(params(name), prec)
    r = rel_err(img, f["image"])
    _report("%s %s" % (name, prec), r)
    assert r.max() < tol


@pytest.mark.parametrize("name", ["fcc_sphere_pvoigt.npz", "bct_sphere_pearson.npz", "sc_sphere_mod2lor_iso.npz"])
def test_peak_shapes_256_against_oracle_random_pixels(oracle_lib, name):
    """The other peak shapes at 256x256 (the FP64 sums): 2,000 random pixels against the
    oracle, which equals the reference bit for bit on the 64x64 fixtures."""
    p = dict(params(name), GridPoints=128)
    e, img = _calc(p, "mixed")
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    rng = np.random.default_rng(4)
    idx = rng.choice(img.size, 2000, replace=False)
    ref = o.image_points(idx % 256 - 128, idx // 256 - 128, nthreads=16)
    r = rel_err(img[idx], ref)
    _report("%s 256^2 vs oracle" % name, r)
    assert r.max() < 1e-12


@pytest.mark.parametrize("prec,tol", [("ref64", 1e-11), ("mixed", 1e-4)])
@pytest.mark.parametrize("name", ["hex_cyl_zdir.npz", "sq_cyl_zdir.npz", "rect_cyl_zdir.npz", "fcc_cyl_zdir.npz",
                                  "fcc_cyl_iso.npz", "hex_cyl_iso.npz", "sc_cube_iso.npz", "bcc_cube_iso.npz",
                                  "bct_cyl_zdir_lorentz.npz", "fcc_cyl_iso_pearson.npz", "sc_cube_iso_pvoigt.npz"])
def test_cylinders_on_lattices(name, prec, tol):
    """Perfectly oriented cylinders with Gaussian peaks on the hexagonal, square and centred
    rectangular 2D lattices and on FCC (the generic kernel's cylinder z-dir branch, served by
    the BCT kernel), and isotropic cylinders on FCC and the hexagonal lattice (the lattice
    sums of the spheres with the cylinder's P(q) and <F>^2), and isotropic homogeneous cubes
    on SC and BCC (<F>^2 = P(q)), against the reference's own images; tolerances as for the
    BCT sets."""
    f = fixture(name)
    e, img = _calc(params(name), prec)
    r = rel_err(img, f["image"])
    _report("%s %s" % (name, prec), r)
    assert r.max() < tol


@pytest.mark.parametrize("name", ["fd3m_sphere.npz", "ia3d_sphere.npz", "pn3m_sphere_iso.npz"])
def test_cubic_lattices_256_against_oracle_random_pixels(oracle_lib, name):
    """Gaussian peaks on the Fd3m, Ia3d and Pn3m lattices at 256x256 in mixed precision: 2,000
    random pixels against the oracle (bit-exact with the reference on the 64x64 fixtures)."""
    p = dict(params(name), GridPoints=128)
    e, img = _calc(p, "mixed")
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    rng = np.random.default_rng(8)
    idx = rng.choice(img.size, 2000, replace=False)
    ref = o.image_points(idx % 256 - 128, idx // 256 - 128, nthreads=16)
    r = rel_err(img[idx], ref)
    _report("%s 256^2 vs oracle" % name, r)
    assert r.max() < 1e-4


def test_bcc_512_against_oracle_random_pixels(oracle_lib):
    p = dict(params("bcc_sphere.npz"), GridPoints=256)
    e, img = _calc(p, "mixed")
    o = oracle_lib.OracleEngine(p)
    o.prepare()
    rng = np.random.default_rng(3)
    idx = rng.choice(img.size, 3000, replace=False)
    ref = o.image_points(idx % 512 - 256, idx // 512 - 256, nthreads=16)
    r = rel_err(img[idx], ref)
    _report("bcc 512^2 vs oracle", r)
    assert r.max() < 1e-4
    a = img.reshape(512, 512)[1:, 1:]
    assert rel_err(a, a[::-1, ::-1]).max() < 1e-6
