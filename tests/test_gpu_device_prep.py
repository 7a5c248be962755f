"""Device preparation of batched lattice sets (csx_set_device_prep).

The tables that one workgroup per set builds on the device from the set's rotated reciprocal
bases -- the reflection table latpar3, the Friedel-pair records in their |u_z| segments, the
|q| bin table, the isotropic list and the per-set scalars derived from them -- equal the host's
(prepare + the staging tables) bit for bit, for every lattice kind the batch serves; so do the
objective sums (graph and eager rounds) and whole batched fits.
"""
import numpy as np
import pytest

from helpers import manifest

from crystalscatter_amd import Context, SasCalc, _lib, nm_batch, nm_eval

pytestmark = pytest.mark.gpu

BCT_KEYS = ["EditRadius", "Length", "uca,ucb", "ucc", "EditDomainSize", "EditAzi"]
BCT_LO = np.array([2, 10, 15, 15, 100, 2.0])
BCT_HI = np.array([6, 60, 40, 80, 400, 10.0])
SPH_KEYS = ["EditRadius", "uca,ucb", "ucc", "EditDomainSize", "EditAzi"]
SPH_LO = np.array([4, 20, 20, 100, 2.0])
SPH_HI = np.array([10, 40, 50, 400, 10.0])


def _base(name):
    p = manifest()["fixtures"][name]["params"]
    return dict(p[0] if isinstance(p, list) else p)


def _fits(params, lo, hi, nfits, n, rng):
    import torch
    fits, keep = [], []
    for _ in range(nfits):
        p = dict(params, GridPoints=n)
        base = SasCalc(device=-1, params=p)
        base.set("ignore_cbs", 1)
        w = 2 * n
        d = torch.tensor(rng.uniform(0.5, 100.0, w * w), dtype=torch.float64, device="cuda")
        keep.append(d)
        x0 = lo + (hi - lo) * rng.uniform(0.2, 0.8, len(lo))
        fits.append(dict(engine=base, d_data=d.data_ptr(), w=w, h=w, x0=x0, lo=lo, hi=hi, bstop=2))
    return fits, keep


def _run(fits, keys, X, precision, device_prep, graph=True):
    ctx = Context(0)
    ctx.set_device_prep(device_prep)
    L = _lib.lib()
    assert L.csx_set_batch_graph(ctx._h, 1 if graph else 0) == 0
    sums = nm_eval(ctx, fits, keys, X, precision=precision)
    nsets = X.shape[0] * X.shape[1]
    tabs = [[ctx.fit_batch_table(s, w).copy() for w in range(6)] for s in range(nsets)]
    counts = ctx.device_prep_counts()
    ctx.close()
    return sums, tabs, counts


CASES = [
    ("cfg4_bct.npz", {}, BCT_KEYS, BCT_LO, BCT_HI, "mixed"),
    ("cfg4_bct.npz", {}, BCT_KEYS, BCT_LO, BCT_HI, "ref64"),
    ("cfg4_bct.npz", {"HKLmax": 7}, BCT_KEYS, BCT_LO, BCT_HI, "mixed"),
    ("bcc_sphere.npz", {}, SPH_KEYS, SPH_LO, SPH_HI, "mixed"),
    ("sc_sphere_iso.npz", {}, SPH_KEYS, SPH_LO, SPH_HI, "mixed"),
    ("fcc_sphere_iso.npz", {}, SPH_KEYS, SPH_LO, SPH_HI, "mixed"),
    ("bct_sphere.npz", {}, SPH_KEYS, SPH_LO, SPH_HI, "mixed"),
    ("fd3m_sphere.npz", {}, SPH_KEYS, SPH_LO, SPH_HI, "mixed"),
    ("ia3d_sphere.npz", {}, SPH_KEYS, SPH_LO, SPH_HI, "mixed"),
    ("fcc_cyl_iso.npz", {}, BCT_KEYS, np.array([2, 20, 20, 20, 100, 2.0]), np.array([4, 60, 40, 50, 400, 10.0]),
     "mixed"),
]


@pytest.mark.parametrize("name,extra,keys,lo,hi,precision", CASES,
                         ids=["%s%s-%s" % (c[0][:-4], "-hkl7" if c[1] else "", c[5]) for c in CASES])
def test_device_tables_equal_host_tables(name, extra, keys, lo, hi, precision):
    rng = np.random.default_rng(5)
    params = dict(_base(name), **extra)
    nfits, nv = 8, 16
    fits, keep = _fits(params, lo, hi, nfits, 16, rng)
    X = lo + (hi - lo) * rng.uniform(0.0, 1.0, (nfits, nv, len(lo)))
    print("case", name, extra, precision, "sets", nfits * nv, flush=True)
    s_dev, t_dev, c_dev = _run(fits, keys, X, precision, True)
    s_host, t_host, c_host = _run(fits, keys, X, precision, False)
    s_eager, t_eager, _ = _run(fits, keys, X, precision, True, graph=False)
    print("device-prepared sets %d, checked roundings %d" % c_dev, flush=True)
    assert c_dev[0] == nfits * nv and c_host == (0, 0)
    assert np.all(np.isfinite(s_host)) and np.all(s_host < 1e20)
    np.testing.assert_array_equal(s_dev.view(np.uint64), s_host.view(np.uint64))
    np.testing.assert_array_equal(s_eager.view(np.uint64), s_host.view(np.uint64))
    names = ["latpar3", "pairs", "pairs64", "lut", "iso", "scalars"]
    for s in range(nfits * nv):
        for w in range(6):
            a, b = t_dev[s][w], t_host[s][w]
            assert a.size == b.size and np.array_equal(a.view(np.uint8), b.view(np.uint8)), (s, names[w])
            assert np.array_equal(t_eager[s][w].view(np.uint8), b.view(np.uint8)), (s, names[w])
    # the tables are not trivial: pairs in several segments, a filled bin table
    sc = t_host[0][5]
    assert sc[0] > 12 and 0 < sc[2] < sc[3]


@pytest.mark.parametrize("name,extra,keys,lo,hi", [("fcc_sphere_hkl0.npz", {}, SPH_KEYS, SPH_LO, SPH_HI),
                                                   ("cfg4_bct.npz", {"HKLmax": 0}, BCT_KEYS, BCT_LO, BCT_HI)],
                         ids=["fcc-hkl0", "bct-hkl0"])
def test_sets_without_reflections(name, extra, keys, lo, hi):
    """A lattice set without reflections (HKLmax 0) asked for device preparation stages its
    empty tables from the host; its sums equal those of the host path bit for bit."""
    rng = np.random.default_rng(6)
    params = dict(_base(name), **extra)
    fits, keep = _fits(params, lo, hi, 4, 16, rng)
    X = lo + (hi - lo) * rng.uniform(0.0, 1.0, (4, 8, len(lo)))
    s_dev, t_dev, c_dev = _run(fits, keys, X, "mixed", True)
    s_host, t_host, c_host = _run(fits, keys, X, "mixed", False)
    assert c_dev == (0, 0) and c_host == (0, 0)
    assert np.all(np.isfinite(s_host)) and np.all(s_host > 0)
    np.testing.assert_array_equal(s_dev.view(np.uint64), s_host.view(np.uint64))
    for s in range(X.shape[0] * X.shape[1]):
        for w in range(6):
            assert np.array_equal(t_dev[s][w].view(np.uint8), t_host[s][w].view(np.uint8)), (s, w)


def test_batched_fits_identical_with_device_preparation():
    """Whole batched fits (BCT cylinders, mixed precision, corrected simplex): every iterate,
    hence the final vertex, value, iteration and evaluation counts, is the same whether the
    host or the device builds the lattice tables."""
    rng = np.random.default_rng(9)
    params = _base("cfg4_bct.npz")
    fits, keep = _fits(params, BCT_LO, BCT_HI, 6, 16, rng)
    out = {}
    for dev in (True, False):
        ctx = Context(0)
        ctx.set_device_prep(dev)
        res, stats = nm_batch(ctx, fits, BCT_KEYS, stp=0.1, maxit=30, ftol=1e-8, precision="mixed", mode="corrected")
        out[dev] = res
        if dev:
            assert ctx.device_prep_counts()[0] >= stats["evals"] > 6 * 30
        ctx.close()
    for (xa, fa, ia, ea), (xb, fb, ib, eb) in zip(out[True], out[False]):
        np.testing.assert_array_equal(xa, xb)
        assert fa == fb and ia == ib and ea == eb


def test_stage_points_images_equal_one_stage_per_image():
    """csx_engine_stage_points (native staging of many parameter points, lattice tables built on
    the device) renders the same images, bit for bit, as staging each image through its own
    engine with host-built tables."""
    import torch
    from crystalscatter_amd import render_batch
    rng = np.random.default_rng(21)
    params = dict(_base("cfg4_bct.npz"), GridPoints=24)
    X = BCT_LO + (BCT_HI - BCT_LO) * rng.uniform(0.0, 1.0, (24, 6))
    base = SasCalc(device=-1, params=params)
    base.set("ignore_cbs", 1)
    base.prepare()
    x0, x1, y0, y1 = base.block()
    npx = (x1 - x0) * (y1 - y0)
    ctx = Context(0)
    L = _lib.lib()
    out = torch.empty((len(X), npx), dtype=torch.float64, device="cuda")
    assert L.csx_fit_batch_begin(ctx._h) == 0
    base.stage_points(ctx, BCT_KEYS, X, out.data_ptr(), npx)
    assert L.csx_render_batch_run(ctx._h, 1) == 0
    assert ctx.device_prep_counts()[0] == len(X)
    engines = []
    for x in X:
        e = SasCalc(device=-1, params=params)
        e.set("ignore_cbs", 1)
        for k, v in zip(BCT_KEYS, x):
            for name in k.split(","):
                e.set(name, v)
        engines.append(e)
    ctx2 = Context(0)
    ctx2.set_device_prep(False)
    ref = render_batch(ctx2, engines, precision="mixed")
    torch.cuda.synchronize()
    assert torch.isfinite(ref).all() and (ref > 0).any()
    assert torch.equal(out, ref)
