"""Batched Nelder-Mead fits on the GPU (csx_nm_batch): every fit follows exactly the path of
the same simplex run one fit at a time on the per-fit objective (csx_engine_fit)."""
import numpy as np
import pytest

from helpers import manifest

from crystalscatter_amd import Context, SasCalc, nm_batch, simplex

pytestmark = pytest.mark.gpu

KEYS = ["EditRadius", "Length", "uca,ucb", "ucc", "EditDomainSize", "EditAzi"]


def _problem(k, p, n=16):
    import torch
    p = dict(p, GridPoints=n, ucb=p["uca"])
    e = SasCalc(device=0, params=p)
    e.set("ignore_cbs", 1)
    img = e.doCalculation("ref64")
    w = 2 * n
    scale = 1e4 / img.max()  # data in counts; the fits know the scale (I0)
    data = img * scale
    p["I0"] = scale
    e.set("I0", scale)
    truth = np.array([p[k.split(",")[0]] for k in KEYS])
    lo = np.array([2, 10, 15, 15, 100, 2.0])
    hi = np.array([6, 60, 40, 80, 400, 10.0])
    x0 = np.clip(truth * (1 + 0.08 * np.cos(np.arange(6) + k)), lo * 1.001, hi * 0.999)
    d = torch.tensor(data, dtype=torch.float64, device="cuda")
    return e, p, d, w, x0, lo, hi


@pytest.mark.parametrize("mode", ["reference", "corrected"])
def test_batched_nm_equals_one_fit_at_a_time(mode):
    import torch
    params = manifest()["fixtures"]["cfg4_bct.npz"]["params"][:4]
    probs = [_problem(k, p) for k, p in enumerate(params)]
    fits = []
    for e, p, d, w, x0, lo, hi in probs:
        base = SasCalc(device=-1, params=p)
        base.set("ignore_cbs", 1)
        if mode == "reference":  # the zeroed first entry (the radius) must stay inside the box
            lo = lo.copy()
            lo[0] = 0.0
        fits.append(dict(engine=base, d_data=d.data_ptr(), w=w, h=w, x0=x0, lo=lo, hi=hi, bstop=2))
    ctx = Context(0)
    res, stats = nm_batch(ctx, fits, KEYS, stp=0.1, maxit=25, ftol=1e-8, precision="mixed", mode=mode)
    assert stats["rounds"] > 25
    for (e, p, d, w, x0, lo, hi), fd, (xb, fb, itb, evb) in zip(probs, fits, res):
        lo = fd["lo"]
        e.setFitData(w, w, d.cpu().numpy())

        def f(x):
            for k, v in zip(KEYS, x):
                for name in k.split(","):
                    e.set(name, v)
            return e.doFitCalculation(bstop=2, border=0, precision="mixed")[0]

        xs, fs, its, evs = simplex(f, x0, lo, hi, stp=0.1, maxit=25, ftol=1e-8, mode=mode)
        np.testing.assert_array_equal(xb, xs)
        assert fb == fs and itb == its and evb == evs
        torch.cuda.synchronize()


def test_batched_nm_lowers_the_objective():
    """Noise-free targets: the fits move downhill from the initial simplex (corrected mode;
    the reference's zeroed first entry leaves its vertices outside these bounds)."""
    params = manifest()["fixtures"]["cfg4_bct.npz"]["params"][:3]
    probs = [_problem(k, p, n=24) for k, p in enumerate(params)]
    fits = []
    for e, p, d, w, x0, lo, hi in probs:
        base = SasCalc(device=-1, params=p)
        base.set("ignore_cbs", 1)
        fits.append(dict(engine=base, d_data=d.data_ptr(), w=w, h=w, x0=x0, lo=lo, hi=hi))
    ctx = Context(0)
    start, _ = nm_batch(ctx, fits, KEYS, stp=0.05, maxit=0, ftol=1e-12, precision="mixed", mode="corrected")
    res, stats = nm_batch(ctx, fits, KEYS, stp=0.05, maxit=200, ftol=1e-12, precision="mixed", mode="corrected")
    for (x0r, f0, _, _), (x, f, it, ev) in zip(start, res):
        assert np.isfinite(f) and f < 0.5 * f0, (f0, f)


def test_nm_eval_equals_per_point_objective():
    """csx_nm_eval at explicit points (one batched launch over fits x points) equals the
    per-fit objective evaluated one point at a time; out-of-box points get 1e20."""
    from crystalscatter_amd import nm_eval
    params = manifest()["fixtures"]["cfg4_bct.npz"]["params"][:3]
    probs = [_problem(k, p) for k, p in enumerate(params)]
    fits = []
    for e, p, d, w, x0, lo, hi in probs:
        base = SasCalc(device=-1, params=p)
        base.set("ignore_cbs", 1)
        fits.append(dict(engine=base, d_data=d.data_ptr(), w=w, h=w, x0=x0, lo=lo, hi=hi, bstop=2))
    rng = np.random.default_rng(3)
    X = np.stack([np.stack([np.clip(x0 * (1 + rng.uniform(-0.1, 0.1, 6)), lo, hi) for _ in range(4)])
                  for (_, _, _, _, x0, lo, hi) in probs])
    X[1, 2, 0] = 100.0  # outside the box
    got = nm_eval(Context(0), fits, KEYS, X, precision="mixed")
    for i, (e, p, d, w, x0, lo, hi) in enumerate(probs):
        e.setFitData(w, w, d.cpu().numpy())
        for v in range(4):
            if i == 1 and v == 2:
                assert got[i, v] == 1e20
                continue
            for k, val in zip(KEYS, X[i, v]):
                for name in k.split(","):
                    e.set(name, val)
            ref = e.doFitCalculation(bstop=2, border=0, precision="mixed")[0]
            assert got[i, v] == ref, (i, v, got[i, v], ref)
