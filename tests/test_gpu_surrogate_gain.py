"""What the trained surrogate network gains over the fits without it (config 5 in small).

The network is trained on the GPU, by the library's BF16 kernels, on patterns the exact kernel
renders (256 BCT-cylinder patterns of 64x64, parameters drawn like the fit workload).  Sixteen
fits of Poisson targets at hidden parameters then run twice from an uninformed start (the
centre of the box):
- without the network: the batched simplex, 100 iterations;
- with it: the network ranks candidate starts, the exact objective scores the best 8 per fit
  in one batched launch, and the simplex runs 30 iterations from the best, its trial points
  screened by the network (csx_nm_batch_screened).
With the network the fits reach the plain fits' median chi^2 with well under half the exact
objective evaluations; every final value is the exact objective at the final point."""
import numpy as np
import pytest

from crystalscatter_amd import Context, SasCalc, nm_batch, nm_eval
from crystalscatter_amd import surrogate as sg

pytestmark = pytest.mark.gpu


def _problems(nf, n, seed):
    import torch
    rng = np.random.default_rng(seed)
    nrng = np.random.default_rng(seed + 1)
    out = []
    for x in sg.draw_parameters(rng, nf):
        p = sg.params_dict(x, n // 2)
        e = SasCalc(device=0, params=p)
        e.set("ignore_cbs", 1)
        img = e.image2d(e.doCalculation("ref64"))
        x0, x1, y0, y1 = e.bounds()
        ii, jj = np.meshgrid(np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
        stop = ii * ii + jj * jj < 16
        good = np.isfinite(img) & (img > 0) & ~stop
        scale = 1e4 / img[good].max()
        data = nrng.poisson(np.where(good, img * scale, 0.0)).astype(np.float64)
        data[stop] = 0
        base = SasCalc(device=-1, params=dict(p, I0=scale))
        base.set("ignore_cbs", 1)
        base.prepare()
        out.append(dict(engine=base, data=torch.tensor(data.ravel(), device="cuda"), w=x1 - x0, h=y1 - y0,
                        truth=x, scale=scale))
    return out


def test_surrogate_gain_over_plain_fits():
    data = sg.generate(256, grid_points=32, px_per_pattern=2048, seed=7)
    net = sg.MLP(device=0, max_batch=16384, seed=1)
    net.set_features(*sg.physics_features())
    hist = sg.train(net, data, 1500, batch=16384, log_every=1499)
    assert hist[-1][1] < 0.1 * hist[0][1]
    probs = _problems(16, 64, 11)
    centre = 0.5 * (sg.FIT_LO + sg.FIT_HI)
    ctx = Context(0)
    fits = [dict(engine=p["engine"], d_data=p["data"].data_ptr(), w=p["w"], h=p["h"], x0=centre, lo=sg.FIT_LO,
                 hi=sg.FIT_HI) for p in probs]
    plain, _ = nm_batch(ctx, fits, sg.FIT_KEYS, stp=0.1, maxit=100, ftol=1e-6, mode="corrected")
    chi_p = np.array([r[1] for r in plain])
    ev_p = sum(r[3] for r in plain)
    rng = np.random.default_rng(99)
    fpx = [sg.FitPixels(p["engine"], p["data"], p["scale"], 256, rng) for p in probs]
    cands, _ = sg.propose_starts(net, fpx, data.mu, data.sigma, [centre] * len(probs), rng, top=8)
    exact = nm_eval(ctx, fits, sg.FIT_KEYS, cands)
    for f, C, k in zip(fits, cands, exact.argmin(axis=1)):
        f["x0"] = C[k]
    screen = sg.SurrogateScreen(net, fpx, data.mu, data.sigma)
    sur, st = nm_batch(ctx, fits, sg.FIT_KEYS, stp=0.02, maxit=30, ftol=1e-6, mode="corrected", screen=screen,
                       margin=0.05)
    chi_s = np.array([r[1] for r in sur])
    ev_s = sum(r[3] for r in sur) + exact.size
    print("plain: median chi2 %.2f, %d evaluations; surrogate: median chi2 %.2f, %d evaluations, %d screened out"
          % (np.median(chi_p), ev_p, np.median(chi_s), ev_s, st["pruned"]), flush=True)
    assert screen.calls > 0
    assert np.median(chi_s) <= 1.01 * np.median(chi_p)
    assert ev_s < 0.5 * ev_p
    # every final value is the exact objective at the final point
    again = nm_eval(ctx, fits, sg.FIT_KEYS, np.array([[r[0]] for r in sur]))[:, 0]
    np.testing.assert_array_equal(again, chi_s)
