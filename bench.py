#!/usr/bin/env python3
"""Benchmark: FCC-sphere 1024x1024 patterns per second on one MI355X.

A step is one full pattern of the headline workload (sphere form factor x FCC
lattice factor with Gaussian peaks, Gaussian domain orientation, combination
key 0x00020500): R=8 nm, Schulz sigma=0.08, a=30 nm, HKLmax=5 (340 anisotropic
reflections), Debye-Scherrer width from a 200 nm domain, Debye-Waller 1.5 nm,
angular width 5 deg, q in [-1,1] nm^-1.  Tables are uploaded to HBM before the
timed region; each step renders all 1048576 pixels into an HBM buffer (FP64,
the reference's IDX order).  The image is checked against the FP64 CPU oracle
on 4096 sampled pixels (max relative error reported, must stay <= 1e-4).

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--precision mixed]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "1024² patterns/s at ≤1e-4 rel err vs FP64; 2D fits/s surrogate vs plain"

CFG3 = dict(LType=5, ComboBoxParticle=0, Ordis=0, ComboBoxPeak=1, ComboBoxInterior=0, EditRadius=8,
            EditSigma=0.08, uca=30, ucb=30, ucc=30, EditDomainSize=200, RadButDebyeScherrer=1,
            EditDebyeWaller=1.5, EditAzi=5, HKLmax=5, GridPoints=512, EditQmax=1, I0=1, Base=0)

# Roofline accounting of the dominant kernel, the per-pixel pass k_image_lat_hot<K_FCC, mixed>
# of the two-pass lattice render (its radial pre-pass k_radial and the fix pass for pixels
# that need an FP64 path are timed with the whole render, render_ms_per_launch).
#
# roofline.achieved counts the work the kernel's algorithm executes in its FP32 pair loop:
# one pair evaluation covers reflections +hkl and -hkl for one pixel, at 40 FP32 FLOP
# (FMA = 2): q^.u 3, |q^ x u|^2 5, acos^2/sin^2 polynomial 19, y^2 1, radial exponent 4,
# first angular exponent 2, second exponent 3, sum of the two exp2 1, weighting 2; plus
# 3 transcendentals (2 exp2, 1 sqrt).  The number of pair evaluations per launch comes
# from the kernel's own counters (csx_render_stats).  The per-pixel FP64 work (q, sphere
# factors, S(q)) is not counted, so achieved is a lower bound.
EXEC_FLOP_PER_PAIR_EVAL, EXEC_T_PER_PAIR_EVAL = 40, 3
#
# Reference-equivalent rate (reported beside it): the reference formula's operations per
# pixel (add/sub/mul/div = 1 FLOP; exp, log, sin, cos, atan, acos, sqrt = 1 transcendental,
# pow = log + exp = 2), closed-form branch of the sphere factors, iso = 0:
#   per anisotropic reflection (Source/sc_gpu_pSph_pGau_lFCC_oGau.h:103-125): cosine 7,
#     radial Gaussian 10, angular factor 9, accumulation 3 = 29 FLOP; acos + 2 exp = 3 T
#   per pixel outside the loop: q 10 FLOP + 1 T, sphere P(q) 65 + 12 T
#     (Source/sc_lib_formpq_partSphere.h:58-74), <F>^2 31 + 9 T
#     (Source/sc_lib_formfq_partSphere.h:8-52), width, S(q) and the affine map 36 + 2 T
# = 340 * 29 + 142 = 10002 FLOP and 340 * 3 + 24 = 1044 T per pixel.  The kernel skips
# reflections whose terms are below 2^-32 of their weight, so this rate can exceed the
# FP32 peak: it measures the speed-up over the reference algorithm, not ALU use.
ALG_FLOP_PER_REFL, ALG_T_PER_REFL = 29, 3
ALG_FLOP_PER_PX_EXTRA, ALG_T_PER_PX_EXTRA = 142, 24
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector, dense (packed dual-issue counted)
PEAK_TRANS_TOPS = 9.83        # 256 CU x 64 lanes x 2.4 GHz / 4 (quarter-rate transcendentals)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_fcc1024_mixed.txt")


def measured_traffic():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


# Batched 2D fits: BCT lattice of cylinders, z-axis orientation (key 0x01020806); the fit
# parameters, their box and the fixed keys are the package's.
from crystalscatter_amd.surrogate import BCT_BASE as CFG4_BASE, FIT_HI, FIT_KEYS, FIT_LO  # noqa: E402


def fit_problems(nfits, seed=1234, noise_seed=5678):
    """The fit workload: hidden true parameters drawn uniformly (R in [2,6] nm, L in [10,60] nm,
    a in [15,40] nm, c/a in [1,2], domain in [100,400] nm, angular peak width EditAzi in
    [2,10] deg), detector N in {128,256,384,512}, q in [-1,1] nm^-1; target = FP64 image
    at the truth scaled to 1e4 counts at the brightest unmasked pixel (the fits know this
    scale as I0), Poisson noise, a circular beam stop of radius 8 px (data 0: excluded by
    the FQS rule); start = truth perturbed by up to +-10 % per parameter."""
    import torch
    from crystalscatter_amd import SasCalc
    rng = np.random.default_rng(seed)
    nrng = np.random.default_rng(noise_seed)
    probs = []
    for k in range(nfits):
        R, L, a, ca = rng.uniform(2, 6), rng.uniform(10, 60), rng.uniform(15, 40), rng.uniform(1, 2)
        dom, azi = rng.uniform(100, 400), rng.uniform(2, 10)
        n = int(rng.choice([128, 256, 384, 512]))
        p = dict(CFG4_BASE, EditRadius=R, Length=L, uca=a, ucb=a, ucc=a * ca, EditDomainSize=dom, EditAzi=azi,
                 GridPoints=n // 2)
        e = SasCalc(device=0, params=p)
        e.set("ignore_cbs", 1)
        img = e.image2d(e.doCalculation("ref64"))
        x0, x1, y0, y1 = e.bounds()
        ii, jj = np.meshgrid(np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
        stop = ii * ii + jj * jj < 8 * 8
        good = np.isfinite(img) & (img > 0) & ~stop
        scale = 1e4 / img[good].max()
        lam = np.where(good, img * scale, 0.0)  # NaN / non-positive model pixels carry no data
        data = nrng.poisson(lam).astype(np.float64)
        data[stop] = 0.0
        truth = np.array([R, L, a, a * ca, dom, azi])
        start = np.clip(truth * (1 + rng.uniform(-0.1, 0.1, 6)), FIT_LO * 1.001, FIT_HI * 0.999)
        base = SasCalc(device=-1, params=dict(p, I0=scale))  # the count scale is known
        base.set("ignore_cbs", 1)
        base.prepare()
        probs.append(dict(engine=base, data=torch.tensor(data.ravel(), dtype=torch.float64, device="cuda"),
                          w=x1 - x0, h=y1 - y0, truth=truth, x0=start, n=n, scale=scale))
    return probs


def run_fits(args, probs=None):
    """Config 4: the batched simplex (corrected mode) over the fit workload; returns its line."""
    import torch
    from crystalscatter_amd import Context, nm_batch
    t0 = time.perf_counter()
    if probs is None:
        probs = fit_problems(args.nfits)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    fits = [dict(engine=p["engine"], d_data=p["data"].data_ptr(), w=p["w"], h=p["h"], x0=p["x0"], lo=FIT_LO,
                 hi=FIT_HI) for p in probs]
    ctx = Context(0)
    ctx.set_device_prep(not args.host_prep)
    if args.eager:
        from crystalscatter_amd import _lib
        _lib.lib().csx_set_batch_graph(ctx._h, 0)
    t0 = time.perf_counter()
    res, st = nm_batch(ctx, fits, FIT_KEYS, stp=0.05, maxit=args.maxit, ftol=1e-6, precision=args.precision, mode="corrected",
                       threads=args.cpu_threads)
    rounds = st["rounds"]
    wall = time.perf_counter() - t0
    px = sum(p["w"] * p["h"] for p in probs)
    evals = np.array([r[3] for r in res])
    iters = np.array([r[2] for r in res])
    rel0 = np.array([np.abs(p["x0"] / p["truth"] - 1) for p in probs])
    rel1 = np.array([np.abs(r[0][:6] / p["truth"] - 1) for p, r in zip(probs, res)])
    chi2 = np.array([r[1] for r in res])
    line = {
        "metric": "2D fits/s (batched Nelder-Mead, exact kernel)",
        "value": len(probs) / wall,
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": int(rounds),
        "warmup": 0,
        "ms_per_step": 1e3 * wall / max(1, rounds),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "%d independent fits, BCT cylinders pCyl_pGau_lBCT_oZDi (key 0x01020806), "
                               "6 free parameters, detectors 128..512, Poisson targets, 8 px beam stop, "
                               "maxit %d, corrected simplex" % (len(probs), args.maxit),
                   "precision": args.precision, "pixels_per_round_mean": px,
                   "lattice_tables": "host" if args.host_prep else "device",
                   "detectors": {str(n): int(sum(1 for p in probs if p["n"] == n)) for n in (128, 256, 384, 512)}},
        "rounds": int(rounds),
        "chi2_median": float(np.median(chi2)),
        "evals_mean": float(evals.mean()),
        "iters_mean": float(iters.mean()),
        "median_rel_param_err_start": float(np.median(rel0)),
        "median_rel_param_err_final": float(np.median(rel1)),
        "target_generation_s": t_gen,
        "host_prep_ms": st["prep_ms"], "staging_ms": st["stage_ms"], "launch_ms": st["launch_ms"],
        "batched_run_ms": st["run_ms"],
    }
    return line


def bench_fits(args):
    print(json.dumps(run_fits(args)), flush=True)


# Matrix-core work per training row of the surrogate, algorithmic (70 inputs, not the padded
# 128): forward 2*(70*256 + 3*256^2 + 256), backward the same for the weight gradients plus
# 2*3*256^2 + 2*256 for the activation gradients (none for the inputs).
MLP_FWD_FLOP = 2 * (70 * 256 + 3 * 256 * 256 + 256)
MLP_TRAIN_FLOP = 2 * MLP_FWD_FLOP + 2 * 3 * 256 * 256 + 2 * 256
PEAK_BF16_TFLOPS = 2500.0


def run_surrogate(args, probs=None):
    """Surrogate workload: simulate training patterns with the exact kernel, train the network
    (BF16 MFMA forward/backward/Adam), measure its held-out log10 error, then run the same
    fits from an uninformed start (the centre of the box) twice: plain batched Nelder-Mead
    with maxit iterations, and surrogate-ranked candidates (the best `rescore` per fit scored
    by the exact objective in one batched launch) followed by maxit_surrogate exact
    iterations from the best.  Every point the fits accept is scored by the exact kernel."""
    import torch
    from crystalscatter_amd import Context, nm_batch, nm_eval
    from crystalscatter_amd import surrogate as sg
    t0 = time.perf_counter()
    data = sg.generate(args.train_patterns, grid_points=128, px_per_pattern=args.train_px, seed=7)
    held = sg.generate(64, grid_points=128, px_per_pattern=4096, seed=8, mu=data.mu, sigma=data.sigma)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    net = sg.MLP(device=0, max_batch=args.batch, seed=1)
    net.set_features(*sg.physics_features())
    p0 = net.get_params()
    sg.train(net, data, 20, batch=args.batch)  # warm-up (kernels loaded, buffers touched)
    net.set_params(p0)                         # back to the initial weights, Adam state reset
    net.synchronize()
    t0 = time.perf_counter()
    hist = sg.train(net, data, args.train_steps, batch=args.batch, log_every=max(1, args.train_steps // 10))
    net.synchronize()
    t_train = time.perf_counter() - t0
    rows_per_s = args.train_steps * args.batch / t_train
    pred = sg.predict_log10(net, held.par, held.sets, held.q, data.mu, data.sigma)
    err = (pred - (held.t * data.sigma + data.mu))[held.inrange]  # the decades the fits see
    rms_log10 = float(torch.sqrt((err * err).mean()))
    med_abs_log10 = float(err.abs().median())

    if probs is None:
        probs = fit_problems(args.nfits)
    torch.cuda.synchronize()
    centre = 0.5 * (FIT_LO + FIT_HI)
    ctx = Context(0)
    fits = [dict(engine=p["engine"], d_data=p["data"].data_ptr(), w=p["w"], h=p["h"], x0=centre, lo=FIT_LO,
                 hi=FIT_HI) for p in probs]
    t0 = time.perf_counter()
    res_p, st_p = nm_batch(ctx, fits, FIT_KEYS, stp=0.1, maxit=args.maxit, ftol=1e-6, precision=args.precision, mode="corrected",
                           threads=args.cpu_threads)
    t_plain = time.perf_counter() - t0
    # warm-up of the proposal path (its torch kernels load on first use), on two fits
    wrng = np.random.default_rng(5)
    wfpx = [sg.FitPixels(p["engine"], p["data"], p["scale"], args.surrogate_px, wrng) for p in probs[:2]]
    sg.propose_starts(net, wfpx, data.mu, data.sigma, [centre] * 2, wrng, top=args.rescore)
    torch.cuda.synchronize()
    rng = np.random.default_rng(99)
    t0 = time.perf_counter()
    fpx = [sg.FitPixels(p["engine"], p["data"], p["scale"], args.surrogate_px, rng) for p in probs]
    cands, sval = sg.propose_starts(net, fpx, data.mu, data.sigma, [centre] * len(probs), rng, top=args.rescore)
    torch.cuda.synchronize()
    t_prop = time.perf_counter() - t0
    # network rows of the proposal: x_init + 128 global draws, then 4 best + 4 x 32 local
    # draws per fit (propose_starts' defaults), surrogate_px pixels each
    prop_rows = len(probs) * (1 + 128 + 4 + 4 * 32) * args.surrogate_px
    # the exact objective scores the surrogate's best candidates, in one batched launch
    t0 = time.perf_counter()
    exact = nm_eval(ctx, fits, FIT_KEYS, cands, precision=args.precision, threads=args.cpu_threads)
    t_rescore = time.perf_counter() - t0
    pick = exact.argmin(axis=1)
    for f, C, k in zip(fits, cands, pick):
        f["x0"] = C[k]
    t0 = time.perf_counter()
    if args.maxit_surrogate > 0:
        res_s, st_s = nm_batch(ctx, fits, FIT_KEYS, stp=0.02, maxit=args.maxit_surrogate, ftol=1e-6, mode="corrected",
                               precision=args.precision, threads=args.cpu_threads)
    else:  # no simplex: the exactly scored best candidate is the result
        res_s = [(C[k], float(exact[i, k]), 0, args.rescore) for i, (C, k) in enumerate(zip(cands, pick))]
        st_s = {"rounds": 1}
    t_sur = time.perf_counter() - t0 + t_prop + t_rescore
    chi_p = np.array([r[1] for r in res_p])
    chi_s = np.array([r[1] for r in res_s])
    err_p = np.array([np.abs(r[0] / p["truth"] - 1) for p, r in zip(probs, res_p)])
    err_s = np.array([np.abs(r[0] / p["truth"] - 1) for p, r in zip(probs, res_s)])
    line = {
        "metric": "2D fits/s, surrogate-started vs plain (batched Nelder-Mead, exact kernel scoring)",
        "value": len(probs) / t_sur,
        "unit": "fits/s",
        "n_gpus": 1,
        "steps": int(st_s["rounds"]),
        "warmup": 0,
        "ms_per_step": 1e3 * t_sur / max(1, st_s["rounds"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network, f32 exact kernel",
        "data": "synthetic",
        "config": {"workload": "%d fits of BCT cylinders (key 0x01020806) from the centre of the box; surrogate "
                               "MLP 70->4x256 SiLU->1 trained on %d patterns of 256x256 (%d px kept each)"
                               % (len(probs), args.train_patterns, args.train_px),
                   "maxit_plain": args.maxit, "maxit_surrogate": args.maxit_surrogate},
        "plain_fits_per_s": len(probs) / t_plain,
        "speedup_vs_plain": t_plain / t_sur,
        "chi2_median_plain": float(np.median(chi_p)), "chi2_median_surrogate": float(np.median(chi_s)),
        "frac_fits_surrogate_chi2_le_plain": float(np.mean(chi_s <= chi_p * (1 + 1e-3))),
        "median_rel_param_err_plain": float(np.median(err_p)),
        "median_rel_param_err_surrogate": float(np.median(err_s)),
        "surrogate_rms_log10_err_heldout": rms_log10, "surrogate_median_abs_log10_err_heldout": med_abs_log10,
        "heldout": "64 patterns of 256x256 (seed 8), pixels within %g decades of each pattern's maximum"
                   % sg.EVAL_DECADES,
        "train_loss_history": hist, "training_rows": len(data), "training_rows_per_s": rows_per_s,
        "training_s": t_train, "data_generation_s": t_data, "proposal_s": t_prop, "exact_rescore_s": t_rescore,
        "rescored_candidates_per_fit": args.rescore, "surrogate_px_per_fit": args.surrogate_px,
        "proposal_network_rows": prop_rows,
        "proposal_inference_tflops": prop_rows * MLP_FWD_FLOP / t_prop / 1e12,
        "proposal_inference": "fused inference kernel (csx_mlp_predict); the rate includes the pixel "
                              "subsampling and candidate draws on the host",
        "start_chi2_median": float(np.median(exact.min(axis=1))),
        "roofline": {"bound": "mfma", "achieved": rows_per_s * MLP_TRAIN_FLOP / 1e12, "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": rows_per_s * MLP_TRAIN_FLOP / 1e12 / PEAK_BF16_TFLOPS,
                     "traffic": None, "kernel": "training step (embed + 4 fwd + 3 bwd + 4 weight-gradient GEMMs)",
                     "flop_per_row": MLP_TRAIN_FLOP},
    }
    return line


def bench_surrogate(args):
    print(json.dumps(run_surrogate(args)), flush=True)


def bench_bct(args, nsets=8):
    """The BCT-cylinder kernel (key 0x01020806, mixed precision) at 512x512 on nsets random
    fit parameter sets (seed 1234, drawn like the fit workload): per set the kernel time
    (HIP events, best of 5), the guard counters and the maximum relative error against the
    FP64 reference mode on every pixel."""
    import ctypes
    import torch
    from crystalscatter_amd import SasCalc, _lib
    from crystalscatter_amd import surrogate as sg
    L = _lib.lib()
    xs = sg.draw_parameters(np.random.default_rng(1234), nsets)
    out = torch.empty(512 * 512, dtype=torch.float64, device="cuda")
    per = []
    for k in range(nsets):
        e = SasCalc(device=0, params=sg.params_dict(xs[k], 256))
        e.set("ignore_cbs", 1)
        e.prepare()
        ctx = L.csx_engine_ctx(e._h)
        e.doCalculationDevice(out.data_ptr(), "ref64")
        L.csx_synchronize(ctx)
        ref = out.cpu().numpy().copy()
        e.doCalculationDevice(out.data_ptr(), "mixed")
        L.csx_set_timing(ctx, 1)
        ms = []
        for _ in range(5):
            e.doCalculationDevice(out.data_ptr(), "mixed")
            ms.append(e.last_kernel_ms())
        L.csx_set_timing(ctx, 0)
        L.csx_synchronize(ctx)
        img = out.cpu().numpy()
        ok = np.isfinite(ref) & (ref != 0)
        rel = float((np.abs(img[ok] - ref[ok]) / np.abs(ref[ok])).max())
        st = (ctypes.c_uint64 * 5)()
        s = e.scalars()
        g = _lib.Grid(s.xmin, s.xmax, s.ymin, s.ymax)
        assert L.csx_render_stats(ctx, ctypes.byref(g), 1, ctypes.c_void_p(out.data_ptr()), st) == 0
        per.append(dict(params=[round(float(v), 3) for v in xs[k]], kernel_us=1e3 * min(ms), max_rel_err=rel,
                        pair_evals=int(st[0]), fp64_window_pixels=int(st[1]), fp64_full_pixels=int(st[4]),
                        near_one_pixels=int(st[2])))
        e.close()
    us = np.array([p["kernel_us"] for p in per])
    line = {
        "metric": "BCT-cylinder 512^2 patterns/s (mixed, guarded), random fit parameter sets",
        "value": 1e6 / us.mean(), "unit": "patterns/s", "n_gpus": 1, "steps": 5 * nsets, "warmup": 1,
        "ms_per_step": us.mean() / 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 exponents, f32 exp2", "data": "synthetic",
        "config": {"workload": "%d BCT-cylinder parameter sets (key 0x01020806) at 512x512, q in [-1,1] nm^-1"
                               % nsets},
        "max_rel_err_vs_fp64": max(p["max_rel_err"] for p in per), "sets": per,
    }
    print(json.dumps(line), flush=True)


AI_IDS = ["EditRadius=0.1", "EditSigma=0.2", "uca=0.05", "ucb=1", "ucc=1", "I0=0.3", "EditDomainSize=0.4",
          "EditDebyeWaller=0.5"]


def bench_aidata(args):
    """Training images of the console's AI-data loop for one parameter line: FCC spheres at
    128x128 with random variation of eight parameters (tpvPerformRandom), noise, a random beam
    stop, the 128x128 r,phi image of each and its FFT modulus image, rand() from seed 1.  Timed end to end (host
    random variation and staging, one batched render, the post-processing on the device), and
    against one render launch per image of the same parameter sets."""
    import ctypes
    import torch
    from crystalscatter_amd import Context, Random, SasCalc, _lib, aidata
    L = _lib.lib()
    n = args.nimages
    bx, by = (int(v) for v in args.beam.split(","))
    p = dict(CFG3, GridPoints=64, BeamPosX=bx, BeamPosY=by)
    opts = aidata.AIOptions(noise=True, bs=(0.1, 0.1, 0.5, 0.5), rphi_size=128, fft_size=128, fft_type=2,
                            fft_swap=True, fft_scale=True)
    ctx = Context(0)
    e = SasCalc(device=-1, params=p)
    e.set("ignore_cbs", 1)
    # warm-up: one call of the same size (device buffers, the pinned staging arena and the
    # allocator's cache reach their steady-state sizes)
    aidata.generate(ctx, e, AI_IDS, n, Random(2), opts, args.precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = aidata.generate(ctx, e, AI_IDS, n, Random(1), opts, args.precision)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # the batched render alone, and one render launch per image, on the same parameter sets
    imgs = torch.empty((n, 128 * 128), dtype=torch.float64, device="cuda")
    from crystalscatter_amd import tpv_images
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    tpv_images(ctx, e, AI_IDS, n, Random(1), args.precision, out=imgs)
    t_batch = time.perf_counter() - t1
    k_batch_ms = L.csx_last_kernel_ms(ctx._h)
    one = SasCalc(device=0, params=p)
    one.set("ignore_cbs", 1)
    rng = Random(1)
    buf = torch.empty(128 * 128, dtype=torch.float64, device="cuda")
    one.doCalculationDevice(buf.data_ptr(), args.precision)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for k in range(n):
        one.tpvPerformRandom(AI_IDS, rng)
        one.doCalculationDevice(buf.data_ptr(), args.precision)
    L.csx_synchronize(L.csx_engine_ctx(one._h))
    t_seq = time.perf_counter() - t2
    ref_last = imgs[n - 1]
    single_vs_batch = float(((buf - ref_last).abs() / ref_last.abs().clamp_min(1e-300)).max())
    cpu = check = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # checker only: the CPU baseline leg
        import aidata_oracle as AO
        m = 32
        e2 = SasCalc(device=-1, params=p)
        e2.set("ignore_cbs", 1)
        rng2 = Random(1)
        keys = ["EditRadius", "EditSigma", "uca", "ucb", "ucc", "I0", "width_zuf", "dw_zuf"]
        res = []
        tc = time.perf_counter()
        for k in range(m):
            e2.tpvPerformRandom(AI_IDS, rng2)
            q = dict(p)
            q.update({key: e2.tpv_value(key) for key in keys})
            o = oracle.OracleEngine(q)
            o.prepare()
            res.append(AO.post_image(o.image(nthreads=args.cpu_threads), o.bounds(), (bx, by), rng2, opts))
        tcpu = time.perf_counter() - tc
        # the timed call's first images against the same CPU sequence (scattering image after
        # noise and beam stop, r,phi and FFT images)
        sc = out["scat"][:m].cpu().numpy()
        rp = out["rphi"][:m].cpu().numpy()
        ff = out["fft"][:m].cpu().numpy()
        e_sc = e_sp = e_rp = e_ff = 0.0
        for k in range(min(m, n)):
            img = res[k][0]
            ok = np.isfinite(img) & (img != 0)
            e_sc = max(e_sc, float((np.abs(sc[k][ok] - img[ok]) / np.abs(img[ok])).max()))
            e_sp = max(e_sp, float(np.abs(sc[k][ok] - img[ok]).max() / np.abs(img[ok]).max()))
            e_rp = max(e_rp, float(np.abs(rp[k] - res[k][2]).max() / max(1e-300, np.abs(res[k][2]).max())))
            e_ff = max(e_ff, float(np.abs(ff[k] - res[k][3]).max() / max(1e-300, np.abs(res[k][3]).max())))
        check = {"images": min(m, n), "scat_max_rel": e_sc, "scat_max_rel_to_peak": e_sp,
                 "rphi_max_rel_to_peak": e_rp, "fft_max_rel_to_peak": e_ff,
                 "note": "per-pixel relative errors of the noisy image include pixels that the noise "
                         "draws close to zero",
                 "against": "oracle/aidata_oracle.post_image on the oracle image, same rand() sequence"}
        cpu = {"value": m / tcpu, "unit": "images/s", "cores": args.cpu_threads, "kind": "port",
               "sample": "%d images of the same workload: FP64 oracle image (%d threads) and the CPU "
                         "post-processing restatement, %.2f s" % (m, args.cpu_threads, tcpu)}
    line = {
        "metric": "AI training images/s (random variation + 128^2 FCC render + noise + beam stop + r,phi + FFT)",
        "value": n / dt, "unit": "images/s", "n_gpus": 1, "steps": n, "warmup": n,
        "ms_per_step": 1e3 * dt / n, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {"workload": "%d images, FCC spheres (key 0x00020500) 128x128, beam at (%d, %d), tpv %s, noise, "
                               "beam stop, r,phi 128x128, FFT modulus 128x128" % (n, bx, by, ";".join(AI_IDS))},
        "render_batch_images_per_s": n / t_batch, "render_batch_kernel_ms": k_batch_ms,
        "one_launch_per_image_images_per_s": n / t_seq, "single_vs_batched_max_rel": single_vs_batch,
        "beam": [bx, by], "oracle_check": check,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


CFG1 = dict(LType=12, ComboBoxParticle=0, Ordis=7, ComboBoxInterior=0, EditRadius=5, EditSigma=0.1,
            GridPoints=64, EditQmax=2, I0=1, Base=0, iso=0)
CFG2 = dict(LType=12, ComboBoxParticle=1, Ordis=0, ComboBoxInterior=0, EditRadius=3, Length=40,
            EditSigma=0.06, SigmaL=0.06, EditDbeta=10, theta=30, phi=45, GridPoints=256, EditQmax=1.5,
            I0=1, Base=0)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _rate(eng, out, precision, steps):
    """Back-to-back device renders of one uploaded parameter set: (patterns/s, image)."""
    import torch
    eng.doCalculationDevice(out.data_ptr(), precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.doCalculationDevice(out.data_ptr(), precision)
    torch.cuda.synchronize()
    return steps / (time.perf_counter() - t0), out.cpu().numpy()


def _err(a, b):
    r = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    return float(r.max()), float(np.sqrt(np.mean(r * r)))


def run_config1(args):
    """Config 1: isotropic Schulz spheres, no lattice, 128x128, FP64 (reference mode): rate and
    error against the reference's own image (tests/golden/cfg1_sphere.npz)."""
    import torch
    from crystalscatter_amd import SasCalc
    eng = SasCalc(device=0, params=CFG1)
    eng.prepare()
    out = torch.empty(128 * 128, dtype=torch.float64, device="cuda")
    rate, img = _rate(eng, out, "ref64", 500)
    ref = np.load(os.path.join(GOLDEN, "cfg1_sphere.npz"))["image"]
    mx, rms = _err(img, ref)
    return {"workload": "pSph_lNon_oIso (key 0x00000C07) 128x128, q in [-2,2], R 5 nm, sigma 0.1, FP64",
            "precision": "ref64", "patterns_per_s": rate, "max_rel_err_vs_reference": mx,
            "rms_rel_err_vs_reference": rms, "checked_pixels": int(ref.size),
            "reference": "the reference's own FP64 image (oracle/_ref, tests/golden/cfg1_sphere.npz)"}


def run_config2(args):
    """Config 2: Gaussian-oriented cylinders, 512x512, physics mode (converged orientation
    average) with the FP32 quadrature kernel: rate and error against an independent nested
    adaptive quadrature at 96 pixels (tests/golden/cfg2_tight.npz); the FP64 Romberg kernel's
    rate beside it."""
    import torch
    from crystalscatter_amd import SasCalc
    eng = SasCalc(device=0, params=dict(CFG2, fix_orcase=1))
    eng.prepare()
    out = torch.empty(512 * 512, dtype=torch.float64, device="cuda")
    rate, img = _rate(eng, out, "mixed", 20)
    t = np.load(os.path.join(GOLDEN, "cfg2_tight.npz"))
    x0, x1, y0, y1 = eng.bounds()
    got = img[(t["ihex"] - x0) + (x1 - x0) * (t["i"] - y0)]
    mx, rms = _err(got, t["value"])
    rate64, _ = _rate(eng, out, "ref64", 2)
    return {"workload": "pCyl_lNon_oGau (key 0x01000C00) 512x512, q in [-1.5,1.5], R 3, L 40, sigma 0.06, "
                        "Gaussian orientation 10 deg about theta 30, phi 45; physics mode",
            "precision": "mixed (FP32 quadrature 64x64 nodes, FP64 series)", "patterns_per_s": rate,
            "max_rel_err_vs_tight_quadrature": mx, "rms_rel_err_vs_tight_quadrature": rms, "checked_pixels": 96,
            "ref64_romberg_patterns_per_s": rate64,
            "reference": "scipy nested adaptive quadrature, epsrel 1e-10 (tests/golden/make_cfg2_tight.py)"}


def _host_cpu():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(threads):
    """FP64 oracle (bit-exact restatement of the reference CPU path) on the host cores, with
    the oracle/reference time ratio measured on the build machine (oracle/cpu_calibration.json)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker only: the CPU baseline leg
    o = oracle.OracleEngine(CFG3)
    o.prepare()
    t0 = time.perf_counter()
    o.image(nthreads=threads)
    dt = time.perf_counter() - t0
    cal = None
    try:
        with open(os.path.join(ROOT, "oracle", "cpu_calibration.json")) as fh:
            cal = json.load(fh)
    except OSError:
        pass
    line = {"value": 1.0 / dt, "unit": "patterns/s", "cores": threads, "kind": "port",
            "sample": "one full 1024x1024 pattern of the benchmark workload, FP64 oracle "
                      "(oracle/csx_oracle.c, bit-exact with the reference CPU path), %d threads, %.2f s" % (threads, dt),
            "host_cpus": os.cpu_count(), "host_cpu_model": _host_cpu()}
    if cal:
        ratio = cal["oracle_over_reference"]
        line["oracle_over_reference_time"] = ratio
        line["reference_estimate_value"] = ratio / dt
        line["calibration"] = "oracle time / compiled reference calc time, median over %d cases on %s (%d CPUs)" % (
            len(cal["rows"]), cal["machine"], cal["cpus"])
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "ref64"])
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("CSX_BENCH_CPU_THREADS", "16")))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="fcc", choices=["fcc", "fits", "surrogate", "bct", "aidata"],
                    help="fcc: the 1024^2 FCC pattern benchmark (default); fits: batched 2D fits; "
                         "surrogate: train the surrogate network, then fits with and without it; "
                         "bct: the BCT-cylinder kernel on random fit parameter sets")
    ap.add_argument("--nfits", type=int, default=512)
    ap.add_argument("--nimages", type=int, default=1024, help="images of the aidata workload")
    ap.add_argument("--beam", default="0,0", help="aidata workload: BeamPosX,BeamPosY in pixels")
    ap.add_argument("--maxit", type=int, default=100)
    ap.add_argument("--eager", action="store_true", help="fits: launch each round's steps one by one, not as a graph")
    ap.add_argument("--host-prep", action="store_true",
                    help="fits: build the lattice tables of every vertex on the host instead of the device")
    ap.add_argument("--maxit-surrogate", type=int, default=0,
                    help="exact simplex iterations after a surrogate start (0: the best re-scored candidate)")
    ap.add_argument("--train-patterns", type=int, default=200000)
    ap.add_argument("--train-px", type=int, default=1024, help="pixels kept per training pattern")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=65536, help="training rows per step")
    ap.add_argument("--rescore", type=int, default=8, help="surrogate candidates per fit scored exactly")
    ap.add_argument("--surrogate-px", type=int, default=256, help="pixels per fit in the surrogate objective")
    ap.add_argument("--no-configs", action="store_true",
                    help="default workload only: skip the other configurations (configs 1, 2, 4, 5)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed region, copy 256 MiB device to device once (a plain copy of "
                         "known size, for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (pixels shard across GPUs "
              "without exchange; see the sharding notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "fits":
        bench_fits(args)
        return
    if args.workload == "surrogate":
        bench_surrogate(args)
        return
    if args.workload == "aidata":
        return bench_aidata(args)
    if args.workload == "bct":
        bench_bct(args)
        return

    import ctypes
    import torch
    from crystalscatter_amd import SasCalc, _lib

    eng = SasCalc(device=0, params=CFG3)
    eng.prepare()
    L = _lib.lib()
    ctx = L.csx_engine_ctx(eng._h)
    # A dedicated torch stream (default priority) shared with the library, so that the HIP
    # events below bracket exactly the kernels of the timed steps.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    assert stream.cuda_stream != 0
    assert L.csx_set_stream(ctx, ctypes.c_void_p(stream.cuda_stream)) == 0
    out = torch.empty(1024 * 1024, dtype=torch.float64, device="cuda")
    s = eng.scalars()
    nrefl = s.peakmax2
    prec = {"mixed": 1, "ref64": 0}[args.precision]

    assert L.csx_set_timing(ctx, 0) == 0   # no library events inside the timed region
    for _ in range(args.warmup):
        eng.doCalculationDevice(out.data_ptr(), args.precision)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        eng.doCalculationDevice(out.data_ptr(), args.precision)
    ev1.record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    region_ms = ev0.elapsed_time(ev1) / args.steps   # back-to-back launches, HIP events on the stream

    # Average duration of one launch (roofline): the same K launches, each bracketed by its
    # own pair of HIP events on the launch stream.
    pairs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
             for _ in range(args.steps)]
    for a, b in pairs:
        a.record(stream)
        eng.doCalculationDevice(out.data_ptr(), args.precision)
        b.record(stream)
    torch.cuda.synchronize()
    kernel_ms = sum(a.elapsed_time(b) for a, b in pairs) / args.steps

    # Average duration of the dominant kernel alone, the per-pixel image pass (for this
    # workload the hot pass of the two-pass lattice render): HIP events that the library
    # records on its launch stream around that launch, over back-to-back renders.
    nk = min(args.steps, _lib.PASS_RING)
    assert L.csx_set_timing(ctx, 1) == 0
    for _ in range(nk):
        eng.doCalculationDevice(out.data_ptr(), args.precision)
    pass_arr = (ctypes.c_double * nk)()
    got = L.csx_pass_times(ctx, nk, pass_arr)
    assert L.csx_set_timing(ctx, 0) == 0
    assert got == nk, got
    pass_ms = sum(pass_arr) / nk

    if args.copy_calibration:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # Work counters of one launch (separate, untimed run of the same kernel).
    stats = (ctypes.c_uint64 * 5)()
    g = _lib.Grid(s.xmin, s.xmax, s.ymin, s.ymax)
    assert L.csx_render_stats(ctx, ctypes.byref(g), prec, ctypes.c_void_p(out.data_ptr()), stats) == 0
    pair_evals, guard_px, near_one_px, two_waves, full_px = (int(v) for v in stats)

    img = out.cpu().numpy()
    # Every pixel against the GPU's FP64 reference mode (itself within 1e-12 of the
    # reference's own pixels), its rate, and the end-to-end rate of doCalculation:
    # preparation + upload + render + copy of the image to host memory.
    full_max = full_rms = rate64 = e2e = None
    if args.precision == "mixed":
        out64 = torch.empty_like(out)
        rate64, ref64 = _rate(eng, out64, "ref64", 5)
        full_max, full_rms = _err(img, ref64)
        del out64
    t0 = time.perf_counter()
    n_e2e = 10
    for _ in range(n_e2e):
        eng.prepare()
        eng.doCalculation(args.precision)
    e2e = n_e2e / (time.perf_counter() - t0)
    max_rel = rms_rel = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        o = oracle.OracleEngine(CFG3)
        o.prepare()
        rng = np.random.default_rng(0)
        idx = rng.choice(img.size, 4096, replace=False)
        ref = o.image_points(idx % 1024 - 512, idx // 1024 - 512, nthreads=args.cpu_threads)
        rel = np.abs(img[idx] - ref) / np.abs(ref)
        max_rel, rms_rel = float(rel.max()), float(np.sqrt((rel * rel).mean()))
    except Exception as exc:  # the check is reported, never silently passed
        max_rel = "unavailable: %s" % exc

    px = 1024 * 1024
    sec = pass_ms * 1e-3
    alg_flop = px * (nrefl * ALG_FLOP_PER_REFL + ALG_FLOP_PER_PX_EXTRA)
    alg_t = px * (nrefl * ALG_T_PER_REFL + ALG_T_PER_PX_EXTRA)
    exec_tflops = pair_evals * EXEC_FLOP_PER_PAIR_EVAL / sec / 1e12
    exec_tops = pair_evals * EXEC_T_PER_PAIR_EVAL / sec / 1e12
    achieved = exec_tflops
    line = {
        "metric": METRIC,
        "value": args.steps / wall,
        "unit": "patterns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "FCC spheres pSph_pGau_lFCC_oGau (key 0x00020500), 1024x1024 px, q in [-1,1] nm^-1, "
                               "R=8 nm sigma=0.08, a=30 nm, HKLmax=5 (%d reflections), domain 200 nm, DW 1.5 nm, "
                               "azi 5 deg" % nrefl,
                   "precision": args.precision, "global_batch": 1, "pixels": px},
        "max_rel_err_vs_fp64": max_rel,
        "rms_rel_err_vs_fp64": rms_rel,
        "error_check": "max/rms: 4096 random pixels against the CPU oracle (bit-exact restatement of the "
                       "reference); full_pattern_*: all 1048576 pixels against the GPU FP64 reference mode",
        "full_pattern_max_rel_err_vs_fp64": full_max,
        "full_pattern_rms_rel_err_vs_fp64": full_rms,
        "ref64_patterns_per_s": rate64,
        "end_to_end_patterns_per_s": e2e,
        "end_to_end": "prepareCalculation + table upload + render + 8 MiB image copied to host memory, per pattern",
        "kernel_ms": pass_ms,
        "render_ms_per_launch": kernel_ms,
        "region_ms_per_launch": region_ms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": measured_traffic(),
                     "kernel": "k_image_lat_hot<K_FCC, CSX_PREC_MIXED> (the render's per-pixel pass; "
                               "render_ms_per_launch adds the radial pre-pass k_radial and the fix pass "
                               "k_image_lat_fix)",
                     "pair_evals_per_launch": pair_evals, "flop_per_pair_eval": EXEC_FLOP_PER_PAIR_EVAL,
                     "pair_evals_uncut": px * ((nrefl + 1) // 2),
                     "transcendental_rate_tops": exec_tops, "transcendental_peak_tops": PEAK_TRANS_TOPS,
                     "reference_equivalent_tflops": alg_flop / sec / 1e12,
                     "reference_equivalent_transcendental_tops": alg_t / sec / 1e12,
                     "fp64_window_pixels": guard_px, "fp64_full_pixels": full_px, "near_one_pixels": near_one_px,
                     "two_term_waves": two_waves},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_threads)
    if not args.no_configs:
        line["configs"] = run_other_configs(args)
    print(json.dumps(line), flush=True)


def run_other_configs(args):
    """Configs 1, 2, 4 and 5 beside the headline (config 3); a failing one reports its error."""
    configs = {}
    for name, fn in (("config1", run_config1), ("config2", run_config2)):
        t0 = time.perf_counter()
        try:
            configs[name] = fn(args)
        except Exception as exc:  # reported, never silently dropped
            configs[name] = {"error": repr(exc)}
        configs[name]["wall_s"] = time.perf_counter() - t0
        print("[bench] %s done in %.1f s" % (name, configs[name]["wall_s"]), file=sys.stderr, flush=True)
    probs = None
    try:
        t0 = time.perf_counter()
        probs = fit_problems(args.nfits)
        t_gen = time.perf_counter() - t0
        configs["config4"] = run_fits(args, probs)
        configs["config4"]["target_generation_s"] = t_gen
    except Exception as exc:
        configs["config4"] = {"error": repr(exc)}
    print("[bench] config4 done", file=sys.stderr, flush=True)
    try:
        configs["config5"] = run_surrogate(args, probs)
        configs["config5"]["training_size_note"] = (
            "%d training patterns of 256x256 rendered by the exact kernel (the configuration names 200k), "
            "%d random pixels kept per pattern; 3000 Adam steps of %d rows see each kept pixel about once"
            % (args.train_patterns, args.train_px, args.batch))
    except Exception as exc:
        configs["config5"] = {"error": repr(exc)}
    print("[bench] config5 done", file=sys.stderr, flush=True)
    return configs


if __name__ == "__main__":
    main()
