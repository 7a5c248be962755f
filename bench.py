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

# Roofline accounting of the dominant kernel (the mixed-precision image kernel).  Its hot
# loop evaluates Friedel pairs: one pair evaluation covers reflections +hkl and -hkl for one
# pixel.  FP32 operations per pair evaluation in the in-plane single-term loop, FMA = 2:
# q^.u 3, |q^ x u|^2 5, acos^2/sin^2 polynomial 18, y^2 1, radial exponent 4, angular
# exponent 2, exp2 1, weighting 2 = 36 FLOP (+ a select).  The pair count per launch comes
# from the kernel's own counters (csx_render_stats): pairs outside a wave's |q| and angle
# window are never evaluated and are not counted.  The per-pixel FP64 form factors and
# the rare FP64 re-sums are left out of the count (it is a lower bound).
FLOP_PER_PAIR_EVAL = 36
# The reference evaluates every reflection at every pixel: per reflection ~29 FLOP + 3
# libm calls (Source/sc_gpu_pSph_pGau_lFCC_oGau.h:85-142).  Reported beside the roofline as
# the reference-equivalent rate.
REF_FLOP_PER_REFL = 29
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector, dense (packed dual-issue counted)


def cpu_baseline(threads):
    """FP64 oracle (bit-exact restatement of the reference CPU path) on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker only: the CPU baseline leg
    o = oracle.OracleEngine(CFG3)
    o.prepare()
    t0 = time.perf_counter()
    o.image(nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "patterns/s", "cores": threads, "kind": "port",
            "sample": "one full 1024x1024 pattern of the benchmark workload, FP64 oracle "
                      "(oracle/csx_oracle.c, bit-exact with the reference CPU path), %d threads, %.2f s" % (threads, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "ref64"])
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("CSX_BENCH_CPU_THREADS", "16")))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (pixels shard across GPUs "
              "without exchange; see the sharding notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

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
    kernel_ms = ev0.elapsed_time(ev1) / args.steps   # HIP events on the launch stream
    last_launch_ms = eng.last_kernel_ms()

    # Work counters of one launch (separate, untimed run of the same kernel).
    stats = (ctypes.c_uint64 * 4)()
    g = _lib.Grid(s.xmin, s.xmax, s.ymin, s.ymax)
    assert L.csx_render_stats(ctx, ctypes.byref(g), prec, ctypes.c_void_p(out.data_ptr()), stats) == 0
    pair_evals, guard_px, near_one_px, two_waves = (int(v) for v in stats)

    img = out.cpu().numpy()
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
    flops = pair_evals * FLOP_PER_PAIR_EVAL
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    ref_equiv = px * nrefl * REF_FLOP_PER_REFL / (kernel_ms * 1e-3) / 1e12
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
        "kernel_ms": kernel_ms,
        "last_launch_ms": last_launch_ms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "k_image<K_FCC, CSX_PREC_MIXED>",
                     "pair_evals_per_launch": pair_evals, "flop_per_pair_eval": FLOP_PER_PAIR_EVAL,
                     "fp64_guard_pixels": guard_px, "near_one_pixels": near_one_px,
                     "two_term_waves": two_waves,
                     "pair_evals_uncut": px * ((nrefl + 1) // 2),
                     "reference_equivalent_tflops": ref_equiv},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
