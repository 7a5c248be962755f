"""Autotuner of the batched objective's kernel variants (csx_tune_*, include/csx.h).

For each lattice combination the batched fits use (BCT cylinders 0x01020806, FCC spheres
0x00020500) and each detector size class (128^2, 256^2, 384^2, 512^2), it stages a batch of
parameter sets drawn like the fit workload, runs csx_fit_batch_run with every variant
forced (the same code with different register budgets), checks that all variants give the
same sums bit for bit, keeps the fastest (median of the kernel times over repeats), and
writes crystalscatter_amd/tuning.txt, which the library reads at its first csx_create.

    python -m crystalscatter_amd.autotune [--sets 96] [--repeats 5] [--out PATH]
"""
import argparse
import ctypes
import datetime
import os
import sys

import numpy as np

from . import Context, SasCalc, _lib
from . import surrogate as sg

HERE = os.path.dirname(os.path.abspath(__file__))
SIZES = [128, 256, 384, 512]
KINDS = {
    0x01020806: "BCT cylinders, z-direction (batched fits)",
    0x00020500: "FCC spheres, Gaussian orientation",
}
FCC = dict(LType=5, ComboBoxParticle=0, Ordis=0, ComboBoxPeak=1, ComboBoxInterior=0, EditSigma=0.08,
           EditDomainSize=200, RadButDebyeScherrer=1, EditDebyeWaller=1.5, EditAzi=5, HKLmax=5, EditQmax=1,
           I0=1, Base=0)


def _sets(key, n, count, rng):
    if key == 0x01020806:
        return [sg.params_dict(x, n // 2) for x in sg.draw_parameters(rng, count)]
    out = []
    for _ in range(count):
        a = rng.uniform(24, 36)
        out.append(dict(FCC, EditRadius=rng.uniform(6, 9), uca=a, ucb=a, ucc=a, GridPoints=n // 2))
    return out


def _stage(ctx, params, n, data):
    L = _lib.lib()
    mask = _lib.Mask()
    for i in range(16):
        mask.rects[i] = -1
    mask.bstop, mask.border = 4, 0
    keep = []
    ctx._chk(L.csx_fit_batch_begin(ctx._h), "csx_fit_batch_begin")
    for p in params:
        h = SasCalc(device=-1, params=p)
        h.set("ignore_cbs", 1)
        h.prepare()
        carr = np.concatenate([h.table("carr%d%s" % (i, t)) for t in ("p", "f") for i in range(1, 10)])
        l1, l2, l3 = h.table("latpar1"), h.table("latpar2"), h.table("latpar3")
        keep += [carr, l1, l2, l3, h]
        ctx._chk(L.csx_fit_batch_stage(ctx._h, ctypes.byref(h.scalars()),
                                       carr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       l1.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                       l2.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                       l3.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n, n, ctypes.byref(mask),
                                       ctypes.c_void_p(data.data_ptr())), "csx_fit_batch_stage")
    return keep


def measure(key, n, count, repeats, seed=1):
    """Median kernel ms of every variant for `count` sets of key at n x n; the sums must agree."""
    import torch
    L = _lib.lib()
    rng = np.random.default_rng(seed)
    params = _sets(key, n, count, rng)
    data = torch.tensor(rng.uniform(0.01, 10.0, n * n), dtype=torch.float64, device="cuda")
    ctx = Context(0)
    times, ref = [], None
    for v in range(_lib.TUNE_VARIANTS):
        ctx._chk(L.csx_tune_force(ctx._h, v), "csx_tune_force")
        ms = []
        for r in range(repeats + 1):
            keep = _stage(ctx, params, n, data)
            sums = np.zeros(count)
            cnts = np.zeros(count, dtype=np.int64)
            ctx._chk(L.csx_fit_batch_run(ctx._h, 1, sums.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                         cnts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))), "csx_fit_batch_run")
            if r > 0:  # the first run of a variant loads its kernel
                ms.append(L.csx_last_kernel_ms(ctx._h))
            del keep
            if ref is None:
                ref = sums.copy()
            elif not np.array_equal(sums, ref):
                raise RuntimeError("variant %d changed the sums of key 0x%08x at %d^2" % (v, key, n))
        times.append(float(np.median(ms)))
    ctx._chk(L.csx_tune_force(ctx._h, -1), "csx_tune_force")
    return times


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sets", type=int, default=96, help="parameter sets per batch")
    ap.add_argument("--repeats", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.txt"))
    args = ap.parse_args(argv)
    import torch
    dev = torch.cuda.get_device_name(0)
    lines = ["# Kernel variants of the batched objective (csx_fit_batch_run), chosen by",
             "# crystalscatter_amd/autotune.py on %s, %s." % (dev, datetime.date.today().isoformat()),
             "# Variant 0: budget for 4 waves/SIMD; 1: uncapped; 2: 5 waves; 3: 6 waves (same sums).",
             "# %d parameter sets per batch, median kernel ms of %d runs per variant." % (args.sets, args.repeats),
             "# key        class variant   # detector: ms of variants 0..%d" % (_lib.TUNE_VARIANTS - 1)]
    for key, name in KINDS.items():
        lines.append("# %s" % name)
        for sc, n in enumerate(SIZES):
            t = measure(key, n, args.sets, args.repeats)
            best = int(np.argmin(t))
            lines.append("0x%08x %d %d   # %dx%d: %s" % (key, sc, best, n, n, " ".join("%.3f" % x for x in t)))
            print(lines[-1], flush=True)
    with open(args.out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print("wrote", args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
