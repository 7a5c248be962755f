"""Write tests/golden/cfg2_tight.npz: the Gaussian-cylinder orientation average of config 2
(physics mode, 512x512, q on [-1.5, 1.5] nm^-1) at 96 pixels across the detector, by nested
adaptive scipy quadrature (epsrel 1e-10 outside, 1e-11 inside), independent of the kernels
and of the oracle's Romberg.  The transverse factor is the oracle's z-oriented cylinder at
the same q (its pql is 1).  The value at a pixel is their product, the quantity the FP32
quadrature kernel evaluates.

    make -C oracle && python tests/golden/make_cfg2_tight.py
"""
import math
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from helpers import CFG2  # noqa: E402

P = dict(CFG2, fix_orcase=1)


def orientation_average(qx, qy, p=P):
    """(1/norm) int_0^dmax dd sin d e^(-d^2/db^2) (1/2pi) int_0^2pi dc P_rod(arg(d, c)), with the
    rod factor of the Schulz-averaged length and the rotation entries the reference stores as
    float (Source/sc_calc_generic_gpu.cu:1402, :1418-1432, :1634-1652)."""
    from scipy import integrate
    sl, L, db, th = p["SigmaL"], p["Length"], p["EditDbeta"] * np.pi / 180, p["theta"] * np.pi / 180
    zl = (1 - sl * sl) / (sl * sl)
    dmax = float(np.float32(min(db * math.sqrt(math.log(1e5)), np.pi / 2)))
    p11, p13, p21 = float(np.float32(-math.cos(th))), float(np.float32(math.sin(th))), -1.0

    def rod(arg):
        a1 = 1 / (2 * zl * (zl - 1))
        return (a1 / arg ** 2) * (1 - np.cos((zl - 1) * np.arctan(2 * arg)) / (1 + 4 * arg * arg) ** ((zl - 1) / 2))

    def inner(d):
        f = lambda c: rod((qx * (p11 * np.cos(c) * np.sin(d) + p13 * np.cos(d)) + qy * (p21 * np.sin(c) * np.sin(d))
                           + 1e-10) * L / (zl + 1))
        return integrate.quad(f, 0, 2 * np.pi, limit=800, epsrel=1e-11)[0] / (2 * np.pi)

    w = lambda d: np.sin(d) * np.exp(-d * d / db ** 2)
    norm = integrate.quad(w, 0, dmax, epsrel=1e-13)[0]
    return integrate.quad(lambda d: w(d) * inner(d), 0, dmax, limit=400, epsrel=1e-10)[0] / norm


def _one(args):
    ih, ii = args
    import oracle
    o = oracle.OracleEngine(P)
    o.prepare()
    ot = oracle.OracleEngine(dict(P, Ordis=6, theta=0, phi=0))
    ot.prepare()
    qx, qy, qz = o.idx2q(ih, ii)
    return orientation_average(qx, qy) * ot.pixel(qx, qy, qz)


def main():
    rng = np.random.default_rng(20261203)
    pts = [(255, 255), (-256, 100), (40, -200), (0, 255), (120, 7), (-30, -40)]
    pts += [tuple(int(v) for v in rng.integers(-256, 256, 2)) for _ in range(90)]
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        vals = pool.map(_one, pts)
    np.savez_compressed(os.path.join(HERE, "cfg2_tight.npz"), ihex=np.array([p[0] for p in pts], np.int32),
                        i=np.array([p[1] for p in pts], np.int32), value=np.array(vals))
    print("wrote", len(pts), "pixels")


if __name__ == "__main__":
    main()
