"""Write the golden fixtures of tests/golden/ from the reference's own engine.

Runs oracle/_ref/ref_driver (the reference's sc_calc_generic_cpu.cpp and
sc_calc_generic_gpu.cu compiled unchanged by `make -C oracle ref`) on the
parameter sets below and stores inputs and outputs as .npz data, plus the
reference's only golden image (Documentation/Vitess-Sphere.dat) decoded.
Needs the reference checkout at /root/reference; the fixtures themselves are
self-contained and are what the tests read.

    make -C oracle ref && python tests/golden/make_golden.py          # all fixtures
    python tests/golden/make_golden.py cube tpv post lattices peaks lattices3d fits qmodes  # only these
"""
import json
import os
import platform
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
from refbin import read_refbin  # noqa: E402

REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
REF_FIT = os.path.join(ROOT, "oracle", "_ref", "ref_fit")
REF_DOC = "/root/reference/Documentation"

# Parameter sets (reference key names).  Config numbers follow the benchmark brief.
CFG1 = dict(LType=12, ComboBoxParticle=0, Ordis=7, ComboBoxInterior=0, EditRadius=5, EditSigma=0.1,
            GridPoints=64, EditQmax=2, I0=1, Base=0, iso=0)
CFG2 = dict(LType=12, ComboBoxParticle=1, Ordis=0, ComboBoxInterior=0, EditRadius=3, Length=40,
            EditSigma=0.06, SigmaL=0.06, EditDbeta=10, theta=30, phi=45, GridPoints=16, EditQmax=1.5,
            I0=1, Base=0)
CFG3 = dict(LType=5, ComboBoxParticle=0, Ordis=0, ComboBoxPeak=1, ComboBoxInterior=0, EditRadius=8,
            EditSigma=0.08, uca=30, ucb=30, ucc=30, EditDomainSize=200, RadButDebyeScherrer=1,
            EditDebyeWaller=1.5, EditAzi=5, HKLmax=5, GridPoints=512, EditQmax=1, I0=1, Base=0)
CFG4_BASE = dict(LType=8, ComboBoxParticle=1, Ordis=6, ComboBoxPeak=1, ComboBoxInterior=0, theta=0,
                 phi=0, EditSigma=0.06, SigmaL=0.06, RadButDebyeScherrer=1, EditQmax=1, I0=1, Base=0,
                 HKLmax=5)
# Isotropic homogeneous cubes (key 0x04000C07): two edge/width combinations, the first with
# the series/Porod switch inside the image, the second with iso and ifluc terms.
CUBE_A = dict(LType=12, ComboBoxParticle=4, Ordis=7, ComboBoxInterior=0, EditRadius=10,
              EditSigma=0.1, GridPoints=32, EditQmax=2, I0=1, Base=0, iso=0)
CUBE_B = dict(LType=12, ComboBoxParticle=4, Ordis=7, ComboBoxInterior=0, EditRadius=4,
              EditSigma=0.25, GridPoints=24, EditQmax=3, I0=2, Base=0.5, iso=0.25, ifluc=3,
              rfluc=2)
VITESS = dict(LType=12, ComboBoxParticle=0, Ordis=7, EditRadius=16.7, EditSigma=0.0639, EditQmax=2,
              GridPoints=128, I0=1, Base=0, CenterBeam=1, EditPixelX=2.5, EditPixelY=2.5,
              EditPixelNoX=256, EditPixelNoY=256, EditDet=10, EditWavelength=0.154)


def cfg4_params(rng):
    """One config-4 parameter draw: R, L, a, c/a, domain size, angular peak width (EditAzi)."""
    R = rng.uniform(2, 6)
    L = rng.uniform(10, 60)
    a = rng.uniform(15, 40)
    ca = rng.uniform(1, 2)
    dom = rng.uniform(100, 400)
    azi = rng.uniform(2.0, 10.0)
    p = dict(CFG4_BASE)
    p.update(EditRadius=R, Length=L, uca=a, ucb=a, ucc=a * ca, EditDomainSize=dom, EditAzi=azi)
    return p


def run_ref(params, extra=None, driver=REF_DRIVER):
    with tempfile.TemporaryDirectory() as td:
        pf = os.path.join(td, "p.txt")
        with open(pf, "w") as fh:
            for k, v in list(params.items()) + list((extra or {}).items()):
                fh.write("%s=%r\n" % (k, v) if not isinstance(v, str) else "%s=%s\n" % (k, v))
        out = os.path.join(td, "o.bin")
        subprocess.run([driver, pf, out], check=True, capture_output=True)
        return read_refbin(out)


TABLES = ["carr1p", "carr2p", "carr4p", "carr1f", "carr4f", "limq"]


def table_dict(r, with_lattice):
    d = {t: r[t] for t in TABLES}
    for s in ("norm", "orcase", "ucvol", "cosphic", "sinphic", "polPhi"):
        d[s] = r[s]
    if with_lattice:
        d["latpar1"] = r["latpar1"]
        d["latpar2"] = r["latpar2"]
        # columns 1..13 are the ones the reference fills (14..16 are never written)
        d["latpar3"] = r["latpar3"].reshape(-1, 17)[:, :14].copy()
        d["peakmax"] = np.array([r["peakmax1"][0], r["peakmax2"][0]])
    return d


def write_cube(manifest):
    """Cube fixtures: images, the carr4p/carr4f tables, limq and the Porod constant."""
    for name, p in (("cube_a.npz", CUBE_A), ("cube_b.npz", CUBE_B)):
        r = run_ref(p)
        np.savez_compressed(os.path.join(HERE, name), image=r["image"], bounds=r["bounds"],
                            por=r["por"], **table_dict(r, False))
        manifest["fixtures"][name] = {"params": p}


# Spheres with Gaussian peaks on BCC and SC lattices (the reference's generic kernel), and FCC
# with isotropic orientation: 64x64 images with their tables.
LATTICES = [("bcc_sphere", dict(CFG3, LType=4, uca=30, ucb=30, ucc=30, GridPoints=32)),
            ("sc_sphere_iso", dict(CFG3, LType=7, uca=25, ucb=25, ucc=25, Ordis=7, GridPoints=32)),
            ("fcc_sphere_iso", dict(CFG3, Ordis=7, GridPoints=32)),
            ("bct_sphere", dict(CFG3, LType=8, uca=30, ucb=30, ucc=45, GridPoints=32)),
            ("sphere_zdir", dict(CFG1, Ordis=6, GridPoints=32)),
            ("cyl_iso", dict(CFG2, Ordis=7, EditRadius=3, Length=40, GridPoints=32, EditQmax=2)),
            ("cyl_iso_b", dict(CFG2, Ordis=7, EditRadius=5, Length=12, EditSigma=0.1, SigmaL=0.1, GridPoints=32,
                               EditQmax=3, iso=0.2, I0=3, Base=0.1)),
            ("cube_zdir", dict(CUBE_A, Ordis=6)),
            ("cube_zdir_b", dict(CUBE_B, Ordis=6))]


def write_lattices(manifest):
    for name, p in LATTICES:
        r = run_ref(p)
        extra = {}
        if int(p.get("ComboBoxParticle", 0)) == 4 and int(p.get("Ordis", 7)) == 6:
            extra["carr11pm"] = r["carr11pm"].reshape(131, 131)[:121, :121].copy()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), image=r["image"], bounds=r["bounds"],
                            **table_dict(r, True), **extra)
        manifest["fixtures"][name + ".npz"] = {"params": p}


# Spheres on lattices with the other peak shapes of the generic kernel (Source/sc_gpu_generic.h:
# 695-834): Lorentzian (ComboBoxPeak 0), mod. 1 and mod. 2 Lorentzian (2, 3), Pseudo-Voigt (4,
# EditPeakPar = 100 eta) and Pearson VII (5, EditPeakPar = beta; its peak norm is the
# reference's trapezoid iteration), on the four cubic/tetragonal lattices, both orientations.
PEAKS = [("fcc_sphere_lorentz", dict(CFG3, ComboBoxPeak=0, GridPoints=32)),
         ("bcc_sphere_mod1lor", dict(CFG3, LType=4, ComboBoxPeak=2, GridPoints=32)),
         ("sc_sphere_mod2lor_iso", dict(CFG3, LType=7, uca=25, ucb=25, ucc=25, Ordis=7, ComboBoxPeak=3,
                                        GridPoints=32)),
         ("fcc_sphere_pvoigt", dict(CFG3, ComboBoxPeak=4, EditPeakPar=30, GridPoints=32)),
         ("bct_sphere_pearson", dict(CFG3, LType=8, uca=30, ucb=30, ucc=45, ComboBoxPeak=5, EditPeakPar=1.5,
                                     GridPoints=32)),
         ("fcc_sphere_pearson_iso", dict(CFG3, Ordis=7, ComboBoxPeak=5, EditPeakPar=3, EditAzi=20,
                                         GridPoints=32, iso=0.3))]


def write_peaks(manifest):
    for name, p in PEAKS:
        r = run_ref(p)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), image=r["image"], bounds=r["bounds"],
                            **table_dict(r, True))
        manifest["fixtures"][name + ".npz"] = {"params": p}


# Spheres on the other 3D lattices of ButtonHKLClick: Ia3d (cubic group Ia-d), Pn3m (Pn--),
# Im3m (I---) and Fd3m (Fd--, the 8-atom diamond basis), Gaussian and isotropic orientation.
# The reference's parameter check fixes vesicles on Ia3d, Pn3m and Im3m; the driver's calculation
# ignores that rule (bIgnoreCbs), as the fits do.
CUBICS = [("ia3d_sphere", dict(CFG3, LType=9, uca=40, ucb=40, ucc=40, EditRadius=6, GridPoints=32)),
          ("pn3m_sphere_iso", dict(CFG3, LType=10, uca=25, ucb=25, ucc=25, EditRadius=5, Ordis=7, GridPoints=32)),
          ("im3m_sphere_lorentz", dict(CFG3, LType=11, uca=28, ucb=28, ucc=28, ComboBoxPeak=0, GridPoints=32)),
          ("fd3m_sphere", dict(CFG3, LType=17, uca=45, ucb=45, ucc=45, EditRadius=7, GridPoints=32)),
          ("fd3m_sphere_pvoigt_iso", dict(CFG3, LType=17, uca=45, ucb=45, ucc=45, EditRadius=7, Ordis=7,
                                          ComboBoxPeak=4, EditPeakPar=60, HKLmax=7, GridPoints=32)),
          # no reflection: HKLmax 0, and HCP and the orthorhombic lattice, whose systems' condition
          # blocks the reference has commented out (ButtonHKLClick accepts nothing there)
          ("fcc_sphere_hkl0", dict(CFG3, HKLmax=0, GridPoints=32)),
          ("hcp_sphere", dict(CFG3, LType=6, uca=30, ucb=30, ucc=49, ucgamma=120, GridPoints=32)),
          ("ortho_sphere_iso", dict(CFG3, LType=18, uca=30, ucb=35, ucc=40, Ordis=7, ComboBoxPeak=3,
                                    GridPoints=32)),
          # spheres under other orientation distributions (they enter neither the sphere's
          # factors nor the peak sums), spheres on a 2D lattice
          ("fcc_sphere_ordis1", dict(CFG3, Ordis=1, GridPoints=32)),
          ("bcc_sphere_zdir_lorentz", dict(CFG3, LType=4, Ordis=6, ComboBoxPeak=0, GridPoints=32)),
          ("hex_sphere", dict(CFG3, LType=1, uca=30, ucb=30, ucgamma=120, GridPoints=32)),
          # perfectly oriented cylinders with Gaussian peaks on the 2D lattices (hexagonal, square,
          # centred rectangular: the cell's plane basis, l in [-H, 0], multiplicity 1) and on FCC
          ("hex_cyl_zdir", dict(CFG4_BASE, LType=1, EditRadius=4, Length=40, uca=20, ucb=20, ucc=20, ucgamma=120,
                                EditDomainSize=250, EditAzi=5, GridPoints=32)),
          ("sq_cyl_zdir", dict(CFG4_BASE, LType=2, EditRadius=3, Length=30, uca=18, ucb=18, ucc=18, theta=90,
                               EditDomainSize=200, EditAzi=8, GridPoints=32)),
          ("rect_cyl_zdir", dict(CFG4_BASE, LType=3, EditRadius=3, Length=50, uca=16, ucb=26, ucc=20,
                                 EditDomainSize=300, EditAzi=4, HKLmax=4, GridPoints=32)),
          ("fcc_cyl_zdir", dict(CFG4_BASE, LType=5, EditRadius=4, Length=30, uca=30, ucb=30, ucc=30,
                                EditDomainSize=250, EditAzi=5, GridPoints=32)),
          # isotropic cylinders (L/R >= 2) with Gaussian peaks on FCC and on the hexagonal 2D lattice
          ("fcc_cyl_iso", dict(CFG4_BASE, LType=5, Ordis=7, EditRadius=3, Length=30, uca=30, ucb=30, ucc=30,
                               EditDomainSize=250, EditAzi=5, GridPoints=32)),
          ("hex_cyl_iso", dict(CFG4_BASE, LType=1, Ordis=7, EditRadius=4, Length=40, uca=20, ucb=20, ucc=20,
                               ucgamma=120, EditDomainSize=200, EditAzi=6, iso=0.2, GridPoints=32)),
          # isotropic homogeneous cubes with Gaussian peaks (calc_GENERIC sets fq = pq)
          ("sc_cube_iso", dict(CFG3, LType=7, ComboBoxParticle=4, Ordis=7, EditRadius=10, EditSigma=0.1, uca=25,
                               ucb=25, ucc=25, GridPoints=32)),
          ("bcc_cube_iso", dict(CFG3, LType=4, ComboBoxParticle=4, Ordis=7, EditRadius=8, EditSigma=0.15, uca=30,
                                ucb=30, ucc=30, EditQmax=2, iso=0.3, GridPoints=32)),
          # the other peak shapes with cylinders and cubes (FP64 sums)
          ("bct_cyl_zdir_lorentz", dict(CFG4_BASE, ComboBoxPeak=0, EditRadius=4, Length=30, uca=25, ucb=25, ucc=40,
                                        EditDomainSize=250, EditAzi=5, GridPoints=32)),
          ("fcc_cyl_iso_pearson", dict(CFG4_BASE, LType=5, Ordis=7, ComboBoxPeak=5, EditPeakPar=2, EditRadius=3,
                                       Length=30, uca=30, ucb=30, ucc=30, EditDomainSize=250, EditAzi=5,
                                       GridPoints=32)),
          ("sc_cube_iso_pvoigt", dict(CFG3, LType=7, ComboBoxParticle=4, Ordis=7, ComboBoxPeak=4, EditPeakPar=40,
                                      EditRadius=10, EditSigma=0.1, uca=25, ucb=25, ucc=25, GridPoints=32))]


def write_cubics(manifest):
    for name, p in CUBICS:
        r = run_ref(p)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), image=r["image"], bounds=r["bounds"],
                            **table_dict(r, True))
        manifest["fixtures"][name + ".npz"] = {"params": p}


# Random variation (tpvPerformRandom) before each of four FCC images, rand() from seed 1.
TPV_IDS = ["EditRadius=0.1", "EditSigma=0.2", "uca=0.05", "ucb=1", "ucc=1", "I0=0.3", "Base=0.1",
           "EditDomainSize=0.4", "EditDebyeWaller=0.5", "EditAzi=0.3", "phi=1"]
TPV_N = 4


def write_tpv(manifest):
    p = dict(CFG3, GridPoints=16)
    r = run_ref(p, {"tpv": "|".join(TPV_IDS), "tpvn": TPV_N})
    rep = r["tpv_report"]
    rep = rep if isinstance(rep, str) else bytes(rep).decode()
    np.savez_compressed(os.path.join(HERE, "tpv_fcc.npz"), images=r["tpv_images"].reshape(TPV_N, -1),
                        bounds=r["bounds"], report=np.frombuffer(rep.encode(), np.uint8))
    manifest["fixtures"]["tpv_fcc.npz"] = {"params": p, "tpv": TPV_IDS, "n": TPV_N, "seed": 1}


# Post-processing: r,phi resampling (with its log scaling) and scaleIntensity of FCC and
# sphere images, from the reference's own routines.
POST_CASES = [("fcc_rphi", dict(CFG3, GridPoints=32), {"rphi": 128, "bsx": 0, "bsy": 0, "scaleint": 0,
                                                        "fft": 2, "fftswap": 1, "fftscale": 1}),
              ("fcc_rphi_off", dict(CFG3, GridPoints=32), {"rphi": 128, "bsx": 10, "bsy": -5, "rphiscale": 1,
                                                            "fft": 3, "fftclip": 1, "fftswap": 1, "fftcutx": 64,
                                                            "fftcuty": 32}),
              ("fcc_rphi_clip", dict(CFG3, GridPoints=24), {"rphi": 128, "bsx": 3, "bsy": 4, "rphiclip": 1,
                                                             "rphiclip40": 1, "fft": 1, "fftforward": 1}),
              ("sphere_lin", dict(CFG1, GridPoints=32), {"scaleint": 1})]


# q modes and image layouts: the 1D mode (use1d: EditQsteps pixels, qx = (qmax - qmin)/zmax * ihex)
# and ExpandImage from the first quadrant and from the right half, on an anisotropic pattern
# (FCC spheres, so that the mirror fill is not hidden by a symmetric image).
QMODE_FCC = dict(CFG3, GridPoints=16, EditQmax=1.2)


def write_qmodes(manifest):
    arrays, meta = {}, {}
    cases = {
        "q1_expand": dict(QMODE_FCC, RadioButtonQ1=1, RadioButtonQ4=0, ExpandImage=1),
        "q1_block": dict(QMODE_FCC, RadioButtonQ1=1, RadioButtonQ4=0, ExpandImage=0),
        "q2_expand": dict(QMODE_FCC, RadioButtonQ2=1, RadioButtonQ4=0, ExpandImage=1),
        "q2_block": dict(QMODE_FCC, RadioButtonQ2=1, RadioButtonQ4=0, ExpandImage=0),
        "d1_sphere": dict(CFG1, use1d=1, EditQsteps=100, EditQmin=0.1, EditQmax=2),
        "d1_fcc": dict(QMODE_FCC, use1d=1, EditQsteps=160, EditQmin=0.0, EditQmax=1.2),
    }
    for name, p in cases.items():
        r = run_ref(p)
        arrays[name] = r["image"]
        arrays[name + "_bounds"] = r["bounds"]
        meta[name] = p
    np.savez_compressed(os.path.join(HERE, "qmodes.npz"), **arrays)
    manifest["fixtures"]["qmodes.npz"] = {"cases": meta}


# Reference 2D simplex fits (oracle/_ref/ref_fit: doSimplexFit2D over the reference engine).
# Each case: the parameters at the start, the hash entries (name, used, min, max), the simplex
# arguments and the true parameters that made the target image.  Bounds are the key table's
# (Source/sc_calc_generic.cpp:36-138; -10000..10000 where it gives none, as
# SC_CalcCons::limitsOfParamValue reads them) unless a case narrows them on purpose.
def _fit_cases():
    rng = np.random.default_rng(1234)
    bct_true = cfg4_params(rng)
    bct_true["GridPoints"] = 16
    bct_start = dict(bct_true, EditRadius=bct_true["EditRadius"] * 1.04, Length=bct_true["Length"] * 0.95,
                     uca=bct_true["uca"] * 1.01, ucb=bct_true["uca"] * 1.01, ucc=bct_true["ucc"] * 0.99,
                     EditDomainSize=bct_true["EditDomainSize"] * 1.1, EditAzi=bct_true["EditAzi"] * 0.9)
    table = dict(EditAzi=(-10000, 10000), EditDomainSize=(0, 5000), EditRadius=(0, 1000), Length=(0, 1000),
                 uca=(0.01, 1000), ucb=(0.01, 1000), ucc=(0.01, 1000), I0=(0, 100000), Base=(-10000, 10000),
                 EditSigma=(-10000, 10000))
    used6 = ["EditAzi", "EditDomainSize", "EditRadius", "Length", "uca", "ucc"]
    cube_true = dict(CUBE_A, GridPoints=16)
    cube_start = dict(cube_true, EditRadius=11.0, EditSigma=0.12)
    return {
        # six fitted entries and two fixed ones in the hash (mmax 8 > numFit 6)
        "bct_fixed": dict(true=bct_true, start=bct_start, used=used6, fixed=["I0", "ucb"], bounds=table,
                          stp=0.05, maxit=40, ftol=1e-10, bstop=0, border=0),
        # three fitted entries only (mmax = numFit), with the beam-stop mask
        "bct_three": dict(true=bct_true, start=bct_start, used=["EditDomainSize", "EditRadius", "Length"],
                          fixed=[], bounds=table, stp=0.1, maxit=40, ftol=1e-10, bstop=4, border=0),
        # first fitted entry bounded away from 0: the zeroed vertex 0 puts every other
        # vertex out of bounds
        "bct_outside": dict(true=bct_true, start=bct_start, used=["EditRadius", "Length"], fixed=[],
                            bounds=dict(table, EditRadius=(2, 8), Length=(10, 60)), stp=0.05, maxit=10,
                            ftol=1e-10, bstop=0, border=0),
        # isotropic cube, two fitted entries, two fixed ones
        "cube": dict(true=cube_true, start=cube_start, used=["EditRadius", "EditSigma"], fixed=["Base", "I0"],
                     bounds=table, stp=0.05, maxit=60, ftol=1e-12, bstop=0, border=2),
    }


def write_fits(manifest):
    arrays, meta = {}, {}
    for name, c in _fit_cases().items():
        target = run_ref(c["true"])["image"]
        gp = c["true"]["GridPoints"]
        w = 2 * gp
        with tempfile.TemporaryDirectory() as td:
            fd = os.path.join(td, "fit.bin")
            target.astype("<f8").tofile(fd)
            extra = dict(fitdata=fd, fitw=w, fith=w, bstop=c["bstop"], border=c["border"], stp=c["stp"],
                         maxit=c["maxit"], ftol=c["ftol"], threads=1)
            for i, k in enumerate(c["used"] + c["fixed"]):
                lo, hi = c["bounds"][k]
                extra["fitpar%d" % i] = "%s,%d,%r,%r" % (k, int(k in c["used"]), float(lo), float(hi))
            r = run_ref(c["start"], extra, driver=REF_FIT)
        arrays[name + "_target"] = target
        arrays[name + "_fitres"] = r["fitres"]
        arrays[name + "_fitstart"] = r["fitstart"]
        meta[name] = dict(start=c["start"], true=c["true"], used=c["used"], fixed=c["fixed"],
                          bounds={k: c["bounds"][k] for k in c["used"] + c["fixed"]}, stp=c["stp"],
                          maxit=c["maxit"], ftol=c["ftol"], bstop=c["bstop"], border=c["border"], w=w,
                          repetitions=int(r["repetitions"][0]), numImgCalc=int(r["numImgCalc"][0]),
                          aborted=int(r["aborted"][0]), retinfo=r["retinfo"], log=r["log"])
    np.savez_compressed(os.path.join(HERE, "simplex_fits.npz"), **arrays)
    manifest["fixtures"]["simplex_fits.npz"] = {"driver": "oracle/ref_fit.cpp", "cases": meta}


def write_post(manifest):
    arrays, cases = {}, {}
    for name, p, extra in POST_CASES:
        r = run_ref(p, extra)
        arrays[name + "_image"] = r["image"]
        arrays[name + "_bounds"] = r["bounds"]
        for k in ("rphi", "rphi_log", "scaled", "fft"):
            if k in r:
                arrays["%s_%s" % (name, k)] = r[k]
        cases[name] = {"params": p, "post": extra}
    np.savez_compressed(os.path.join(HERE, "post.npz"), **arrays)
    manifest["fixtures"]["post.npz"] = {"cases": cases}


def main():
    if not os.path.exists(REF_DRIVER):
        sys.exit("build the reference first: make -C oracle ref")
    only = sys.argv[1:]  # e.g. "cube": write only those fixtures, keep the manifest's others
    if only:
        with open(os.path.join(HERE, "manifest.json")) as fh:
            manifest = json.load(fh)
        if "cube" in only:
            write_cube(manifest)
        if "tpv" in only:
            write_tpv(manifest)
        if "post" in only:
            write_post(manifest)
        if "lattices" in only:
            write_lattices(manifest)
        if "peaks" in only:
            write_peaks(manifest)
        if "lattices3d" in only:
            write_cubics(manifest)
        if "fits" in only:
            write_fits(manifest)
        if "qmodes" in only:
            write_qmodes(manifest)
        with open(os.path.join(HERE, "manifest.json"), "w") as fh:
            json.dump(manifest, fh, indent=1, default=lambda o: o.tolist() if hasattr(o, "tolist") else str(o))
        return
    manifest = {"generator": "tests/golden/make_golden.py", "driver": "oracle/ref_driver.cpp",
                "compiler": subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0],
                "libc": platform.libc_ver(), "fixtures": {}}

    # Reference golden image, decoded (Documentation/Vitess-Sphere.dat: 19-byte header + f64).
    raw = open(os.path.join(REF_DOC, "Vitess-Sphere.dat"), "rb").read()
    np.savez_compressed(os.path.join(HERE, "vitess_sphere.npz"), header=np.frombuffer(raw[:19], np.uint8),
                        image=np.frombuffer(raw[19:], "<f8"))
    manifest["fixtures"]["vitess_sphere.npz"] = {"source": "Documentation/Vitess-Sphere.dat",
                                                 "params": VITESS}

    # Config 1: sphere, 128x128, FP64.
    r = run_ref(CFG1)
    np.savez_compressed(os.path.join(HERE, "cfg1_sphere.npz"), image=r["image"], bounds=r["bounds"],
                        **table_dict(r, False))
    manifest["fixtures"]["cfg1_sphere.npz"] = {"params": CFG1}

    # Config 2: Gaussian-oriented cylinders (reference mode), 32x32 + tables.
    r = run_ref(CFG2)
    c11 = r["carr11pm"].reshape(131, 131)[:121, :121].copy()
    np.savez_compressed(os.path.join(HERE, "cfg2_cylgauss.npz"), image=r["image"], carr11pm=c11,
                        carr3p=r["carr3p"] if "carr3p" in r else np.zeros(0), **table_dict(r, False))
    manifest["fixtures"]["cfg2_cylgauss.npz"] = {"params": CFG2}

    # Config 3: FCC spheres.  64x64 full image at the same q range, and 4096 sampled pixels of
    # the full 1024x1024 image.
    small = dict(CFG3, GridPoints=32)
    r = run_ref(small)
    np.savez_compressed(os.path.join(HERE, "cfg3_fcc_g32.npz"), image=r["image"], **table_dict(r, True))
    manifest["fixtures"]["cfg3_fcc_g32.npz"] = {"params": small}
    r = run_ref(CFG3, {"threads": 8})
    rng = np.random.default_rng(20261202)
    idx = np.sort(rng.choice(r["image"].size, 4096, replace=False))
    np.savez_compressed(os.path.join(HERE, "cfg3_fcc_1024_samples.npz"), idx=idx.astype(np.int64),
                        values=r["image"][idx], bounds=r["bounds"])
    manifest["fixtures"]["cfg3_fcc_1024_samples.npz"] = {"params": CFG3, "sample": "4096 pixels, seed 20261202",
                                                        "calc_ms_8threads": float(r["calc_ms"][0])}

    # Config 4: BCT cylinders, six parameter draws at GridPoints 16 (32x32), plus the fit
    # objective with every mask mode on one of them.
    rng = np.random.default_rng(1234)
    sets = []
    arrays = {}
    for k in range(6):
        p = cfg4_params(rng)
        p["GridPoints"] = 16
        r = run_ref(p)
        sets.append(p)
        arrays["image_%d" % k] = r["image"]
        for t, v in table_dict(r, True).items():
            arrays["%s_%d" % (t, k)] = v
    p = dict(sets[0], BeamPosX=16, BeamPosY=16)
    data = np.random.default_rng(5678).uniform(0.01, 10.0, 32 * 32)
    fits = []
    with tempfile.TemporaryDirectory() as td:
        fd = os.path.join(td, "fit.bin")
        data.astype("<f8").tofile(fd)
        for bstop, border, rect in [(0, 0, None), (4, 0, None), (0, 2, None), (-1, 0, None), (0, 0, (3, 4, 9, 12))]:
            extra = dict(mode="fit", fitdata=fd, fitw=32, fith=32, bstop=bstop, border=border)
            if rect:
                extra["nofit0"] = ",".join(map(str, rect))
            r = run_ref(p, extra)
            fits.append(dict(bstop=bstop, border=border, rect=rect, fqs=float(r["fqs"][0]),
                             cnt=int(r["cnt"][0]), nancnt=int(r["nancnt"][0])))
            arrays["fitfqs_b%d_r%d_%s" % (bstop, border, "rect" if rect else "none")] = r["fitfqs"]
    arrays["fitdata"] = data
    np.savez_compressed(os.path.join(HERE, "cfg4_bct.npz"), **arrays)
    manifest["fixtures"]["cfg4_bct.npz"] = {"params": sets, "fit_params": p, "fits": fits,
                                            "fitdata": "uniform(0.01,10) seed 5678, 32x32"}

    write_cube(manifest)
    write_tpv(manifest)
    write_post(manifest)
    write_lattices(manifest)
    write_peaks(manifest)
    write_cubics(manifest)
    write_fits(manifest)
    write_qmodes(manifest)

    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, default=lambda o: o.tolist() if hasattr(o, "tolist") else str(o))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
