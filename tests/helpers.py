"""Shared test helpers: golden fixtures and the benchmark parameter sets."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as fh:
        return json.load(fh)


def fixture(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def params(name):
    return dict(manifest()["fixtures"][name]["params"])


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


# Benchmark configurations (reference keys).  Config 3 is the headline workload.
CFG1 = dict(LType=12, ComboBoxParticle=0, Ordis=7, ComboBoxInterior=0, EditRadius=5, EditSigma=0.1,
            GridPoints=64, EditQmax=2, I0=1, Base=0, iso=0)
CFG2 = dict(LType=12, ComboBoxParticle=1, Ordis=0, ComboBoxInterior=0, EditRadius=3, Length=40,
            EditSigma=0.06, SigmaL=0.06, EditDbeta=10, theta=30, phi=45, GridPoints=256, EditQmax=1.5,
            I0=1, Base=0)
CFG3 = dict(LType=5, ComboBoxParticle=0, Ordis=0, ComboBoxPeak=1, ComboBoxInterior=0, EditRadius=8,
            EditSigma=0.08, uca=30, ucb=30, ucc=30, EditDomainSize=200, RadButDebyeScherrer=1,
            EditDebyeWaller=1.5, EditAzi=5, HKLmax=5, GridPoints=512, EditQmax=1, I0=1, Base=0)
