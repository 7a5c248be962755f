"""The 2D simplex against the reference's own runs of doSimplexFit2D.

tests/golden/simplex_fits.npz holds four fits that the reference ran itself
(oracle/ref_fit.cpp: SasCalc_SimplexFit2D::doSimplexFit2D, Source/sc_simplexfit2d.cpp:254-611,
over the reference engine, as the console's autofit calls it): the final vertex of every hash
entry, the iteration and evaluation counts and the progress log, whose lines give each
objective value in evaluation order (%lf) with its pixel count.  The oracle's simplex
(csxo_simplex) and the product's (csx_nm_run_cb, mode "reference") run the same fits on the
oracle's FP64 objective and must follow the same path: every evaluated value, the counts and
the final vertex bit for bit.  The final vertex depends only on the comparisons along the way,
so a single different decision (or a wrong step, factor, copy or shrink) changes it.
"""
import json
import os
import re

import numpy as np
import pytest

from helpers import GOLDEN

FIX = np.load(os.path.join(GOLDEN, "simplex_fits.npz"))
with open(os.path.join(GOLDEN, "manifest.json")) as fh:
    CASES = json.load(fh)["fixtures"]["simplex_fits.npz"]["cases"]

_NUM = re.compile(r"^(FQS\((\d+)\)|amotry\(simplex (reflection|expansion|contraction)\)) = (.*)$")


def reference_events(case):
    """The evaluations of the reference's log, in order: (value or None if logged as text,
    pixel count or None, in bounds).  Initial vertices past numFit+1 are copies, not
    evaluations (Source/sc_simplexfit2d.cpp:420-428)."""
    nfit, mmax = len(case["used"]), len(case["used"]) + len(case["fixed"])
    out, init = [], True
    for line in case["log"].splitlines():
        if line.startswith("@REP"):
            init = False
            continue
        m = _NUM.match(line)
        if not m:
            continue
        if init and m.group(2) is not None and int(m.group(2)) > min(mmax, nfit + 1):
            continue
        rest = m.group(4)
        if rest.startswith(("MIN Index", "MAX Index")):
            out.append((None, None, False))
            continue
        tok = rest.split()
        px = re.match(r"\((\d+)px/", rest[len(tok[0]):].strip()) if not rest.startswith("(") else None
        val = float(tok[0]) if not rest.startswith("(") else None
        out.append((val, int(px.group(1)) if px else None, True))
    return out


class OracleObjective:
    """fehlerquadratsumme on the oracle engine: every hash entry set, prepared, doFitCalculation
    (1e20 when the preparation fails, Source/sc_calc_generic_cpu.cpp:1341)."""

    def __init__(self, oracle_lib, case, name):
        self.eng = oracle_lib.OracleEngine(case["start"])
        self.keys = case["used"] + case["fixed"]
        self.target = FIX[name + "_target"]
        self.case = case
        self.trace = []

    def __call__(self, x):
        self.eng.set(dict(zip(self.keys, x)))
        try:
            self.eng.prepare()
        except RuntimeError:
            self.trace.append((1e20, 0))
            return 1e20
        w = self.case["w"]
        s, cnt, _, _ = self.eng.fit_fqs(self.target, w, w, bstop=self.case["bstop"], border=self.case["border"])
        self.trace.append((s, cnt))
        return s


def start_vectors(case):
    keys = case["used"] + case["fixed"]
    x0 = np.array([case["start"].get(k, {"I0": 1.0, "Base": 0.0}.get(k)) for k in keys], dtype=np.float64)
    lo = np.array([case["bounds"][k][0] for k in keys], dtype=np.float64)
    hi = np.array([case["bounds"][k][1] for k in keys], dtype=np.float64)
    return x0, lo, hi


def check_path(name, case, trace, x, iters, evals):
    ref = reference_events(case)
    evaluated = [e for e in ref if e[2]]
    assert evals == case["numImgCalc"] == len(evaluated) == len(trace)
    assert iters == case["repetitions"]
    for k, ((v, px, _), (f, cnt)) in enumerate(zip(evaluated, trace)):
        if v is not None:  # %lf: six decimals
            assert abs(f - v) <= 5e-7 + 1e-12 * abs(v), (name, k, f, v)
        if px is not None:
            assert cnt == px, (name, k, cnt, px)
    np.testing.assert_array_equal(x, FIX[name + "_fitres"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_simplex_follows_reference(oracle_lib, name):
    case = CASES[name]
    x0, lo, hi = start_vectors(case)
    np.testing.assert_array_equal(x0, FIX[name + "_fitstart"])
    f = OracleObjective(oracle_lib, case, name)
    x, fb, it, nev = oracle_lib.simplex(f, x0, len(case["used"]), lo, hi, stp=case["stp"], maxit=case["maxit"],
                                        ftol=case["ftol"])
    check_path(name, case, f.trace, x, it, nev)


@pytest.mark.parametrize("name", sorted(CASES))
def test_product_simplex_follows_reference(oracle_lib, name):
    from crystalscatter_amd import simplex
    case = CASES[name]
    x0, lo, hi = start_vectors(case)
    f = OracleObjective(oracle_lib, case, name)
    x, fb, it, nev = simplex(f, x0, lo, hi, stp=case["stp"], maxit=case["maxit"], ftol=case["ftol"],
                             nfree=len(case["used"]), mode="reference")
    check_path(name, case, f.trace, x, it, nev)


def test_reference_log_shows_the_quirks():
    """What the fixtures show of the reference's simplex beyond the textbook one: vertex 0's
    first entry is 0 after the clearing loop (the result of a fit that never moves vertex 0),
    and with more hash entries than fitted ones vertex numFit+1 is evaluated and copied on."""
    c = CASES["bct_outside"]
    assert FIX["bct_outside_fitres"][0] == 0.0 and c["numImgCalc"] == 1
    np.testing.assert_array_equal(FIX["bct_outside_fitres"][1:], FIX["bct_outside_fitstart"][1:])
    c = CASES["bct_fixed"]
    lines = c["log"].splitlines()
    assert lines[7].split("=")[1] == lines[8].split("=")[1]  # FQS(8) is a copy of FQS(7)
