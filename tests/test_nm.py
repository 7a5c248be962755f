"""The resumable Nelder-Mead of the batched fit driver against the oracle's restatement of
doSimplexFit2D (Source/sc_simplexfit2d.cpp:254-611) on host objectives: same iterates,
bit for bit (CPU only); the initial simplex and first trial point traced by hand; the
corrected mode on a quadratic."""
import numpy as np
import pytest

from crystalscatter_amd import simplex


def _cases():
    quad = lambda c, w: (lambda x: float(((x - c) ** 2 * w).sum()))
    rosen = lambda x: float(sum(100 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2))
    return [
        ("quadratic3", quad(np.array([1.5, -2.0, 0.75]), np.array([1.0, 3.0, 0.5])), [1.0, 1.0, 1.0], [-10] * 3, [10] * 3, 0.5, 300, 1e-12),
        ("quadratic6", quad(np.arange(6) * 0.3 + 1, np.linspace(1, 4, 6)), [2.0] * 6, [0.1] * 6, [9] * 6, 0.1, 500, 1e-10),
        ("rosenbrock4", rosen, [-1.2, 1.0, -0.5, 0.8], [-5] * 4, [5] * 4, 0.2, 400, 1e-9),
        ("bounded", quad(np.array([0.9, 0.1]), np.array([1.0, 1.0])), [0.5, 0.5], [0.4, 0.4], [0.6, 0.6], 0.5, 50, 1e-12),
        ("zero_start", quad(np.array([0.3, -0.2]), np.array([2.0, 1.0])), [0.0, 0.0], [0.0, -1.0], [1.0, 1.0], 0.1, 100, 1e-10),
        ("angle_exempt", quad(np.array([10.0, 1.0]), np.array([1.0, 1.0])), [30.0, 2.0], [0.0, 0.5], [360.0, 5.0], 0.3, 100, 1e-10),
    ]


@pytest.mark.parametrize("nfixed", [0, 2])
@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_nm_matches_oracle_simplex(oracle_lib, case, nfixed):
    """Reference mode against the oracle, with and without fixed hash entries (which the
    objective receives: trial points carry them at their column sums)."""
    name, f0, x0, lo, hi, stp, maxit, ftol = case
    n = len(x0)
    f = (lambda x: f0(x[:n]) + 0.01 * float(((x[n:] - 2.0) ** 2).sum())) if nfixed else f0
    x0 = list(x0) + [1.0, 3.0][:nfixed]
    lo = list(lo) + [0.0] * nfixed
    hi = list(hi) + [100.0] * nfixed
    x, fb, it, nev = simplex(f, x0, lo, hi, stp=stp, maxit=maxit, ftol=ftol, nfree=n)
    xo, fo, ito, nevo = oracle_lib.simplex(f, x0, n, lo, hi, stp=stp, maxit=maxit, ftol=ftol)
    np.testing.assert_array_equal(x, xo)
    assert fb == fo and it == ito and nev == nevo


def _trace(x0, lo, hi, stp, nfree, mode, maxit=1):
    seen = []
    def f(x):
        seen.append(x.copy())
        return float(((x[:nfree] - 10.0) ** 2).sum())
    simplex(f, x0, lo, hi, stp=stp, maxit=maxit, ftol=0.0, nfree=nfree, mode=mode)
    return seen


def test_reference_initial_simplex_by_hand():
    """Source/sc_simplexfit2d.cpp:339-428 for x0 = (2, 3, 5), the first two entries fitted
    (mmax 3), stp 0.1: rootn1 = sqrt(4) - 1 = 1, mroot2 = 1/(3 sqrt 2), steps (0.2, 0.3, 0),
    pp = 4 step mroot2, q = step mroot2.  The clearing loop zeroes vertex 0's first entry, y[0]
    is taken at x0, vertices 1..3 are built from the zeroed vertex 0, and vertex 3 = numFit+1 is
    evaluated.  The first reflection (fac = -1, factors 2/3 and 5/3 over mmax) moves the worst
    vertex and carries the fixed entry at its column sum 3 * 5."""
    m2 = 1.0 / (3 * np.sqrt(2.0))
    q = np.array([0.2, 0.3]) * m2
    pp = np.array([0.2, 0.3]) * 4 * m2
    seen = _trace([2.0, 3.0, 5.0], [-100.0] * 3, [100.0] * 3, 0.1, 2, "reference")
    v0 = np.array([0.0, 3.0, 5.0])
    v1 = np.array([0.0 + q[0] + pp[0], 3.0 + q[1], 5.0])
    v2 = np.array([0.0 + q[0], 3.0 + q[1] + pp[1], 5.0])
    v3 = np.array([0.0 + q[0], 3.0 + q[1], 5.0])
    np.testing.assert_array_equal(seen[0], [2.0, 3.0, 5.0])
    np.testing.assert_allclose(seen[1], v1, rtol=1e-15)
    np.testing.assert_allclose(seen[2], v2, rtol=1e-15)
    np.testing.assert_allclose(seen[3], v3, rtol=1e-15)
    # y[0] belongs to x0, not to the zeroed vertex 0; the worst of the three is reflected
    f = lambda x: float(((x[:2] - 10.0) ** 2).sum())
    y = [f(np.array([2.0, 3.0])), f(v1), f(v2)]
    worst = [v0, v1, v2][int(np.argmax(y))]
    psum = v0 + v1 + v2
    refl = psum[:2] * (2.0 / 3.0) - worst[:2] * (5.0 / 3.0)
    np.testing.assert_allclose(seen[4][:2], refl, rtol=1e-14)
    assert seen[4][2] == 15.0


def test_corrected_initial_simplex_by_hand():
    """Corrected mode: vertex 0 kept, steps over the two fitted entries (rootn1 = sqrt(3) - 1,
    mroot2 = 1/(2 sqrt 2)), no extra vertex, fixed entry at its start value in trials."""
    r1, m2 = np.sqrt(3.0) - 1, 1.0 / (2 * np.sqrt(2.0))
    st = np.array([0.2, 0.3])
    q, pp = st * r1 * m2, st * (r1 + 2) * m2
    seen = _trace([2.0, 3.0, 5.0], [-100.0] * 3, [100.0] * 3, 0.1, 2, "corrected")
    np.testing.assert_array_equal(seen[0], [2.0, 3.0, 5.0])
    np.testing.assert_allclose(seen[1], [2.0 + q[0] + pp[0], 3.0 + q[1], 5.0], rtol=1e-15)
    np.testing.assert_allclose(seen[2], [2.0 + q[0], 3.0 + q[1] + pp[1], 5.0], rtol=1e-15)
    assert all(s[2] == 5.0 for s in seen)


def test_corrected_mode_converges_with_fixed_entries():
    c = np.array([1.5, -2.0, 0.75])
    f = lambda x: float(((x[:3] - c) ** 2).sum()) + (0.0 if x[3] == 7.0 else 1e6)
    x, fb, it, nev = simplex(f, [1.0, 1.0, 1.0, 7.0], [-10] * 4, [10] * 4, stp=0.5, maxit=2000, ftol=1e-14,
                             nfree=3, mode="corrected")
    assert np.abs(x[:3] - c).max() < 1e-5 and fb < 1e-10 and x[3] == 7.0


def test_nm_converges_on_quadratic():
    c = np.array([1.5, -2.0, 0.75])
    x, fb, it, nev = simplex(lambda x: float(((x - c) ** 2).sum()), [1.0, 1.0, 1.0], [-10] * 3, [10] * 3,
                             stp=0.5, maxit=2000, ftol=1e-14)
    assert np.abs(x - c).max() < 1e-5 and fb < 1e-10
