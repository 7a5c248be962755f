"""The resumable Nelder-Mead of the batched fit driver against the oracle's restatement of
doSimplexFit2D (Source/sc_simplexfit2d.cpp:254-611) on host objectives: same iterates,
bit for bit (CPU only)."""
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


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_nm_matches_oracle_simplex(oracle_lib, case):
    name, f, x0, lo, hi, stp, maxit, ftol = case
    x, fb, it, nev = simplex(f, x0, lo, hi, stp=stp, maxit=maxit, ftol=ftol)
    xo, fo, ito, nevo = oracle_lib.simplex(f, x0, len(x0), lo, hi, stp=stp, maxit=maxit, ftol=ftol)
    np.testing.assert_array_equal(x, xo)
    assert fb == fo and it == ito and nev == nevo


def test_nm_converges_on_quadratic():
    c = np.array([1.5, -2.0, 0.75])
    x, fb, it, nev = simplex(lambda x: float(((x - c) ** 2).sum()), [1.0, 1.0, 1.0], [-10] * 3, [10] * 3,
                             stp=0.5, maxit=2000, ftol=1e-14)
    assert np.abs(x - c).max() < 1e-5 and fb < 1e-10
