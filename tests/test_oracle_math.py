"""Pins of the CPU oracle that do not come from the reference's outputs:
closed forms, textbook integrals, brute-force enumerations and invariants.
Each is chosen so that a plausible slip (dropped term, wrong sign, swapped
index, wrong normalisation) fails it."""
import math

import numpy as np
import pytest
from scipy import integrate, special

from helpers import CFG1, rel_err


def _schulz_avg(fun, R, z):
    """Schulz average over radii with weight r^z exp(-(z+1) r/R) (mean R), the size
    distribution the reference's closed forms average over."""
    w = lambda r: np.exp(z * np.log(r / R) - (z + 1) * (r / R - 1))
    lo, hi = R * 0.3, R * 2.2
    num = integrate.quad(lambda r: w(r) * fun(r), lo, hi, limit=500, epsrel=1e-12)[0]
    den = integrate.quad(w, lo, hi, limit=500, epsrel=1e-12)[0]
    return num / den


def _sphere_amp(x):
    return 3 * (np.sin(x) - x * np.cos(x)) / x ** 3


def test_sphere_pq_matches_schulz_average(oracle_lib):
    R, sig = 5.0, 0.1
    z = (1 - sig * sig) / (sig * sig)
    e = oracle_lib.OracleEngine(dict(CFG1, EditRadius=R, EditSigma=sig))
    e.prepare()
    limq4 = e.table("limq")[3]
    # series branch (q < 0.4 limq4): Taylor series of the average, exit at 1e-4 relative change
    for q in (0.05, 0.2, 0.35 * limq4):
        v = e.pixel(q - 1e-9, 0, 0)
        assert abs(v / _schulz_avg(lambda r: _sphere_amp(q * r) ** 2, R, z) - 1) < 1e-5
    # closed form (Kotlarchyk-Chen) at large qR, where the reference's pq3 exponent quirk fades
    for q in (5.0, 8.0):
        v = e.pixel(q - 1e-9, 0, 0)
        assert abs(v / _schulz_avg(lambda r: _sphere_amp(q * r) ** 2, R, z) - 1) < 1e-6


def test_sphere_monodisperse_limit_is_rayleigh(oracle_lib):
    """sigma -> 0 gives [3(sin x - x cos x)/x^3]^2 on the series branch.  (The closed form
    does not reach it: the reference's pq3 term uses the exponent zr-5 where the average
    has (zr-5)/2, and that bias survives the limit; the fixtures pin that branch.)"""
    R = 5.0
    e = oracle_lib.OracleEngine(dict(CFG1, EditRadius=R, EditSigma=0.0005))
    e.prepare()
    for x in (2.0, 4.0, 5.763):
        assert x / R < 0.4 * e.table("limq")[3]
        v = e.pixel(x / R - 1e-9, 0, 0)
        assert abs(v / _sphere_amp(x) ** 2 - 1) < 1e-4  # series exit rule: relative change < 1e-4


CUBE = dict(LType=12, ComboBoxParticle=4, Ordis=7, ComboBoxInterior=0, GridPoints=16, EditQmax=1,
            I0=1, Base=0, iso=0)


def _cube_orient_avg(x):
    """Orientation average of [sinc(x nx) sinc(x ny) sinc(x nz)]^2: the form factor of a cube
    of half-edge 1 at q = x, by a 96x96 Gauss-Legendre rule over one octant of directions."""
    t, wt = np.polynomial.legendre.leggauss(96)
    th = (t + 1) * np.pi / 4
    w = wt * np.pi / 4
    TH, PH = np.meshgrid(th, th, indexing="ij")
    W = np.outer(w, w) * np.sin(TH)
    nx, ny, nz = np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)
    f = (np.sinc(x * nx / np.pi) * np.sinc(x * ny / np.pi) * np.sinc(x * nz / np.pi)) ** 2
    return float((W * f).sum() / (np.pi / 2))


def test_cube_series_is_the_schulz_average_of_the_orientation_average(oracle_lib):
    """Below the series radius, P(q) of the isotropic cube is the number-weighted Schulz
    average of the orientation-averaged monodisperse cube form factor."""
    R, sig = 10.0, 0.1
    z = (1 - sig * sig) / (sig * sig)
    e = oracle_lib.OracleEngine(dict(CUBE, EditRadius=R, EditSigma=sig))
    e.prepare()
    limq4 = e.table("limq")[3]
    for q in (0.02, 0.1, 0.25, 0.5 * limq4):
        v = e.pixel(q - 1e-9, 0, 0)
        assert abs(v / _schulz_avg(lambda r: _cube_orient_avg(q * r), R, z) - 1) < 2e-5, q


def test_cube_monodisperse_series_coefficients(oracle_lib):
    """sigma -> 0: P(q) = 1 - (qR)^2/3 + 11 (qR)^4/225 - ... (Guinier with Rg^2 = R^2 for the
    half-edge R, and the fourth-order term of exp(-x^2/3)(1 - x^4 <n^4 sum>/90))."""
    R = 3.0
    e = oracle_lib.OracleEngine(dict(CUBE, EditRadius=R, EditSigma=0.0005))
    e.prepare()
    c = e.table("carr4p")
    assert abs(c[1] / (-R ** 2 / 3) - 1) < 1e-5
    assert abs(c[2] / (11 * R ** 4 / 225) - 1) < 1e-5
    for x in (1.0, 2.5, 4.0):
        assert x / R < e.table("limq")[3]
        assert abs(e.pixel(x / R - 1e-9, 0, 0) / _cube_orient_avg(x) - 1) < 1e-4


def test_cube_porod_constant_is_the_schulz_average(oracle_lib):
    """por/q^4 is the Porod law 2 pi S/(V^2 q^4) = (3 pi/4)/(r q)^4 averaged over the Schulz
    distribution: por = (3 pi/4) <r^-4>."""
    for R, sig in ((10.0, 0.1), (4.0, 0.25)):
        z = (1 - sig * sig) / (sig * sig)
        e = oracle_lib.OracleEngine(dict(CUBE, EditRadius=R, EditSigma=sig))
        e.prepare()
        # full-range Schulz average (the r^-4 weight reaches below the helper's 0.3 R cut)
        w = lambda r: np.exp(z * np.log(r / R) - (z + 1) * (r / R - 1))
        num = integrate.quad(lambda r: w(r) * 0.75 * np.pi / r ** 4, 0, 12 * R, limit=500, epsrel=1e-12)[0]
        den = integrate.quad(w, 0, 12 * R, limit=500, epsrel=1e-12)[0]
        assert abs(e.scalar("por") / (num / den) - 1) < 1e-9
        q = 3.0 * e.table("limq")[3]
        assert abs(e.pixel(q - 1e-9, 0, 0) / (e.scalar("por") / q ** 4) - 1) < 1e-14


def test_cube_other_interiors_have_unit_form_factor(oracle_lib):
    e = oracle_lib.OracleEngine(dict(CUBE, EditRadius=5, EditSigma=0.1, ComboBoxInterior=1, I0=2, iso=0.5,
                                     Base=0.25, ifluc=3, rfluc=2))
    e.prepare()
    for q in (0.1, 0.7):
        assert e.pixel(q - 1e-9, 0, 0) == 0.25 + 2 * (1 + 0.5) + 3 / (1 + q * q * 4)


def test_center_pixel_is_i0_times_one_plus_iso(oracle_lib):
    e = oracle_lib.OracleEngine(dict(CFG1, I0=7.5, iso=0.25, Base=0.5))
    e.prepare()
    x0, x1, y0, y1 = e.bounds()
    img = e.image().reshape(y1 - y0, x1 - x0)
    assert abs(img[-y0, -x0] - (0.5 + 7.5 * 1.25)) < 1e-12


@pytest.mark.parametrize("a,tol", [(0.002, 3e-4), (0.005, 3e-4), (1.0, 1e-4), (10.0, 1e-8), (100.0, 1e-12)])
def test_gaussnorm3_is_the_sine_weighted_gaussian_integral(oracle_lib, a, tol):
    exact = integrate.quad(lambda x: np.exp(-a * x * x) * np.sin(x), 0, np.pi, epsrel=1e-14)[0]
    assert abs(oracle_lib.gaussnorm3(a) / exact - 1) < tol


def test_reference_gamma_accuracy(oracle_lib):
    for x in (0.1, 0.3, 0.5, 0.9, 1.0):
        assert abs(oracle_lib.gamma(x) / math.gamma(x) - 1) < 1e-10
    for x in (1.5, 2.5, 7.3, 50.5, 154.75):
        assert abs(oracle_lib.gamma(x) / math.gamma(x) - 1) < 3e-7
    for n in (2, 5, 10, 20):
        assert oracle_lib.gamma(float(n)) == math.factorial(n - 1)
    # reflection formula for negative non-integers
    assert abs(oracle_lib.gamma(-0.5) / math.gamma(-0.5) - 1) < 1e-9


def _lattice_engine(oracle_lib, ltype, a=30.0, c=30.0, H=5):
    p = dict(LType=ltype, ComboBoxParticle=0 if ltype == 5 else 1, Ordis=0 if ltype == 5 else 6,
             ComboBoxPeak=1, EditRadius=4, Length=20, EditSigma=0.08, uca=a, ucb=a, ucc=c,
             RadButDebyeScherrer=1, EditDomainSize=200, EditAzi=5, HKLmax=H, GridPoints=8,
             theta=0, phi=0)
    e = oracle_lib.OracleEngine(p)
    e.prepare()
    return e


def test_fcc_reflection_lists_brute_force(oracle_lib):
    e = _lattice_engine(oracle_lib, 5)
    L1 = e.table("latpar1").reshape(-1, 6)[1:]
    L2 = e.table("latpar2").reshape(-1, 6)[1:]
    r = range(0, 6)
    iso = {(h, k, l) for h in r for k in r for l in r if h % 2 == k % 2 == l % 2 and (h, k, l) != (0, 0, 0)}
    r = range(-5, 6)
    ani = {(h, k, l) for h in r for k in r for l in r if h % 2 == k % 2 == l % 2 and (h, k, l) != (0, 0, 0)}
    assert set(map(tuple, L1[:, 1:4])) == iso and len(L1) == 53
    assert set(map(tuple, L2[:, 1:4])) == ani and len(L2) == 340
    assert (L2[:, 5] == 16).all() and (L1[:, 5] == 16).all()   # |F|^2 of the 4-atom basis
    # cubic multiplicities: {100}: 6, {110}: 12, {111}: 8, {hk0}: 24, {hhl}: 24, {hkl}: 48
    mult = {tuple(x[1:4]): x[4] for x in L1}
    assert mult[(1, 1, 1)] == 8 and mult[(2, 0, 0)] == 6 and mult[(2, 2, 0)] == 12
    assert mult[(3, 1, 1)] == 24 and mult[(4, 2, 0)] == 24 and mult[(5, 3, 1)] == 48


def test_bct_reflection_lists_brute_force(oracle_lib):
    e = _lattice_engine(oracle_lib, 8, a=25.0, c=40.0)
    L2 = e.table("latpar2").reshape(-1, 6)[1:]
    r = range(-5, 6)
    ani = {(h, k, l) for h in r for k in r for l in r if (h + k + l) % 2 == 0 and (h, k, l) != (0, 0, 0)}
    assert set(map(tuple, L2[:, 1:4])) == ani and len(L2) == 664
    assert (L2[:, 5] == 4).all()


def _cubic_engine(oracle_lib, ltype, H=6):
    p = dict(LType=ltype, ComboBoxParticle=0, Ordis=0, ComboBoxPeak=1, EditRadius=4, EditSigma=0.08, uca=30,
             ucb=30, ucc=30, RadButDebyeScherrer=1, EditDomainSize=200, EditAzi=5, HKLmax=H, GridPoints=8)
    e = oracle_lib.OracleEngine(p)
    e.prepare()
    return e


def test_fd3m_reflections_are_those_of_the_diamond_structure(oracle_lib):
    """Fd3m (cubic group Fd--): the accepted reflections with a nonzero structure factor are
    exactly those where the diamond basis' |F|^2 = |F_fcc|^2 |1 + i^(h+k+l)|^2 is nonzero (all
    odd, or all even with h+k+l = 4n), and |F|^2 is 32 (all odd) or 64."""
    e = _cubic_engine(oracle_lib, 17)
    L2 = e.table("latpar2").reshape(-1, 6)[1:]
    got = {tuple(x[1:4]) for x in L2 if x[5] != 0}
    r = range(-6, 7)
    want = {(h, k, l) for h in r for k in r for l in r
            if (h % 2 == k % 2 == l % 2) and (h % 2 == 1 or (h + k + l) % 4 == 0) and (h, k, l) != (0, 0, 0)}
    assert got == want
    for x in L2:
        if x[5]:
            assert x[5] == (32 if x[1] % 2 else 64)


def test_ia3d_reflection_conditions(oracle_lib):
    """Ia3d (group 230): hkl: h+k+l = 2n; 0kl: k, l = 2n; hhl: 2h+l = 4n; h00: h = 4n
    (International Tables, with the cyclic permutations), checked by brute force."""
    e = _cubic_engine(oracle_lib, 9)
    L2 = e.table("latpar2").reshape(-1, 6)[1:]
    r = range(-6, 7)

    def ok(h, k, l):
        if (h + k + l) % 2:
            return False
        v = (h, k, l)
        for z in range(3):
            a, b, c = v[z], v[(z + 1) % 3], v[(z + 2) % 3]
            if a == 0 and (b % 2 or c % 2):
                return False
            if b == c and a != 0 and ((2 * b + a) % 4 or a % 2):
                return False
            if b == 0 and c == 0 and a % 4:
                return False
        return True

    want = {(h, k, l) for h in r for k in r for l in r if (h, k, l) != (0, 0, 0) and ok(h, k, l)}
    assert set(map(tuple, L2[:, 1:4])) == want
    assert (L2[:, 5] == 4).all()   # body-centred two-atom basis, h+k+l even


def test_2d_hexagonal_lattice_reciprocal_vectors(oracle_lib):
    """Hexagonal 2D lattice (a = b, gamma = 120 deg): every (h, k) with l = 0 is accepted with
    multiplicity 1, and |q_hk0| = (4 pi / (sqrt(3) a)) sqrt(h^2 + k^2 + h k) -- the hexagonal
    d-spacing, which a sign slip in the plane cell's reciprocal basis would change to h^2 + k^2 - h k."""
    p = dict(LType=1, ComboBoxParticle=1, Ordis=6, ComboBoxPeak=1, EditRadius=3, Length=30, EditSigma=0.06,
             SigmaL=0.06, uca=20, ucb=20, ucc=20, ucgamma=120, RadButDebyeScherrer=1, EditDomainSize=200,
             EditAzi=5, HKLmax=4, GridPoints=8, theta=0, phi=0)
    e = oracle_lib.OracleEngine(p)
    e.prepare()
    L2 = e.table("latpar2").reshape(-1, 6)[1:]
    L3 = e.table("latpar3").reshape(-1, 17)[1:]
    assert len(L2) == 9 * 9 * 5 - 1 and (L2[:, 4] == 1).all() and (L2[:, 5] == 1).all()
    sel = L2[:, 3] == 0
    h, k = L2[sel, 1].astype(float), L2[sel, 2].astype(float)
    expect = 4 * np.pi / (np.sqrt(3) * 20) * np.sqrt(h * h + k * k + h * k)
    assert rel_err(L3[sel, 5], expect).max() < 2e-6


@pytest.mark.parametrize("ltype,a,c", [(5, 30.0, 30.0), (8, 25.0, 40.0)])
def test_reflection_vectors_have_lattice_lengths(oracle_lib, ltype, a, c):
    e = _lattice_engine(oracle_lib, ltype, a, c)
    L2 = e.table("latpar2").reshape(-1, 6)[1:]
    L3 = e.table("latpar3").reshape(-1, 17)[1:]
    h, k, l = L2[:, 1], L2[:, 2], L2[:, 3]
    expect = 2 * np.pi * np.sqrt((h * h + k * k) / a ** 2 + l * l / c ** 2)
    assert rel_err(L3[:, 5], expect).max() < 2e-6
    assert rel_err(np.sqrt(L3[:, 2] ** 2 + L3[:, 3] ** 2 + L3[:, 4] ** 2), expect).max() < 2e-6


@pytest.mark.parametrize("ltype,a,c", [(5, 30.0, 30.0), (8, 25.0, 40.0)])
def test_rotated_reciprocal_basis_is_a_rotation_of_the_cell(oracle_lib, ltype, a, c):
    """corotations returns the reciprocal basis a*, b*, c* after a rotation: R R^T equals
    the reciprocal metric diag(1/a^2, 1/a^2, 1/c^2) of the orthogonal cells."""
    e = _lattice_engine(oracle_lib, ltype, a, c)
    R = e.table("ri").reshape(3, 3)
    G = R @ R.T
    np.testing.assert_allclose(G, np.diag([1 / a ** 2, 1 / a ** 2, 1 / c ** 2]), rtol=0, atol=2e-7 / a ** 2)
    assert abs(abs(np.linalg.det(R)) - 1 / (a * a * c)) < 1e-6 / (a * a * c)


def test_fcc_image_point_symmetric(oracle_lib):
    from helpers import CFG3
    e = oracle_lib.OracleEngine(dict(CFG3, GridPoints=24))
    e.prepare()
    x0, x1, y0, y1 = e.bounds()
    img = e.image(nthreads=4).reshape(y1 - y0, x1 - x0)
    # (ihex, i) -> (-ihex, -i): rows/cols 1.. mirror onto each other around index -x0
    a = img[1:, 1:]
    assert rel_err(a, a[::-1, ::-1]).max() < 1e-12


def test_orientation_norm_is_the_weight_integral(oracle_lib):
    p = dict(LType=12, ComboBoxParticle=1, Ordis=0, EditRadius=3, Length=40, EditSigma=0.06, SigmaL=0.06,
             EditDbeta=10, theta=30, phi=45, GridPoints=8, EditQmax=1.5)
    e = oracle_lib.OracleEngine(p)
    e.prepare()
    db = 10 * np.pi / 180
    hi = float(np.float32(db * np.sqrt(np.log(1e5))))
    exact = integrate.quad(lambda x: np.sin(x) * np.exp(-x * x / db ** 2), 0, hi, epsrel=1e-14)[0]
    assert abs(e.scalar("norm") / exact - 1) < 1e-4  # Romberg relative tolerance 1e-4


def test_gaussian_orientation_average_physics_mode(oracle_lib):
    """With fix_orcase=1 the nested Romberg evaluates the orientation average of the
    longitudinal rod factor; compare with an independent 2-D quadrature."""
    R, L, sr, sl, db_deg, th = 3.0, 40.0, 0.06, 0.06, 10.0, 30.0
    p = dict(LType=12, ComboBoxParticle=1, Ordis=0, EditRadius=R, Length=L, EditSigma=sr, SigmaL=sl,
             EditDbeta=db_deg, theta=th, phi=45, GridPoints=8, EditQmax=1.5, I0=1, Base=0, fix_orcase=1)
    e = oracle_lib.OracleEngine(p)
    e.prepare()
    zl = (1 - sl * sl) / (sl * sl)
    db = db_deg * np.pi / 180
    dmax = float(np.float32(db * np.sqrt(np.log(1e5))))
    t = th * np.pi / 180
    p11, p13, p21 = -np.cos(t), np.sin(t), -1.0

    def rod(arg):
        a1 = 1 / (2 * zl * (zl - 1))
        return (a1 / arg ** 2) * (1 - np.cos((zl - 1) * np.arctan(2 * arg)) / (1 + 4 * arg * arg) ** ((zl - 1) / 2))

    norm = integrate.quad(lambda d: np.sin(d) * np.exp(-d * d / db ** 2), 0, dmax, epsrel=1e-13)[0]
    for qx, qy in ((0.3, 0.1), (0.05, 0.6), (0.7, -0.4)):
        def inner(d):
            f = lambda c: rod((qx * (p11 * np.cos(c) * np.sin(d) + p13 * np.cos(d)) + qy * (p21 * np.sin(c) * np.sin(d)) + 1e-10) * L / (zl + 1))
            return integrate.quad(f, 0, 2 * np.pi, limit=400, epsrel=1e-10)[0] / (2 * np.pi)
        pl = integrate.quad(lambda d: np.sin(d) * np.exp(-d * d / db ** 2) * inner(d), 0, dmax, limit=200, epsrel=1e-9)[0] / norm
        q = math.hypot(qx, qy)
        # transverse factor from the same engine: the z-dir cylinder without lattice has pql = 1
        et = oracle_lib.OracleEngine(dict(p, Ordis=6, theta=0, phi=0))
        et.prepare()
        pt = et.pixel(q, 0, 0) if False else et.pixel(qx, qy, 1e-20)
        v = e.pixel(qx, qy, 1e-20)
        assert abs(v / (pl * pt) - 1) < 5e-4, (qx, qy, v, pl * pt)


def test_simplex_finds_quadratic_minimum(oracle_lib):
    target = np.array([1.5, -2.0, 0.75])
    f = lambda x: float(((x - target) ** 2 * np.array([1.0, 3.0, 0.5])).sum())
    x, fb, it, nev = oracle_lib.simplex(f, [1.0, 1.0, 1.0], 3, [-10] * 3, [10] * 3, stp=0.5, maxit=2000, ftol=1e-14)
    assert np.abs(x - target).max() < 1e-5 and fb < 1e-10


def test_simplex_bounds_return_1e20(oracle_lib):
    seen = []
    f = lambda x: (seen.append(x.copy()), float((x ** 2).sum()))[1]
    oracle_lib.simplex(f, [0.5, 0.5], 2, [0.4, 0.4], [0.6, 0.6], stp=0.5, maxit=5, ftol=1e-12)
    assert all((s >= 0.4).all() and (s <= 0.6).all() for s in seen)
