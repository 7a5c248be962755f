"""Product host preparation (libcsx.so, host-only engine: no GPU needed) against the
reference's golden tables, and the parameter interface."""
import numpy as np
import pytest

from helpers import fixture, manifest, params

from crystalscatter_amd import CsxError, SasCalc

TABLES = ["carr1p", "carr2p", "carr4p", "carr1f", "carr4f", "limq"]


def _host(p, ignore_cbs=True):
    e = SasCalc(device=-1, params=p)
    e.set("ignore_cbs", 1 if ignore_cbs else 0)
    e.prepare()
    return e


@pytest.mark.parametrize("name", ["cfg1_sphere.npz", "cfg2_cylgauss.npz", "cfg3_fcc_g32.npz", "bcc_sphere.npz",
                                  "sc_sphere_iso.npz", "fcc_sphere_iso.npz", "bct_sphere.npz", "sphere_zdir.npz",
                                  "cyl_iso.npz", "cyl_iso_b.npz", "cube_zdir.npz", "cube_zdir_b.npz",
                                  "fcc_sphere_lorentz.npz", "bcc_sphere_mod1lor.npz", "sc_sphere_mod2lor_iso.npz",
                                  "fcc_sphere_pvoigt.npz", "bct_sphere_pearson.npz", "fcc_sphere_pearson_iso.npz",
                                  "ia3d_sphere.npz", "pn3m_sphere_iso.npz", "im3m_sphere_lorentz.npz",
                                  "fd3m_sphere.npz", "fd3m_sphere_pvoigt_iso.npz", "fcc_sphere_hkl0.npz",
                                  "hcp_sphere.npz", "ortho_sphere_iso.npz", "fcc_sphere_ordis1.npz",
                                  "bcc_sphere_zdir_lorentz.npz", "hex_sphere.npz", "hex_cyl_zdir.npz",
                                  "sq_cyl_zdir.npz", "rect_cyl_zdir.npz", "fcc_cyl_zdir.npz", "fcc_cyl_iso.npz",
                                  "hex_cyl_iso.npz", "sc_cube_iso.npz", "bcc_cube_iso.npz",
                                  "bct_cyl_zdir_lorentz.npz", "fcc_cyl_iso_pearson.npz", "sc_cube_iso_pvoigt.npz"])
def test_tables_match_reference(name):
    f = fixture(name)
    e = _host(params(name))
    for t in TABLES:
        np.testing.assert_array_equal(e.table(t), f[t], err_msg=t)
    s = e.scalars()
    assert s.norm == f["norm"][0] and s.orcase == f["orcase"][0]
    assert s.cosphic == f["cosphic"][0] and s.sinphic == f["sinphic"][0]
    if "latpar3" in f and int(params(name).get("LType", 12)) != 12:
        assert (s.peakmax1, s.peakmax2) == tuple(f["peakmax"])
        np.testing.assert_array_equal(e.table("latpar1")[6:], f["latpar1"][6:])
        np.testing.assert_array_equal(e.table("latpar2")[6:], f["latpar2"][6:])
        np.testing.assert_array_equal(e.table("latpar3").reshape(-1, 17)[1:, 1:14], f["latpar3"][1:, 1:14])
        sphno = {3: 2, 4: 2, 5: 4, 7: 1, 8: 2, 9: 2, 10: 1, 11: 2, 17: 8}.get(int(params(name)["LType"]), 1) if s.peakmax1 else 1
        assert s.sphno_cubevol == f["ucvol"][0] * sphno
    if "carr11pm" in f:
        np.testing.assert_array_equal(e.table("carr11pm").reshape(131, 131)[:121, :121], f["carr11pm"])


@pytest.mark.parametrize("name", ["cube_a.npz", "cube_b.npz"])
def test_cube_tables_match_reference(name):
    """Isotropic cubes: P(q) and <F>^2 coefficients, limq and the Porod constant."""
    f = fixture(name)
    e = _host(params(name))
    for t in TABLES:
        np.testing.assert_array_equal(e.table(t), f[t], err_msg=t)
    s = e.scalars()
    assert s.por == f["por"][0] and s.norm == f["norm"][0] and s.cs == 0
    # the z-only sums are cached per thread: a second preparation at another size and the
    # same width, and one after a different width, give the same tables as a fresh one
    p = params(name)
    for other in (dict(p, EditRadius=p["EditRadius"] * 1.5), dict(p, EditSigma=0.3), p):
        e2 = _host(other)
        np.testing.assert_array_equal(e2.table("carr4f"), _host_fresh_carr4f(other))
    np.testing.assert_array_equal(e2.table("carr4f"), f["carr4f"])


def _host_fresh_carr4f(p):
    """carr4f computed in a fresh process (empty per-thread cache)."""
    import json
    import subprocess
    import sys
    code = ("import json,sys,numpy as np;from crystalscatter_amd import SasCalc;"
            "e=SasCalc(device=-1,params=json.loads(sys.argv[1]));e.set('ignore_cbs',1);e.prepare();"
            "sys.stdout.write(json.dumps(e.table('carr4f').tolist()))")
    out = subprocess.run([sys.executable, "-c", code, json.dumps(p)], check=True, capture_output=True, text=True,
                         cwd=__import__("os").path.dirname(__import__("os").path.dirname(__file__)))
    return np.array(json.loads(out.stdout.replace("Infinity", "1e999").replace("NaN", "null")), dtype=float)


def test_bct_sets_match_reference():
    f = fixture("cfg4_bct.npz")
    for k, p in enumerate(manifest()["fixtures"]["cfg4_bct.npz"]["params"]):
        e = _host(p)
        for t in TABLES:
            np.testing.assert_array_equal(e.table(t), f["%s_%d" % (t, k)], err_msg="%s/%d" % (t, k))
        np.testing.assert_array_equal(e.table("latpar3").reshape(-1, 17)[1:, 1:14], f["latpar3_%d" % k][1:, 1:14])
        assert e.scalars().sphno_cubevol == f["ucvol_%d" % k][0] * 2


def test_key_table_defaults_and_unknown_key():
    e = SasCalc(device=-1)
    # defaults of the reference key table (Source/sc_calc_generic.cpp:36-138)
    assert e.get("uca") == 30 and e.get("EditAzi") == 180 and e.get("I0") == 10000
    assert e.get("GridPoints") == 64 and e.get("HKLmax") == 5 and e.get("SigmaL") == 0.06
    with pytest.raises(KeyError):
        e.set("NoSuchKey", 1)


def test_parameter_checks_like_the_reference():
    p = params("cfg3_fcc_g32.npz")
    e = SasCalc(device=-1, params=dict(p, uca=0.005))
    with pytest.raises(CsxError, match="uca<=0.01"):
        e.prepare()
    # cylinder + BCT is rejected by isCbsValid unless the check is skipped
    p4 = manifest()["fixtures"]["cfg4_bct.npz"]["params"][0]
    e = SasCalc(device=-1, params=p4)
    with pytest.raises(CsxError, match="INVALID CBParticle"):
        e.prepare()
    e.set("ignore_cbs", 1)
    e.prepare()


def test_unsupported_combination_is_an_error():
    e = SasCalc(device=-1, params=dict(LType=12, ComboBoxParticle=5, Ordis=7))  # ellipsoid
    with pytest.raises(CsxError):
        e.prepare()
    e = SasCalc(device=-1, params=dict(LType=12, ComboBoxParticle=4, Ordis=6, ComboBoxInterior=1))  # core/shell
    with pytest.raises(CsxError, match="homogeneous"):                                           # oriented cube
        e.prepare()
    e = SasCalc(device=-1, params=dict(LType=12, ComboBoxParticle=1, Ordis=7, EditRadius=5, Length=8))  # L/R < 2
    with pytest.raises(CsxError, match="L/R"):
        e.prepare()
