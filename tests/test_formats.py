"""On-disk formats (crystalscatter_amd/formats.py) against the reference's own files in
Documentation/ (Vitess-Sphere.*): the golden .dat is rebuilt byte for byte, the .csv writer
reproduces the reference's first lines, the .nofit and .ini readers read what they hold."""
import os

import numpy as np

from helpers import fixture, manifest

from crystalscatter_amd import SasCalc
from crystalscatter_amd import formats as fm

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    f = fixture("vitess_sphere.npz")
    head = bytes(f["header"]).decode()
    bounds = tuple(int(v) for v in head.split())
    return head, bounds, f["image"]


def test_dat_writer_rebuilds_the_reference_file(tmp_path):
    head, bounds, img = _golden()
    p = tmp_path / "v.dat"
    fm.write_dat(p, img, bounds)
    raw = p.read_bytes()
    assert raw[:19].decode() == head and len(raw) == 19 + 8 * img.size
    assert raw[19:] == np.ascontiguousarray(img, dtype="<f8").tobytes()
    b, back = fm.read_dat(p)
    assert b == bounds and np.array_equal(back, img)


def test_dat_header_widths():
    # fields are right-aligned in 4 characters; wider numbers widen the header
    assert "%4d %4d %4d %4d" % (-8, 8, 0, 16) == "  -8    8    0   16"


def test_csv_writer_matches_the_reference_text(tmp_path):
    _, bounds, img = _golden()
    p = tmp_path / "v.csv"
    fm.write_csv(p, img, bounds)
    ref = open(os.path.join(GOLD, "vitess_sphere_head.csv")).read().splitlines()
    got = p.read_text().splitlines()
    assert got[:3] == ref[:3]
    b, back = fm.read_csv(p)
    assert b == bounds and np.allclose(back, img, rtol=1e-5)


def test_nofit_reader():
    rects = fm.read_nofit(os.path.join(GOLD, "vitess_sphere.dat.nofit"))
    assert rects == [(0, 0, -1, -1)] * 4
    assert fm.read_nofit("/nonexistent.nofit") == [(-1, -1, -1, -1)] * 4


def test_ini_reader_vitess():
    params, flags, _ = fm.read_ini(os.path.join(GOLD, "vitess_sphere.ini"))
    m = manifest()["fixtures"]["vitess_sphere.npz"]["params"]
    for k in ("LType", "ComboBoxParticle", "Ordis", "EditRadius", "EditSigma", "EditQmax", "I0", "Base",
              "EditDet", "EditWavelength"):
        assert params[k] == m[k], k
    assert params["EditPixelX"] == 2.5 and params["EditPixelY"] == 2.5   # metres -> mm
    assert params["GridPoints"] == 128 and params["HKLmax"] == 1
    assert params["RadioButtonQ4"] == 1 and params["ExpandImage"] == 0   # Q4 forces no expansion
    assert params["CheckBoxTwinned"] == 0                                # LType 12 override
    assert flags["EditSigma"] is False and flags["uca"] is True          # FITFLAG_...=00 / =10
    assert "EditQmaxPreset" not in params                                # radio group skipped


def test_engine_takes_the_ini():
    e = SasCalc(device=-1)
    flags, unused = e.load_ini(os.path.join(GOLD, "vitess_sphere.ini"))
    assert e.get("EditRadius") == 16.7 and e.get("GridPoints") == 128
    assert "SigX" in unused                                              # not an engine key
