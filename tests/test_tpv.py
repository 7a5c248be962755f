"""Random variation for training data (tpvPerformRandom) on the host: the rand() sequence,
the varied parameters and the report, against the reference's own run (tests/golden/tpv_fcc.npz,
four FCC images, each after one tpvPerformRandom call, rand() from seed 1)."""
import ctypes

import numpy as np
import pytest

from helpers import fixture, manifest

from crystalscatter_amd import CsxError, Random, SasCalc

TPV_KEYS = ["EditRadius", "EditSigma", "uca", "ucb", "ucc", "I0", "Base", "width_zuf", "dw_zuf", "EditAzi", "phi"]


def test_random_is_the_c_library_sequence():
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    r = Random(1)
    assert [r.next() for _ in range(1000)] == [libc.rand() for _ in range(1000)]
    libc.srand(987654)
    z = Random(987654).zuf(500)
    ref = np.array([((libc.rand() / 2147483647 * 200.0) - 100.0) / 100.0 for _ in range(500)])
    assert np.array_equal(z, ref) and z.min() >= -1 and z.max() <= 1


def test_tpv_images_match_the_reference_run(oracle_lib):
    f = fixture("tpv_fcc.npz")
    m = manifest()["fixtures"]["tpv_fcc.npz"]
    e = SasCalc(device=-1, params=m["params"])
    e.set("ignore_cbs", 1)
    rng = Random(m["seed"])
    reports = []
    for k in range(m["n"]):
        reports.append(e.tpvPerformRandom(m["tpv"], rng))
        q = dict(m["params"])
        q.update({key: e.tpv_value(key) for key in TPV_KEYS})
        o = oracle_lib.OracleEngine(q)
        o.prepare()
        np.testing.assert_array_equal(o.image(nthreads=4), f["images"][k], err_msg="image %d" % k)
    assert "#".join(reports) + "#" == bytes(f["report"]).decode()


def test_tpv_semantics():
    p = dict(LType=12, ComboBoxParticle=1, Ordis=6, EditRadius=3, Length=40, I0=1, Base=0.5, GridPoints=16)
    e = SasCalc(device=-1, params=p)
    rng = Random(3)
    rep = e.tpvPerformRandom(["EditRadius=0.1", "Length=0,2", "ucb=1", "nosuch=3"], rng)
    assert [l.split(" = ")[0] for l in rep.split("\r\n") if l] == ["EditRadius", "Length", "ucb"]
    assert e.get("EditRadius") == 3 and e.tpv_value("EditRadius") != 3      # the set value is kept
    assert abs(e.tpv_value("Length") / 40 - 1) <= 0.2                          # decimal comma read
    assert e.tpv_value("ucb") == e.get("uca")
    first = e.tpv_value("EditRadius")
    e.tpvPerformRandom(["EditRadius=0.1"], rng)                                # varies the first value
    assert abs(e.tpv_value("EditRadius") / 3 - 1) <= 0.1 and e.tpv_value("EditRadius") != first
    assert e.tpv_value("Length") == 40                                         # not listed: as set
    e.set("EditRadius", 5)                                                     # a setter wins until the next call
    assert e.tpv_value("EditRadius") == 5
    assert e.tpvPerformRandom([], None) == "TPV: clear"
    assert e.tpv_value("EditRadius") == 5
    # disks: the length is kept above the diameter (cpu:1139-1140)
    d = SasCalc(device=-1, params=dict(p, ComboBoxParticle=2, Length=5, EditRadius=3))
    d.tpvPerformRandom(["EditRadius=0.01"], Random(1))
    assert d.tpv_value("Length") == 5 + 2 * d.tpv_value("EditRadius")
    with pytest.raises(CsxError):
        e.tpvPerformRandom(["EditRadius=abc"], rng)
