"""The batched objective's tuning table (csx_tune_*): set/get per combination and size class,
and the text format the autotuner writes (CPU only: no kernel runs)."""
from crystalscatter_amd import _lib

BCT, FCC = 0x01020806, 0x00020500


def test_set_get_and_bounds():
    L = _lib.lib()
    keep = [L.csx_tune_get(BCT, sc) for sc in range(_lib.TUNE_SIZES)]
    try:
        assert L.csx_tune_set(BCT, 2, 3) == 0
        assert L.csx_tune_get(BCT, 2) == 3
        assert L.csx_tune_set(BCT, -1, 1) == 0
        assert [L.csx_tune_get(BCT, sc) for sc in range(_lib.TUNE_SIZES)] == [1] * _lib.TUNE_SIZES
        assert L.csx_tune_set(BCT, 0, _lib.TUNE_VARIANTS) == -1
        assert L.csx_tune_set(BCT, _lib.TUNE_SIZES, 0) == -1
        assert L.csx_tune_set(0x12345678, 0, 0) == -1
    finally:
        for sc, v in enumerate(keep):
            L.csx_tune_set(BCT, sc, v)


def test_load_text_format(tmp_path):
    L = _lib.lib()
    keep = [L.csx_tune_get(FCC, sc) for sc in range(_lib.TUNE_SIZES)]
    f = tmp_path / "t.txt"
    f.write_text("# comment\n0x00020500 0 2   # 128x128: 1.0 0.9\n0x00020500 3 1\nbad line\n0x00020500 9 1\n")
    try:
        assert L.csx_tune_load(str(f).encode()) == 2
        assert L.csx_tune_get(FCC, 0) == 2 and L.csx_tune_get(FCC, 3) == 1
        assert L.csx_tune_load(str(tmp_path / "missing.txt").encode()) == -1
    finally:
        for sc, v in enumerate(keep):
            L.csx_tune_set(FCC, sc, v)
