"""csx_fit_fqs_noalloc (the objective with its scratch in a static device array) against
csx_fit_fqs, through the C ABI: sums, counts and per-pixel terms bit for bit, on two calls in
a row on one context (the second with a smaller fit array, on a new upload), in both
precisions; and the size limit turned into an error on the host."""
import ctypes

import numpy as np
import pytest

from helpers import manifest

from crystalscatter_amd import SasCalc, _lib

pytestmark = pytest.mark.gpu


def _mask(bstop=0, border=0):
    m = _lib.Mask()
    for i in range(16):
        m.rects[i] = -1
    m.rects[0:4] = [3, 4, 9, 12]
    m.bstop, m.border = bstop, border
    return m


def _objective(L, fn, ctx, w, h, mask, prec, d_data, d_fqs):
    s, cnt, nan = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
    rc = fn(ctx, w, h, ctypes.byref(mask), prec, ctypes.c_void_p(d_data.data_ptr()), ctypes.byref(s),
            ctypes.byref(cnt), ctypes.byref(nan), ctypes.c_void_p(d_fqs.data_ptr()))
    assert rc == 0, L
    return s.value, cnt.value, nan.value, d_fqs.cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1])  # CSX_PREC_REF64, CSX_PREC_MIXED
def test_noalloc_objective_equals_csx_fit_fqs(prec):
    import torch
    L = _lib.lib()
    p = manifest()["fixtures"]["cfg4_bct.npz"]["params"][1]
    e = SasCalc(device=0, params=dict(p, GridPoints=64, BeamPosX=10, BeamPosY=-6))
    e.set("ignore_cbs", 1)
    rng = np.random.default_rng(21)
    ctx = L.csx_engine_ctx(e._h)
    for n, bstop in ((64, 5), (24, 0)):  # 128x128, then 48x48 on a new upload
        e.set("GridPoints", n)
        e.prepare()
        w = 2 * n
        d_data = torch.tensor(rng.uniform(0.01, 10.0, w * w), dtype=torch.float64, device="cuda")
        f1 = torch.full((w * w,), -1.0, dtype=torch.float64, device="cuda")
        f2 = torch.full((w * w,), -2.0, dtype=torch.float64, device="cuda")
        mask = _mask(bstop=bstop, border=2)
        a = _objective(L, L.csx_fit_fqs, ctx, w, w, mask, prec, d_data, f1)
        b = _objective(L, L.csx_fit_fqs_noalloc, ctx, w, w, mask, prec, d_data, f2)
        print("n=%d prec=%d: fqs %.17g cnt %d / %.17g cnt %d" % (n, prec, a[0], a[1], b[0], b[1]), flush=True)
        assert a[1] > 0 and a[:3] == b[:3]
        np.testing.assert_array_equal(a[3], b[3])


def test_noalloc_rejects_more_than_its_scratch():
    import torch
    L = _lib.lib()
    e = SasCalc(device=0, params=dict(LType=12, ComboBoxParticle=0, Ordis=7, EditRadius=5, EditSigma=0.1,
                                      GridPoints=1040, EditQmax=2, I0=1, Base=0))
    e.prepare()
    ctx = L.csx_engine_ctx(e._h)
    d = torch.zeros(16, dtype=torch.float64, device="cuda")  # never read: the call is refused first
    s, cnt, nan = ctypes.c_double(), ctypes.c_int64(), ctypes.c_int64()
    rc = L.csx_fit_fqs_noalloc(ctx, 2080, 2080, ctypes.byref(_mask()), 1, ctypes.c_void_p(d.data_ptr()),
                               ctypes.byref(s), ctypes.byref(cnt), ctypes.byref(nan), None)
    assert rc == -1
    buf = ctypes.create_string_buffer(256)
    L.csx_last_error(ctx, buf, 256)
    assert b"16384" in buf.value
