"""The two-pass lattice render (mixed precision): the hot pass and the fix pass that finishes
the pixels needing an FP64 path (conditioning guard, near-one pairs with a centre beam), and
the per-pass timing ring.  Every image is checked against the FP64 reference mode, which is
pinned to the reference elsewhere."""
import ctypes

import numpy as np
import pytest

from helpers import CFG3, manifest, rel_err

from crystalscatter_amd import SasCalc, _lib

pytestmark = pytest.mark.gpu


def _render_stats(p, prec):
    import torch
    e = SasCalc(device=0, params=p)
    e.set("ignore_cbs", 1)
    e.prepare()
    L = _lib.lib()
    ctx = L.csx_engine_ctx(e._h)
    s = e.scalars()
    g = _lib.Grid(s.xmin, s.xmax, s.ymin, s.ymax)
    out = torch.empty((s.xmax - s.xmin) * (s.ymax - s.ymin), dtype=torch.float64, device="cuda")
    st = (ctypes.c_uint64 * 5)()
    assert L.csx_render_stats(ctx, ctypes.byref(g), {"ref64": 0, "mixed": 1}[prec], ctypes.c_void_p(out.data_ptr()),
                              st) == 0
    torch.cuda.synchronize()
    return out.cpu().numpy(), [int(v) for v in st]


CASES = {
    # centre beam: pixel q from the detector geometry, near-one pairs left to the fix pass
    "fcc_centre_beam": dict(CFG3, GridPoints=48, CenterBeam=1, EditPixelX=2.5, EditPixelY=2.5, EditPixelNoX=96,
                            EditPixelNoY=96, EditDet=10, EditWavelength=0.154),
    # narrow peaks: the conditioning guard re-sums pixels between them in FP64
    "fcc_narrow": dict(CFG3, GridPoints=64, EditDomainSize=2000, EditAzi=20),
    "bct_fit_set": dict(manifest()["fixtures"]["cfg4_bct.npz"]["params"][0], GridPoints=64),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_two_pass_against_fp64(name):
    p = CASES[name]
    mixed, st = _render_stats(p, "mixed")
    ref, _ = _render_stats(p, "ref64")
    r = rel_err(mixed, ref)
    print("%s: max rel %.3e, pair evals %d, guard re-sums %d, near-one %d, full sums %d" %
          (name, r.max(), st[0], st[1], st[2], st[4]), flush=True)
    assert np.isfinite(mixed).all()
    assert r.max() <= 1e-4


def test_fix_pass_is_exercised():
    """At least one case sends pixels through the fix pass (guard re-sums or centre-beam
    near-one pairs), so the test above covers it."""
    total = 0
    for name in ("fcc_centre_beam", "fcc_narrow"):
        _, st = _render_stats(CASES[name], "mixed")
        total += st[1] + st[4] + (st[2] if name == "fcc_centre_beam" else 0)
    assert total > 0


def test_pass_timing_ring():
    import torch
    e = SasCalc(device=0, params=dict(CFG3, GridPoints=64))
    e.prepare()
    L = _lib.lib()
    ctx = L.csx_engine_ctx(e._h)
    out = torch.empty(128 * 128, dtype=torch.float64, device="cuda")
    assert L.csx_set_timing(ctx, 1) == 0
    for _ in range(3):
        e.doCalculationDevice(out.data_ptr(), "mixed")
    ms = (ctypes.c_double * 8)()
    n = L.csx_pass_times(ctx, 8, ms)
    assert L.csx_set_timing(ctx, 0) == 0
    assert n == 3
    assert all(0.0 < ms[k] < 1000.0 for k in range(n))
    assert L.csx_pass_times(ctx, 8, ms) == 0  # the ring restarts with timing
