"""Surrogate network of the fits over the C ABI (csx_mlp_*, include/csx.h).

The network maps a normalised parameter set and a pixel's (qx, qy) to the normalised
log10-intensity of that pixel.  Forward, backward and Adam run in the library's BF16 MFMA
kernels; torch provides device memory and the stream.  There is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import CsxError


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


class MLP:
    """One network on one device; buffers for up to max_batch rows per call."""

    NPAR, NFREQ = _lib.MLP_NPAR, _lib.MLP_NFREQ

    def __init__(self, device=0, max_batch=65536, seed=1):
        import torch
        self._L = _lib.lib()
        h = ctypes.c_void_p()
        rc = self._L.csx_mlp_create(int(device), int(max_batch), int(seed), ctypes.byref(h))
        if rc != 0:
            raise CsxError("csx_mlp_create failed (%d)" % rc)
        self._h = h
        self.device = torch.device("cuda", device)
        self.max_batch = max_batch
        self.nparams = int(self._L.csx_mlp_nparams())
        self._stream = torch.cuda.Stream(device=self.device)
        self._chk(self._L.csx_mlp_set_stream(self._h, ctypes.c_void_p(self._stream.cuda_stream)), "set_stream")

    def close(self):
        if getattr(self, "_h", None):
            self._L.csx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            b = ctypes.create_string_buffer(512)
            self._L.csx_mlp_last_error(self._h, b, 512)
            raise CsxError("csx_mlp_%s failed (%d): %s" % (what, rc, b.value.decode(errors="replace")))

    @property
    def stream(self):
        return self._stream

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        self._chk(self._L.csx_mlp_get_params(self._h, p.ctypes.data_as(_lib._PF)), "get_params")
        return p

    def set_params(self, p):
        pa = np.ascontiguousarray(p, dtype=np.float32)
        self._chk(self._L.csx_mlp_set_params(self._h, pa.ctypes.data_as(_lib._PF)), "set_params")

    def get_features(self):
        """(F [NFREQ, 3], pidx [NFREQ], affine [NPAR, 2]) of the input phases."""
        F = np.empty((self.NFREQ, 3), dtype=np.float32)
        pi = np.empty(self.NFREQ, dtype=np.int32)
        af = np.empty((self.NPAR, 2), dtype=np.float32)
        self._chk(self._L.csx_mlp_get_features(self._h, F.ctypes.data_as(_lib._PF), pi.ctypes.data_as(_lib._PI32),
                                               af.ctypes.data_as(_lib._PF)), "get_features")
        return F, pi, af

    def set_features(self, F, pidx, affine):
        F = np.ascontiguousarray(F, dtype=np.float32).reshape(self.NFREQ, 3)
        pi = np.ascontiguousarray(pidx, dtype=np.int32).reshape(self.NFREQ)
        af = np.ascontiguousarray(affine, dtype=np.float32).reshape(self.NPAR, 2)
        self._chk(self._L.csx_mlp_set_features(self._h, F.ctypes.data_as(_lib._PF), pi.ctypes.data_as(_lib._PI32),
                                               af.ctypes.data_as(_lib._PF)), "set_features")

    def get_grads(self):
        g = np.empty(self.nparams, dtype=np.float32)
        self._chk(self._L.csx_mlp_get_grads(self._h, g.ctypes.data_as(_lib._PF)), "get_grads")
        return g

    def _rows(self, par, sets, q):
        n = int(q.shape[0])
        if not (1 <= n <= self.max_batch):
            raise ValueError("1..max_batch rows per call")
        for t, dt in ((par, "float32"), (sets, "int32"), (q, "float32")):
            if t.device != self.device or str(t.dtype) != "torch." + dt or not t.is_contiguous():
                raise ValueError("inputs must be contiguous %s tensors on %s" % (dt, self.device))
        if par.dim() != 2 or par.shape[1] != self.NPAR or q.shape[-1] != 2 or sets.shape[0] != n:
            raise ValueError("par is (nsets, 6), sets (n,), q (n, 2)")
        return n, int(par.shape[0])

    def _in_use(self, *ts):
        """The model's stream reads/writes these torch tensors: keep the caching allocator from
        handing their memory to the caller's stream before that work is done."""
        for t in ts:
            t.record_stream(self._stream)

    def forward(self, par, sets, q, out=None):
        """y[i] = network(par[sets[i]], q[i]); device tensors; returns a float32 tensor."""
        import torch
        n, ns = self._rows(par, sets, q)
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=self.device)
        self._stream.wait_stream(torch.cuda.current_stream(self.device))
        self._in_use(par, sets, q, out)
        self._chk(self._L.csx_mlp_forward(self._h, n, ns, _p(par), _p(sets), _p(q), _p(out)), "forward")
        torch.cuda.current_stream(self.device).wait_stream(self._stream)
        return out

    def predict(self, par, sets, q, out=None):
        """forward() bit for bit from the fused inference kernel, any number of rows."""
        import torch
        n = int(q.shape[0])
        if n < 1:
            raise ValueError("at least one row")
        for t, dt in ((par, "float32"), (sets, "int32"), (q, "float32")):
            if t.device != self.device or str(t.dtype) != "torch." + dt or not t.is_contiguous():
                raise ValueError("inputs must be contiguous %s tensors on %s" % (dt, self.device))
        if par.dim() != 2 or par.shape[1] != self.NPAR or q.shape[-1] != 2 or sets.shape[0] != n:
            raise ValueError("par is (nsets, 6), sets (n,), q (n, 2)")
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=self.device)
        self._stream.wait_stream(torch.cuda.current_stream(self.device))
        self._in_use(par, sets, q, out)
        self._chk(self._L.csx_mlp_predict(self._h, n, int(par.shape[0]), _p(par), _p(sets), _p(q), _p(out)), "predict")
        torch.cuda.current_stream(self.device).wait_stream(self._stream)
        return out

    def _target(self, target, n):
        if (target.device != self.device or str(target.dtype) != "torch.float32" or not target.is_contiguous()
                or target.numel() != n):
            raise ValueError("target must be a contiguous float32 tensor of n values on %s" % self.device)

    def gradient(self, par, sets, q, target):
        """Loss 1/n sum (y - target)^2 (synchronous) and the gradient in the model."""
        import torch
        n, ns = self._rows(par, sets, q)
        self._target(target, n)
        loss = ctypes.c_float()
        self._stream.wait_stream(torch.cuda.current_stream(self.device))
        self._in_use(par, sets, q, target)
        self._chk(self._L.csx_mlp_gradient(self._h, n, ns, _p(par), _p(sets), _p(q), _p(target), ctypes.byref(loss)),
                  "gradient")
        return loss.value

    def train_step(self, par, sets, q, target, lr, want_loss=False):
        import torch
        n, ns = self._rows(par, sets, q)
        self._target(target, n)
        loss = ctypes.c_float()
        self._stream.wait_stream(torch.cuda.current_stream(self.device))
        self._in_use(par, sets, q, target)
        self._chk(self._L.csx_mlp_train_step(self._h, n, ns, _p(par), _p(sets), _p(q), _p(target), ctypes.c_float(lr),
                                             ctypes.byref(loss) if want_loss else None), "train_step")
        return loss.value if want_loss else None

    def synchronize(self):
        self._chk(self._L.csx_mlp_synchronize(self._h), "synchronize")


# ---- the fits' parameter space ---------------------------------------------------------
# Free parameters of the BCT-cylinder fits and their box (the fit driver's bounds): radius R,
# length L, a (= uca = ucb), c (= ucc), domain size and angular peak width EditAzi.
FIT_KEYS = ["EditRadius", "Length", "uca,ucb", "ucc", "EditDomainSize", "EditAzi"]
FIT_LO = np.array([2.0, 10.0, 15.0, 15.0, 100.0, 2.0])
FIT_HI = np.array([6.0, 60.0, 40.0, 80.0, 400.0, 10.0])
# Fixed keys of the BCT-cylinder combination (key 0x01020806): z-oriented cylinders on a BCT
# lattice with Gaussian peaks, Debye-Scherrer widths, q in [-1, 1] nm^-1.
BCT_BASE = dict(LType=8, ComboBoxParticle=1, Ordis=6, ComboBoxPeak=1, ComboBoxInterior=0, theta=0, phi=0,
                EditSigma=0.06, SigmaL=0.06, RadButDebyeScherrer=1, EditQmax=1, I0=1, Base=0, HKLmax=5)


def physics_features(seed=5):
    """Input phases for the BCT-cylinder fits.  Bragg rings sit at |q| proportional to 1/a
    and 1/c, the cylinders' form factor oscillates in |q| R: eight harmonics of |q| a, of
    |q| c and of |q| R, plus eight plain Fourier features of (qx, qy) for the anisotropy."""
    rng = np.random.default_rng(seed)
    F = np.zeros((32, 3))
    pidx = np.full(32, -1, dtype=np.int32)
    lattice = [0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 2.0, 2.5]
    for k, m in enumerate(lattice):
        F[k, 2], pidx[k] = m, 2           # |q| a
        F[8 + k, 2], pidx[8 + k] = m, 3   # |q| c
    for k, m in enumerate([0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 4.0, 5.0]):
        F[16 + k, 2], pidx[16 + k] = m, 0  # |q| R
    for k in range(8):
        mag, ang = 0.5 * 32.0 ** rng.uniform(), 2 * np.pi * rng.uniform()
        F[24 + k, :2] = mag * np.cos(ang), mag * np.sin(ang)
    affine = np.stack([0.5 * (FIT_LO + FIT_HI), 0.5 * (FIT_HI - FIT_LO)], axis=1)
    return F, pidx, affine


def normalize(x):
    """Fit parameters -> [-1, 1]^6 (the network's parameter inputs)."""
    return 2.0 * (np.asarray(x, dtype=np.float64) - FIT_LO) / (FIT_HI - FIT_LO) - 1.0


def draw_parameters(rng, n):
    """n parameter vectors drawn like the fit workload: R, L, a, c/a, domain, width uniform."""
    R, L = rng.uniform(2, 6, n), rng.uniform(10, 60, n)
    a, ca = rng.uniform(15, 40, n), rng.uniform(1, 2, n)
    dom, azi = rng.uniform(100, 400, n), rng.uniform(2, 10, n)
    return np.stack([R, L, a, a * ca, dom, azi], axis=1)


def params_dict(x, grid_points):
    R, L, a, c, dom, azi = (float(v) for v in x)
    return dict(BCT_BASE, EditRadius=R, Length=L, uca=a, ucb=a, ucc=c, EditDomainSize=dom, EditAzi=azi,
                GridPoints=grid_points)


def pixel_q(engine, idx):
    """(qx, qy) of pixels at IDX positions of an engine's image (midpoint grid):
    ihex = xmin + idx % W (-> qy), i = ymin + idx // W (-> qx), q = qmax * index / zmax."""
    s = engine.scalars()
    W = s.xmax - s.xmin
    ihex = s.xmin + idx % W
    i = s.ymin + idx // W
    f = s.qmax / s.zmax
    return i * f, ihex * f


# Decades below a pattern's maximum that the targets resolve.  The fits see counts up to
# ~1e4 at the brightest pixel and drop pixels without counts, so intensities more than ~5
# decades below the maximum never enter an objective; targets are floored at 6 decades.
FLOOR_DECADES = 6.0
EVAL_DECADES = 5.0


class TrainingSet:
    """Pixels of simulated patterns on the device: set index, (qx, qy), normalised log10 I
    (floored at FLOOR_DECADES below the pattern's maximum), and whether the pixel lies within
    EVAL_DECADES of the maximum (the range the fits' objective sees)."""

    def __init__(self, par, sets, q, t, inrange, mu, sigma, npatterns, px_per_pattern):
        self.par, self.sets, self.q, self.t, self.inrange = par, sets, q, t, inrange
        self.mu, self.sigma = mu, sigma
        self.npatterns, self.px_per_pattern = npatterns, px_per_pattern

    def __len__(self):
        return int(self.t.numel())


def generate(npatterns, grid_points=128, px_per_pattern=2048, seed=7, device=0, mu=None, sigma=None,
             chunk_bytes=1 << 30):
    """Render npatterns BCT-cylinder patterns with the exact kernel (mixed precision, I0 = 1)
    at parameters drawn like the fit workload, and keep px_per_pattern random pixels of each
    with a positive, finite intensity (the first that many of 2*px_per_pattern draws).
    Targets are (max(log10 I, log10 I_max - FLOOR_DECADES) - mu) / sigma; mu and sigma come
    from the data unless given.  Patterns go in chunks of chunk_bytes of images: one native
    staging call (lattice tables built on the device) and one batched render per chunk."""
    import torch
    from . import Context, SasCalc, _lib, _prec
    rng = np.random.default_rng(seed)
    xs = draw_parameters(rng, npatterns)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    sets, qs, ts, rs = [], [], [], []
    e = SasCalc(device=-1, params=params_dict(xs[0], grid_points))
    e.set("ignore_cbs", 1)
    e.prepare()
    x0, x1, y0, y1 = e.block()
    npx = (x1 - x0) * (y1 - y0)
    chunk = max(1, min(npatterns, chunk_bytes // (8 * npx)))
    buf = torch.empty((chunk, npx), dtype=torch.float64, device="cuda")
    ctx = Context(device)
    L = _lib.lib()
    dev = torch.device("cuda")
    for k0 in range(0, npatterns, chunk):
        kn = min(chunk, npatterns - k0)
        ctx._chk(L.csx_fit_batch_begin(ctx._h), "csx_fit_batch_begin")
        e.stage_points(ctx, FIT_KEYS, xs[k0:k0 + kn], buf.data_ptr(), npx)
        ctx._chk(L.csx_render_batch_run(ctx._h, _prec("mixed")), "csx_render_batch_run")
        img = buf[:kn]
        top = torch.log10(torch.nan_to_num(img, nan=0.0, posinf=0.0).amax(dim=1))
        idx = torch.randint(0, npx, (kn, 2 * px_per_pattern), device=dev, generator=gen)
        v = torch.gather(img, 1, idx)
        ok = torch.isfinite(v) & (v > 0)
        keep = ok & (torch.cumsum(ok, dim=1) <= px_per_pattern)
        row = torch.arange(kn, device=dev, dtype=torch.int64).unsqueeze(1).expand_as(idx)[keep]
        idx, v = idx[keep], v[keep]
        qx, qy = pixel_q(e, idx.to(torch.float64))
        sets.append((row + k0).to(torch.int32))
        qs.append(torch.stack([qx, qy], dim=1).to(torch.float32))
        lv = torch.log10(v)
        ts.append(torch.maximum(lv, top[row] - FLOOR_DECADES))
        rs.append(lv >= top[row] - EVAL_DECADES)
        torch.cuda.current_stream().synchronize()  # the chunk buffer is reused by the next batch
    e.close()
    ctx.close()
    t = torch.cat(ts)
    if mu is None:
        mu, sigma = float(t.mean()), float(t.std())
    par = torch.from_numpy(normalize(xs).astype(np.float32)).cuda()
    return TrainingSet(par, torch.cat(sets), torch.cat(qs).contiguous(), ((t - mu) / sigma).to(torch.float32),
                       torch.cat(rs), mu, sigma, npatterns, px_per_pattern)


def train(net, data, steps, batch=65536, lr=2e-3, lr_end=1e-4, seed=0, log_every=0):
    """Adam over random minibatches of the training set; cosine decay from lr to lr_end.
    Returns the losses at the logged steps (synchronous only there)."""
    import torch
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    n = len(data)
    hist = []
    for s in range(steps):
        idx = torch.randint(0, n, (batch,), device="cuda", generator=gen)
        rate = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * s / max(1, steps - 1)))
        want = bool(log_every) and (s % log_every == 0 or s == steps - 1)
        loss = net.train_step(data.par, data.sets[idx].contiguous(), data.q[idx].contiguous(), data.t[idx].contiguous(),
                              float(rate), want_loss=want)
        if want:
            hist.append((s, loss))
    return hist


def predict_log10(net, data_par, sets, q, mu, sigma):
    """log10 I predicted for rows (device tensors), in one fused inference launch."""
    return net.predict(data_par, sets.contiguous(), q.contiguous()) * sigma + mu


class FitPixels:
    """A subsample of one fit's usable pixels (data > 0, the pixels the objective counts):
    (qx, qy) and log10(data / scale), where scale is the fit's known count scale (I0)."""

    def __init__(self, engine, data, scale, npx, rng):
        import torch
        d = data if data.dim() == 1 else data.reshape(-1)
        good = torch.nonzero(d > 0).squeeze(1)
        pick = torch.from_numpy(rng.choice(good.numel(), min(npx, good.numel()), replace=False)).cuda()
        idx = good[pick]
        qx, qy = pixel_q(engine, idx.to(torch.float64))
        self.q = torch.stack([qx, qy], dim=1).to(torch.float32).contiguous()
        self.logd = (torch.log10(d[idx]) - np.log10(scale)).to(torch.float32)
        self.n = int(idx.numel())


def surrogate_objectives(net, fits_px, cands, mu, sigma):
    """Surrogate FQS (mean over each fit's pixel subsample of (log10 I_pred - log10 I_data)^2)
    for cands[f] (k x 6 parameter vectors, the same k for every fit) of every fit f.  Rows are
    built with whole-tensor operations: fit-major, then candidate, then pixel.  Returns an
    (nfits, k) array."""
    import torch
    nf = len(fits_px)
    C = np.asarray(cands, dtype=np.float64)
    k = C.shape[1]
    npx = min(fp.n for fp in fits_px)
    Q = torch.stack([fp.q[:npx] for fp in fits_px])                    # nf x npx x 2
    T = torch.stack([fp.logd[:npx] for fp in fits_px])                 # nf x npx
    par = torch.from_numpy(normalize(C.reshape(-1, 6)).astype(np.float32)).cuda()
    sets = torch.arange(nf * k, dtype=torch.int32, device="cuda").repeat_interleave(npx)
    q = Q[:, None].expand(nf, k, npx, 2).reshape(-1, 2).contiguous()
    pred = predict_log10(net, par, sets, q, mu, sigma).reshape(nf, k, npx)
    r2 = (pred - T[:, None, :]) ** 2
    return r2.mean(dim=2).double().cpu().numpy()


def propose_starts(net, fits_px, mu, sigma, x_init, rng, n_global=128, n_keep=4, n_local=32, shrink=0.1, top=8):
    """Candidate starting points for the exact fits, ranked by the surrogate objective: the
    best of n_global draws over the box (plus x_init), then n_local draws in a box of +-shrink
    of the range around each of the n_keep best.  Returns (candidates [nfits, top, 6],
    their surrogate objectives [nfits, top]); the exact objective picks among them."""
    nf = len(fits_px)
    cand = np.stack([np.vstack([x_init[f][None, :], draw_parameters(rng, n_global)]) for f in range(nf)])
    cand = np.clip(cand, FIT_LO, FIT_HI)
    vals = surrogate_objectives(net, fits_px, cand, mu, sigma)
    best = np.take_along_axis(cand, np.argsort(vals, axis=1)[:, :n_keep, None], axis=1)   # nf x keep x 6
    span = shrink * (FIT_HI - FIT_LO)
    pts = best.repeat(n_local, axis=1) + rng.uniform(-1, 1, (nf, n_keep * n_local, 6)) * span
    loc = np.clip(np.concatenate([best, pts], axis=1), FIT_LO, FIT_HI)
    lv = surrogate_objectives(net, fits_px, loc, mu, sigma)
    order = np.argsort(lv, axis=1)[:, :top]
    return np.take_along_axis(loc, order[:, :, None], axis=1), np.take_along_axis(lv, order, axis=1)


class SurrogateScreen:
    """Screen of the batched simplex's posted points (nm_batch(screen=...)): the network's
    objective at each point, the mean over its fit's pixel subsample (FitPixels) of the squared
    log10 residual.  The fits' first six parameter entries are FIT_KEYS; every point of a
    round goes through the network in one pass."""

    def __init__(self, net, fits_px, mu, sigma):
        import torch
        self.net, self.mu, self.sigma = net, mu, sigma
        self.npx = min(fp.n for fp in fits_px)
        self.Q = torch.stack([fp.q[:self.npx] for fp in fits_px])      # nf x npx x 2
        self.T = torch.stack([fp.logd[:self.npx] for fp in fits_px])   # nf x npx
        self.calls = 0
        self.points = 0

    def __call__(self, fit_ids, X):
        import torch
        P = len(fit_ids)
        f = torch.from_numpy(np.asarray(fit_ids, dtype=np.int64)).cuda()
        par = torch.from_numpy(normalize(np.asarray(X)[:, :6]).astype(np.float32)).cuda()
        sets = torch.arange(P, dtype=torch.int32, device="cuda").repeat_interleave(self.npx)
        q = self.Q[f].reshape(-1, 2).contiguous()
        pred = predict_log10(self.net, par, sets, q, self.mu, self.sigma).reshape(P, self.npx)
        self.calls += 1
        self.points += P
        return ((pred - self.T[f]) ** 2).mean(dim=1).double().cpu().numpy()
