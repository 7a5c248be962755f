"""CPU oracle of the surrogate network (csx_mlp_*, include/csx.h).

TEST INFRASTRUCTURE ONLY: the checker for the GPU network's forward pass and gradients.
Only tests/ may import it; the product never does.

The network has no counterpart in the reference: it is the learned part of the fits.
Forward: x (70 inputs, zero-padded to 128) -> 4 x [z = W h + b, h = silu(z)] (256 wide) ->
y = w . h + b.  Loss 1/n sum (y - t)^2.  `bf16=True` rounds to bfloat16 at the points where
the GPU stores bfloat16 (inputs, weights, Z, H, dZ) and accumulates in float64, so the only
remaining differences to the GPU are its FP32 accumulation order and its fast exp/sincos.
The backward pass is pinned against central finite differences in float64
(tests/test_mlp_oracle.py).
"""
import numpy as np

NPAR, NFREQ, IN, IN_PAD, HID, LAYERS = 6, 32, 70, 128, 256, 4


def layer_in(l):
    return IN_PAD if l == 0 else HID


def offsets():
    """(w_off, b_off) per hidden layer, then (w_out, b_out), and the total count."""
    off, out = 0, []
    for l in range(LAYERS):
        w = off
        off += HID * layer_in(l)
        b = off
        off += HID
        out.append((w, b))
    wo = off
    off += HID
    bo = off
    off += 1
    return out, (wo, bo), off


def nparams():
    return offsets()[2]


def used_params():
    """Parameters that are not padding: 70 inputs instead of 128 in the first layer."""
    return HID * IN + HID + (LAYERS - 1) * (HID * HID + HID) + HID + 1


def to_bf16(a):
    """Round float64/float32 values to the nearest bfloat16 (ties to even), as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    if nan.any():
        out[nan] = np.nan
    return out


def unpack(p):
    lay, (wo, bo), n = offsets()
    p = np.asarray(p, dtype=np.float64)
    assert p.size == n
    Ws = [p[w:w + HID * layer_in(l)].reshape(HID, layer_in(l)) for l, (w, b) in enumerate(lay)]
    bs = [p[b:b + HID] for (w, b) in lay]
    return Ws, bs, p[wo:wo + HID], p[bo]


def features(par, sets, q, freq, pidx, affine):
    """Rows of inputs: par[sets[i]] (6), sin(a_f) (32), cos(a_f) (32), zeros, with phases
    a_f = 2 pi (F_f0 qx + F_f1 qy) + F_f2 |q| v_f, v_f = offset_j + scale_j par_j for
    j = pidx_f >= 0, else 1."""
    par = np.asarray(par, dtype=np.float64).reshape(-1, NPAR)
    q = np.asarray(q, dtype=np.float64).reshape(-1, 2)
    F = np.asarray(freq, dtype=np.float64).reshape(NFREQ, 3)
    pidx = np.asarray(pidx).reshape(NFREQ)
    affine = np.asarray(affine, dtype=np.float64).reshape(NPAR, 2)
    n = q.shape[0]
    P = par[np.asarray(sets)]
    X = np.zeros((n, IN_PAD))
    X[:, :NPAR] = P
    qa = np.sqrt(q[:, 0] ** 2 + q[:, 1] ** 2)
    a = 2 * np.pi * (q @ F[:, :2].T)
    for f in range(NFREQ):
        j = int(pidx[f])
        v = affine[j, 0] + affine[j, 1] * P[:, j] if j >= 0 else 1.0
        a[:, f] += F[f, 2] * qa * v
    X[:, NPAR:NPAR + NFREQ] = np.sin(a)
    X[:, NPAR + NFREQ:IN] = np.cos(a)
    return X


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, bf16=False):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs, wo, bo = unpack(p)
    h = r(X)
    cache = [h]
    zs = []
    for l in range(LAYERS):
        z = h @ r(Ws[l]).T + bs[l]
        hn = silu(z)
        zs.append(r(z))
        h = r(hn)
        cache.append(h)
    y = h @ wo + bo
    return y, (cache, zs)


def loss_and_grad(p, X, t, bf16=False):
    """Loss 1/n sum (y - t)^2 and its gradient (flat, like the parameter vector)."""
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs, wo, bo = unpack(p)
    y, (hs, zs) = forward(p, X, bf16)
    n = X.shape[0]
    d = y - np.asarray(t, dtype=np.float64)
    loss = float((d * d).sum() / n)
    dy = 2.0 * d / n
    lay, (wo_off, bo_off), npar = offsets()
    g = np.zeros(npar)
    g[wo_off:wo_off + HID] = hs[LAYERS].T @ dy
    g[bo_off] = dy.sum()
    dz = r(np.outer(dy, wo) * silu_grad(zs[LAYERS - 1]))
    for l in range(LAYERS - 1, -1, -1):
        w, b = lay[l]
        g[w:w + HID * layer_in(l)] = (dz.T @ hs[l]).ravel()
        g[b:b + HID] = dz.sum(axis=0)
        if l == 0:
            break
        dh = dz @ r(Ws[l])
        dz = r(dh * silu_grad(zs[l - 1]))
    return loss, g
