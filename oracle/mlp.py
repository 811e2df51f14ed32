"""CPU oracle of the warm-start network and its features (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may use this
module; the product never does.

The reference has no learned component (its contact duals restart from zero
at every detection, src/solver/contact_solver.cu:243).  What this file
restates is the build's own definition of the network and its features, in
FP64 NumPy:
  * MLP 10 -> 128 -> 128 -> 128 -> 2 with SiLU, flat parameter layout
    W1[128][10] b1 W2[128][128] b2 W3[128][128] b3 W4[2][128] b4;
  * loss L = 1/(2n) sum |y - t|^2, its gradient by backpropagation;
  * Adam (Kingma and Ba, ICLR 2015; beta1 0.9, beta2 0.999, eps 1e-8);
  * optional BF16 rounding at the points where the device kernels round
    (products read BF16 operands, accumulate in FP32), to compare bit-near;
  * the 10 per-contact features.
Pinned by: central finite differences of the loss (gradient), the closed form
of Adam's first step (w - lr g / (|g| + eps)), and hand-computed features of a
single contact (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 10, 128, 2
OW1 = 0
OB1 = OW1 + H * IN
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H * H
OW4 = OB3 + H
OB4 = OW4 + OUT * H
NPARAM = OB4 + OUT


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    if nan.any():
        out[nan] = np.nan
    return out


def unpack(w: np.ndarray) -> list[tuple[np.ndarray, np.ndarray]]:
    w = np.asarray(w, np.float64)
    return [(w[OW1:OB1].reshape(H, IN), w[OB1:OW2]), (w[OW2:OB2].reshape(H, H), w[OB2:OW3]),
            (w[OW3:OB3].reshape(H, H), w[OB3:OW4]), (w[OW4:OB4].reshape(OUT, H), w[OB4:NPARAM])]


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def silu(a):
    return a / (1.0 + np.exp(-a))


def dsilu(a):
    s = 1.0 / (1.0 + np.exp(-a))
    return s * (1.0 + a * (1.0 - s))


def forward(w, x, emulate_bf16: bool = False):
    """Returns (y, cache).  With emulate_bf16 the operands of every product are
    rounded to BF16 as on the device (weights, inputs, hidden activations)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    L = unpack(w)
    h = r(x)
    cache = [h]
    pre = []
    for W, b in L[:3]:
        a = h @ r(W).T + b
        pre.append(a)
        h = r(silu(a))
        cache.append(h)
    W4, b4 = L[3]
    y = h @ r(W4).T + b4
    return y, (cache, pre)


def loss_and_grad(w, x, t, emulate_bf16: bool = False):
    """L = 1/(2n) sum |y - t|^2 and dL/dw in the flat layout."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    n = x.shape[0]
    y, (cache, pre) = forward(w, x, emulate_bf16)
    e = y - t
    L = unpack(w)
    loss = 0.5 * np.sum(e * e) / n
    g = r(e / n)                       # dL/dy
    grads = [None] * 4
    grads[3] = (g.T @ cache[3], g.sum(axis=0))
    dh = g @ r(L[3][0])
    for l in (2, 1, 0):
        da = r(dh * dsilu(pre[l]))
        grads[l] = (da.T @ cache[l], da.sum(axis=0))
        if l > 0:
            dh = da @ r(L[l][0])
    return loss, pack(grads)


class Adam:
    def __init__(self, w, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.w = np.asarray(w, np.float64).copy()
        self.m = np.zeros_like(self.w)
        self.v = np.zeros_like(self.w)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        self.w = self.w - self.lr * mh / (np.sqrt(vh) + self.eps)
        return self.w


def contact_features(h, dhat, mu, n_dyn, x, x0, Uc, p_static, vel, ints, reals, Kc, prev=None):
    """The 10 features of each contact at the start of a detection window.

    ints[c] = (type, prim0, prim1, v0, v1, v2, v3), reals[c] = (bary[4],
    normal[3], h_cN, dist) as the detector reports them; Kc[c] = K_c (3);
    p_static: velocities of the static vertices (p on [n_dyn, n_all));
    prev = (keys[k, 3] of the previous window, its duals[k, 3]) or None.
    Returns (features[nc, 10], tangents[nc, 3], match[nc])."""
    s = h / dhat
    nc = ints.shape[0]
    f = np.zeros((nc, 10))
    tans = np.zeros((nc, 3))
    match = -np.ones(nc, np.int64)
    lut = {}
    if prev is not None:
        for j, key in enumerate(map(tuple, np.asarray(prev[0], np.int64))):
            lut[key] = j
    for c in range(nc):
        vs = ints[c, 3:7]
        b = reals[c, 0:4]
        nn = reals[c, 4:7]
        u = np.array(Kc[c], np.float64)
        vr = np.zeros(3)
        for k in range(4):
            i = vs[k]
            pk = (x[i] - x0[i] + Uc[i]) / h if i < n_dyn else p_static[i]
            u += b[k] * pk
            vr += b[k] * vel[i]
        un = nn @ u
        ut = u - un * nn
        utn = np.linalg.norm(ut)
        if utn > 1e-30:
            t = ut / utn
        else:
            a = np.zeros(3)
            a[int(np.argmin(np.abs(nn)))] = 1.0
            cr = np.cross(nn, a)
            t = cr / np.linalg.norm(cr)
        vn = nn @ vr
        vt = vr - vn * nn
        j = lut.get((int(ints[c, 0]), int(ints[c, 1]), int(ints[c, 2])), -1)
        ln = lt = 0.0
        if j >= 0:
            lam = np.asarray(prev[1][j], np.float64)
            ln, lt = s * (nn @ lam), s * (t @ lam)
        f[c] = [reals[c, 7] / dhat, s * un, s * utn, s * vn, s * np.linalg.norm(vt), mu, ints[c, 0],
                1.0 if j >= 0 else 0.0, ln, lt]
        tans[c] = t
        match[c] = j
    return f, tans, match
