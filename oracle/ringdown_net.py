"""CPU oracle of the ringdown network (test infrastructure only: the tests compare the HIP
training kernels with it; nothing in the product calls it).

Forward and backward of the network in include/oaring.h, in float64, with every matrix-product
operand rounded to BF16 (round to nearest even) exactly where the GPU kernel rounds it and every
stored intermediate rounded to FP32, so the GPU gradients must agree to FP32 accumulation error.
With ``rnd=False`` the same code is plain float64 calculus; the tests pin that version with
central finite differences, which pins the formulas the rounded version shares.

The network has no reference counterpart (the reference extrapolates tails only with the matrix
pencil, packages/oa-fdtd/src/oa_fdtd/tail.py:39-98).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

C, CIN, NL, NF, NH, NO, NP, LAST = 64, 32, 6, 512, 256, 256, 16, 8
OFF_WIN = 0
OFF_BIN = OFF_WIN + C * CIN
OFF_WL = OFF_BIN + C
LSTRIDE = C * 2 * C + C
OFF_W1 = OFF_WL + NL * LSTRIDE
OFF_B1 = OFF_W1 + NH * NF
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NO * NH
NPARAM = OFF_B2 + NO


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even) through the float32 bit pattern."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@dataclass
class Batch:
    U: np.ndarray      # (32, S*T)
    amp: np.ndarray    # (S,)
    kap: np.ndarray    # (S, 4, F) complex
    tgt: np.ndarray    # (S, 4, F) complex
    t0: np.ndarray     # (S,)
    omega: np.ndarray  # (F,)
    dt: float
    sig0: float
    wmax: float
    T: int


def _views(p: np.ndarray):
    win = p[OFF_WIN:OFF_BIN].reshape(C, CIN)
    bin_ = p[OFF_BIN:OFF_WL]
    layers = []
    for l in range(NL):
        o = OFF_WL + l * LSTRIDE
        layers.append((p[o:o + 2 * C * C].reshape(C, 2 * C), p[o + 2 * C * C:o + LSTRIDE]))
    w1 = p[OFF_W1:OFF_B1].reshape(NH, NF)
    b1 = p[OFF_B1:OFF_W2]
    w2 = p[OFF_W2:OFF_B2].reshape(NO, NH)
    b2 = p[OFF_B2:NPARAM]
    return win, bin_, layers, w1, b1, w2, b2


def _shift(x: np.ndarray, T: int, d: int) -> np.ndarray:
    """x(t - d) per sample (0 before the start); x is (rows, S*T)."""
    r, n = x.shape
    y = np.zeros_like(x).reshape(r, n // T, T)
    y[:, :, d:] = x.reshape(r, n // T, T)[:, :, :T - d]
    return y.reshape(r, n)


def _unshift(x: np.ndarray, T: int, d: int) -> np.ndarray:
    """x(t + d) per sample (0 past the end)."""
    r, n = x.shape
    y = np.zeros_like(x).reshape(r, n // T, T)
    y[:, :, :T - d] = x.reshape(r, n // T, T)[:, :, d:]
    return y.reshape(r, n)


def _feat(x6: np.ndarray, T: int) -> np.ndarray:
    S = x6.shape[1] // T
    last = x6.reshape(C, S, T)[:, :, T - LAST:]          # (C, S, LAST)
    return last.transpose(2, 0, 1).reshape(LAST * C, S)  # row j*64 + c


def _poles(o: np.ndarray, b: Batch):
    """sigma, omega, c per (sample, port, pole) from the raw outputs o (256, S)."""
    S = o.shape[1]
    r = o.T.reshape(S, 4, NP, 4)
    u, v, pp, qq = r[..., 0], r[..., 1], r[..., 2], r[..., 3]
    sp = np.where(u > 30.0, u, np.log1p(np.exp(np.minimum(u, 30.0))))
    sig = -sp * b.sig0
    om = b.wmax / (1.0 + np.exp(-v))
    c = (pp + 1j * qq) * b.amp[:, None, None]
    return sig, om, c, u, v


def _g(s: np.ndarray, b: Batch):
    """g(s, f) = e^{-j w t0} dt / (1 - e^{(s - jw) dt}) and g'(s, f), shapes (S, 4, NP, F)."""
    w = b.omega[None, None, None, :]
    z = np.exp((s[..., None] - 1j * w) * b.dt)
    e = np.exp(-1j * w * b.t0[:, None, None, None])
    g = e * b.dt / (1.0 - z)
    return g, g * b.dt * z / (1.0 - z)


def pole_loss(o: np.ndarray, b: Batch):
    """Mean |kap P - tgt|^2 and its gradient with respect to o (256, S)."""
    S = o.shape[1]
    F = b.omega.size
    sig, om, c, u, v = _poles(o, b)
    s = sig + 1j * om
    ga, ha = _g(s, b)
    gb, hb = _g(np.conj(s), b)
    cc = c[..., None]
    P = np.sum(cc * ga + np.conj(cc) * gb, axis=2)  # (S, 4, F)
    y = b.kap * P
    r = y - b.tgt
    n = S * 4 * F
    loss = float(np.sum(np.abs(r) ** 2) / n)
    e = 2.0 * np.conj(b.kap) * r  # dL/dP as Re(conj(e) dP)
    e = e[:, :, None, :]
    dP_dsig = cc * ha + np.conj(cc) * hb
    dP_dom = 1j * cc * ha - 1j * np.conj(cc) * hb
    dP_dcr = ga + gb
    dP_dci = 1j * ga - 1j * gb
    red = lambda d: np.sum(np.real(np.conj(e) * d), axis=3) / n  # noqa: E731
    g_sig, g_om, g_cr, g_ci = red(dP_dsig), red(dP_dom), red(dP_dcr), red(dP_dci)
    su, sv = 1.0 / (1.0 + np.exp(-u)), 1.0 / (1.0 + np.exp(-v))
    d = np.stack([g_sig * (-b.sig0 * su), g_om * (b.wmax * sv * (1 - sv)),
                  g_cr * b.amp[:, None, None], g_ci * b.amp[:, None, None]], axis=-1)
    return loss, d.reshape(S, NO).T.copy(), y


def f32(a: np.ndarray) -> np.ndarray:
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def forward(p: np.ndarray, b: Batch, rnd: bool = True):
    """Loss, predictions (S, 4, F) and the cache for :func:`backward`. With ``rnd`` the
    matrix-product operands are rounded to BF16 and every product, bias, tanh and residual
    result to FP32, as the GPU kernels store them."""
    q = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    r = f32 if rnd else (lambda a: a)
    win, bin_, layers, w1, b1, w2, b2 = _views(np.asarray(p, dtype=np.float64))
    T = b.T
    xs = [r(np.tanh(r(r(q(win) @ q(b.U)) + bin_[:, None])))]
    for l, (w, bb) in enumerate(layers):
        x = xs[-1]
        inp = np.concatenate([x, _shift(x, T, 1 << l)], axis=0)
        xs.append(r(x + r(np.tanh(r(r(q(w) @ q(inp)) + bb[:, None])))))
    f = _feat(xs[-1], T)
    h = r(np.tanh(r(r(q(w1) @ q(f)) + b1[:, None])))
    o = r(r(q(w2) @ q(h)) + b2[:, None])
    loss, d_o, y = pole_loss(o, b)
    return loss, y, dict(xs=xs, f=f, h=h, o=o, d_o=d_o)


def backward(p: np.ndarray, b: Batch, cache: dict, rnd: bool = True) -> np.ndarray:
    q = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    r = f32 if rnd else (lambda a: a)
    win, bin_, layers, w1, b1, w2, b2 = _views(np.asarray(p, dtype=np.float64))
    T = b.T
    xs, f, h, d_o = cache["xs"], cache["f"], cache["h"], r(cache["d_o"])
    g = np.zeros(NPARAM)
    g[OFF_W2:OFF_B2] = (q(d_o) @ q(h).T).ravel()
    g[OFF_B2:NPARAM] = d_o.sum(axis=1)
    dh = r(q(w2).T @ q(d_o))
    dz1 = r(dh * (1.0 - h * h))
    g[OFF_W1:OFF_B1] = (q(dz1) @ q(f).T).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(axis=1)
    df = r(q(w1).T @ q(dz1))  # (512, S)
    S = f.shape[1]
    gx = np.zeros_like(xs[-1])
    gx.reshape(C, S, T)[:, :, T - LAST:] = df.reshape(LAST, C, S).transpose(1, 2, 0)
    for l in range(NL - 1, -1, -1):
        w, _ = layers[l]
        d = 1 << l
        x, xn = xs[l], xs[l + 1]
        th = r(xn - x)
        dz = r(gx * (1.0 - th * th))
        inp = np.concatenate([x, _shift(x, T, d)], axis=0)
        o = OFF_WL + l * LSTRIDE
        g[o:o + 2 * C * C] = (q(dz) @ q(inp).T).ravel()
        g[o + 2 * C * C:o + LSTRIDE] = dz.sum(axis=1)
        dzs = np.concatenate([dz, _unshift(dz, T, d)], axis=0)  # (128, N)
        wt = np.concatenate([w[:, :C].T, w[:, C:].T], axis=1)  # (64 ci, 128 = tap*64 + co)
        gx = r(gx + r(q(wt) @ q(dzs)))
    dz0 = r(gx * (1.0 - xs[0] ** 2))
    g[OFF_WIN:OFF_BIN] = (q(dz0) @ q(b.U).T).ravel()
    g[OFF_BIN:OFF_WL] = dz0.sum(axis=1)
    return g
