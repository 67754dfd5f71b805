"""Passive multi-pole Debye fits (non-negative pole strengths on a fixed log-spaced pole set).

Mirrors ``packages/oa-fdtd/src/oa_fdtd/materials/debye.py:53-152``. Convention
``exp(+j w t)``: ``eps = eps' - j eps''``; a fit is ``eps_inf + sum_k d_k / (1 + j w tau_k)``.
The pole count grows from ``min_poles`` until the permittivity (0.2 %) and loss-tangent (2 %)
limits hold over the band.
"""

from __future__ import annotations

import math
from collections.abc import Callable
from dataclasses import dataclass

import numpy as np
import numpy.typing as npt
from scipy.optimize import nnls

FloatArray = npt.NDArray[np.float64]
ComplexArray = npt.NDArray[np.complex128]
TWO_PI = 2.0 * math.pi


@dataclass(frozen=True)
class DebyeFit:
    eps_inf: float
    delta: FloatArray
    tau: FloatArray
    f_lo: float
    f_hi: float
    err_eps: float
    err_tand: float

    @property
    def poles(self) -> tuple[tuple[float, float], ...]:
        """``(delta, tau)`` pairs as :class:`Region` takes them."""
        return tuple((float(d), float(t)) for d, t in zip(self.delta, self.tau, strict=True))

    def eval(self, freqs: FloatArray) -> ComplexArray:
        jw = 1j * TWO_PI * np.asarray(freqs, dtype=np.float64)
        out = np.full(jw.shape, self.eps_inf, dtype=np.complex128)
        for d, t in zip(self.delta, self.tau, strict=True):
            out = out + d / (1.0 + jw * t)
        return out


def djordjevic_sarkar_eps(
    freqs: FloatArray, dk: float, df: float, f_ref: float, f_low: float, f_high: float
) -> ComplexArray:
    """Wideband Debye (Djordjevic-Sarkar) permittivity matched to ``dk``/``df`` at ``f_ref``
    (Djordjevic et al., IEEE Trans. EMC 43(4), 2001; debye.py:53-68)."""
    w1, w2, wr = TWO_PI * f_low, TWO_PI * f_high, TWO_PI * f_ref
    c = dk * df / (math.atan(wr / w1) - math.atan(wr / w2))
    eps_inf = dk - 0.5 * c * math.log((w2 * w2 + wr * wr) / (w1 * w1 + wr * wr))
    w = TWO_PI * np.asarray(freqs, dtype=np.float64)
    re = eps_inf + 0.5 * c * np.log((w2 * w2 + w * w) / (w1 * w1 + w * w))
    im = c * (np.arctan(w / w1) - np.arctan(w / w2))
    return np.asarray(re - 1j * im, dtype=np.complex128)


def fit_debye(
    freqs: FloatArray,
    eps: ComplexArray,
    f_lo: float,
    f_hi: float,
    *,
    tol_eps: float = 2e-3,
    tol_tand: float = 2e-2,
    max_poles: int = 12,
    min_poles: int = 3,
    margin_decades: float = 1.0,
) -> DebyeFit:
    """Fewest-pole passive fit meeting both tolerances over the band (debye.py:80-128)."""
    f = np.asarray(freqs, dtype=np.float64)
    e = np.asarray(eps, dtype=np.complex128)
    keep = (f >= f_lo * (1 - 1e-9)) & (f <= f_hi * (1 + 1e-9))
    f, e = f[keep], e[keep]
    if f.size < 4 or not 0 < f_lo < f_hi:
        raise ValueError("fit needs 0 < f_lo < f_hi and at least four samples in the band")
    w = TWO_PI * f
    tand = -e.imag / e.real
    w_re = 1.0 / e.real
    w_im = (tol_eps / tol_tand) / e.real
    best: DebyeFit | None = None
    best_score = math.inf
    for n in range(min_poles, max_poles + 1):
        fp = np.geomspace(f_lo * 10.0**-margin_decades, f_hi * 10.0**margin_decades, n)
        tau = 1.0 / (TWO_PI * fp)
        x = w[:, None] * tau[None, :]
        den = 1.0 + x * x
        a_re = np.hstack([np.ones((f.size, 1)), 1.0 / den])
        a_im = np.hstack([np.zeros((f.size, 1)), x / den])
        a = np.vstack([a_re * w_re[:, None], a_im * w_im[:, None]])
        rhs = np.concatenate([(e.real - 1.0) * w_re, -e.imag * w_im])
        sol, _ = nnls(a, rhs, maxiter=50 * a.shape[1])
        trial = DebyeFit(1.0 + float(sol[0]), sol[1:].copy(), tau, f_lo, f_hi, 0.0, 0.0)
        ev = trial.eval(f)
        err_eps = float(np.max(np.abs(ev.real - e.real) / e.real))
        err_tand = float(np.max(np.abs(-ev.imag / ev.real - tand) / np.maximum(tand, 1e-12)))
        trial = DebyeFit(trial.eps_inf, trial.delta, tau, f_lo, f_hi, err_eps, err_tand)
        score = err_eps / tol_eps + err_tand / tol_tand
        if score < best_score:
            best, best_score = trial, score
        if err_eps <= tol_eps and err_tand <= tol_tand:
            return trial
    assert best is not None
    return best


def fit_djordjevic_sarkar(
    dk: float,
    df: float,
    f_ref: float,
    f_lo: float,
    f_hi: float,
    f_low: float = 1e4 / TWO_PI,
    f_high: float = 1e12 / TWO_PI,
    n_samples: int = 400,
) -> DebyeFit:
    f = np.geomspace(f_lo, f_hi, n_samples)
    return fit_debye(f, djordjevic_sarkar_eps(f, dk, df, f_ref, f_low, f_high), f_lo, f_hi)


def fit_callable(
    fn: Callable[[FloatArray], ComplexArray], f_lo: float, f_hi: float, n_samples: int = 400
) -> DebyeFit:
    f = np.geomspace(f_lo, f_hi, n_samples)
    return fit_debye(f, fn(f), f_lo, f_hi)
