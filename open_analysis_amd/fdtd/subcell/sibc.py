"""Finite-slab surface impedance and its passive R-L ladder fit.

Mirrors ``packages/oa-fdtd/src/oa_fdtd/subcell/sibc.py:29-115``. ``Zs = k/sigma coth(k t)`` with
``k = sqrt(j w mu sigma)`` (convention ``exp(+j w t)``), times a roughness factor, is fitted by
``R0 + sum_p R_p j w tau_p / (1 + j w tau_p)`` with non-negative ``R_p`` on a fixed log-spaced
pole set. Each section's bilinear recursion ``v <- R_p c_p (J - J_prev) + g_p v`` is
unconditionally stable. The fit uses SciPy's NNLS, the same solver the reference calls, so the
ladder coefficients follow the same problem statement.
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
MU0 = 4e-7 * math.pi
TWO_PI = 2.0 * math.pi


def slab_surface_impedance(
    freqs: FloatArray, sigma: float, thickness: float, mu_r: float = 1.0
) -> ComplexArray:
    """Surface impedance [ohm/sq] of a slab shorted on its far side (sibc.py:29-40)."""
    w = TWO_PI * np.maximum(np.asarray(freqs, dtype=np.float64), 1e-30)
    k = np.sqrt(1j * w * MU0 * mu_r * sigma)
    return np.asarray(k / (sigma * np.tanh(k * thickness)), dtype=np.complex128)


def hammerstad_jensen(freqs: FloatArray, sigma: float, rms: float, mu_r: float = 1.0) -> FloatArray:
    """Roughness factor ``1 + (2/pi) atan(1.4 (rms/delta)^2)`` (sibc.py:43-49)."""
    f = np.asarray(freqs, dtype=np.float64)
    if rms <= 0:
        return np.ones_like(f)
    w = TWO_PI * np.maximum(f, 1e-30)
    skin = np.sqrt(2.0 / (w * MU0 * mu_r * sigma))
    return np.asarray(1.0 + (2.0 / math.pi) * np.arctan(1.4 * (rms / skin) ** 2))


@dataclass(frozen=True)
class SibcFit:
    r0: float
    rp: FloatArray
    tau: FloatArray
    f_lo: float
    f_hi: float
    err: float

    def eval(self, freqs: FloatArray) -> ComplexArray:
        jw = 1j * TWO_PI * np.asarray(freqs, dtype=np.float64)
        z = np.full(jw.shape, self.r0, dtype=np.complex128)
        for r, t in zip(self.rp, self.tau, strict=True):
            z = z + r * (jw * t) / (1.0 + jw * t)
        return z

    def recursion(self, dt: float) -> tuple[FloatArray, FloatArray]:
        """Bilinear ``(c_p, gamma_p)`` with ``x = 2 tau / dt`` (sibc.py:71-74)."""
        x = 2.0 * self.tau / dt
        return x / (1.0 + x), (x - 1.0) / (x + 1.0)


def fit_surface_impedance(
    sigma: float,
    thickness: float,
    f_lo: float,
    f_hi: float,
    *,
    rms: float = 0.0,
    roughness: Callable[[FloatArray], FloatArray] | None = None,
    n_poles: int = 16,
    mu_r: float = 1.0,
) -> SibcFit:
    """Passive ladder fit of ``Zs(f) K(f)`` over ``[f_lo, f_hi]`` (sibc.py:77-115).

    Sample set: 240 log-spaced points over ``[f_lo/5, f_hi]`` plus two DC-pinning points;
    poles log-spaced over ``[f_lo/10, 5 f_hi]``; rows weighted by ``1/|Z|``, the reactance rows
    by a further 0.2 so the loss leads.
    """
    if not 0 < f_lo < f_hi:
        raise ValueError("fit band needs 0 < f_lo < f_hi")
    f = np.concatenate(([f_lo * 1e-3, f_lo * 1e-2], np.geomspace(f_lo / 5.0, f_hi, 240)))
    kf = roughness(f) if roughness is not None else hammerstad_jensen(f, sigma, rms, mu_r)
    z = slab_surface_impedance(f, sigma, thickness, mu_r) * kf
    w = TWO_PI * f
    tau = 1.0 / (TWO_PI * np.geomspace(0.1 * f_lo, 5.0 * f_hi, n_poles))
    x = w[:, None] * tau[None, :]
    den = 1.0 + x * x
    a_re = np.hstack([np.ones((f.size, 1)), x * x / den])
    a_im = np.hstack([np.zeros((f.size, 1)), x / den])
    wt = 1.0 / np.abs(z)
    a = np.vstack([a_re * wt[:, None], 0.2 * a_im * wt[:, None]])
    rhs = np.concatenate([z.real * wt, 0.2 * z.imag * wt])
    sol, _ = nnls(a, rhs, maxiter=100 * a.shape[1])
    fit = SibcFit(float(sol[0]), sol[1:].copy(), tau, f_lo, f_hi, 0.0)
    band = (f >= f_lo * (1 - 1e-9)) & (f <= f_hi * (1 + 1e-9))
    err = float(np.max(np.abs(fit.eval(f[band]).real - z[band].real) / z[band].real))
    return SibcFit(fit.r0, fit.rp, fit.tau, f_lo, f_hi, err)
