"""Matrix-pencil continuation of a truncated probe series and its closed-form DFT.

Mirrors ``packages/oa-fdtd/src/oa_fdtd/tail.py:39-98``. The last ``window`` samples are fitted
by ``y_n = sum_m c_m z_m^n`` (matrix pencil on the decimated window, amplitudes by least
squares); the DFT of the continuation from sample ``S`` on is
``sum_m c_m e^{(s_m - j w)(t_S - t_ref)} e^{-j w t_ref} dt / (1 - e^{(s_m - j w) dt})``.
Non-decaying and far out-of-band poles are dropped and counted.

The factor ``e^{-j w t_ref}`` is required because the running DFT uses absolute time stamps
while the fit's exponentials are referred to ``t_ref``; the reference's closed form
(tail.py:89-98, and its module docstring) leaves it out, which rotates the continuation's phase
by ``w t_ref``. This module keeps the factor; the tests pin it against the direct sum.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import numpy.typing as npt

FloatArray = npt.NDArray[np.float64]
ComplexArray = npt.NDArray[np.complex128]


@dataclass(frozen=True)
class TailFit:
    s: ComplexArray
    c: ComplexArray
    t_ref: float
    n_dropped: int
    residual: float


def fit_tail(y: FloatArray, dt: float, t_first: float, *, f_band: float, window: int = 6000,
             max_poles: int = 40, sv_tol: float = 1e-4, samples: int = 700) -> TailFit | None:
    """Fit the last ``window`` samples (sample ``n`` at ``t_first + n dt``); ``None`` when the
    tail is below numerical noise (tail.py:39-86)."""
    x = np.asarray(y, dtype=np.float64)
    w = min(window, x.size)
    if x.size == 0:
        return None
    scale = float(np.max(np.abs(x)))
    seg = x[x.size - w :]
    if scale == 0.0 or float(np.max(np.abs(seg))) < 1e-9 * scale:
        return None
    dec_max = max(int(1.0 / (6.0 * f_band * dt)), 1)
    dec = int(np.clip(w // samples, 1, dec_max))
    seg = seg[::-1][::dec][::-1]  # decimate, keeping the newest sample
    n = seg.size
    if n < 16:
        return None
    t0 = t_first + (x.size - 1 - dec * (n - 1)) * dt
    h = dec * dt
    ell = n // 3
    hank = seg[np.arange(n - ell)[:, None] + np.arange(ell + 1)[None, :]]
    _, sv, vh = np.linalg.svd(hank, full_matrices=False)
    m = min(max_poles, int(np.count_nonzero(sv > sv_tol * sv[0])))
    if m < 1:
        return None
    v = vh[:m].conj().T
    z = np.linalg.eigvals(np.linalg.pinv(v[:-1]) @ v[1:])
    s = np.log(z.astype(np.complex128)) / h
    keep = (s.real < -1e-3 / (n * h)) & (np.abs(s.imag) < 4.0 * np.pi * f_band)
    dropped = int(s.size - np.count_nonzero(keep))
    s = s[keep]
    if s.size == 0:
        return None
    vand = np.exp(np.outer(np.arange(n) * h, s))
    coef, *_ = np.linalg.lstsq(vand, seg.astype(np.complex128), rcond=None)
    res = float(np.linalg.norm(vand @ coef - seg) / max(np.linalg.norm(seg), 1e-300))
    return TailFit(s=s.astype(np.complex128), c=coef.astype(np.complex128), t_ref=t0,
                   n_dropped=dropped, residual=res)


def extrapolate_dft(fit: TailFit, dt: float, t_next: float, freqs: FloatArray) -> ComplexArray:
    """``sum_{n >= S} y_n e^{-j w t_n} dt`` of the fitted continuation (tail.py:89-98)."""
    w = 2.0 * np.pi * np.asarray(freqs, dtype=np.float64)
    a = fit.s[None, :] - 1j * w[:, None]
    num = fit.c[None, :] * np.exp(a * (t_next - fit.t_ref) - 1j * w[:, None] * fit.t_ref) * dt
    return np.sum(num / (1.0 - np.exp(a * dt)), axis=1)
