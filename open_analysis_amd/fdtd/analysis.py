"""Post-processing helpers: matrix-pencil mode extraction and power-wave S-parameters
(mirrors ``packages/oa-fdtd/src/oa_fdtd/analysis.py:14-55``)."""

from __future__ import annotations

import numpy as np
import numpy.typing as npt

FloatArray = npt.NDArray[np.float64]
ComplexArray = npt.NDArray[np.complex128]


def matrix_pencil(y: FloatArray, dt: float, order: int, decimate: int = 1,
                  pencil: float = 0.4) -> tuple[FloatArray, FloatArray]:
    """Non-negative frequencies [Hz] (ascending) and damping rates [1/s] of ``order`` complex
    exponentials in ``y``."""
    x = np.asarray(y[::decimate], dtype=np.float64)
    h = dt * decimate
    n = x.size
    ell = int(pencil * n)
    hank = x[np.arange(n - ell)[:, None] + np.arange(ell + 1)[None, :]]
    _, _, vh = np.linalg.svd(hank, full_matrices=False)
    v = vh[:order].conj().T
    z = np.linalg.eigvals(np.linalg.pinv(v[:-1]) @ v[1:])
    s = np.log(z.astype(np.complex128)) / h
    f = s.imag / (2.0 * np.pi)
    pos = f >= 0
    order_ix = np.argsort(f[pos])
    return f[pos][order_ix], (-s.real[pos])[order_ix]


def power_wave_s(v: ComplexArray, vs: ComplexArray, r_ohm: FloatArray) -> ComplexArray:
    """``S[i, j, f]`` from port voltages ``v[i, j, f]`` (port ``j`` driven by ``vs[j]``)."""
    n = v.shape[0]
    s = np.zeros_like(v)
    for j in range(n):
        a_j = vs[j] / (2.0 * np.sqrt(r_ohm[j]))
        for i in range(n):
            s[i, j] = (v[i, j] / np.sqrt(r_ohm[i]) - (a_j if i == j else 0.0)) / a_j
    return s
