"""CPU oracle of plane-wave de-embedding (test infrastructure only: the tests compare the HIP
FP64 entry points ``oafd_gamma_from_planes`` / ``oafd_smooth_gamma`` / ``oafd_shift_reference``
/ ``oafd_wave_sparams`` with these functions; the product never imports this module).

A NumPy restatement of packages/oa-fdtd/src/oa_fdtd/ports/deembed.py: ``gamma_from_planes``
(:24-53), ``smooth_gamma`` (:55-82), ``_unequal_residual`` / ``_newton_step`` (:85-101),
``shift_reference`` (:104-121) and ``wave_sparams`` (:264-338). Pinned by closed forms in
tests/test_deembed.py: a line wave ``A e^{-g z} + B e^{g z}`` sampled at three planes gives back
``g`` (equal and unequal spacings, phases past pi), a shifted reference multiplies ``S_ij`` by
``e^{g_i L_i} e^{g_j L_j}``, waves synthesised from a known S matrix give it back, and the smooth
fit reproduces a gamma that lies in its basis.
"""

from __future__ import annotations

import numpy as np


def gamma_from_planes(v1, v2, v3, spacing, spacing_b=None):
    """Propagation constant from three plane voltages (deembed.py:24-53). As in the reference,
    unequal spacings run 8 finite-difference Newton steps from the equal-spacing value; g = 0
    also solves that residual, and forward and backward waves of comparable size can lead there
    (the reference keeps its spacings within 10 % and its phases in 0.3 to 2.2 rad)."""
    db = spacing if spacing_b is None else spacing_b
    dm = 0.5 * (spacing + db)
    c = (v1 + v3) / (2.0 * v2)
    g0 = np.arccosh(c.astype(np.complex128)) / dm
    g = np.where(g0.real < 0, -g0, g0)
    if abs(db - spacing) > 1e-9 * spacing:
        for _ in range(8):
            g = _newton_step(g, v1, v2, v3, spacing, db)
    g = np.where(g.real < 0, -g, g)
    g = np.where(g.imag < 0, np.conj(g), g)
    beta = np.unwrap(g.imag * dm) / dm
    return g.real + 1j * beta


def _unequal_residual(g, v1, v2, v3, da, db):
    ca, sa, cb, sb = np.cosh(g * da), np.sinh(g * da), np.cosh(g * db), np.sinh(g * db)
    return (v1 - v2 * ca) * sb + (v3 - v2 * cb) * sa


def _newton_step(g, v1, v2, v3, da, db):
    h = 1e-6 * np.maximum(np.abs(g), 1.0)
    f0 = _unequal_residual(g, v1, v2, v3, da, db)
    df = (_unequal_residual(g + h, v1, v2, v3, da, db) - f0) / h
    ok = np.abs(df) > 0
    return np.where(ok, g - f0 / np.where(ok, df, 1.0), g)


def smooth_gamma(gamma, freqs, spacing, min_phase, fit_mask=None):
    """Low-frequency smooth fit of an ill-conditioned gamma (deembed.py:55-82)."""
    ok = np.abs(gamma.imag) * spacing >= min_phase
    use = ok if fit_mask is None else ok & fit_mask
    if use.sum() < 6 or ok.all():
        return gamma
    fn = np.asarray(freqs, np.float64) / float(np.max(freqs))
    a_basis = np.column_stack([np.sqrt(fn), fn, fn**2])
    b_basis = np.column_stack([fn, fn**2, fn**3])
    ca, *_ = np.linalg.lstsq(a_basis[use], gamma.real[use], rcond=None)
    cb, *_ = np.linalg.lstsq(b_basis[use], gamma.imag[use], rcond=None)
    fit = a_basis @ ca + 1j * (b_basis @ cb)
    return np.where(ok, gamma, fit)


def shift_reference(s, gamma, lengths):
    """``S'_ij = S_ij e^{g_i L_i} e^{g_j L_j}`` (deembed.py:104-121); gamma is (n, F)."""
    n = s.shape[0]
    fac = [np.exp(np.asarray(gamma[i]) * lengths[i]) for i in range(n)]
    out = np.empty_like(s)
    for i in range(n):
        for j in range(n):
            out[i, j] = s[i, j] * fac[i] * fac[j]
    return out


def wave_sparams(v, spacing, spacing_b, ref, port_r, s_port=None, min_phase=0.3, freqs=None,
                 floor_phase=0.04, fit_mask=None):
    """Plane-wave S matrix (deembed.py:264-338). ``v`` is (n ports, 3 planes, n excitations, F):
    the three modal voltages of port i when port j is driven."""
    n, _, _, nf = v.shape
    s = np.zeros((n, n, nf), np.complex128)
    a = np.zeros((n, n, nf), np.complex128)
    b = np.zeros((n, n, nf), np.complex128)
    gammas = []
    for i in range(n):
        for j in range(n):
            v1, v2, v3 = v[i, 0, j], v[i, 1, j], v[i, 2, j]
            g = gamma_from_planes(v1, v2, v3, spacing[i], spacing_b[i])
            if freqs is not None:
                g = smooth_gamma(g, freqs, spacing[i], min_phase, fit_mask)
            if j == 0:
                gammas.append(g)
            d_mid = (v1 - v2 * np.cosh(g * spacing[i])) / np.sinh(g * spacing[i])
            a[i, j] = 0.5 * (v2 + d_mid) * np.exp(-g * ref[i]) / np.sqrt(port_r[i])
            b[i, j] = 0.5 * (v2 - d_mid) * np.exp(g * ref[i]) / np.sqrt(port_r[i])
    for k in range(nf):
        am, bm = a[:, :, k], b[:, :, k]
        diag = np.abs(np.diag(am))
        off = np.abs(am - np.diag(np.diag(am))).max() if n > 1 else 0.0
        if n == 1 or off > 0.3 * diag.min() or not np.all(np.isfinite(am)):
            s[:, :, k] = bm / np.diag(am)[None, :]
        else:
            s[:, :, k] = np.linalg.solve(am.T, bm.T).T
    if s_port is not None:
        thr = min_phase if freqs is None else floor_phase
        weak = np.zeros(nf, dtype=bool)
        for i, g in enumerate(gammas):
            weak |= np.abs(g.imag) * spacing[i] < thr
        if freqs is None:
            s[:, :, weak] = s_port[:, :, weak]
        else:
            for i in range(n):
                s[i, i, weak] = s_port[i, i, weak]
            noisy = np.abs(s).max(axis=(0, 1)) > 1.02
            s[:, :, noisy] = s_port[:, :, noisy]
    return s, np.array(gammas)
