"""Plane-wave de-embedding in FP64 on the GPU, with the reference's API
(packages/oa-fdtd/src/oa_fdtd/ports/deembed.py): :func:`gamma_from_planes`,
:func:`smooth_gamma`, :func:`shift_reference`, :class:`PlaneSet` and :func:`wave_sparams` run on
the device through ``oafd_gamma_from_planes`` / ``oafd_smooth_gamma`` /
``oafd_shift_reference`` / ``oafd_wave_sparams``; there is no host fallback.

The reference places its three planes with ``add_wave_planes`` on a mode port (modal voltage
probes). Mode ports are not part of this engine, so :func:`add_voltage_planes` places three
voltage probes (E-field line integrals, ``weight="line"``) across the line behind a lumped port
instead; on a quasi-TEM line they carry the same forward and backward waves.
"""

from __future__ import annotations

import ctypes as C
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np
import numpy.typing as npt

from . import _lib
from .model import Model, ModelError, Probe

FloatArray = npt.NDArray[np.float64]
ComplexArray = npt.NDArray[np.complex128]


def _c(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def gamma_from_planes(v1: ComplexArray, v2: ComplexArray, v3: ComplexArray, spacing: float,
                      spacing_b: float | None = None) -> ComplexArray:
    """Propagation constant [1/m] from the voltages at three planes (deembed.py:24-53): the
    cosh formula for equal spacings, Newton iterations for unequal ones, ``Re g >= 0``, the phase
    unwrapped along the last axis (frequency)."""
    a1, a2, a3 = (_c(v) for v in (v1, v2, v3))
    if not (a1.shape == a2.shape == a3.shape) or a1.ndim not in (1, 2):
        raise ValueError("v1, v2, v3 must share a (F,) or (rows, F) shape")
    rows = a1.reshape(-1, a1.shape[-1])
    out = np.empty_like(rows)
    db = spacing if spacing_b is None else spacing_b
    _lib.check(_lib.lib().oafd_gamma_from_planes(rows.shape[0], rows.shape[1], _p(a1), _p(a2), _p(a3),
                                                 float(spacing), float(db), _p(out)), "gamma_from_planes")
    return out.reshape(a1.shape)


def smooth_gamma(gamma: ComplexArray, freqs: FloatArray, spacing: float, min_phase: float,
                 fit_mask: npt.NDArray[np.bool_] | None = None) -> ComplexArray:
    """Smooth fit of the ill-conditioned low-frequency part of gamma (deembed.py:55-82)."""
    g = _c(gamma).copy()
    rows = g.reshape(-1, g.shape[-1])
    f = np.ascontiguousarray(freqs, np.float64)
    mask = None if fit_mask is None else np.ascontiguousarray(fit_mask, np.uint8)
    _lib.check(_lib.lib().oafd_smooth_gamma(rows.shape[0], rows.shape[1], _p(rows), _p(f), float(spacing),
                                            float(min_phase), _p(mask)), "smooth_gamma")
    return rows.reshape(g.shape)


def shift_reference(s: ComplexArray, gamma: Sequence[ComplexArray] | ComplexArray,
                    lengths: Sequence[float]) -> ComplexArray:
    """Move the reference plane of port ``i`` inwards by ``lengths[i]`` [m]:
    ``S'_ij = S_ij e^{g_i L_i} e^{g_j L_j}`` (deembed.py:104-121). ``gamma`` is one array for all
    ports or one per port; ``s`` is (N, N, F)."""
    sm = _c(s)
    n, _, nf = sm.shape
    if np.ndim(gamma[0]) == 0:
        g = np.repeat(_c(np.asarray(gamma))[None, :], n, axis=0)
    else:
        g = _c(np.stack([np.asarray(x) for x in gamma]))
    ln = np.ascontiguousarray(lengths, np.float64)
    if g.shape != (n, nf) or ln.shape != (n,):
        raise ValueError("gamma must be (F,) or N arrays of (F,), lengths (N,)")
    out = np.empty_like(sm)
    _lib.check(_lib.lib().oafd_shift_reference(n, nf, _p(sm), _p(g), _p(ln), _p(out)), "shift_reference")
    return out


@dataclass(frozen=True)
class PlaneSet:
    """Three voltage planes of one port, outer to inner (deembed.py:124-136): ``spacing`` [m]
    separates the first two (``spacing_b`` the last two when unequal); ``ref`` [m] is the signed
    distance from the middle plane to the reference plane, positive inwards."""

    port: str
    probes: tuple[str, str, str]
    spacing: float
    ref: float
    spacing_b: float | None = None


def add_voltage_planes(model: Model, port: str, positions: Sequence[float],
                       lo: tuple[float, float, float], hi: tuple[float, float, float],
                       axis: int, along: int, ref_position: float | None = None) -> PlaneSet:
    """Add three voltage probes across the line behind lumped port ``port`` and return their
    :class:`PlaneSet`. ``positions`` are the planes' coordinates along axis ``along``, ordered
    from the port inwards; each probe integrates E along ``axis`` over the box ``lo``..``hi``
    with its ``along`` coordinate replaced by the plane's. The reference plane is the port's
    plane unless ``ref_position`` is given."""
    lp = next((p for p in model.ports if p.name == port), None)
    if lp is None:
        raise ModelError("FDTD_BAD_PORT", f"no port named {port!r}")
    pos = [float(x) for x in positions]
    if len(pos) != 3:
        raise ModelError("FDTD_UNEQUAL_PLANES", "three plane positions are needed")
    port_x = 0.5 * (lp.lo[along] + lp.hi[along])
    step = 1.0 if pos[1] > port_x else -1.0
    d1, d2 = (pos[1] - pos[0]) * step, (pos[2] - pos[1]) * step
    if d1 <= 0 or d2 <= 0:
        raise ModelError("FDTD_UNEQUAL_PLANES", "plane positions must run from the port inwards")
    names = tuple(f"V:{port}:{k}" for k in range(3))
    for nm, x in zip(names, pos, strict=True):
        a, b = list(lo), list(hi)
        a[along] = b[along] = x
        model.probes.append(Probe(nm, "E", axis, tuple(a), tuple(b), "line"))
    ref_x = port_x if ref_position is None else float(ref_position)
    return PlaneSet(port, names, d1, (ref_x - pos[1]) * step,  # type: ignore[arg-type]
                    None if abs(d1 - d2) <= 1e-12 * d1 else d2)


def wave_sparams(probes: dict[str, ComplexArray], planes: Sequence[PlaneSet], port_r: FloatArray,
                 s_port: ComplexArray | None = None, min_phase: float = 0.3,
                 freqs: FloatArray | None = None, floor_phase: float = 0.04,
                 fit_mask: npt.NDArray[np.bool_] | None = None) -> tuple[ComplexArray, list[ComplexArray]]:
    """S matrix (N, N, F) from the waves on each port's planes (deembed.py:264-338): batch member
    j of every probe is the run that drives port j. Returns ``(S, [gamma_i])``."""
    n = len(planes)
    nf = probes[planes[0].probes[0]].shape[1]
    v = np.empty((n, 3, n, nf), np.complex128)
    for i, pl in enumerate(planes):
        for p, nm in enumerate(pl.probes):
            if probes[nm].shape[0] < n:
                raise ValueError(f"probe {nm} has {probes[nm].shape[0]} runs, {n} needed")
            v[i, p] = probes[nm][:n]
    sp = np.array([pl.spacing for pl in planes], np.float64)
    sb = np.array([pl.spacing if pl.spacing_b is None else pl.spacing_b for pl in planes], np.float64)
    ref = np.array([pl.ref for pl in planes], np.float64)
    r = np.ascontiguousarray(port_r, np.float64)
    spo = None if s_port is None else _c(s_port)
    fq = None if freqs is None else np.ascontiguousarray(freqs, np.float64)
    mask = None if fit_mask is None else np.ascontiguousarray(fit_mask, np.uint8)
    s = np.empty((n, n, nf), np.complex128)
    g = np.empty((n, nf), np.complex128)
    _lib.check(_lib.lib().oafd_wave_sparams(n, nf, _p(v), _p(sp), _p(sb), _p(ref), _p(r), _p(spo), _p(fq),
                                            _p(mask), float(min_phase), float(floor_phase), _p(s), _p(g)),
               "wave_sparams")
    return s, [g[i] for i in range(n)]
