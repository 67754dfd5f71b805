"""CFS-CPML coefficients and correction boxes (Roden and Gedney, MOTL 27(5), 2000).

Mirrors ``packages/oa-fdtd/src/oa_fdtd/pml.py:26-146``. Per slab and curl term the derivative
``d`` becomes ``d/kappa + psi`` with ``psi <- b psi + c d``; psi storage exists only inside the
slabs. The HIP engine applies the correction inside its fused update; the oracle applies it as a
separate pass after the bulk update, exactly as the reference engines do.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import numpy.typing as npt

from .grid import EPS0, ETA0, MU0, Grid

FloatArray = npt.NDArray[np.float64]
Box = tuple[tuple[int, int], tuple[int, int], tuple[int, int]]

# Curl terms (corrected component c, derivative axis a, source component b, sign):
#   E_c += sign * cb   * d_a H_b        H_c += sign * (dt/mu0) * d_a E_b
# (pml.py:26-41; the signs are those of curl H and -curl E on the Yee lattice).
E_TERMS: tuple[tuple[int, int, int, float], ...] = (
    (0, 1, 2, +1.0),
    (0, 2, 1, -1.0),
    (1, 2, 0, +1.0),
    (1, 0, 2, -1.0),
    (2, 0, 1, +1.0),
    (2, 1, 0, -1.0),
)
H_TERMS: tuple[tuple[int, int, int, float], ...] = tuple((c, a, b, -s) for c, a, b, s in E_TERMS)


@dataclass(frozen=True)
class CpmlSpec:
    """Layer parameters (pml.py:45-52)."""

    cells: int = 10
    m: float = 3.0
    kappa_max: float = 8.0
    r0: float = 1e-8
    f_low: float = 1e8


@dataclass(frozen=True)
class Correction:
    """``F_c += sign * k * (ek * d + psi)`` over ``box`` (storage ``[lo, hi)`` for x, y, z)."""

    kind: str
    c: int
    a: int
    b: int
    sign: float
    box: Box
    side: int = 0  # 0 = low wall of axis a, 1 = high wall

    @property
    def shape(self) -> tuple[int, int, int]:
        """psi shape ``(nz, ny, nx)`` of one batch member (pml.py:66-69)."""
        return (
            self.box[2][1] - self.box[2][0],
            self.box[1][1] - self.box[1][0],
            self.box[0][1] - self.box[0][0],
        )


@dataclass(frozen=True)
class AxisCoeffs:
    """Per-axis coefficient arrays over the storage index: E-type at nodes, H-type at cells."""

    eb: FloatArray
    ec: FloatArray
    ek: FloatArray
    hb: FloatArray
    hc: FloatArray
    hk: FloatArray


@dataclass(frozen=True)
class Pml:
    corrections: tuple[Correction, ...]
    coeffs: dict[int, AxisCoeffs]


def _grade(depth: FloatArray, thick: float, spec: CpmlSpec, dt: float) -> tuple[FloatArray, ...]:
    """Polynomial-graded sigma/kappa, linearly graded alpha, and the recursion coefficients
    ``b = exp(-(sigma/kappa + alpha) dt/eps0)``, ``c = sigma (b - 1) / (kappa (sigma + kappa
    alpha))``, ``k = 1/kappa - 1`` (pml.py:90-100)."""
    rho = np.clip(depth / thick, 0.0, 1.0)
    poly = rho**spec.m
    sigma_max = -(spec.m + 1.0) * math.log(spec.r0) / (2.0 * ETA0 * thick)
    sigma = sigma_max * poly
    kappa = 1.0 + (spec.kappa_max - 1.0) * poly
    inside = rho > 0.0
    alpha = np.where(inside, 2.0 * math.pi * spec.f_low * EPS0 * (1.0 - rho), 0.0)
    b = np.exp(-(sigma / kappa + alpha) * dt / EPS0)
    den = kappa * (sigma + kappa * alpha)
    safe = np.where(den > 0.0, den, 1.0)
    c = np.where(den > 0.0, sigma * (b - 1.0) / safe, 0.0)
    b = np.where(inside, b, 0.0)
    return b, c, 1.0 / kappa - 1.0


def build_pml(grid: Grid, walls: tuple[str, ...], spec: CpmlSpec, dt: float) -> Pml:
    """Coefficients and correction boxes of every ``"cpml"`` wall (pml.py:103-146)."""
    corrections: list[Correction] = []
    coeffs: dict[int, AxisCoeffs] = {}
    for a in range(3):
        sides = [s for s in (0, 1) if walls[2 * a + s] == "cpml"]
        if not sides:
            continue
        n, npml = grid.n[a], spec.cells
        if 2 * npml >= n:
            raise ValueError(f"axis {'xyz'[a]}: {n} cells cannot hold two {npml}-cell CPML layers")
        ln = grid.length(a)
        node_x = np.zeros(ln)
        cell_x = np.zeros(ln)
        node_ok = np.zeros(ln, dtype=bool)
        cell_ok = np.zeros(ln, dtype=bool)
        node_x[1 : n + 2] = grid.lines[a]
        node_ok[1 : n + 2] = True
        cell_x[1 : n + 1] = grid.centers(a)
        cell_ok[1 : n + 1] = True
        arr = {k: np.zeros(ln) for k in ("eb", "ec", "ek", "hb", "hc", "hk")}
        for s in sides:
            if s == 0:
                inner = grid.lines[a][npml]
                outer = grid.lines[a][0]
                dn = np.where(node_ok, inner - node_x, 0.0)
                dc = np.where(cell_ok, inner - cell_x, 0.0)
                e_rng = (1, npml + 1)
                h_rng = (1, npml + 1)
            else:
                inner = grid.lines[a][n - npml]
                outer = grid.lines[a][n]
                dn = np.where(node_ok, node_x - inner, 0.0)
                dc = np.where(cell_ok, cell_x - inner, 0.0)
                e_rng = (n - npml + 2, n + 2)
                h_rng = (n - npml + 1, n + 1)
            thick = abs(inner - outer)
            for pre, depth, rng in (("e", dn, e_rng), ("h", dc, h_rng)):
                b, c, k = _grade(depth, thick, spec, dt)
                lo, hi = rng
                arr[pre + "b"][lo:hi] = b[lo:hi]
                arr[pre + "c"][lo:hi] = c[lo:hi]
                arr[pre + "k"][lo:hi] = k[lo:hi]
            for kind, terms, rng in (("E", E_TERMS, e_rng), ("H", H_TERMS, h_rng)):
                for c_, a_, b_, sign in terms:
                    if a_ != a:
                        continue
                    box = list(grid.valid(kind, c_))
                    box[a] = rng
                    corrections.append(Correction(kind, c_, a, b_, sign, tuple(box), s))  # type: ignore[arg-type]
        coeffs[a] = AxisCoeffs(**arr)
    return Pml(tuple(corrections), coeffs)


def mu_dt(dt: float) -> float:
    """``dt / mu0``."""
    return dt / MU0
