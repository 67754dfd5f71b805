"""Non-uniform Yee mesh and the storage layout shared by the HIP engine and the oracle.

Mirrors ``packages/oa-fdtd/src/oa_fdtd/grid.py:30-152`` (the reference ``Grid``): the same
storage convention, so flat indices produced by :func:`compile_model` mean the same cell in
both engines.

Storage convention (one array per field component, C order, x fastest)::

    (B, nz + 2, ny + 2, pitch),   pitch = ceil((nx + 2) / 32) * 32

Along every axis the storage index ``I = i + 1`` holds primal node ``i`` for components that
are node-type along that axis and dual cell ``i`` for cell-type ones; ``I = 0`` and
``I = n + 1`` are ghost layers (PMC mirrors). ``E_a`` is cell-type along ``a``; ``H_a`` is
node-type along ``a`` (``grid.py:3-10``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import numpy.typing as npt

C0 = 299792458.0
MU0 = 4e-7 * np.pi
EPS0 = 1.0 / (MU0 * C0 * C0)
ETA0 = MU0 * C0
LANE = 32  # row padding in elements (grid.py:24)

FloatArray = npt.NDArray[np.float64]


def _is_node(kind: str, comp: int, axis: int) -> bool:
    """Whether component ``kind``/``comp`` sits on primal nodes along ``axis``."""
    return (axis != comp) if kind == "E" else (axis == comp)


@dataclass(frozen=True)
class Grid:
    """Mesh lines ``x, y, z`` in metres (strictly increasing, >= 2 lines each)."""

    x: FloatArray
    y: FloatArray
    z: FloatArray

    def __post_init__(self) -> None:
        for name in ("x", "y", "z"):
            v = np.asarray(getattr(self, name), dtype=np.float64)
            ok = v.ndim == 1 and v.size >= 2 and bool(np.all(v[1:] > v[:-1]))
            if not ok:
                raise ValueError(f"grid lines {name} must be 1-D, increasing, with >= 2 entries")
            object.__setattr__(self, name, v)

    # ---- sizes
    @property
    def lines(self) -> tuple[FloatArray, FloatArray, FloatArray]:
        return (self.x, self.y, self.z)

    @property
    def n(self) -> tuple[int, int, int]:
        """Cells per axis ``(nx, ny, nz)``."""
        return (self.x.size - 1, self.y.size - 1, self.z.size - 1)

    @property
    def cells(self) -> int:
        nx, ny, nz = self.n
        return nx * ny * nz

    @property
    def pitch(self) -> int:
        """Padded row length in elements (grid.py:59-61)."""
        w = self.n[0] + 2
        return ((w + LANE - 1) // LANE) * LANE

    @property
    def shape(self) -> tuple[int, int, int]:
        """Storage shape ``(nz + 2, ny + 2, pitch)`` without the batch axis (grid.py:64-66)."""
        return (self.n[2] + 2, self.n[1] + 2, self.pitch)

    def length(self, axis: int) -> int:
        """Storage extent along ``axis`` (grid.py:68-70)."""
        return self.pitch if axis == 0 else self.n[axis] + 2

    # ---- 1-D metric arrays
    def d(self, axis: int) -> FloatArray:
        """Primal cell sizes [m]."""
        return np.diff(self.lines[axis])

    def centers(self, axis: int) -> FloatArray:
        v = self.lines[axis]
        return 0.5 * (v[:-1] + v[1:])

    def coords(self, axis: int, node: bool) -> FloatArray:
        return self.lines[axis] if node else self.centers(axis)

    def inv_primal(self, axis: int) -> FloatArray:
        """``1/d_i`` stored at ``I = i + 1``; zero elsewhere (grid.py:84-88)."""
        n = self.n[axis]
        out = np.zeros(self.length(axis))
        out[1 : n + 1] = 1.0 / self.d(axis)
        return out

    def inv_dual(self, axis: int) -> FloatArray:
        """``2/(d_{i-1}+d_i)`` at node ``I = i + 1``, the missing cell mirrored at the walls
        (grid.py:90-99)."""
        n = self.n[axis]
        d = self.d(axis)
        left = np.concatenate(([d[0]], d))  # d_{i-1} with d_{-1} := d_0
        right = np.concatenate((d, [d[-1]]))  # d_i with d_n := d_{n-1}
        out = np.zeros(self.length(axis))
        out[1 : n + 2] = 2.0 / (left + right)
        return out

    def node_weights(self, axis: int) -> FloatArray:
        """Dual length of every node, half cells at the walls (grid.py:101-107)."""
        n = self.n[axis]
        d = self.d(axis)
        out = np.zeros(self.length(axis))
        out[1 : n + 2] = 0.5 * (np.concatenate(([0.0], d)) + np.concatenate((d, [0.0])))
        return out

    def cell_weights(self, axis: int) -> FloatArray:
        """Primal length of every cell at ``I = i + 1`` (grid.py:109-112)."""
        out = np.zeros(self.length(axis))
        out[1 : self.n[axis] + 1] = self.d(axis)
        return out

    def dt(self, courant: float = 0.95, c_max: float = C0) -> float:
        """Courant-limited time step of the non-uniform mesh [s] (grid.py:114-117)."""
        inv2 = 0.0
        for a in range(3):
            dmin = float(np.min(self.d(a)))
            inv2 += 1.0 / (dmin * dmin)
        return float(courant / (c_max * np.sqrt(inv2)))

    def valid(self, kind: str, comp: int) -> tuple[tuple[int, int], ...]:
        """Storage ranges ``[lo, hi)`` per axis (x, y, z) that hold real edges (grid.py:119-129)."""
        return tuple(
            (1, self.n[a] + 2) if _is_node(kind, comp, a) else (1, self.n[a] + 1) for a in range(3)
        )

    def points(
        self, kind: str, comp: int, lo: tuple[float, float, float], hi: tuple[float, float, float]
    ) -> tuple[npt.NDArray[np.intp], tuple[int, int, int]]:
        """Flat storage indices (z, y, x order) of a component's lattice points in a closed box,
        and the per-axis counts ``(nx, ny, nz)``; a degenerate box snaps to the nearest point
        (grid.py:131-152)."""
        picks: list[npt.NDArray[np.intp]] = []
        for a in range(3):
            c = self.coords(a, _is_node(kind, comp, a))
            tol = 1e-6 * float(np.min(self.d(a)))
            sel = np.nonzero((c >= lo[a] - tol) & (c <= hi[a] + tol))[0]
            if sel.size == 0 and hi[a] - lo[a] <= 2.0 * tol:
                sel = np.array([int(np.argmin(np.abs(c - 0.5 * (lo[a] + hi[a]))))])
            if sel.size == 0:
                raise ValueError(f"box selects no {kind}{'xyz'[comp]} point along {'xyz'[a]}")
            picks.append((sel + 1).astype(np.intp))
        ix, iy, iz = picks
        zs, ys, xs = self.shape
        flat = (iz[:, None, None] * ys + iy[None, :, None]) * xs + ix[None, None, :]
        return flat.ravel().astype(np.intp), (ix.size, iy.size, iz.size)
