"""Edge classification of a copper box: PEC interior and surface-impedance face edges.

Mirrors ``packages/oa-fdtd/src/oa_fdtd/subcell/edges.py:37-112``. For an edge ``E_c`` at storage
index ``p`` with ``(c, a, b)`` cyclic, the E update circulates ``H_b`` at ``p`` and ``p - s_a`` and
``H_a`` at ``p`` and ``p - s_b``. With the outward normal ``n`` (metal to dielectric) the
impedance condition ``E_t = Zs n x H_t`` reads:

- face normal ``+a``: ``E_c = +Zs H_b[p]``;   face normal ``-a``: ``E_c = -Zs H_b[p - s_a]``;
- face normal ``+b``: ``E_c = -Zs H_a[p]``;   face normal ``-b``: ``E_c = +Zs H_a[p - s_b]``.

An edge on two faces averages its two outside samples (sign divided by the face count).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import numpy.typing as npt

from ..grid import Grid

IntArray = npt.NDArray[np.intp]


@dataclass
class ConductorEdges:
    """Per component: ``(k, j, i)`` arrays on the ``(nz+1, ny+1, nx+1)`` edge lattice."""

    interior: list[tuple[IntArray, IntArray, IntArray]]
    surface: list[tuple[IntArray, IntArray, IntArray]]
    face_a: list[npt.NDArray[np.int8]]
    face_b: list[npt.NDArray[np.int8]]


def classify_box(g: Grid, lo: tuple[float, ...], hi: tuple[float, ...]) -> ConductorEdges:
    """Split the E edges inside the closed box into interior and face edges (edges.py:37-75)."""
    interior, surface, fa_out, fb_out = [], [], [], []
    empty = tuple(np.zeros(0, np.intp) for _ in range(3))
    for c in range(3):
        a, b = (c + 1) % 3, (c + 2) % 3
        sel: list[IntArray] = []
        flags: list[npt.NDArray[np.int8]] = []
        for ax in range(3):
            node = ax != c
            co = g.coords(ax, node)
            tol = 1e-6 * float(np.min(g.d(ax)))
            ids = np.nonzero((co >= lo[ax] - tol) & (co <= hi[ax] + tol))[0]
            f = np.zeros(ids.size, np.int8)
            if node:
                f[np.abs(co[ids] - hi[ax]) <= tol] = 1
                f[np.abs(co[ids] - lo[ax]) <= tol] = -1  # a zero-width box counts as low
            sel.append(ids)
            flags.append(f)
        if min(s.size for s in sel) == 0:
            interior.append(empty)
            surface.append(empty)
            fa_out.append(np.zeros(0, np.int8))
            fb_out.append(np.zeros(0, np.int8))
            continue
        ii, jj, kk = sel
        shape = (kk.size, jj.size, ii.size)
        K = np.broadcast_to(kk[:, None, None], shape).ravel()
        J = np.broadcast_to(jj[None, :, None], shape).ravel()
        I = np.broadcast_to(ii[None, None, :], shape).ravel()
        # face flags of the transverse axes a and b, broadcast over the box
        def _bc(ax: int) -> npt.NDArray[np.int8]:
            f = flags[ax]
            if ax == 0:
                return np.broadcast_to(f[None, None, :], shape).ravel()
            if ax == 1:
                return np.broadcast_to(f[None, :, None], shape).ravel()
            return np.broadcast_to(f[:, None, None], shape).ravel()

        fa, fb = _bc(a), _bc(b)
        on = (fa != 0) | (fb != 0)
        interior.append((K[~on], J[~on], I[~on]))
        surface.append((K[on], J[on], I[on]))
        fa_out.append(fa[on].copy())
        fb_out.append(fb[on].copy())
    return ConductorEdges(interior, surface, fa_out, fb_out)


def strides(g: Grid) -> tuple[int, int, int]:
    """Flat storage strides of x, y, z."""
    return (1, g.shape[2], g.shape[2] * g.shape[1])


def outside_samples(
    g: Grid,
    c: int,
    k: IntArray,
    j: IntArray,
    i: IntArray,
    fa: npt.NDArray[np.int8],
    fb: npt.NDArray[np.int8],
) -> tuple[npt.NDArray[np.int64], npt.NDArray[np.int32], npt.NDArray[np.int64], npt.NDArray[np.float64]]:
    """Flat edge index, H components ``(ne, 2)``, flat H indices ``(ne, 2)`` and signed weights
    ``(ne, 2)`` of the outside H samples (edges.py:83-112). Slot 0 serves the face normal to
    ``a`` (samples ``H_b``), slot 1 the face normal to ``b`` (samples ``H_a``)."""
    a, b = (c + 1) % 3, (c + 2) % 3
    st = strides(g)
    zs, ys, xs = g.shape
    p = (((k + 1) * ys + (j + 1)) * xs + (i + 1)).astype(np.int64)
    ne = p.size
    nface = (fa != 0).astype(np.float64) + (fb != 0).astype(np.float64)
    div = np.maximum(nface, 1.0)
    hcomp = np.empty((ne, 2), np.int32)
    hflat = np.empty((ne, 2), np.int64)
    hsign = np.zeros((ne, 2), np.float64)
    hcomp[:, 0] = b
    hcomp[:, 1] = a
    hflat[:, 0] = np.where(fa > 0, p, p - st[a])
    hflat[:, 1] = np.where(fb > 0, p, p - st[b])
    hsign[:, 0] = np.where(fa > 0, 1.0, np.where(fa < 0, -1.0, 0.0)) / div
    hsign[:, 1] = np.where(fb > 0, -1.0, np.where(fb < 0, 1.0, 0.0)) / div
    return p, hcomp, hflat, hsign
