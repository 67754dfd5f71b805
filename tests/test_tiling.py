"""The fused step's tile plan (computed by the library's host planner, no GPU): every E and H
cell of the domain is owned by exactly one tile, and interior tiles, whose kernel has no bounds
checks and no boundary code, read only inside the storage and never touch a CPML slab or a wall.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from open_analysis_amd.fdtd import _lib

pytestmark = pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="library not built")
TXE = 63
SHAPES = [(4, 1), (4, 2), (4, 4), (8, 1), (8, 2), (12, 1), (16, 1)]
GRIDS = [
    # nx, ny, nz, batch, npml (xlo xhi ylo yhi zlo zhi)
    (150, 30, 26, 1, (4, 4, 0, 0, 4, 4)),
    (24, 20, 22, 2, (4, 4, 0, 0, 0, 0)),
    (160, 80, 48, 2, (8, 8, 8, 8, 8, 8)),
    (256, 256, 256, 1, (10, 10, 10, 10, 10, 10)),
    (200, 130, 37, 3, (0, 6, 5, 0, 3, 0)),
    (127, 64, 20, 1, (0, 0, 0, 0, 0, 0)),
    (400, 200, 100, 1, (10, 10, 10, 10, 10, 10)),
]


def plan(nx, ny, nz, b, npml, nw, rpw, zslab=True):
    kb = np.zeros(4096, np.int32)
    tiles = np.zeros(4 * 2_000_000, np.int32)
    info = np.zeros(9, np.int32)
    npm = np.asarray(npml, np.int32)
    rc = _lib.lib().oafd_plan(nx, ny, nz, b, npm.ctypes.data, nw, rpw, int(zslab), kb.ctypes.data,
                              kb.size, tiles.ctypes.data, tiles.size, info.ctypes.data)
    assert rc == 0
    nch = int(info[6])
    return kb[: nch + 1].copy(), tiles[: 4 * info[7]].reshape(-1, 4).copy(), info.copy()


def all_tiles(kb, gen, info, b):
    """(tx, ty, chunk, b, kind) with kind 0 general, 1 interior, 2 z-slab kernel."""
    tx0, ty0, zc0, nxi, nyi, nzc = (int(v) for v in info[:6])
    nch = int(info[6])
    out = [tuple(t) + (0,) for t in gen]
    zs = [0, nch - 1] if int(info[8]) == 2 else []
    for bb in range(b):
        for zc in list(range(zc0, zc0 + nzc)) + zs:
            for ty in range(ty0, ty0 + nyi):
                for tx in range(tx0, tx0 + nxi):
                    out.append((tx, ty, zc, bb, 1 if zc not in zs else 2))
    return out


@pytest.mark.parametrize("zslab", [True, False])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("grid", GRIDS)
def test_tiles_cover_every_cell_once(grid, shape, zslab) -> None:
    nx, ny, nz, b, npml = grid
    nw, rpw = shape
    tye = nw * rpw - 1
    kb, gen, info = plan(nx, ny, nz, b, npml, nw, rpw, zslab)
    assert kb[0] == 1 and kb[-1] == nz + 2 and np.all(np.diff(kb) > 0)
    e_cov = np.zeros((b, nz + 2, ny + 2, nx + 2), np.int16)
    h_cov = np.zeros((b, nz + 2, ny + 2, nx + 2), np.int16)
    X, Y, Z = nx + 2, ny + 2, nz + 2  # the storage pitch is at least nx + 2
    for tx, ty, zc, bb, kind in all_tiles(kb, gen, info, b):
        i0, j0 = 1 + tx * TXE, 1 + ty * tye
        k0, k1 = int(kb[zc]), int(kb[zc + 1])
        ie = slice(i0, min(i0 + TXE, nx + 2))
        je = slice(j0, min(j0 + tye, ny + 2))
        e_cov[bb, k0:k1, je, ie] += 1
        hi = slice(0 if i0 == 1 else i0, min(i0 + TXE, nx + 2))
        hj = slice(0 if j0 == 1 else j0, min(j0 + tye, ny + 2))
        h_cov[bb, k0:k1, hj, hi] += 1
        if kind:
            xl, xh, yl, yh, zl, zh = npml
            # H cells with halo outside every x/y slab and wall; E halo loads inside the storage
            assert i0 - 1 >= xl + 1 and i0 + TXE - 1 <= nx - xh
            assert j0 - 1 >= yl + 1 and j0 + tye - 1 <= ny - yh
            assert i0 + TXE < X and j0 + tye < Y
            if kind == 1:  # interior chunks also avoid the z slabs and walls
                assert k0 - 1 >= zl + 1 and k1 - 1 <= nz - zh and k1 + 1 <= Z
            else:  # z-slab kernel: exactly the first or the last chunk
                assert zslab and (k0 == 1 or k1 == nz + 2)
    dom = (slice(None), slice(1, nz + 2), slice(1, ny + 2), slice(1, nx + 2))
    assert np.all(e_cov[dom] == 1)
    assert e_cov.sum() == e_cov[dom].sum()
    hdom = (slice(None), slice(1, nz + 2), slice(0, ny + 2), slice(0, nx + 2))
    assert np.all(h_cov[hdom] == 1)
    assert h_cov.sum() == h_cov[hdom].sum()
    if grid[0] >= 150 and sum(npml[:2]) < 60:
        assert info[5] > 0, "large grids must get interior tiles"
