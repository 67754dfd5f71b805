"""CPU oracle of compile_model's edge-wise core (test infrastructure only: the tests compare
the native entry points ``oafd_compile_edge_average`` / ``oafd_compile_material_table`` /
``oafd_compile_ade_edges`` with these NumPy functions, bit for bit; the product calls the native
core).

NumPy restatements of packages/oa-fdtd/src/oa_fdtd/model.py: the area-weighted edge average
(:300-321), the merged material table with np.unique on rounded complex keys (:688-722) and
the Debye edge lists (:331-367). Pinned against the reference's golden compile outputs
(tests/golden, ``test_golden``) and by closed forms in tests/test_compile_core.py: a uniform
medium averages to itself, an interface edge gets the length-weighted mean of its cells, and
the table of a two-material grid has exactly those two rows in (eps, sigma) order.
"""

from __future__ import annotations

import numpy as np

from open_analysis_amd.fdtd.grid import EPS0  # the constant, not the computation


def edge_average(g, comp, arrays):
    """Area-weighted mean of the (up to) four cells around every ``comp`` edge, on the
    ``(nz+1, ny+1, nx+1)`` lattice ``[k-1, j-1, i-1]`` (model.py:300-321). The weights are the
    primal lengths of the two transverse axes; ghost cells weigh zero."""
    n = g.n
    p, q = [a for a in range(3) if a != comp]
    wts = [g.cell_weights(a)[: n[a] + 2] for a in range(3)]
    out_shape = (n[2] + 1, n[1] + 1, n[0] + 1)
    num = [np.zeros(out_shape) for _ in arrays]
    den = np.zeros(out_shape)
    for dp in (0, 1):
        for dq in (0, 1):
            shift = [0, 0, 0]
            shift[p], shift[q] = dp, dq
            sl = tuple(slice(1 - shift[a], n[a] + 2 - shift[a]) for a in (2, 1, 0))
            wp = wts[p][1 - dp : n[p] + 2 - dp]
            wq = wts[q][1 - dq : n[q] + 2 - dq]
            shp_p = [1, 1, 1]
            shp_p[2 - p] = wp.size
            shp_q = [1, 1, 1]
            shp_q[2 - q] = wq.size
            w = wp.reshape(shp_p) * wq.reshape(shp_q)
            for acc, arr in zip(num, arrays, strict=True):
                acc += w * arr[sl]
            den += w
    return [a / den for a in num]


def material_table(edge_eps, edge_sig, pec):
    """Sorted unique ``(round(eps, 12), round(sigma, 12))`` pairs over the non-PEC edges of all
    components (model.py:688-722); returns the table and per-component row indices (1-based)."""
    per_u, per_inv = [], []
    for c in range(3):
        keep = ~pec[c]
        key = np.round(edge_eps[c][keep], 12) + 1j * np.round(edge_sig[c][keep], 12)
        u, inv = np.unique(key, return_inverse=True)
        per_u.append(u)
        per_inv.append(inv.ravel())
    uni, merge = np.unique(np.concatenate(per_u), return_inverse=True)
    merge = merge.ravel()
    rows = []
    off = 0
    for u, inv in zip(per_u, per_inv, strict=True):
        rows.append(merge[off : off + u.size][inv] + 1)
        off += u.size
    return uni.real.copy(), uni.imag.copy(), rows


def index_arrays(g, pec, rows):
    """The uint16 storage index arrays from the table rows (model.py:716-722)."""
    nx, ny, nz = g.n
    out = []
    for c in range(3):
        lattice = np.zeros(pec[c].shape, dtype=np.uint16)
        lattice[~pec[c]] = rows[c].astype(np.uint16)
        arr = np.zeros(g.shape, dtype=np.uint16)
        arr[1 : nz + 2, 1 : ny + 2, 1 : nx + 2] = lattice
        out.append(arr)
    return out


def ade_lists(g, dt, taus, edge_del, pec, idx, cb_tab):
    """Dispersive edge lists: any pole strength > 1e-12 on a non-PEC edge (model.py:331-367)."""
    if taus.size == 0:
        return None
    comps, flats, betas, cbes = [], [], [], []
    for c in range(3):
        total = np.zeros_like(edge_del[c][0])
        for d in edge_del[c]:
            total += d
        mask = (total > 1e-12) & ~pec[c]
        k, j, i = np.nonzero(mask)
        if k.size == 0:
            continue
        zs, ys, xs = g.shape
        flats.append((((k + 1) * ys + (j + 1)) * xs + (i + 1)).astype(np.int64))
        betas.append(np.stack([EPS0 * edge_del[c][pi][mask] / (taus[pi] + 0.5 * dt)
                               for pi in range(taus.size)]))
        cbes.append(cb_tab[idx[c][k + 1, j + 1, i + 1]])
        comps.append(c)
    if not comps:
        return None
    return comps, flats, betas, cbes
