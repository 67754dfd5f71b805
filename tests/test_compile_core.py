"""The native core of compile_model (edge averaging, material table and index arrays, Debye edge
lists; host C++) against its NumPy formulation (oracle/compile_core.py), bit for bit, and the
oracle against closed forms. CPU only: the native core runs on the host."""

from __future__ import annotations

import numpy as np
import pytest

import oracle.compile_core as OC
from open_analysis_amd.fdtd import _lib
from open_analysis_amd.fdtd import model as M
from open_analysis_amd.fdtd import workloads as W
from open_analysis_amd.fdtd.grid import Grid

pytestmark = pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="library not built")


def _models():
    import cases
    from golden_io import OWN_CLASSES

    from open_analysis_amd.fdtd import vias

    yield "gpu_cpu", cases.gpu_cpu_problem(OWN_CLASSES, 2)
    yield "interior_block", cases.interior_block(OWN_CLASSES)
    yield "debye_sibc_line", cases.debye_sibc_line(OWN_CLASSES)
    yield "seam", cases.tile_seam(OWN_CLASSES, 128, ("cpml", "cpml", "pmc", "pmc", "pmc", "pmc"))
    yield "microstrip", W.by_name("microstrip", reduced=True)
    yield "dispersive_microstrip", W.by_name("dispersive_microstrip", reduced=True)
    yield "stripline_stack", W.by_name("stripline_stack", reduced=True)
    yield "via", vias.build(vias.random_spec(np.random.default_rng(3), small=True))


def _oracle_index(g, edge_eps, edge_sig, pec):
    eu, su, rows = OC.material_table(edge_eps, edge_sig, pec)
    if eu.size + 1 > M.MAX_TABLE:
        raise M.ModelError("FDTD_TOO_MANY_MATERIALS", "table")
    return eu, su, OC.index_arrays(g, pec, rows)


def _oracle_ade(g, dt, taus, edge_del, pec, idx, cb_tab):
    r = OC.ade_lists(g, dt, taus, edge_del, pec, idx, cb_tab)
    if r is None:
        return None
    comps, flats, betas, cbes = r
    h = 0.5 * dt / taus
    return M.AdeData(kp=(1.0 - h) / (1.0 + h), tau=taus.copy(), comp=comps, flat=flats, beta=betas, cbe=cbes)


@pytest.mark.parametrize("name", ["gpu_cpu", "interior_block", "debye_sibc_line", "seam", "microstrip",
                                  "dispersive_microstrip", "stripline_stack", "via"])
def test_native_compile_equals_numpy_bitwise(name: str, monkeypatch) -> None:
    """compile_model with the native core and with the NumPy formulation of the same three steps
    produce identical arrays: index arrays, tables, coefficients, Debye lists, injection amps."""
    m = dict(_models())[name]
    a = M.compile_model(m)
    monkeypatch.setattr(M, "_edge_average", OC.edge_average)
    monkeypatch.setattr(M, "_material_index", _oracle_index)
    monkeypatch.setattr(M, "_compile_ade", _oracle_ade)
    b = M.compile_model(m)
    for c in range(3):
        assert np.array_equal(a.idx[c], b.idx[c]), c
    for k in ("eps_tab", "sig_tab", "ca_tab", "cb_tab"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert (a.ade is None) == (b.ade is None)
    if a.ade is not None:
        assert a.ade.comp == b.ade.comp and np.array_equal(a.ade.kp, b.ade.kp)
        for x, y in zip(a.ade.flat + a.ade.beta + a.ade.cbe, b.ade.flat + b.ade.beta + b.ade.cbe, strict=True):
            assert x.dtype == y.dtype and np.array_equal(x, y)
        for c in range(3):
            assert np.array_equal(a.edge_poles[c], b.edge_poles[c])
    for ca, cb in zip(a.channels, b.channels, strict=True):
        assert np.array_equal(ca["amps"], cb["amps"])


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_native_edge_average_equals_numpy_on_graded_mesh(comp: int) -> None:
    """Random cell values on a graded mesh: every edge average has the NumPy bits."""
    rng = np.random.default_rng(comp)
    g = Grid(np.cumsum(np.r_[0, rng.uniform(0.5, 1.5, 17)]) * 1e-4,
             np.cumsum(np.r_[0, rng.uniform(0.5, 1.5, 11)]) * 1e-4,
             np.cumsum(np.r_[0, rng.uniform(0.5, 1.5, 9)]) * 1e-4)
    nx, ny, nz = g.n
    arrays = [rng.uniform(1, 5, (nz + 2, ny + 2, nx + 2)) for _ in range(3)]
    for x, y in zip(M._edge_average(g, comp, arrays), OC.edge_average(g, comp, arrays), strict=True):
        assert np.array_equal(x, y)


def test_oracle_edge_average_closed_forms() -> None:
    """A uniform medium averages to itself; an x edge on the interface between two y layers of
    widths dy0, dy1 takes (dy0 e0 + dy1 e1) / (dy0 + dy1)."""
    g = Grid(np.linspace(0, 4e-3, 5), np.array([0.0, 1e-3, 3e-3]), np.linspace(0, 2e-3, 3))
    nx, ny, nz = g.n
    uni = np.full((nz + 2, ny + 2, nx + 2), 3.7)
    assert np.allclose(OC.edge_average(g, 0, [uni])[0], 3.7, rtol=1e-15)
    lay = np.ones((nz + 2, ny + 2, nx + 2))
    lay[:, 2, :] = 5.0  # the second y cell (1 mm .. 3 mm)
    ex = OC.edge_average(g, 0, [lay])[0]
    # x edge at y = 1 mm (lattice j = 1), interior in z: cells y0 (1 mm) and y1 (2 mm)
    assert np.isclose(ex[1, 1, 1], (1e-3 * 1.0 + 2e-3 * 5.0) / 3e-3, rtol=1e-14)


def test_oracle_material_table_two_materials() -> None:
    """Two materials give exactly two rows, sorted by (eps, sigma), and the PEC edges row 0."""
    shape = (3, 4, 5)
    e = [np.where(np.arange(60).reshape(shape) % 2 == 0, 4.4, 1.0) for _ in range(3)]
    s = [np.where(np.arange(60).reshape(shape) % 2 == 0, 0.02, 0.0) for _ in range(3)]
    pec = [np.zeros(shape, bool) for _ in range(3)]
    pec[1][0] = True
    eu, su, rows = OC.material_table(e, s, pec)
    assert np.array_equal(eu, [1.0, 4.4]) and np.array_equal(su, [0.0, 0.02])
    assert np.array_equal(rows[0], np.where(np.arange(60) % 2 == 0, 2, 1))
    assert rows[1].size == 60 - 20


def test_native_table_overflow_raises() -> None:
    """More distinct averaged materials than the uint16 table holds: FDTD_TOO_MANY_MATERIALS."""
    g = Grid(np.linspace(0, 1e-2, 41), np.linspace(0, 1e-2, 41), np.linspace(0, 1e-2, 41))
    lat = (41, 41, 41)
    rng = np.random.default_rng(0)
    e = [rng.uniform(1, 10, lat) for _ in range(3)]
    s = [np.zeros(lat) for _ in range(3)]
    pec = [np.zeros(lat, bool) for _ in range(3)]
    with pytest.raises(M.ModelError) as ei:
        M._material_index(g, e, s, pec)
    assert ei.value.code == "FDTD_TOO_MANY_MATERIALS"
