"""Problem description and its compilation into the arrays the engines consume.

Mirrors ``packages/oa-fdtd/src/oa_fdtd/model.py:28-768`` (``Model`` and its parts,
``compile_model`` -> ``Compiled``). The compiled arrays use the reference's storage layout
(:mod:`.grid`) and its material-table convention: every E edge stores a ``uint16`` index into a
table of edge-averaged ``(eps, sigma)`` pairs, index 0 being PEC; ``ca``/``cb`` are rebuilt from
the table. Lumped ports are a loading conductivity plus a soft source (Piket-May 1994).

Quasi-TEM mode ports (``ModePort``) need the reference's 2-D static mode solver, which is not
part of the time-stepping path; compiling one raises ``FDTD_UNSUPPORTED``.
"""

from __future__ import annotations

import math
from collections.abc import Callable
from dataclasses import dataclass, field
from typing import Any

import numpy as np
import numpy.typing as npt

from .grid import C0, EPS0, Grid
from .pml import CpmlSpec, Pml, build_pml

FloatArray = npt.NDArray[np.float64]
WALL_KINDS = ("pec", "pmc", "cpml")
MAX_TABLE = 65535


class ModelError(ValueError):
    """Bad problem description; ``code`` is a stable ``FDTD_*`` identifier (model.py:28-33)."""

    def __init__(self, code: str, message: str, hint: str = "") -> None:
        super().__init__(f"{code}: {message}")
        self.code, self.message, self.hint = code, message, hint


@dataclass(frozen=True)
class Waveform:
    """Excitation ``s(t)`` (model.py:37-64).

    ``gauss_mod``: ``exp(-((t-t0)/tau)^2) sin(2 pi f0 (t-t0))``, ``tau = 2 sqrt(ln 10)/(pi bw)``.
    ``gauss``: the envelope alone, ``tau = sqrt(ln 10)/(pi bw)``.
    ``gauss_deriv`` (extension, no reference counterpart): the time derivative of the
    ``gauss`` pulse scaled to unit peak, ``-sqrt(2e) u exp(-u^2)`` with ``u = (t-t0)/tau``; it
    has no DC content. ``t0 = delay_tau * tau`` in every kind.
    """

    kind: str = "gauss_mod"
    f0: float = 0.0
    bandwidth: float = 0.0
    delay_tau: float = 5.0

    @property
    def tau(self) -> float:
        if self.bandwidth <= 0:
            raise ModelError("FDTD_BAD_WAVEFORM", "bandwidth must be positive")
        if self.kind == "gauss_mod":
            return 2.0 * math.sqrt(math.log(10.0)) / (math.pi * self.bandwidth)
        if self.kind in ("gauss", "gauss_deriv"):
            return math.sqrt(math.log(10.0)) / (math.pi * self.bandwidth)
        raise ModelError("FDTD_BAD_WAVEFORM", f"unknown waveform kind {self.kind!r}")

    def __call__(self, t: FloatArray) -> FloatArray:
        tau = self.tau
        u = (np.asarray(t, dtype=np.float64) - self.delay_tau * tau) / tau
        env = np.exp(-(u * u))
        if self.kind == "gauss":
            return env
        if self.kind == "gauss_deriv":
            return -math.sqrt(2.0 * math.e) * u * env
        return env * np.sin(2.0 * math.pi * self.f0 * u * tau)


@dataclass(frozen=True)
class Region:
    """Material box (cells whose centre lies inside) or PEC box (edges inside)."""

    lo: tuple[float, float, float]
    hi: tuple[float, float, float]
    eps_r: float = 1.0
    sigma: float = 0.0
    pec: bool = False
    poles: tuple[tuple[float, float], ...] = ()  # Debye (delta_eps, tau [s])


@dataclass(frozen=True)
class Conductor:
    """Copper box: PEC interior, surface-impedance faces (model.py:79-97)."""

    lo: tuple[float, float, float]
    hi: tuple[float, float, float]
    sigma: float = 5.8e7
    thickness: float | None = None
    roughness_rms: float = 0.0
    k_func: Callable[[FloatArray], FloatArray] | None = None
    model: str = "thick_sibc"
    edge_frac: float = 0.0


@dataclass(frozen=True)
class LumpedPort:
    name: str
    axis: int
    lo: tuple[float, float, float]
    hi: tuple[float, float, float]
    r_ohm: float = 50.0
    sign: float = 1.0


@dataclass(frozen=True)
class ModePort:
    """Quasi-TEM mode port (API-compatible placeholder; compiling one is unsupported)."""

    name: str
    axis: int
    position: float
    lo: tuple[float, float]
    hi: tuple[float, float]
    signals: tuple[tuple[tuple[float, float], tuple[float, float]], ...]
    mode: int = 0
    direction: int = 1
    sign: float = 1.0


@dataclass(frozen=True)
class SoftSource:
    name: str
    axis: int
    lo: tuple[float, float, float]
    hi: tuple[float, float, float]
    amplitude: float = 1.0


@dataclass(frozen=True)
class Probe:
    name: str
    kind: str  # "E" | "H"
    axis: int
    lo: tuple[float, float, float]
    hi: tuple[float, float, float]
    weight: str = "sum"  # "sum" | "line" | "mean"


@dataclass
class Model:
    """A complete engine-level problem (model.py:157-214)."""

    grid: Grid
    waveform: Waveform
    frequencies: FloatArray
    regions: list[Region] = field(default_factory=list)
    conductors: list[Conductor] = field(default_factory=list)
    walls: tuple[str, str, str, str, str, str] = ("pec",) * 6  # type: ignore[assignment]
    cpml: CpmlSpec = field(default_factory=CpmlSpec)
    ports: list[Any] = field(default_factory=list)
    sources: list[SoftSource] = field(default_factory=list)
    mode_probes: list[ModePort] = field(default_factory=list)
    probes: list[Probe] = field(default_factory=list)
    drive: FloatArray | None = None
    courant: float = 0.95
    excite: tuple[str, ...] | None = None
    f_band: tuple[float, float] | None = None

    @property
    def band(self) -> tuple[float, float]:
        if self.f_band is not None:
            return self.f_band
        f = np.asarray(self.frequencies, dtype=np.float64).ravel()
        return (float(f.min()), float(f.max()))

    @property
    def channels(self) -> int:
        return len(self.ports) + len(self.sources)

    def excited_indices(self) -> list[int]:
        """Driven ports in batch order (model.py:194-202)."""
        if self.excite is None:
            return list(range(len(self.ports)))
        names = [p.name for p in self.ports]
        unknown = [q for q in self.excite if q not in names]
        if unknown:
            raise ModelError("FDTD_BAD_DRIVE", f"excite names unknown ports {unknown}")
        return [names.index(q) for q in self.excite]

    def drive_matrix(self) -> FloatArray:
        """``(B, channels)`` drive amplitudes (model.py:204-214)."""
        if self.drive is not None:
            d = np.atleast_2d(np.asarray(self.drive, dtype=np.float64))
            if d.shape[1] != self.channels:
                raise ModelError("FDTD_BAD_DRIVE", "drive needs one column per port and source")
            return d
        if self.ports:
            rows = self.excited_indices()
            d = np.zeros((len(rows), self.channels))
            for r, p in enumerate(rows):
                d[r, p] = 1.0
            d[:, len(self.ports) :] = 1.0
            return d
        return np.ones((1, self.channels))


@dataclass
class AdeData:
    """Debye ADE edge lists, one entry per E component with dispersive edges."""

    kp: FloatArray
    tau: FloatArray
    comp: list[int]
    flat: list[npt.NDArray[np.int64]]
    beta: list[FloatArray]  # (P, ne)
    cbe: list[FloatArray]  # (ne,)


@dataclass
class SibcData:
    comp: npt.NDArray[np.int32]
    flat: npt.NDArray[np.int64]
    hcomp: npt.NDArray[np.int32]
    hflat: npt.NDArray[np.int64]
    hsign: FloatArray
    mat: npt.NDArray[np.int32]
    r0: FloatArray
    rp: FloatArray
    cp: FloatArray
    gp: FloatArray


@dataclass
class Compiled:
    """Everything the engines need, as plain host arrays (model.py:245-269)."""

    grid: Grid
    dt: float
    idx: list[npt.NDArray[np.uint16]]
    eps_tab: FloatArray
    sig_tab: FloatArray
    ca_tab: FloatArray
    cb_tab: FloatArray
    walls: tuple[str, ...]
    pml: Pml
    channels: list[dict[str, Any]]
    probes: list[dict[str, Any]]
    ports: list[dict[str, Any]]
    drive: FloatArray
    frequencies: FloatArray
    ade: AdeData | None = None
    sibc: SibcData | None = None
    probe_sums: dict[str, list[str]] = field(default_factory=dict)
    # per-component dispersive strength per pole on the edge lattice, kept for the HIP engine's
    # extended material table: list over c of (P, nz+1, ny+1, nx+1) float arrays or None
    edge_poles: list[FloatArray] | None = None

    @property
    def batch(self) -> int:
        return int(self.drive.shape[0])


# ---------------------------------------------------------------------------------------------
def _pole_taus(model: Model) -> FloatArray:
    """Sorted union of the relaxation times of all dispersive regions (model.py:272-275)."""
    taus = {round(float(t), 18) for r in model.regions if not r.pec for _, t in r.poles}
    return np.array(sorted(taus), dtype=np.float64)


def _cell_arrays(model: Model, taus: FloatArray) -> tuple[FloatArray, FloatArray, list[FloatArray]]:
    """Per-cell eps, sigma and pole strengths on ``(nz+2, ny+2, nx+2)`` (model.py:278-297)."""
    g = model.grid
    nx, ny, nz = g.n
    eps = np.ones((nz + 2, ny + 2, nx + 2))
    sig = np.zeros_like(eps)
    dels = [np.zeros_like(eps) for _ in range(taus.size)]
    cen = [g.centers(a) for a in range(3)]
    for r in model.regions:
        if r.pec:
            continue
        sel = [np.nonzero((cen[a] >= r.lo[a]) & (cen[a] <= r.hi[a]))[0] + 1 for a in range(3)]
        if min(s.size for s in sel) == 0:
            continue
        box = (slice(sel[2][0], sel[2][-1] + 1), slice(sel[1][0], sel[1][-1] + 1),
               slice(sel[0][0], sel[0][-1] + 1))  # cell centres are monotone: a contiguous run
        eps[box] = r.eps_r
        sig[box] = r.sigma
        strengths = {round(float(t), 18): d for d, t in r.poles}
        for pi, t in enumerate(taus):
            dels[pi][box] = strengths.get(round(float(t), 18), 0.0)
    return eps, sig, dels


def _ptrs(arrays: list[np.ndarray]) -> Any:
    import ctypes as C

    return (C.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])


def _edge_average(g: Grid, comp: int, arrays: list[FloatArray]) -> list[FloatArray]:
    """Area-weighted mean of the (up to) four cells around every ``comp`` edge, on the
    ``(nz+1, ny+1, nx+1)`` lattice ``[k-1, j-1, i-1]`` (model.py:300-321); the weights are the
    primal lengths of the two transverse axes, ghost cells weigh zero. Native
    (``oafd_compile_edge_average``), with the bits of the NumPy formulation."""
    from . import _lib

    n = g.n
    wts = [np.ascontiguousarray(g.cell_weights(a)[: n[a] + 2]) for a in range(3)]
    cells = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    out = [np.empty((n[2] + 1, n[1] + 1, n[0] + 1)) for _ in arrays]
    nn = np.array(n, np.int32)
    _lib.check(_lib.lib().oafd_compile_edge_average(
        nn.ctypes.data, comp, wts[0].ctypes.data, wts[1].ctypes.data, wts[2].ctypes.data, len(cells),
        _ptrs(cells), _ptrs(out)), "compile_model")
    return out


def _compile_ade(
    g: Grid,
    dt: float,
    taus: FloatArray,
    edge_del: list[list[FloatArray]],
    pec: list[npt.NDArray[np.bool_]],
    idx: list[npt.NDArray[np.uint16]],
    cb_tab: FloatArray,
) -> AdeData | None:
    """Dispersive edge lists: any pole strength > 1e-12 on a non-PEC edge (model.py:331-367).
    Native (``oafd_compile_ade_edges``), C order, with the bits of the NumPy formulation."""
    if taus.size == 0:
        return None
    import ctypes as C

    from . import _lib

    lib = _lib.lib()
    nn = np.array(g.n, np.int32)
    tau = np.ascontiguousarray(taus, np.float64)
    cbt = np.ascontiguousarray(cb_tab, np.float64)
    comps, flats, betas, cbes = [], [], [], []
    for c in range(3):
        dels = [np.ascontiguousarray(d, np.float64) for d in edge_del[c]]
        pc = np.ascontiguousarray(pec[c]).view(np.uint8)
        ne = C.c_int64(0)
        args = (nn.ctypes.data, g.pitch, taus.size, _ptrs(dels), pc.ctypes.data, tau.ctypes.data, float(dt),
                float(EPS0), idx[c].ctypes.data, cbt.ctypes.data)
        _lib.check(lib.oafd_compile_ade_edges(*args, 0, C.byref(ne), None, None, None), "compile_model")
        if ne.value == 0:
            continue
        flat = np.empty(ne.value, np.int64)
        beta = np.empty((taus.size, ne.value))
        cbe = np.empty(ne.value)
        _lib.check(lib.oafd_compile_ade_edges(*args, ne.value, C.byref(ne), flat.ctypes.data, beta.ctypes.data,
                                              cbe.ctypes.data), "compile_model")
        comps.append(c)
        flats.append(flat)
        betas.append(beta)
        cbes.append(cbe)
    if not comps:
        return None
    h = 0.5 * dt / taus
    return AdeData(kp=(1.0 - h) / (1.0 + h), tau=taus.copy(), comp=comps, flat=flats,
                   beta=betas, cbe=cbes)


def _compile_conductors(
    model: Model, g: Grid, pec: list[npt.NDArray[np.bool_]], dt: float
) -> SibcData | None:
    """PEC copper interiors and the surface-impedance edge list (model.py:370-466)."""
    if not model.conductors:
        return None
    from .subcell.edges import classify_box, outside_samples
    from .subcell.sibc import SibcFit, fit_surface_impedance

    freqs = np.asarray(model.frequencies, dtype=np.float64)
    wf = model.waveform
    top = wf.f0 + 0.5 * wf.bandwidth if wf.kind == "gauss_mod" else wf.bandwidth
    f_hi = float(max(model.band[1], top))
    if model.f_band is not None:
        f_lo = max(f_hi * 1e-3, 1e5)
    else:
        f_lo = max(float(np.min(freqs)) * 0.5, f_hi * 1e-3, 1e5)
    fit_of: dict[tuple[float, float, float, int], int] = {}
    fits: list[SibcFit] = []
    boxes = []
    for cd in model.conductors:
        edges = classify_box(g, cd.lo, cd.hi)
        for c in range(3):
            pec[c][edges.interior[c]] = True
        if cd.model == "pec":
            for c in range(3):
                pec[c][edges.surface[c]] = True
            continue
        if cd.model != "thick_sibc":
            raise ModelError("FDTD_BAD_CONDUCTOR", f"conductor model {cd.model!r}")
        thick = cd.thickness if cd.thickness is not None else min(
            h - lo for lo, h in zip(cd.lo, cd.hi, strict=True))
        if thick <= 0:
            raise ModelError("FDTD_BAD_CONDUCTOR", "conductor box has zero thickness")
        key = (cd.sigma, float(thick), cd.roughness_rms, id(cd.k_func))
        if key not in fit_of:
            fit_of[key] = len(fits)
            fits.append(fit_surface_impedance(cd.sigma, thick, f_lo, f_hi,
                                              rms=cd.roughness_rms, roughness=cd.k_func))
        boxes.append((edges, fit_of[key]))
    parts: list[tuple[npt.NDArray[Any], ...]] = []
    for edges, m in boxes:
        for c in range(3):
            k, j, i = edges.surface[c]
            if k.size == 0:
                continue
            free = ~pec[c][k, j, i]  # interiors, walls and PEC regions take precedence
            if not free.any():
                continue
            k, j, i = k[free], j[free], i[free]
            p, hc, hf, hs = outside_samples(g, c, k, j, i, edges.face_a[c][free],
                                            edges.face_b[c][free])
            parts.append((np.full(p.size, c), p, hc, hf, hs, np.full(p.size, m), k, j, i))
    if not parts:
        return None
    cols = [np.concatenate([pt[n] for pt in parts]) for n in range(9)]
    key = cols[0].astype(np.int64) * (1 << 44) + cols[1]
    # an edge shared by several boxes belongs to the last box that lists it
    rev_first = np.unique(key[::-1], return_index=True)[1]
    keep = np.sort(key.size - 1 - rev_first)
    cols = [v[keep] for v in cols]
    order = np.argsort(cols[0].astype(np.int64) * (1 << 44) + cols[1], kind="stable")
    cols = [v[order] for v in cols]
    for c in range(3):
        m = cols[0] == c
        pec[c][cols[6][m], cols[7][m], cols[8][m]] = True  # bulk writes 0, SIBC overwrites
    tau = fits[0].tau
    for f in fits[1:]:
        if f.tau.shape != tau.shape or not np.allclose(f.tau, tau):
            raise ModelError("FDTD_BAD_CONDUCTOR", "conductors need one common pole set")
    n_p = max(f.rp.size for f in fits)
    rp = np.zeros((len(fits), n_p))
    for mi, f in enumerate(fits):
        rp[mi, : f.rp.size] = f.rp
    cp, gp = fits[0].recursion(dt)
    return SibcData(
        comp=cols[0].astype(np.int32),
        flat=cols[1].astype(np.int64),
        hcomp=cols[2].astype(np.int32),
        hflat=cols[3].astype(np.int64),
        hsign=cols[4].astype(np.float64),
        mat=cols[5].astype(np.int32),
        r0=np.array([f.r0 for f in fits]),
        rp=rp,
        cp=cp,
        gp=gp,
    )


def _material_index(
    g: Grid, edge_eps: list[FloatArray], edge_sig: list[FloatArray], pec: list[npt.NDArray[np.bool_]]
) -> tuple[FloatArray, FloatArray, list[npt.NDArray[np.uint16]]]:
    """Sorted unique ``(round(eps, 12), round(sigma, 12))`` pairs over the non-PEC edges of all
    components and the uint16 storage index arrays (row + 1, 0 = PEC) (model.py:688-722).
    Native (``oafd_compile_material_table``): the table np.unique gives on the complex keys."""
    import ctypes as C

    from . import _lib

    nn = np.array(g.n, np.int32)
    e = [np.ascontiguousarray(a, np.float64) for a in edge_eps]
    s_ = [np.ascontiguousarray(a, np.float64) for a in edge_sig]
    p = [np.ascontiguousarray(a).view(np.uint8) for a in pec]
    idx = [np.empty(g.shape, dtype=np.uint16) for _ in range(3)]
    cap = MAX_TABLE - 1
    te, ts = np.empty(cap), np.empty(cap)
    rows = C.c_int64(0)
    _lib.check(_lib.lib().oafd_compile_material_table(
        nn.ctypes.data, g.pitch, _ptrs(e), _ptrs(s_), _ptrs(p), cap, te.ctypes.data, ts.ctypes.data,
        _ptrs(idx), C.byref(rows)), "compile_model")
    if rows.value < 0:
        raise ModelError("FDTD_TOO_MANY_MATERIALS",
                         f"more than {cap} averaged materials exceed the uint16 table",
                         "Use fewer distinct materials or a more uniform grid.")
    return te[: rows.value].copy(), ts[: rows.value].copy(), idx


def compile_model(model: Model, dtype: str = "float32") -> Compiled:
    """Indices, tables, CPML, channels and probes (model.py:501-768). Raises ModelError."""
    del dtype  # the compiled arrays are precision independent (host float64)
    g = model.grid
    nx, ny, nz = g.n
    if len(model.walls) != 6 or any(w not in WALL_KINDS for w in model.walls):
        raise ModelError("FDTD_BAD_WALLS", f"walls must be six of {WALL_KINDS}")
    if any(isinstance(p, ModePort) for p in model.ports) or model.mode_probes:
        raise ModelError("FDTD_UNSUPPORTED", "mode ports are not supported by this engine",
                         "Use lumped ports.")
    eps_min = min([1.0] + [r.eps_r for r in model.regions if not r.pec])
    dt = g.dt(model.courant, C0 / math.sqrt(eps_min))
    uniform = not model.regions and not model.ports and not model.conductors
    taus = _pole_taus(model)
    lat = (nz + 1, ny + 1, nx + 1)

    edge_eps: list[FloatArray] = []
    edge_sig: list[FloatArray] = []
    edge_del: list[list[FloatArray]] = []
    pec: list[npt.NDArray[np.bool_]] = []
    if not uniform:
        eps_c, sig_c, del_c = _cell_arrays(model, taus)
    for c in range(3):
        if uniform:
            edge_eps.append(np.ones(lat))
            edge_sig.append(np.zeros(lat))
            edge_del.append([])
        else:
            e, s, *ds = _edge_average(g, c, [eps_c, sig_c, *del_c])
            edge_eps.append(e)
            edge_sig.append(s)
            edge_del.append(ds)
        ok = np.zeros(lat, dtype=bool)
        (x0, x1), (y0, y1), (z0, z1) = g.valid("E", c)
        ok[z0 - 1 : z1 - 1, y0 - 1 : y1 - 1, x0 - 1 : x1 - 1] = True
        pec.append(~ok)
    if not uniform:
        del eps_c, sig_c, del_c
    # static part of the Debye polarisation current joins eps (model.py:534-537)
    for c in range(3):
        for pi, tau in enumerate(taus):
            edge_eps[c] += edge_del[c][pi] * (dt / (2.0 * float(tau) + dt))
    # tangential E on every non-PMC outer plane is PEC (model.py:538-549)
    for a in range(3):
        for side in (0, 1):
            if model.walls[2 * a + side] == "pmc":
                continue
            plane = 0 if side == 0 else g.n[a]
            for c in range(3):
                if c == a:
                    continue
                sl: list[Any] = [slice(None)] * 3
                sl[2 - a] = plane
                pec[c][tuple(sl)] = True
    for r in model.regions:  # PEC boxes (model.py:550-561)
        if not r.pec:
            continue
        for c in range(3):
            ids = []
            for a in (2, 1, 0):
                co = g.coords(a, a != c)
                tol = 1e-6 * float(np.min(g.d(a)))
                ids.append(np.nonzero((co >= r.lo[a] - tol) & (co <= r.hi[a] + tol))[0])
            if min(v.size for v in ids) == 0:
                continue
            pec[c][np.ix_(*ids)] = True

    sibc = _compile_conductors(model, g, pec, dt)

    channels: list[dict[str, Any]] = []
    ports_info: list[dict[str, Any]] = []
    probes: list[dict[str, Any]] = []
    for port in model.ports:
        if port.r_ohm <= 0:
            raise ModelError("FDTD_BAD_PORT", f"port {port.name}: r_ohm must be positive")
        flat, counts = g.points("E", port.axis, port.lo, port.hi)
        zs, ys, xs = g.shape
        k, rem = np.divmod(flat, ys * xs)
        j, i = np.divmod(rem, xs)
        ijk = (i, j, k)
        a = port.axis
        t0, t1 = [t for t in range(3) if t != a]
        length = g.d(a)[ijk[a] - 1]
        area = g.node_weights(t0)[ijk[t0]] * g.node_weights(t1)[ijk[t1]]
        n_ser = counts[a]
        total_len = float(length.sum()) / (flat.size / n_ser)
        total_area = float(area.sum()) / n_ser
        sig_eq = total_len / (port.r_ohm * total_area)
        edge_sig[a][k - 1, j - 1, i - 1] += sig_eq
        ports_info.append({"name": port.name, "axis": a, "r_ohm": port.r_ohm, "sign": port.sign,
                           "flat": flat, "channel": len(channels), "sig_eq": sig_eq,
                           "total_len": total_len, "total_area": total_area})
        channels.append({"comp": a, "flat": flat, "port": len(ports_info) - 1,
                         "sign": port.sign})
        probes.append({"name": f"V:{port.name}", "kind": "E", "comp": a, "flat": flat,
                       "weights": length * area / total_area})
    for src in model.sources:
        flat, _ = g.points("E", src.axis, src.lo, src.hi)
        channels.append({"comp": src.axis, "flat": flat, "amp": src.amplitude})
    for pr in model.probes:
        flat, _ = g.points(pr.kind, pr.axis, pr.lo, pr.hi)
        zs, ys, xs = g.shape
        k, rem = np.divmod(flat, ys * xs)
        j, i = np.divmod(rem, xs)
        ijk = (i, j, k)
        if pr.weight == "sum":
            w = np.ones(flat.size)
        elif pr.weight == "mean":
            w = np.full(flat.size, 1.0 / flat.size)
        elif pr.weight == "line":
            ax = pr.axis
            w = (g.d(ax)[ijk[ax] - 1] if pr.kind == "E" else g.node_weights(ax)[ijk[ax]])
            w = np.asarray(w, dtype=np.float64)
        else:
            raise ModelError("FDTD_BAD_PROBE", f"probe {pr.name}: weight {pr.weight!r}")
        probes.append({"name": pr.name, "kind": pr.kind, "comp": pr.axis, "flat": flat,
                       "weights": w})

    eps_u, sig_u, idx = _material_index(g, edge_eps, edge_sig, pec)
    eps_tab = np.concatenate(([0.0], eps_u))
    sig_tab = np.concatenate(([0.0], sig_u))
    ep = eps_tab * EPS0
    ep_safe = np.where(ep > 0, ep, 1.0)
    loss = sig_tab * dt / (2.0 * ep_safe)
    ca_tab = np.where(eps_tab > 0, (1.0 - loss) / (1.0 + loss), 0.0)
    cb_tab = np.where(eps_tab > 0, dt / ep_safe / (1.0 + loss), 0.0)

    ade = _compile_ade(g, dt, taus, edge_del, pec, idx, cb_tab)
    edge_poles = None
    if ade is not None:
        edge_poles = [np.stack(edge_del[c]) if edge_del[c] else np.zeros((0,) + lat)
                      for c in range(3)]

    for ch in channels:  # injection amplitudes per unit waveform (model.py:730-748)
        flat = ch["flat"]
        mat = idx[int(ch["comp"])].ravel()[flat]
        if "port" in ch:
            pinfo = ports_info[ch["port"]]
            ch["amps"] = ch["sign"] * cb_tab[mat] * pinfo["sig_eq"] / pinfo["total_len"]
        else:
            ch["amps"] = ch["amp"] * np.ones(flat.size)
        ch["waveform"] = model.waveform
    pml = build_pml(g, model.walls, model.cpml, dt)
    return Compiled(
        grid=g, dt=dt, idx=idx, eps_tab=eps_tab, sig_tab=sig_tab, ca_tab=ca_tab, cb_tab=cb_tab,
        walls=tuple(model.walls), pml=pml, channels=channels, probes=probes, ports=ports_info,
        drive=model.drive_matrix(), frequencies=np.asarray(model.frequencies, dtype=np.float64),
        ade=ade, sibc=sibc, edge_poles=edge_poles,
    )
