"""Seeded differential via transitions of uneven size: the many-geometry workload.

Each geometry is a uniform 0.05 mm grid, 96..192 x 96 x 96..128 cells with an 8-cell CPML on
every face. A dielectric stack (three layers, permittivities drawn from U(3.0, 4.6)) carries two
PEC ground planes with square antipads; two square PEC via barrels join a top edge-coupled pair
(over the upper plane) to a bottom pair (under the lower plane). Four 50-ohm z-directed lumped
ports sit at the far ends of the four traces and every port is driven once (drive = I4, batch
4). The barrel side is the conformal-radius square of a round drill, ``a = r / 0.5902``
(the reference's barrel rule, packages/oa-fdtd/src/oa_fdtd/subcell/pcb.py:187-200), rounded to
whole cells.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .grid import Grid
from .model import LumpedPort, Model, Region, Waveform
from .pml import CpmlSpec

D = 0.05e-3
NPML = 8
SQUARE_EQ_RADIUS = 0.5902


@dataclass(frozen=True)
class ViaSpec:
    """Geometry of one transition, in cells unless stated."""

    nx: int
    ny: int
    nz: int
    eps: tuple[float, float, float]
    drill_r: float      # [m]
    antipad: int        # half side of the square antipad
    pitch: int          # via centre spacing (y)
    trace_w: int
    h_top: int          # top traces above the upper plane
    core: int           # plane spacing
    h_bot: int          # bottom traces below the lower plane
    npml: int = NPML

    def descriptor(self) -> np.ndarray:
        """12 numbers describing the geometry (sizes in mm)."""
        return np.array([self.nx * D * 1e3, self.ny * D * 1e3, self.nz * D * 1e3, *self.eps,
                         self.drill_r * 1e3, self.antipad * D * 1e3, self.pitch * D * 1e3,
                         self.trace_w * D * 1e3, (self.h_top + self.h_bot) * D * 1e3,
                         self.core * D * 1e3])


def random_spec(rng: np.random.Generator, small: bool = False) -> ViaSpec:
    """One seeded geometry; ``small`` draws 40..56 x 40 x 36..44 grids with a 4-cell CPML (the
    same structure at a size the CPU oracle steps quickly)."""
    if small:
        nx = int(rng.choice([40, 48, 56]))
        nz = int(rng.choice([36, 44]))
        npml, sizes = 4, dict(antipad=(3, 5), pitch=(10, 13), trace_w=(2, 4), h=(2, 4))
    else:
        nx = int(rng.choice([96, 128, 160, 192]))
        nz = int(rng.choice([96, 112, 128]))
        npml, sizes = NPML, dict(antipad=(4, 8), pitch=(16, 25), trace_w=(3, 6), h=(4, 9))
    ny = 40 if small else 96
    h_top, h_bot = (int(rng.integers(*sizes["h"])) for _ in range(2))
    core_max = nz - 2 * npml - 12 - h_top - h_bot
    return ViaSpec(
        nx=nx, ny=ny, nz=nz, eps=tuple(float(e) for e in rng.uniform(3.0, 4.6, size=3)),
        drill_r=float(rng.uniform(0.05e-3, 0.1e-3)), antipad=int(rng.integers(*sizes["antipad"])),
        pitch=int(rng.integers(*sizes["pitch"])), trace_w=int(rng.integers(*sizes["trace_w"])),
        h_top=h_top, core=int(rng.integers(6, max(core_max, 7))), h_bot=h_bot, npml=npml)


def build(spec: ViaSpec, freqs: np.ndarray | None = None, pulse: str = "gauss_mod") -> Model:
    """``pulse``: ``gauss_mod`` (10.5 GHz centre, 21 GHz bandwidth) or ``gauss`` (baseband,
    -20 dB at 21 GHz, half as long)."""
    s = spec
    g = Grid(np.arange(s.nx + 1) * D, np.arange(s.ny + 1) * D, np.arange(s.nz + 1) * D)
    f = np.linspace(1e9, 20e9, 96) if freqs is None else np.asarray(freqs)
    wf = Waveform("gauss_mod", 10.5e9, 21e9) if pulse == "gauss_mod" else Waveform("gauss", bandwidth=21e9)
    m = Model(g, wf, f, walls=("cpml",) * 6, courant=0.99)
    npml = s.npml
    m.cpml = CpmlSpec(cells=npml, m=3.0, r0=1e-6)
    # z levels (cells): bottom traces, lower plane, upper plane, top traces
    zb = (s.nz - (s.h_bot + s.core + s.h_top)) // 2
    zg1 = zb + s.h_bot
    zg2 = zg1 + s.core
    zt = zg2 + s.h_top
    z_cuts = [npml + 2, zg1, zg2, s.nz - npml - 2]
    for li in range(3):
        m.regions.append(Region((-1.0, -1.0, z_cuts[li] * D), (1.0, 1.0, z_cuts[li + 1] * D),
                                s.eps[li]))
    xv = s.nx // 2
    yc = s.ny // 2
    ys = (yc - s.pitch // 2, yc - s.pitch // 2 + s.pitch)
    ra = s.antipad
    # ground planes: full sheets minus the two square antipads
    for zp in (zg1, zg2):
        z = zp * D
        m.regions.append(Region((0.0, 0.0, z), ((xv - ra) * D, s.ny * D, z), pec=True))
        m.regions.append(Region(((xv + ra) * D, 0.0, z), (s.nx * D, s.ny * D, z), pec=True))
        m.regions.append(Region(((xv - ra) * D, 0.0, z), ((xv + ra) * D, (ys[0] - ra) * D, z),
                                pec=True))
        m.regions.append(Region(((xv - ra) * D, (ys[0] + ra) * D, z),
                                ((xv + ra) * D, (ys[1] - ra) * D, z), pec=True))
        m.regions.append(Region(((xv - ra) * D, (ys[1] + ra) * D, z),
                                ((xv + ra) * D, s.ny * D, z), pec=True))
    side = max(1, int(round(s.drill_r / SQUARE_EQ_RADIUS / D)))
    lo = side // 2
    x_in = npml + 6
    x_out = s.nx - npml - 6
    for n, y in zip(("1", "2"), ys, strict=True):
        # barrel, pads at both trace levels, traces
        m.regions.append(Region(((xv - lo) * D, (y - lo) * D, zb * D),
                                ((xv - lo + side) * D, (y - lo + side) * D, zt * D), pec=True))
        w0, w1 = y - s.trace_w // 2, y - s.trace_w // 2 + s.trace_w
        m.regions.append(Region((x_in * D, w0 * D, zt * D), ((xv + lo + 1) * D, w1 * D, zt * D),
                                pec=True))
        m.regions.append(Region(((xv - lo - 1) * D, w0 * D, zb * D), (x_out * D, w1 * D, zb * D),
                                pec=True))
        yp = (w0 + w1) * 0.5 * D
        m.ports.append(LumpedPort(f"t{n}", 2, (x_in * D, yp, zg2 * D), (x_in * D, yp, zt * D), 50.0))
        m.ports.append(LumpedPort(f"b{n}", 2, (x_out * D, yp, zb * D), (x_out * D, yp, zg1 * D),
                                  50.0))
    return m


def geometry_set(n: int, seed: int = 11, freqs: np.ndarray | None = None, small: bool = False,
                 pulse: str = "gauss_mod") -> tuple[list[Model], list[ViaSpec]]:
    """``n`` seeded transitions (uneven grids)."""
    rng = np.random.default_rng(seed)
    specs = [random_spec(rng, small) for _ in range(n)]
    return [build(s, freqs, pulse) for s in specs], specs
