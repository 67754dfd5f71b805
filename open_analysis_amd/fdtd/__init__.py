"""oa-fdtd engine-level API (mirrors ``packages/oa-fdtd/src/oa_fdtd``)."""

from .grid import C0, EPS0, ETA0, MU0, Grid
from .model import (
    Compiled,
    Conductor,
    LumpedPort,
    Model,
    ModelError,
    ModePort,
    Probe,
    Region,
    SoftSource,
    Waveform,
    compile_model,
)
from .pml import CpmlSpec
from .run import RunResult, make_engine, port_voltage_factor, run, source_spectrum, sparams

__all__ = [
    "C0", "EPS0", "ETA0", "MU0", "Compiled", "Conductor", "CpmlSpec", "Grid", "LumpedPort",
    "ModePort", "Model", "ModelError", "Probe", "Region", "RunResult", "SoftSource", "Waveform",
    "compile_model", "make_engine", "port_voltage_factor", "run", "source_spectrum", "sparams",
]
