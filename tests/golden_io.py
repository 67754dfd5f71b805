"""Loading of the committed golden vectors (written by tests/golden/make_golden.py from the
reference oa-fdtd NumPy engine) and building the matching problems with this package."""

from __future__ import annotations

from pathlib import Path
from typing import Any

import numpy as np

from open_analysis_amd.fdtd import grid as G
from open_analysis_amd.fdtd import model as M
from open_analysis_amd.fdtd.pml import CpmlSpec

GOLDEN = Path(__file__).resolve().parent / "golden"

OWN_CLASSES: dict[str, Any] = {
    "Grid": G.Grid, "Model": M.Model, "Waveform": M.Waveform, "Region": M.Region,
    "Conductor": M.Conductor, "LumpedPort": M.LumpedPort, "SoftSource": M.SoftSource,
    "Probe": M.Probe, "CpmlSpec": CpmlSpec,
}


def load(name: str) -> dict[str, np.ndarray]:
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def build(name: str) -> tuple[Any, int]:
    import cases

    builder, steps = cases.CASES[name]
    return builder(OWN_CLASSES), steps


def field_errors(fields: dict[str, np.ndarray], gold: dict[str, np.ndarray]) -> dict[str, float]:
    """Per component: max |sample difference| and |sum of squares difference|, relative to the
    largest sample (resp. sum of squares) among the three components of the same kind, so a
    component that is zero by symmetry is judged on the scale of its field."""
    out = {}
    for kind in ("E", "H"):
        names = [kind + a for a in "xyz"]
        scale = max(max(float(np.abs(gold[f"field_{n}_val"]).max()) for n in names), 1e-300)
        sq_scale = max(max(float(gold[f"field_{n}_sq"]) for n in names), 1e-300)
        for n in names:
            f = fields[n].ravel()
            pick = gold[f"field_{n}_pick"]
            out[n] = float(np.abs(f[pick] - gold[f"field_{n}_val"]).max()) / scale
            out[n + "_sq"] = abs(float((f * f).sum()) - float(gold[f"field_{n}_sq"])) / sq_scale
    return out
