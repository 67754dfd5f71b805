"""Reduced-precision field storage against an FP64 run of the same grid.

The FP16 and BF16 forms store the six field components in 16 bits and update them in FP32
registers; the running DFT stays FP64 with compensated summation. This module runs a model in
each storage form and reports every S-parameter entry's worst |dB| and phase error over the band
against the FP64 run, plus each run's wall time.

    python -m open_analysis_amd.fdtd.precision --workload stripline_stack --steps 30000 \\
        --out precision.json
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from typing import Any

import numpy as np

from .model import Model, compile_model
from .run import run, sparams

FLOOR_DB = -40.0


def entry_errors(s: np.ndarray, ref: np.ndarray, cols: list[int], names: list[str],
                 floor_db: float = FLOOR_DB) -> dict[str, dict[str, float]]:
    """Worst |20 log10|S| - 20 log10|S_ref|| [dB] and |arg(S / S_ref)| [deg] over the band for
    every driven column, counting frequencies where |S_ref| is above ``floor_db``."""
    out: dict[str, dict[str, float]] = {}
    for j in cols:
        for i in range(s.shape[0]):
            r, t = ref[i, j], s[i, j]
            db_r = 20.0 * np.log10(np.abs(r))
            keep = db_r > floor_db
            if not np.any(keep):
                continue
            ddb = np.abs(20.0 * np.log10(np.abs(t[keep])) - db_r[keep])
            dph = np.abs(np.angle(t[keep] / r[keep], deg=True))
            out[f"S({names[i]},{names[j]})"] = {
                "max_db": float(ddb.max()), "max_deg": float(dph.max()),
                "ref_min_db": float(db_r[keep].min()), "ref_max_db": float(db_r[keep].max())}
    return out


def storage_errors(model: Model, steps: int, dtypes: tuple[str, ...] = ("float32", "float16", "bfloat16"),
                   reference: str = "float64", engine: Any = "hip",
                   log: Any = None) -> dict[str, Any]:
    """Run ``model`` for ``steps`` steps in the ``reference`` storage form and in each of
    ``dtypes``; return the per-entry errors and the wall time of each run."""
    names = [p.name for p in model.ports]
    cols = model.excited_indices()
    c = compile_model(model)
    walls: dict[str, float] = {}
    smat: dict[str, np.ndarray] = {}
    for dt in (reference, *dtypes):
        last = [time.perf_counter()]

        def progress(done: int, dt: str = dt) -> None:
            now = time.perf_counter()
            if log is not None and now - last[0] > 20.0:
                last[0] = now
                print(f"[precision] {dt}: {done}/{steps} steps", file=log, flush=True)

        t0 = time.perf_counter()
        res = run(model, engine=engine, dtype=dt, end_db=None, max_steps=steps,
                  progress=progress, compiled=c)
        walls[dt] = time.perf_counter() - t0
        smat[dt] = sparams(model, res)
        if log is not None:
            print(f"[precision] {dt} done in {walls[dt]:.1f} s", file=log, flush=True)
    ref = smat[reference]
    f = np.asarray(model.frequencies)
    report: dict[str, Any] = {
        "steps": steps, "reference": reference, "band_hz": [float(f.min()), float(f.max())],
        "n_freq": int(f.size), "floor_db": FLOOR_DB, "wall_s": walls, "variants": {}}
    for dt in dtypes:
        errs = entry_errors(smat[dt], ref, cols, names)
        report["variants"][dt] = {
            "entries": errs,
            "worst_db": max(e["max_db"] for e in errs.values()),
            "worst_deg": max(e["max_deg"] for e in errs.values())}
    return report


def main(argv: list[str] | None = None) -> int:
    from . import workloads

    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", default="stripline_stack")
    ap.add_argument("--reduced", action="store_true", help="the workload's small grid")
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--dtypes", default="float32,float16,bfloat16")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    model = workloads.by_name(a.workload, reduced=a.reduced)
    rep = storage_errors(model, a.steps, tuple(a.dtypes.split(",")), log=sys.stderr)
    rep["workload"] = a.workload + (" (reduced)" if a.reduced else "")
    text = json.dumps(rep, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text + "\n")
    print(text, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
