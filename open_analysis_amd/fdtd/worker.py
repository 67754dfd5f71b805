"""``python -m open_analysis_amd.fdtd.worker job.json``: the output side of the reference's
adapter worker (packages/oa-fdtd/src/oa_fdtd/adapter/worker.py:1-60) on the HIP engine.

The reference worker reads an ``EMProblem`` (PCB ingest, which this package does not carry);
this one reads the engine-level job file of :mod:`open_analysis_amd.fdtd.cli`. It solves with
every excitation batched (batch member k drives port k) and writes the worker's two files next
to the job:

- ``ports.npz``: ``freqs`` (F,), ``s`` (F, N, N), ``z0`` (N,), ``names`` (N,) (worker.py:47);
- ``run.json``: the run statistics of ``run_meta`` (adapter/sim.py:204-220) plus
  ``fdtd_version`` and ``port_names`` (worker.py:48-51).

Progress lines are ``step <n>`` on stdout (at most one a second, and the last step), then one
JSON line ``{"steps", "stopped_by", "engine"}``. Failures print one ErrorInfo-shaped JSON line
(``code``, ``message``, ``path``, ``hint``) and exit with status 2.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path
from typing import Any

import numpy as np

from .cli import model_from_job
from .model import ModelError, compile_model
from .run import run, sparams

VERSION = "oa-fdtd 0.1.0 (open_analysis_amd HIP engine)"


def solve_job(path: Path, out=sys.stdout) -> dict[str, Any]:
    """Solve the job at ``path`` and write ``ports.npz`` and ``run.json`` beside it."""
    job = json.loads(path.read_text(encoding="utf-8"))
    model = model_from_job(job)
    if not model.ports:
        raise ModelError("FDTD_BAD_JOB", "the worker needs at least one port",
                         "Add a 'ports' list to the job.")
    opts = job.get("run", {})
    t0 = time.perf_counter()
    c = compile_model(model)
    wall_build = time.perf_counter() - t0
    last = [0.0]

    def progress(done: int) -> None:
        now = time.perf_counter()
        if now - last[0] >= 1.0:
            last[0] = now
            print(f"step {done}", file=out, flush=True)

    t0 = time.perf_counter()
    res = run(model, engine=opts.get("engine", "auto"), dtype=opts.get("dtype", "float32"),
              max_steps=int(opts.get("max_steps", 100_000)), end_db=opts.get("end_db", -50.0),
              progress=progress, compiled=c)
    wall_run = time.perf_counter() - t0
    s = np.moveaxis(sparams(model, res), 2, 0)  # (F, N, N), the worker's layout
    names = [p.name for p in model.ports]
    here = path.resolve().parent
    np.savez(here / "ports.npz", freqs=np.asarray(model.frequencies, np.float64), s=s,
             z0=np.array([float(p.r_ohm) for p in model.ports]), names=np.array(names))
    nx, ny, nz = c.grid.n
    meta: dict[str, Any] = {
        "engine": res.engine, "dtype": res.dtype, "steps": res.steps,
        "stopped_by": res.stopped_by, "energy_db": res.energy_db, "dt_s": res.dt,
        "peak_gpu_bytes": res.peak_bytes, "wall_build_s": wall_build, "wall_run_s": wall_run,
        "cells": int(nx * ny * nz), "diagnostics": [], "accel": res.report, "tail": res.tail,
        "fdtd_version": VERSION, "port_names": names,
    }
    (here / "run.json").write_text(json.dumps(meta, indent=1, default=str) + "\n", encoding="utf-8")
    print(f"step {res.steps}", file=out, flush=True)
    return meta


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="open_analysis_amd.fdtd.worker")
    ap.add_argument("job", type=Path)
    args = ap.parse_args(argv)
    try:
        meta = solve_job(args.job)
    except ModelError as exc:
        print(json.dumps({"code": exc.code, "message": exc.message, "path": str(args.job),
                          "hint": exc.hint}))
        return 2
    except Exception as exc:  # AccelError carries an ErrorInfo
        info = getattr(exc, "info", None)
        d = info.to_dict() if info is not None else {
            "code": getattr(exc, "code", "FDTD.ERROR"), "message": getattr(exc, "message", str(exc)),
            "path": str(args.job), "hint": getattr(exc, "hint", "")}
        print(json.dumps(d))
        return 2
    print(json.dumps({"steps": meta["steps"], "stopped_by": meta["stopped_by"],
                      "engine": meta["engine"]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
