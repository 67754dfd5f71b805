"""The job-file CLI and the Touchstone writer (CPU), and a job run end to end (GPU)."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from open_analysis_amd.fdtd import cli
from open_analysis_amd.fdtd.touchstone import format_touchstone, read_touchstone


def _job(tmp: Path, **run: object) -> Path:
    w, h, dy, ny = 0.008, 0.004, 0.5e-3, 40
    job = {
        "grid": {"x": list(np.linspace(0, w, 5)), "y": list(np.arange(ny + 1) * dy),
                 "z": list(np.linspace(0, h, 5))},
        "waveform": {"kind": "gauss_mod", "f0": 1.5e9, "bandwidth": 3e9},
        "frequencies": list(np.linspace(0.5e9, 3e9, 6)),
        "walls": ["pmc", "pmc", "pmc", "pmc", "pec", "pec"],
        "ports": [{"name": "p1", "axis": 2, "lo": [0, 0, 0], "hi": [w, 0, h], "r_ohm": 188.4},
                  {"name": "p2", "axis": 2, "lo": [0, ny * dy, 0], "hi": [w, ny * dy, h],
                   "r_ohm": 188.4}],
        "run": {"engine": "auto", "dtype": "float32", "max_steps": 400, "end_db": None, **run},
        "output": "ports.npz", "touchstone": "ports.s2p",
    }
    p = tmp / "job.json"
    p.write_text(json.dumps(job))
    return p


def test_touchstone_two_port_v11_layout() -> None:
    f = np.array([1e9, 2e9])
    s = np.zeros((2, 2, 2), complex)
    s[:, 0, 0], s[:, 1, 0], s[:, 0, 1], s[:, 1, 1] = 0.1, 0.2 + 0.5j, 0.3, 0.4
    text = format_touchstone(f, s, [50, 50], ["a", "b"], ["hello"])
    lines = text.splitlines()
    assert lines[:4] == ["! hello", "! Port[1] = a", "! Port[2] = b", "# HZ S RI R 50"]
    vals = [float(v) for v in lines[4].split()]
    assert vals[0] == 1e9 and vals[1:9] == [0.1, 0.0, 0.2, 0.5, 0.3, 0.0, 0.4, 0.0]
    f2, s2, z0, names = read_touchstone(text)
    assert np.array_equal(f2, f) and np.array_equal(s2, s) and z0 == [50, 50] and names == ["a", "b"]


def test_touchstone_v20_for_mixed_references() -> None:
    rng = np.random.default_rng(0)
    f = np.linspace(1e9, 5e9, 3)
    s = rng.normal(size=(3, 5, 5)) + 1j * rng.normal(size=(3, 5, 5))
    text = format_touchstone(f, s, [50, 50, 25, 50, 50], list("abcde"))
    assert "[Version] 2.0" in text and "[Reference] 50 50 25 50 50" in text
    assert text.rstrip().endswith("[End]")
    data = [ln for ln in text.splitlines() if ln[:1] in "0123456789 " and ln.strip()]
    assert len(data) == 3 * 5 * 2  # 5 rows of 5 pairs, four pairs per line
    f2, s2, z0, _ = read_touchstone(text)
    assert np.array_equal(f2, f) and np.allclose(s2, s, rtol=1e-11) and z0 == [50, 50, 25, 50, 50]


def test_malformed_job_exits_2_with_error_info(tmp_path: Path, capsys) -> None:
    p = tmp_path / "job.json"
    p.write_text(json.dumps({"grid": {"x": [0, 1]}}))
    assert cli.main(["run", str(p)]) == 2
    err = json.loads(capsys.readouterr().out)
    assert err["code"] == "FDTD_BAD_JOB" and err["path"] == str(p)


def test_cpu_engine_request_exits_2(tmp_path: Path, capsys) -> None:
    assert cli.main(["run", str(_job(tmp_path, engine="numpy"))]) == 2
    assert json.loads(capsys.readouterr().out)["code"] == "ACCEL_NO_CPU_ENGINE"


@pytest.mark.gpu
def test_job_runs_on_the_hip_engine(tmp_path: Path, capsys) -> None:
    from open_analysis_amd.fdtd import run, sparams

    p = _job(tmp_path)
    assert cli.main(["run", str(p)]) == 0
    summary = json.loads(capsys.readouterr().out)
    assert summary["engine"] == "hip" and summary["steps"] == 400
    out = np.load(tmp_path / "ports.npz")
    model = cli.model_from_job(json.loads(p.read_text()))
    res = run(model, engine="hip", dtype="float32", max_steps=400, end_db=None)
    s = sparams(model, res)
    assert np.array_equal(out["s"], s)
    _, s_ts, z0, names = read_touchstone((tmp_path / "ports.s2p").read_text())
    assert names == ["p1", "p2"]
    assert np.allclose(s_ts, np.moveaxis(s, 2, 0), rtol=1e-11, atol=1e-15)


def test_worker_without_ports_exits_2(tmp_path: Path, capsys) -> None:
    from open_analysis_amd.fdtd import worker

    p = _job(tmp_path)
    job = json.loads(p.read_text())
    job["ports"] = []
    p.write_text(json.dumps(job))
    assert worker.main([str(p)]) == 2
    err = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert err["code"] == "FDTD_BAD_JOB" and err["path"] == str(p)


def test_worker_cpu_engine_request_exits_2(tmp_path: Path, capsys) -> None:
    from open_analysis_amd.fdtd import worker

    assert worker.main([str(_job(tmp_path, engine="numpy"))]) == 2
    assert json.loads(capsys.readouterr().out.strip().splitlines()[-1])["code"] == "ACCEL_NO_CPU_ENGINE"


@pytest.mark.gpu
def test_worker_writes_ports_npz_and_run_json(tmp_path: Path, capsys) -> None:
    """The adapter worker's files (worker.py:47-51): ports.npz with s as (F, N, N), z0 and names,
    and run.json with the run_meta fields; S equal to run() + sparams() on the same job."""
    from open_analysis_amd.fdtd import run, sparams, worker

    p = _job(tmp_path)
    assert worker.main([str(p)]) == 0
    lines = capsys.readouterr().out.strip().splitlines()
    assert lines[-2] == "step 400"
    assert json.loads(lines[-1]) == {"steps": 400, "stopped_by": "max_steps", "engine": "hip"}
    out = np.load(tmp_path / "ports.npz")
    assert list(out["names"]) == ["p1", "p2"] and np.array_equal(out["z0"], [188.4, 188.4])
    model = cli.model_from_job(json.loads(p.read_text()))
    res = run(model, engine="hip", dtype="float32", max_steps=400, end_db=None)
    assert out["s"].shape == (6, 2, 2)
    assert np.array_equal(out["s"], np.moveaxis(sparams(model, res), 2, 0))
    assert np.array_equal(out["freqs"], model.frequencies)
    meta = json.loads((tmp_path / "run.json").read_text())
    for k in ("engine", "dtype", "steps", "stopped_by", "energy_db", "dt_s", "peak_gpu_bytes",
              "wall_build_s", "wall_run_s", "cells", "diagnostics", "accel", "tail",
              "fdtd_version", "port_names"):
        assert k in meta, k
    assert meta["cells"] == 4 * 40 * 4 and meta["port_names"] == ["p1", "p2"]
