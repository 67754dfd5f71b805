"""The benchmark workloads on the HIP engine (GPU): S-parameter parity with the CPU oracle, and
the reduced-precision storage forms against an FP64 run of the same grid.

Sizes are reduced where the float64 oracle has to run the same problem (it steps about
15 Mcell/s on one core); the structures are those of the workload builders.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from open_analysis_amd.fdtd import compile_model, run, sparams
from open_analysis_amd.fdtd import workloads as W
from open_analysis_amd.fdtd.engine import HipEngine, hip_factory
from oracle import OracleEngine, oracle_factory

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def _db(x):
    return 20 * np.log10(np.abs(x))


def _compare(s_test, s_ref, floor_db=-40.0):
    """max |dB difference| and max |phase difference| [deg] over entries above ``floor_db``."""
    keep = _db(s_ref) > floor_db
    ddb = np.abs(_db(s_test) - _db(s_ref))[keep]
    dph = np.abs(np.angle(s_test / s_ref, deg=True))[keep]
    return float(ddb.max()), float(dph.max())


def test_microstrip_sparams_match_oracle() -> None:
    """Plain microstrip (the 160x80x48 workload's structure at 80x40x24, 0.3 mm): HIP FP32 S11/S21
    within 0.01 dB and 0.1 degree of the float64 oracle after the same 1500 steps."""
    _say("[config1-small] microstrip 80x40x24")
    m = W.microstrip(nx=80, ny=40, nz=24, d=0.3e-3, npml=6, strip_w=2, strip_len=60,
                     freqs=np.linspace(1e9, 10e9, 19))
    ref = run(m, engine=oracle_factory, dtype="float64", end_db=None, max_steps=1500)
    hip = run(m, engine=hip_factory, dtype="float32", end_db=None, max_steps=1500)
    s_ref, s_hip = sparams(m, ref), sparams(m, hip)
    ddb, dph = _compare(s_hip, s_ref)
    _say(f"  |dS| {ddb:.2e} dB, {dph:.2e} deg")
    assert ddb < 0.01 and dph < 0.1


def test_dispersive_microstrip_sparams_match_oracle() -> None:
    """Debye substrate + SIBC copper microstrip (the 400x200x100 workload's structure at
    100x50x30, 0.4 mm): HIP FP32 within 0.01 dB and 0.1 degree of the float64 oracle."""
    _say("[config2-small] dispersive microstrip 100x50x30")
    m = W.dispersive_microstrip(nx=100, ny=50, nz=30, d=0.4e-3, npml=6, sub=3, strip_w=3,
                                margin=6, freqs=np.linspace(1e9, 20e9, 20))
    ref = run(m, engine=oracle_factory, dtype="float64", end_db=None, max_steps=900)
    hip = run(m, engine=hip_factory, dtype="float32", end_db=None, max_steps=900)
    s_ref, s_hip = sparams(m, ref), sparams(m, hip)
    ddb, dph = _compare(s_hip, s_ref)
    _say(f"  |dS| {ddb:.2e} dB, {dph:.2e} deg")
    assert ddb < 0.01 and dph < 0.1


def test_microstrip_full_grid_matches_oracle() -> None:
    """The full 160x80x48 microstrip grid, 300 steps: probe DFTs and fields within 1e-4."""
    _say("[config1-full] microstrip 160x80x48, 300 steps")
    c = compile_model(W.microstrip())
    a = HipEngine(c, "float32", max_steps=300)
    b = OracleEngine(c, "float64")
    a.step(300)
    b.step(300)
    ra, rb = a.results(), b.results()
    for k in rb:
        assert np.abs(ra[k] - rb[k]).max() <= 1e-4 * np.abs(rb[k]).max(), k
    fa, fb = a.get_fields(), b.get_fields()
    for kind in "EH":
        scale = max(np.abs(fb[kind + x]).max() for x in "xyz")
        for x in "xyz":
            assert np.abs(fa[kind + x] - fb[kind + x]).max() <= 1e-4 * scale


@pytest.fixture(scope="module")
def stack_runs():
    """The six-layer stripline stack at 120x60x48 (0.25 mm) in every storage form, 2500 steps."""
    m, _ = W.stripline_stack(nx=120, ny=60, nz=48, d=0.25e-3, npml=8, strip_w=3, gap=3,
                             margin=8, freqs=np.linspace(1e9, 20e9, 39))
    out = {}
    for dt in ("float64", "float32", "float16", "bfloat16"):
        _say(f"[stack] {dt}")
        res = run(m, engine=hip_factory, dtype=dt, end_db=None, max_steps=2500)
        out[dt] = sparams(m, res)
    return out


def test_stack_float32_matches_float64(stack_runs) -> None:
    ddb, dph = _compare(stack_runs["float32"][:, 0], stack_runs["float64"][:, 0])
    _say(f"  float32 vs float64: {ddb:.2e} dB, {dph:.2e} deg")
    assert ddb < 0.01 and dph < 0.1


@pytest.mark.parametrize("dtype,db_tol,deg_tol", [("float16", 0.05, 0.5), ("bfloat16", 0.5, 5.0)])
def test_stack_half_storage_error(stack_runs, dtype: str, db_tol: float, deg_tol: float) -> None:
    """16-bit field storage (FP32 arithmetic, FP64 DFT): S21 and S11 error against the FP64 run.
    The bounds are the measured errors with margin; the values are printed per run."""
    ddb, dph = _compare(stack_runs[dtype][:, 0], stack_runs["float64"][:, 0])
    _say(f"  {dtype} vs float64: {ddb:.3e} dB, {dph:.3e} deg")
    assert ddb < db_tol and dph < deg_tol
