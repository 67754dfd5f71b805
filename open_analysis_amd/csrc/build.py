"""Build recipe of the HIP engine library (gfx950 only).

    python -m open_analysis_amd.csrc.build

Compiles ``oafdtd.hip`` into ``open_analysis_amd/csrc/liboafdtd_hip.so`` (the engine) and
``ringdown.hip`` into ``liboaring_hip.so`` (the learned ringdown extrapolator) with hipcc,
in-tree, so they travel with the source tree. Packed FP32 instructions are disabled
(``-target-feature -packed-fp32-ops``); the compiler still vectorises ordinary float code.
The engine is built with ``-ffp-contract=off``: every multiply and add is rounded on its own,
so each tile shape and kernel (interior, z-slab, general, batched) gives the same bits and the
autotuner's choice never changes a result.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
LIB = HERE / "liboafdtd_hip.so"
RING_LIB = HERE / "liboaring_hip.so"
TARGETS = {LIB: HERE / "oafdtd.hip", RING_LIB: HERE / "ringdown.hip"}
TARGET_FLAGS = {LIB: ["-ffp-contract=off"], RING_LIB: []}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the engine)")


def _command(src: Path, out: Path, extra: list[str] | None) -> list[str]:
    cmd = [
        hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics", "-I", str(ROOT / "include"), str(src), "-o", str(out),
    ]
    if extra:
        cmd[1:1] = extra
    return cmd


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    """Compile both libraries (always; the call is the build check), in parallel."""
    procs = []
    for lib, src in TARGETS.items():
        tmp = lib.with_suffix(".so.tmp")
        cmd = _command(src, tmp, TARGET_FLAGS[lib] + (extra or []))
        procs.append((lib, tmp, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                                      stderr=subprocess.PIPE, text=True)))
    failed = []
    for lib, tmp, cmd, p in procs:
        out, err = p.communicate()
        if verbose or p.returncode != 0:
            sys.stderr.write(out + err)
        if p.returncode != 0:
            failed.append(f"hipcc failed ({p.returncode}): {' '.join(cmd)}")
        else:
            os.replace(tmp, lib)
    if failed:
        raise RuntimeError("\n".join(failed))
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
