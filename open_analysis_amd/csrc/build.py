"""Build recipe of the HIP engine library (gfx950 only).

    python -m open_analysis_amd.csrc.build

Compiles ``oafdtd.hip`` with hipcc into ``open_analysis_amd/csrc/liboafdtd_hip.so`` (in-tree, so
it travels with the source tree). Packed FP32 instructions are disabled
(``-target-feature -packed-fp32-ops``); the compiler still vectorises ordinary float code.
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
SOURCES = [HERE / "oafdtd.hip"]
DEPS = SOURCES + [HERE / "kernels.hpp", ROOT / "include" / "oafdtd.h"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the engine)")


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    """Compile the library (always; the call is the build check)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [
        hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics", "-I", str(ROOT / "include"),
        *[str(s) for s in SOURCES], "-o", str(tmp),
    ]
    if extra:
        cmd[1:1] = extra
    res = subprocess.run(cmd, capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
