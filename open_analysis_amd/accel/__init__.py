"""Device selection, memory budget, provenance and errors for the HIP engine.

Mirrors the surface of ``packages/oa-accel/src/oa_accel`` that the FDTD path calls
(``policy.py:147-171`` ``select``, ``memory.py:46-86`` ``MemoryBudget``, ``report.py:13-34``
``AccelReport``, ``errors.py:17-34`` ``ErrorInfo``/``AccelError``, ``tolerance.py:35`` the
``fdtd`` tolerance class). The device is discovered through the engine's own C ABI
(``oafd_device_count``/``oafd_device_info``), so no array library is involved.
"""

from __future__ import annotations

import os
from collections.abc import Mapping
from dataclasses import asdict, dataclass, field
from typing import Any, Literal

import numpy as np

GIB = 1024**3
RESERVE_GB = 2.0
DEFAULT_PINNED_GB = 4.0


@dataclass(frozen=True)
class ErrorInfo:
    code: str
    message: str
    path: str | None = None
    hint: str | None = None

    def to_dict(self) -> dict[str, str | None]:
        return asdict(self)


class AccelError(Exception):
    """Raised by device entry points; always carries an :class:`ErrorInfo`."""

    def __init__(self, info: ErrorInfo) -> None:
        super().__init__(f"{info.code}: {info.message}")
        self.info = info


@dataclass
class AccelReport:
    backend: str = "cpu"
    device_name: str = "cpu"
    device_id: int | None = None
    compute_capability: str | None = None
    driver_version: str | None = None
    runtime_version: str | None = None
    library_versions: dict[str, str] = field(default_factory=dict)
    peak_bytes: int = 0
    budget_bytes: int = 0
    fallbacks: list[str] = field(default_factory=list)
    deterministic: bool = True
    untested_on_hardware: bool = False

    def to_dict(self) -> dict[str, Any]:
        return asdict(self)

    def cache_key(self) -> str:
        lib = "oafdtd_hip" if self.backend != "cpu" else "numpy"
        ver = self.library_versions.get(lib, "0").split(".")[0]
        return f"{self.backend}/{lib}{ver}"


def _env_gb(env: Mapping[str, str], key: str) -> float | None:
    raw = env.get(key)
    if raw is None or raw.strip() == "":
        return None
    try:
        val = float(raw)
    except ValueError:
        val = -1.0
    if val < 0:
        raise AccelError(ErrorInfo("ACCEL_BAD_CONFIG", f"{key}={raw!r} is not a non-negative "
                                   "number of GiB.", f"env.{key}", f"Set {key} to e.g. 12."))
    return val


def default_gpu_bytes(total_bytes: int) -> int:
    """``min(0.75 total, total - 2 GiB)`` (memory.py:40-43)."""
    return max(0, int(min(0.75 * total_bytes, total_bytes - RESERVE_GB * GIB)))


@dataclass(frozen=True)
class MemoryBudget:
    gpu_bytes: int
    pinned_bytes: int
    total_bytes: int = 0

    @classmethod
    def from_env(cls, total_bytes: int, *, gpu_gb: float | None = None,
                 env: Mapping[str, str] | None = None) -> MemoryBudget:
        e = os.environ if env is None else env
        gpu = gpu_gb if gpu_gb is not None else _env_gb(e, "OA_GPU_MEM_GB")
        pinned = _env_gb(e, "OA_GPU_PINNED_GB")
        b = default_gpu_bytes(total_bytes) if gpu is None else int(gpu * GIB)
        if total_bytes:
            b = min(b, total_bytes)
        return cls(b, int((DEFAULT_PINNED_GB if pinned is None else pinned) * GIB), total_bytes)

    def check(self, estimate: int, op: str) -> ErrorInfo | None:
        if estimate <= self.gpu_bytes:
            return None
        return ErrorInfo("ACCEL_OUT_OF_MEMORY",
                         f"{op} needs about {estimate / GIB:.2f} GiB on the GPU; the budget is "
                         f"{self.gpu_bytes / GIB:.2f} GiB.", op,
                         f"estimate={estimate} B, budget={self.gpu_bytes} B. Raise "
                         "OA_GPU_MEM_GB or reduce the problem.")

    def require(self, estimate: int, op: str) -> None:
        """Refuse before allocating (memory.py:82-86)."""
        info = self.check(estimate, op)
        if info is not None:
            raise AccelError(info)


@dataclass(frozen=True)
class Problem:
    kind: str = "generic"
    cells: int = 0


@dataclass(frozen=True)
class DevicePolicy:
    device: str | None = None
    problem: Problem = field(default_factory=Problem)
    fallback: Literal["cpu", "error"] = "error"
    mem_gb: float | None = None


@dataclass
class DeviceInfo:
    index: int
    name: str
    arch: str
    total_bytes: int


@dataclass
class Backend:
    """The selected HIP device with its budget (the ``be`` of ``run.make_engine``)."""

    name: str
    device: DeviceInfo
    memory: MemoryBudget
    lib_version: str

    def report(self) -> AccelReport:
        return AccelReport(backend="rocm", device_name=self.device.name,
                           device_id=self.device.index, compute_capability=self.device.arch,
                           library_versions={"oafdtd_hip": self.lib_version},
                           budget_bytes=self.memory.gpu_bytes)


def probe_devices() -> list[DeviceInfo]:
    """Devices visible to the HIP runtime, via the engine library. Raises AccelError when the
    library is not built."""
    from ..fdtd._lib import lib

    L = lib()
    out = []
    import ctypes

    for i in range(L.oafd_device_count()):
        name = ctypes.create_string_buffer(256)
        arch = ctypes.create_string_buffer(64)
        total = ctypes.c_int64(0)
        if L.oafd_device_info(i, name, 256, arch, 64, ctypes.byref(total)) == 0:
            out.append(DeviceInfo(i, name.value.decode(), arch.value.decode(), total.value))
    return out


def select(policy: DevicePolicy | str | None = None) -> Backend:
    """Pick the HIP device (``OA_DEVICE=rocm:N`` or ``policy.device``); raises
    ``AccelError(ACCEL_NO_DEVICE)`` when none is visible (policy.py:147-171)."""
    if isinstance(policy, str) or policy is None:
        policy = DevicePolicy(device=policy)
    spec = policy.device or os.environ.get("OA_DEVICE") or "rocm"
    if spec.startswith("cpu"):
        raise AccelError(ErrorInfo("ACCEL_NO_CPU_ENGINE", "this package has no CPU engine",
                                   "device", "Select a ROCm device."))
    want = 0
    if ":" in spec:
        try:
            want = int(spec.split(":", 1)[1])
        except ValueError as exc:
            raise AccelError(ErrorInfo("ACCEL_BAD_DEVICE", f"bad device {spec!r}")) from exc
    devs = probe_devices()
    if not devs:
        raise AccelError(ErrorInfo("ACCEL_NO_DEVICE", "no ROCm device is visible", "device",
                                   "Run on a machine with an MI355X (gfx950)."))
    if want >= len(devs):
        raise AccelError(ErrorInfo("ACCEL_NO_DEVICE", f"device {want} does not exist"))
    from ..fdtd._lib import lib

    dev = devs[want]
    budget = MemoryBudget.from_env(dev.total_bytes, gpu_gb=policy.mem_gb)
    return Backend("rocm", dev, budget, lib().oafd_version().decode())


# ---- tolerance classes (tolerance.py:35-75)
@dataclass(frozen=True)
class Tolerance:
    rtol: float
    atol: float


TOLERANCES: dict[str, Tolerance] = {"fdtd": Tolerance(rtol=0.0, atol=1e-4)}


@dataclass(frozen=True)
class Comparison:
    ok: bool
    max_abs: float
    rel: float
    limit_abs: float
    limit_rel: float


def compare(gpu: Any, cpu: Any, cls: str = "fdtd", *, scale: float = 1.0) -> Comparison:
    tol = TOLERANCES[cls]
    a, b = np.asarray(gpu), np.asarray(cpu)
    lim = tol.atol * scale
    if a.shape != b.shape:
        return Comparison(False, float("inf"), float("inf"), lim, tol.rtol)
    diff = float(np.max(np.abs(a - b))) if a.size else 0.0
    ref = float(np.max(np.abs(b))) if b.size else 0.0
    rel = diff / ref if ref > 0 else diff
    ok = bool(np.isfinite(diff)) and (diff <= lim or (tol.rtol > 0 and rel <= tol.rtol))
    return Comparison(ok, diff, rel, lim, tol.rtol)
