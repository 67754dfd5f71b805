"""Test configuration: the ``gpu`` marker and shared helpers.

``gpu`` tests need the built HIP library and a visible MI355X; on a machine without a GPU they
are deselected by ``-m "not gpu"`` (and skipped with a reason if run anyway).
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (str(ROOT), str(ROOT / "tests"), str(ROOT / "tests" / "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs the HIP library and an MI355X (gfx950)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


def _gpu_reason() -> str | None:
    try:
        from open_analysis_amd.fdtd._lib import lib

        n = lib().oafd_device_count()
    except Exception as exc:  # library not built or not loadable
        return f"HIP library unavailable: {exc}"
    if n < 1:
        return "no ROCm device visible"
    return None


def pytest_collection_modifyitems(config: pytest.Config, items: list[pytest.Item]) -> None:
    gpu_items = [i for i in items if i.get_closest_marker("gpu")]
    if not gpu_items:
        return
    reason = _gpu_reason()
    if reason:
        for item in gpu_items:
            item.add_marker(pytest.mark.skip(reason=reason))
