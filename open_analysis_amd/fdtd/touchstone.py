"""Touchstone files of S-parameters: real/imaginary pairs at full precision, version 1.1 when
every port has the same real reference impedance and version 2.0 otherwise (the reference's
writer, packages/oa-fdtd/src/oa_fdtd/adapter/touchstone.py:32-75, produces the same files:
port names in ``! Port[n] = name`` comments, 2-port data in S11 S21 S12 S22 order, larger
networks row by row with four pairs per line).
"""

from __future__ import annotations

from collections.abc import Sequence
from pathlib import Path

import numpy as np
import numpy.typing as npt

ComplexArray = npt.NDArray[np.complex128]
FloatArray = npt.NDArray[np.float64]


def _ri(z: complex) -> str:
    return f"{z.real:.12e} {z.imag:.12e}"


def entry_order(n: int) -> list[tuple[int, int]]:
    """(row, column) order of the entries on a frequency's lines."""
    if n == 2:
        return [(0, 0), (1, 0), (0, 1), (1, 1)]
    return [(i, j) for i in range(n) for j in range(n)]


def format_touchstone(freqs: FloatArray, s: ComplexArray, z0: Sequence[float],
                      names: Sequence[str], comments: Sequence[str] = ()) -> str:
    """``s`` is (F, N, N); returns the file text."""
    s = np.asarray(s)
    if s.ndim != 3 or s.shape[1] != s.shape[2] or s.shape[0] != len(freqs):
        raise ValueError("s must be (frequencies, ports, ports)")
    n = s.shape[1]
    if len(z0) != n or len(names) != n:
        raise ValueError("one reference impedance and one name per port")
    out = [f"! {c}" for c in comments]
    out += [f"! Port[{i + 1}] = {name}" for i, name in enumerate(names)]
    single = len({float(z) for z in z0}) == 1
    if single:
        out.append(f"# HZ S RI R {float(z0[0]):.12g}")
    else:
        out += ["[Version] 2.0", "# HZ S RI R 50", f"[Number of Ports] {n}"]
        if n == 2:
            out.append("[Two-Port Data Order] 21_12")
        out += [f"[Number of Frequencies] {len(freqs)}",
                "[Reference] " + " ".join(f"{float(z):.12g}" for z in z0),
                "[Matrix Format] Full", "[Network Data]"]
    order = entry_order(n)
    for k, f in enumerate(freqs):
        pairs = [_ri(complex(s[k, i, j])) for i, j in order]
        head = f"{float(f):.12e}"
        if n <= 2:
            out.append(head + " " + " ".join(pairs))
            continue
        for r in range(n):
            row = pairs[r * n:(r + 1) * n]
            for c in range(0, n, 4):
                lead = head + " " if r == 0 and c == 0 else "  "
                out.append(lead + " ".join(row[c:c + 4]))
    if not single:
        out.append("[End]")
    return "\n".join(out) + "\n"


def write_touchstone(path: str | Path, freqs: FloatArray, s: ComplexArray, z0: Sequence[float],
                     names: Sequence[str], comments: Sequence[str] = ()) -> Path:
    p = Path(path)
    p.write_text(format_touchstone(freqs, s, z0, names, comments), encoding="utf-8")
    return p


def read_touchstone(text: str) -> tuple[FloatArray, ComplexArray, list[float], list[str]]:
    """Parse a file written by :func:`format_touchstone` (RI, Hz): (freqs, s, z0, names)."""
    names: list[str] = []
    z0: list[float] = []
    n_ports = None
    nums: list[float] = []
    for line in text.splitlines():
        t = line.strip()
        if not t:
            continue
        if t.startswith("!"):
            if t.startswith("! Port["):
                names.append(t.split("=", 1)[1].strip())
            continue
        if t.startswith("#"):
            parts = t.split()
            z0 = [float(parts[parts.index("R") + 1])]
            continue
        if t.startswith("["):
            key, _, val = t.partition("]")
            key = key[1:].strip().lower()
            if key == "number of ports":
                n_ports = int(val)
            elif key == "reference":
                z0 = [float(v) for v in val.split()]
            continue
        nums.extend(float(v) for v in t.split())
    n = n_ports or len(names)
    per = 1 + 2 * n * n
    a = np.asarray(nums).reshape(-1, per)
    freqs = a[:, 0]
    vals = a[:, 1::2] + 1j * a[:, 2::2]
    s = np.zeros((len(freqs), n, n), np.complex128)
    for q, (i, j) in enumerate(entry_order(n)):
        s[:, i, j] = vals[:, q]
    if len(z0) == 1:
        z0 = z0 * n
    return freqs, s, z0, names
