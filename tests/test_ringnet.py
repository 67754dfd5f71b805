"""The ringdown network: its CPU oracle against finite differences (CPU), and the HIP training
kernels against that oracle (GPU): the BF16 MFMA product, the loss and every parameter
gradient."""

from __future__ import annotations

import ctypes as C
import re
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import ringdown_net as ref

ROOT = Path(__file__).resolve().parents[1]


def _say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def _batch(S: int, T: int, F: int, seed: int = 0) -> ref.Batch:
    rng = np.random.default_rng(seed)
    dt = 1e-12
    U = np.zeros((32, S * T))
    t = np.arange(T) * 16 * dt
    for s in range(S):
        for i in range(4):
            f0 = rng.uniform(2e9, 15e9)
            U[i, s * T:(s + 1) * T] = np.exp(-t / 2e-10) * np.sin(2 * np.pi * f0 * t + rng.uniform(0, 6))
        U[4 + s % 4, s * T:(s + 1) * T] = 1.0
        U[8:20, s * T:(s + 1) * T] = rng.uniform(-1, 1, size=(12, 1))
    omega = 2 * np.pi * np.linspace(1e9, 20e9, F)
    kap = (rng.normal(size=(S, 4, F)) + 1j * rng.normal(size=(S, 4, F))) * 1e9
    tgt = (rng.normal(size=(S, 4, F)) + 1j * rng.normal(size=(S, 4, F))) * 0.05
    # every value the library receives in FP32 is made FP32-exact, so the oracle sees the same
    f32 = lambda x: np.asarray(x, dtype=np.float32).astype(np.float64)  # noqa: E731
    c32 = lambda z: f32(z.real) + 1j * f32(z.imag)  # noqa: E731
    return ref.Batch(U=f32(U), amp=f32(rng.uniform(0.5, 2.0, size=S)), kap=c32(kap), tgt=c32(tgt),
                     t0=f32(np.full(S, 3001 * dt)), omega=omega, dt=float(f32(dt)),
                     sig0=float(f32(1.0 / (200 * dt))), wmax=float(f32(2 * np.pi * 25e9)), T=T)


def _params(seed: int = 1) -> np.ndarray:
    from open_analysis_amd.ringdown.net import init_params

    p = init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 7)
    p[ref.OFF_W2:ref.NPARAM] += rng.normal(scale=0.05, size=ref.NPARAM - ref.OFF_W2)
    return p


def test_layout_matches_header_and_library() -> None:
    from open_analysis_amd.ringdown import net

    assert net.NPARAM == ref.NPARAM
    header = (ROOT / "include" / "oaring.h").read_text()
    declared = set(re.findall(r"\b(oaring_[a-z_]+)\s*\(", header))
    from open_analysis_amd.ringdown import _lib

    assert declared == set(_lib.EXPORTS)
    if _lib.LIB_PATH.exists():
        L = _lib.lib()
        for name in declared:
            assert hasattr(L, name), name
        assert L.oaring_param_count() == ref.NPARAM
        h = C.c_void_p()
        rc = L.oaring_create(0, 16, 8, C.byref(h))
        assert rc == _lib.E_INVALID and not h.value


def test_bf16_rounding_is_nearest_even() -> None:
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.5 - 2 ** -9, 3.0e-39], np.float32)
    got = ref.bf16(x)
    assert got[0] == 1.0 and got[1] == 1.0  # tie -> even
    assert got[2] == 1.0 + 4 * 2 ** -8       # tie -> even (up)
    assert got[3] == -1.5                    # below the tie
    assert abs(got[4] - 3.0e-39) < 1e-40


def test_oracle_gradient_matches_finite_differences() -> None:
    """The float64 oracle's analytic gradient against central differences, every block."""
    b = _batch(2, 16, 6)
    p = _params()
    loss, _, cache = ref.forward(p, b, rnd=False)
    g = ref.backward(p, b, cache, rnd=False)
    rng = np.random.default_rng(3)
    blocks = [(ref.OFF_WIN, ref.OFF_BIN), (ref.OFF_BIN, ref.OFF_WL)]
    for l in range(ref.NL):
        o = ref.OFF_WL + l * ref.LSTRIDE
        blocks += [(o, o + 2 * ref.C * ref.C), (o + 2 * ref.C * ref.C, o + ref.LSTRIDE)]
    blocks += [(ref.OFF_W1, ref.OFF_B1), (ref.OFF_B1, ref.OFF_W2), (ref.OFF_W2, ref.OFF_B2),
               (ref.OFF_B2, ref.NPARAM)]
    for lo, hi in blocks:
        idx = rng.integers(lo, hi, size=3)
        for i in idx:
            h = 1e-6 * max(1.0, abs(p[i]))
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (ref.forward(pp, b, rnd=False)[0] - ref.forward(pm, b, rnd=False)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-3 * abs(loss)), (lo, i, fd, g[i])


def test_planted_sign_error_in_oracle_is_caught() -> None:
    """A sign flip in the pole-frequency chain rule must fail the finite-difference check."""
    b = _batch(2, 16, 6)
    p = _params()
    loss, _, cache = ref.forward(p, b, rnd=False)
    g = ref.backward(p, b, cache, rnd=False)
    i = ref.OFF_B2 + 1  # v of pole 0, port 0
    h = 1e-6
    pp, pm = p.copy(), p.copy()
    pp[i] += h
    pm[i] -= h
    fd = (ref.forward(pp, b, rnd=False)[0] - ref.forward(pm, b, rnd=False)[0]) / (2 * h)
    assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-12)
    assert abs(fd - (-g[i])) > 1e-5 * max(abs(fd), 1e-12)


@pytest.mark.gpu
def test_gemm_bf16_lane_layout() -> None:
    """C = A B through the MFMA product kernel equals the BF16-rounded product (non-square,
    ragged sizes and a split reduction)."""
    from open_analysis_amd.ringdown import gemm_check

    rng = np.random.default_rng(0)
    for m, n, k in [(16, 16, 32), (64, 64, 32), (70, 33, 45), (128, 96, 1300)]:
        _say(f"[gemm] {m}x{n}x{k}")
        a = rng.normal(size=(m, k)).astype(np.float32)
        b = rng.normal(size=(k, n)).astype(np.float32)
        got = gemm_check(a, b)
        want = ref.bf16(a) @ ref.bf16(b)
        assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max() * np.sqrt(k), (m, n, k)


@pytest.mark.gpu
def test_network_loss_and_gradients_match_oracle() -> None:
    """Stage by stage on shared inputs: the network outputs against the oracle's; the pole layer
    (loss, predicted tails, output gradient) fed the GPU's own outputs; every parameter gradient.
    The outputs carry the BF16 operand roundings, which FP32 and FP64 pre-rounding values can
    resolve differently in a few elements; near a pole's resonance the tail is sensitive to
    such a change, so the pole layer is checked on identical inputs."""
    from open_analysis_amd.ringdown import BatchArrays, RingNet

    S, T, F = 8, 48, 20
    b = _batch(S, T, F, seed=2)
    p = _params(3)
    loss_ref, _, cache = ref.forward(p, b)
    net = RingNet(S, T, F, params=p.astype(np.float32))
    net.set_batch(BatchArrays(U=b.U, amp=b.amp, kap=b.kap, tgt=b.tgt, t0=b.t0, omega=b.omega,
                              dt=b.dt, sig0=b.sig0, wmax=b.wmax))
    loss, pred, raw = net.forward(pred=True, raw=True)
    o_err = np.abs(raw - cache["o"]).max() / np.abs(cache["o"]).max()
    _say(f"[ringnet] loss gpu {loss:.6e} oracle {loss_ref:.6e}; outputs rel err {o_err:.2e}")
    assert o_err < 2e-3
    assert abs(loss - loss_ref) <= 1e-3 * loss_ref
    # pole layer on the GPU's outputs
    l2, d_o, y2 = ref.pole_loss(raw.astype(np.float64), b)
    assert abs(loss - l2) <= 1e-5 * l2
    assert np.abs(pred - y2).max() <= 1e-5 * np.abs(y2).max()
    net.backward()
    g = net.grads().astype(np.float64)
    cache = dict(cache, d_o=d_o)
    g_ref = ref.backward(p, b, cache)
    blocks = {"Win": (ref.OFF_WIN, ref.OFF_BIN), "bin": (ref.OFF_BIN, ref.OFF_WL)}
    for l in range(ref.NL):
        o = ref.OFF_WL + l * ref.LSTRIDE
        blocks[f"W{l + 1}"] = (o, o + 2 * ref.C * ref.C)
        blocks[f"b{l + 1}"] = (o + 2 * ref.C * ref.C, o + ref.LSTRIDE)
    blocks.update({"W_h": (ref.OFF_W1, ref.OFF_B1), "b_h": (ref.OFF_B1, ref.OFF_W2),
                   "W_o": (ref.OFF_W2, ref.OFF_B2), "b_o": (ref.OFF_B2, ref.NPARAM)})
    # per block: relative L2 error, and the worst element against the block's largest. After
    # six backward layers a few BF16 operand roundings resolve differently for FP32 values that
    # differ in their last bits (GPU accumulation order against the oracle's), hence 1e-3 / 1e-2.
    l2, mx = {}, {}
    for name, (lo, hi) in blocks.items():
        d = g[lo:hi] - g_ref[lo:hi]
        l2[name] = np.linalg.norm(d) / max(np.linalg.norm(g_ref[lo:hi]), 1e-30)
        mx[name] = np.abs(d).max() / max(np.abs(g_ref[lo:hi]).max(), 1e-30)
        _say(f"[ringnet] grad {name}: rel L2 {l2[name]:.2e}, worst element {mx[name]:.2e}")
    assert max(l2.values()) < 1e-3, l2
    assert max(mx.values()) < 1e-2, mx
    # determinism: the same batch gives the same bits
    net.forward()
    net.backward()
    assert np.array_equal(net.grads().astype(np.float64), g)


@pytest.mark.gpu
def test_adam_step_matches_formula() -> None:
    from open_analysis_amd.ringdown import BatchArrays, RingNet

    S, T, F = 4, 32, 8
    b = _batch(S, T, F, seed=5)
    p = _params(4).astype(np.float32)
    net = RingNet(S, T, F, params=p)
    net.set_batch(BatchArrays(U=b.U, amp=b.amp, kap=b.kap, tgt=b.tgt, t0=b.t0, omega=b.omega,
                              dt=b.dt, sig0=b.sig0, wmax=b.wmax))
    net.forward()
    net.backward()
    g = net.grads()
    net.adam(1e-3, 1)
    got = net.params()
    m = 0.1 * g
    v = 0.001 * g * g
    want = p - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    assert np.abs(got - want).max() <= 1e-6
