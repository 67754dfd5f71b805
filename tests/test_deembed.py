"""Plane-wave de-embedding: the oracle (oracle/deembed.py) pinned by closed forms on the CPU, the
device FP64 entry points against the oracle on the GPU, and an FDTD microstrip end to end."""

from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import deembed as OD

RNG = np.random.default_rng(11)


def _line(g, a, b, z):
    """V(z) = A e^{-g z} + B e^{g z}: the forward and backward waves of a uniform line."""
    return a * np.exp(-g * z) + b * np.exp(g * z)


def _gamma_true(nf=40, lo=0.04, hi=2.6, d=0.01):
    f = np.linspace(0.5e9, 20e9, nf)
    beta = np.linspace(lo, hi, nf) / d
    alpha = 0.8 + 3.0 * np.sqrt(f / f.max())
    return f, alpha + 1j * beta


# ---- the oracle against closed forms -------------------------------------------------------


@pytest.mark.parametrize("db", [0.01, 0.0108])
def test_oracle_gamma_from_line_waves(db: float) -> None:
    """Three samples of A e^{-gz} + B e^{gz} give back g: equal spacing by the cosh formula for
    any reflection and phases up to 2.6 rad, unequal spacing (within 10 %) by the Newton
    iterations for the reference's working range (0.3 to 2.2 rad, |B| <= 0.3 |A|; with
    comparable waves the iteration can settle on the trivial root g = 0 of its residual)."""
    da = 0.01
    equal = db == da
    _, g = _gamma_true(d=max(da, db), lo=0.04 if equal else 0.3, hi=2.6 if equal else 2.2)
    a = RNG.normal(size=g.size) + 1j * RNG.normal(size=g.size)
    b = (1.5 if equal else 0.3) * np.abs(a) * np.exp(2j * np.pi * RNG.random(g.size))
    v1, v2, v3 = _line(g, a, b, -da), _line(g, a, b, 0.0), _line(g, a, b, db)
    got = OD.gamma_from_planes(v1, v2, v3, da, db)
    assert np.allclose(got, g, rtol=1e-9, atol=0)


def test_oracle_gamma_aliases_past_pi() -> None:
    """Past pi between planes the cosh formula folds the phase back (why the reference caps the
    plane phase at 2.2 rad): beta d = pi + x comes back as pi - x."""
    d = 0.01
    g = np.array([1.0 + 1j * (np.pi + 0.3) / d])
    v = [_line(g, 1.0, 0.2, z) for z in (-d, 0.0, d)]
    got = OD.gamma_from_planes(*v, d)
    assert np.allclose(got.imag * d, np.pi - 0.3) and np.allclose(got.real, 1.0)


def test_oracle_shift_reference_removes_a_line_section() -> None:
    """A matched line of length L has S21 = S12 = e^{-gL}; moving both reference planes inwards by
    L/2 leaves S21 = 1 (shift_reference's closed form)."""
    f, g = _gamma_true()
    length = 0.02
    s = np.zeros((2, 2, f.size), np.complex128)
    s[0, 1] = s[1, 0] = np.exp(-g * length)
    out = OD.shift_reference(s, np.stack([g, g]), [length / 2, length / 2])
    assert np.allclose(out[1, 0], 1.0, rtol=1e-12) and np.allclose(out[0, 0], 0.0)


def _synth(n, s, g, da, db, ref, r, a):
    """Plane voltages (n, 3, n, F) of ports whose waves at the reference planes are
    a (incident, n x n x F: port i, run j) and b = S a."""
    b = np.einsum("ikf,kjf->ijf", s, a)
    v = np.empty((n, 3, n, s.shape[2]), np.complex128)
    for i in range(n):
        amid = a[i] * np.sqrt(r[i]) * np.exp(g[i] * ref[i])
        bmid = b[i] * np.sqrt(r[i]) * np.exp(-g[i] * ref[i])
        for p, z in enumerate((-da[i], 0.0, db[i])):
            v[i, p] = _line(g[i], amid, bmid, z)
    return v


def _case(n, nf=30, off=0.05, unequal=False, seed=0):
    rng = np.random.default_rng(seed)
    f = np.linspace(1e9, 20e9, nf)
    s = 0.15 * (rng.normal(size=(n, n, nf)) + 1j * rng.normal(size=(n, n, nf)))
    a = np.eye(n)[:, :, None] * (1 + 0.1j) + off * (rng.normal(size=(n, n, nf)) + 1j * rng.normal(size=(n, n, nf)))
    da = 0.004 + 0.001 * rng.random(n)
    db = da * (1.08 if unequal else 1.0)
    g = np.stack([1.0 + 2.0 * (f / f.max()) + 1j * np.linspace(0.35, 2.0, nf) / da[i] for i in range(n)])
    ref = -0.002 - 0.003 * rng.random(n)
    r = 50.0 * (1 + rng.random(n))
    return f, s, a, g, da, db, ref, r, _synth(n, s, g, da, db, ref, r, a)


@pytest.mark.parametrize("n", [1, 2, 3])
@pytest.mark.parametrize("unequal", [False, True])
def test_oracle_wave_sparams_recovers_s(n: int, unequal: bool) -> None:
    """Waves synthesised from a known S (incident waves with small cross terms, as the undriven
    ports' own re-reflection) give that S back: S = B A^-1."""
    f, s, a, g, da, db, ref, r, v = _case(n, unequal=unequal, seed=n)
    got, gam = OD.wave_sparams(v, da, db, ref, r)
    if n == 1:  # one port: the driven-wave normalisation, b / a
        assert np.allclose(got, s, rtol=1e-9)
    else:
        assert np.allclose(got, s, rtol=1e-8, atol=1e-11)
    assert np.allclose(gam, g, rtol=1e-9)


def test_oracle_wave_sparams_normalises_by_the_driven_wave_when_ill_conditioned() -> None:
    """Cross terms above 0.3 of the smallest driven wave switch to S_ij = b_ij / a_jj."""
    f, s, a, g, da, db, ref, r, v = _case(2, off=0.6, seed=5)
    got, _ = OD.wave_sparams(v, da, db, ref, r)
    b = np.einsum("ikf,kjf->ijf", s, a)
    want = b / np.stack([a[j, j] for j in range(2)])[None, :, :]
    big = np.abs(a[0, 1]).max() > 0.3 * np.abs(a[0, 0]).min()
    assert big and np.allclose(got, want, rtol=1e-9)


def test_oracle_weak_band_takes_the_port_s() -> None:
    """Below min_phase (no fit) every entry comes from s_port; with a fit only the reflections do
    below floor_phase, and a non-passive plane S (|S| > 1.02) is replaced whole."""
    f, s, a, g, da, db, ref, r, v = _case(2, seed=3)
    s_port = np.full_like(s, 0.5 + 0.5j)
    phase = np.abs(g[0].imag) * da[0]
    got, _ = OD.wave_sparams(v, da, db, ref, r, s_port=s_port, min_phase=0.8)
    weak = (np.abs(g[0].imag) * da[0] < 0.8) | (np.abs(g[1].imag) * da[1] < 0.8)
    assert weak.any() and not weak.all()
    assert np.array_equal(got[:, :, weak], s_port[:, :, weak])
    got2, _ = OD.wave_sparams(v, da, db, ref, r, s_port=s_port, min_phase=0.01, freqs=f, floor_phase=0.8)
    noisy = np.abs(s).max(axis=(0, 1)) > 1.02
    for k in np.flatnonzero(weak & ~noisy):
        assert got2[0, 0, k] == s_port[0, 0, k] and got2[1, 1, k] == s_port[1, 1, k]
        assert np.isclose(got2[1, 0, k], s[1, 0, k], rtol=1e-8)
    assert np.array_equal(got2[:, :, noisy], s_port[:, :, noisy])
    assert phase.min() < 0.8


def test_oracle_smooth_gamma_reproduces_its_basis() -> None:
    """A gamma inside the fit basis (alpha = a1 sqrt(f) + a2 f + a3 f^2, beta = b1 f + b2 f^2 +
    b3 f^3) is reproduced where the plane phase is too small; conditioned points are kept."""
    f = np.linspace(0.2e9, 20e9, 60)
    fn = f / f.max()
    g = (0.5 * np.sqrt(fn) + 1.5 * fn + 0.3 * fn**2) + 1j * (300 * fn + 40 * fn**2 + 5 * fn**3)
    noisy = g + np.where(np.abs(g.imag) * 0.004 < 0.3, 7.0 + 3.0j, 0.0)
    got = OD.smooth_gamma(noisy, f, 0.004, 0.3)
    assert np.allclose(got, g, rtol=1e-9)
    ph = np.sort(np.abs(noisy.imag) * 0.004)[::-1]
    thr = 0.5 * (ph[4] + ph[5])  # only 5 conditioned points: left alone
    assert np.array_equal(OD.smooth_gamma(noisy, f, 0.004, thr), noisy)


# ---- the device entry points against the oracle ----------------------------------------------


@pytest.mark.gpu
@pytest.mark.parametrize("db", [0.01, 0.0108])
def test_device_gamma_from_planes_matches_oracle(db: float) -> None:
    from open_analysis_amd.fdtd import deembed as D

    print(f"[deembed] gamma_from_planes db={db}", flush=True)
    da = 0.01
    _, g = _gamma_true(d=max(da, db), lo=0.3, hi=2.2)
    rows = []
    for k in range(3):
        a = RNG.normal(size=g.size) + 1j * RNG.normal(size=g.size)
        b = 0.3 * np.abs(a) * np.exp(2j * np.pi * RNG.random(g.size))
        rows.append([_line(g * (1 + 0.1 * k), a, b, z) for z in (-da, 0.0, db)])
    v1, v2, v3 = (np.stack([r[p] for r in rows]) for p in range(3))
    got = D.gamma_from_planes(v1, v2, v3, da, db)
    want = np.stack([OD.gamma_from_planes(v1[k], v2[k], v3[k], da, db) for k in range(3)])
    assert np.allclose(got, want, rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_device_smooth_gamma_and_shift_reference_match_oracle() -> None:
    from open_analysis_amd.fdtd import deembed as D

    print("[deembed] smooth_gamma, shift_reference", flush=True)
    f = np.linspace(0.2e9, 20e9, 60)
    fn = f / f.max()
    g = (0.5 * np.sqrt(fn) + 1.5 * fn) + 1j * (300 * fn + 40 * fn**2)
    noisy = g + RNG.normal(size=f.size) * np.where(np.abs(g.imag) * 0.004 < 0.3, 5.0, 0.01)
    mask = np.ones(f.size, bool)
    mask[-5:] = False
    for m in (None, mask):
        assert np.allclose(D.smooth_gamma(noisy, f, 0.004, 0.3, m), OD.smooth_gamma(noisy, f, 0.004, 0.3, m),
                           rtol=1e-10, atol=1e-12)
    s = RNG.normal(size=(3, 3, f.size)) + 1j * RNG.normal(size=(3, 3, f.size))
    gs = [g, 1.1 * g, 0.9 * g]
    assert np.allclose(D.shift_reference(s, gs, [0.001, -0.002, 0.003]),
                       OD.shift_reference(s, np.stack(gs), [0.001, -0.002, 0.003]), rtol=1e-13, atol=0)
    assert np.allclose(D.shift_reference(s, g, [0.001, 0.0, 0.002]),
                       OD.shift_reference(s, np.stack([g, g, g]), [0.001, 0.0, 0.002]), rtol=1e-13, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2, 4])
@pytest.mark.parametrize("mode", ["plain", "ill", "port", "fit"])
def test_device_wave_sparams_matches_oracle(n: int, mode: str) -> None:
    from open_analysis_amd.fdtd import deembed as D

    print(f"[deembed] wave_sparams n={n} mode={mode}", flush=True)
    f, s, a, g, da, db, ref, r, v = _case(n, off=0.6 if mode == "ill" else 0.05, unequal=n == 2, seed=7 + n)
    kw: dict = {}
    if mode in ("port", "fit"):
        kw["s_port"] = 0.2 * (RNG.normal(size=s.shape) + 1j * RNG.normal(size=s.shape))
        kw["min_phase"] = 0.8
    if mode == "fit":
        kw.update(freqs=f, floor_phase=0.6, fit_mask=np.arange(f.size) < f.size - 3)
    want, wg = OD.wave_sparams(v, da, db, ref, r, **kw)
    planes = [D.PlaneSet(f"p{i}", (f"a{i}", f"b{i}", f"c{i}"), float(da[i]), float(ref[i]), float(db[i]))
              for i in range(n)]
    probes = {nm: v[i, p] for i in range(n) for p, nm in enumerate(planes[i].probes)}
    got, gg = D.wave_sparams(probes, planes, r, **kw)
    assert np.allclose(got, want, rtol=1e-10, atol=1e-13)
    assert np.allclose(np.stack(gg), wg, rtol=1e-10, atol=1e-13)


@pytest.mark.gpu
def test_microstrip_plane_waves_de_embed_the_ports() -> None:
    """An FDTD microstrip between two 50-ohm lumped ports (100x30x20 cells, 0.25 mm, 3000 steps,
    both ports driven): three voltage planes behind each port. The plane S (device) equals the
    oracle's on the same probe DFTs, and it removes the port regions: |S11| below -25 dB where
    the lumped-port S11 reaches -18 dB, |S21| within 0.2 dB of the lossless line and its phase
    within 3 degrees of e^{-gamma L} over the conditioned band."""
    from open_analysis_amd.fdtd import run, sparams
    from open_analysis_amd.fdtd import workloads as W
    from open_analysis_amd.fdtd.deembed import add_voltage_planes, wave_sparams

    print("[deembed] microstrip planes", flush=True)
    sys.stdout.flush()
    d = 0.25e-3
    m = W.microstrip(nx=100, ny=30, nz=20, d=d, npml=6, sub=2, strip_w=4, strip_len=80,
                     freqs=np.linspace(1e9, 10e9, 19))
    x0, x1 = 10 * d, 90 * d
    yc, zg, zs = 15 * d, 8 * d, 10 * d
    sp = 8 * d
    pls = [add_voltage_planes(m, "p1", [x0 + 6 * d + k * sp for k in range(3)], (0, yc, zg), (0, yc, zs), 2, 0),
           add_voltage_planes(m, "p2", [x1 - 6 * d - k * sp for k in range(3)], (0, yc, zg), (0, yc, zs), 2, 0)]
    res = run(m, engine="hip", dtype="float64", end_db=None, max_steps=3000)
    s_port = sparams(m, res)
    s, gam = wave_sparams(res.probes, pls, np.array([50.0, 50.0]))
    v = np.stack([np.stack([res.probes[nm][:2] for nm in pl.probes]) for pl in pls])
    want, _ = OD.wave_sparams(v, np.array([p.spacing for p in pls]), np.array([p.spacing for p in pls]),
                              np.array([p.ref for p in pls]), np.array([50.0, 50.0]))
    assert np.allclose(s, want, rtol=1e-10, atol=1e-13)
    ok = np.abs(gam[0].imag) * pls[0].spacing >= 0.3
    db = lambda x: 20 * np.log10(np.abs(x))  # noqa: E731
    assert db(s_port[0, 0]).max() > -20.0  # the lumped ports are not matched to the line
    assert db(s[0, 0]).max() < -25.0 and db(s[1, 1]).max() < -25.0
    assert np.abs(db(s[1, 0][ok])).max() < 0.2
    ph = np.angle(s[1, 0] * np.exp(gam[0] * (x1 - x0)), deg=True)
    assert np.abs(ph[ok]).max() < 3.0
