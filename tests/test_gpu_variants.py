"""Every tridiagonalisation kernel variant through the full dense eigensolver
(runs on the MI355X).

The library picks the hemv variant per trailing-size bucket and the
trailing-update kind by timing them on the device, so an ordinary run only
exercises the winners.  Here each variant is pinned (BSE_HEMV_CODE,
BSE_HER2K; read when a context tunes) and a Hermitian matrix of order 700
(not a multiple of the 64-row tile) is diagonalised in FP64 storage (all
vectors) and in FP32 storage (16 vectors), against LAPACK eigenvalues.
Tolerances: FP64 1e-12 x max|lambda| x sqrt(n), residual 1e-11, orthogonality
1e-10; FP32 storage 1e-5 x max|lambda|.
"""
import ctypes

import numpy as np
import pytest

from bse_amd import _abi

pytestmark = pytest.mark.gpu

N = 700
# chunk length ch in {4, 8, 16} + 32 u2 (u2 = 0 plain, 1 two tile rows per
# step, 2 butterfly column sums, 3 the same over 4 tile columns) + 128 for the
# default cache policy
HEMV_CODES = [ch + 32 * u2 + 128 * t for t in (0, 1) for u2 in range(4) for ch in (4, 8, 16)]


@pytest.fixture(scope="module")
def problem():
    rng = np.random.default_rng(700)
    a = rng.normal(size=(N, N)) + 1j * rng.normal(size=(N, N))
    a = 0.5 * (a + a.conj().T)
    return a, np.linalg.eigvalsh(a)


def _check(a, ref, prec, ev, z):
    scale = np.abs(ref).max()
    if prec == "fp64":
        assert np.abs(ev - ref).max() < 1e-12 * scale * N ** 0.5
        assert np.abs(a @ z - z * ev[None, :]).max() < 1e-11 * scale * N ** 0.5
        assert np.abs(z.conj().T @ z - np.eye(N)).max() < 1e-10
    else:
        assert np.abs(ev - ref).max() < 1e-5 * scale
        assert z.shape == (N, 16)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("code", HEMV_CODES)
def test_hemv_variant(problem, monkeypatch, code, prec):
    print(f"hemv variant {code} {prec} n={N}", flush=True)
    monkeypatch.setenv("BSE_HEMV_CODE", str(code))
    a, ref = problem
    with _abi.Context(0) as ctx:
        ev, z = ctx.eigh_dense(a, prec, -1 if prec == "fp64" else 16)
    _check(a, ref, prec, ev, z)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["valu", "mfma", "mfma2"])
def test_trailing_update_kind(problem, monkeypatch, kind, prec):
    print(f"trailing update {kind} {prec} n={N}", flush=True)
    monkeypatch.setenv("BSE_HER2K", kind)
    a, ref = problem
    with _abi.Context(0) as ctx:
        ev, z = ctx.eigh_dense(a, prec, -1 if prec == "fp64" else 16)
    _check(a, ref, prec, ev, z)


def test_instruction_rate_probe():
    """The instruction-rate probe (a pure arithmetic loop) reports a positive
    rate below the published peak of each form (FP32 MFMA 157.3, FP64 MFMA
    and FP64 FMA 78.6 TFLOP/s); an unknown form is refused."""
    peaks = {"mfma_f32_16x16x4": 157.3, "mfma_f64_16x16x4": 78.6, "fma_f64": 78.6}
    for form, peak in peaks.items():
        r = _abi.probe_rate(form)
        print(f"{form}: {r:.1f} TFLOP/s", flush=True)
        assert 0.05 * peak < r < 1.05 * peak
    lib = _abi.load_library()
    out = ctypes.c_double(0.0)
    assert lib.bse_probe_rate(7, ctypes.byref(out)) != 0
