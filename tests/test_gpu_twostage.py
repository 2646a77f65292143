"""Two-stage Hermitian reduction (dense -> band -> tridiagonal) on the MI355X.

The two-stage path serves eigendecompositions that form at most 20
eigenvectors: all eigenvalues from the tridiagonal, the wanted vectors by
inverse iteration on the band matrix and the first-stage reflectors, the
oscillator strengths of every exciton from Q^H rho carried through the
bulge chase.  It must give what the one-stage path gives: eigenvalues to
1e-12 ||A|| sqrt(n), residuals and orthonormality of the vectors, and the
oracle's results for the BSE systems.  Contexts here force the path on
(BSE_TWO_STAGE=1) or off (=0) for any size.
"""
import os
import pathlib

import numpy as np
import pytest

import bse_amd
from bse_amd import _abi
from parity_util import cluster_overlap_error, cluster_f2_error

pytestmark = pytest.mark.gpu
GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def _context(mode):
    old = os.environ.get("BSE_TWO_STAGE")
    os.environ["BSE_TWO_STAGE"] = mode
    try:
        return _abi.Context(0)
    finally:
        if old is None:
            del os.environ["BSE_TWO_STAGE"]
        else:
            os.environ["BSE_TWO_STAGE"] = old


@pytest.fixture(scope="module")
def ctx2():
    c = _context("1")
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx1():
    c = _context("0")
    yield c
    c.close()


def _herm(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return 0.5 * (a + a.conj().T)


@pytest.mark.parametrize("n,k", [(2, 1), (5, 3), (33, 8), (64, 16), (65, 16), (97, 5), (200, 16),
                                 (333, 16), (1000, 12), (2049, 16)])
def test_two_stage_dense_random(ctx2, n, k):
    print(f"two-stage dense n={n} k={k}", flush=True)
    a = _herm(n, n + 7)
    ev, z = ctx2.eigh_dense(a, "fp64", k)
    ref = np.linalg.eigvalsh(a)
    scale = np.abs(ref).max()
    assert np.abs(ev - ref).max() < 1e-12 * scale * max(1, n) ** 0.5
    assert np.abs(a @ z - z * ev[None, :k]).max() < 1e-11 * scale * max(1, n) ** 0.5
    assert np.abs(z.conj().T @ z - np.eye(k)).max() < 1e-10


def test_two_stage_eigenvalues_only(ctx2):
    a = _herm(700, 3)
    ev, _ = ctx2.eigh_dense(a, "fp64", 0)
    ref = np.linalg.eigvalsh(a)
    assert np.abs(ev - ref).max() < 1e-12 * np.abs(ref).max() * 700 ** 0.5


def test_two_stage_fp32_storage(ctx2):
    a = _herm(600, 11)
    ev, z = ctx2.eigh_dense(a, "fp32", 16)
    ref = np.linalg.eigvalsh(a)
    scale = np.abs(ref).max()
    assert np.abs(ev - ref).max() < 1e-5 * scale
    assert np.abs(z.conj().T @ z - np.eye(16)).max() < 1e-10
    assert np.abs(a @ z - z * ev[None, :16]).max() < 1e-4 * scale


def test_two_stage_structured(ctx2):
    """Already banded, diagonal and exactly degenerate matrices (zero panels,
    split tridiagonals, clusters of equal eigenvalues)."""
    n = 300
    i = np.arange(n)
    lap = np.zeros((n, n), dtype=np.complex128)
    lap[i, i] = 2.0
    lap[i[:-1], i[:-1] + 1] = -1.0
    lap[i[:-1] + 1, i[:-1]] = -1.0
    ev, z = ctx2.eigh_dense(lap, "fp64", 4)
    ref = 2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    assert np.abs(ev - ref).max() < 1e-12 * 4 * n ** 0.5
    assert np.abs(lap @ z - z * ev[None, :4]).max() < 1e-11
    dg = np.diag(np.linspace(-3, 5, n)).astype(np.complex128)
    ev, z = ctx2.eigh_dense(dg, "fp64", 3)
    assert np.abs(ev - np.linspace(-3, 5, n)).max() < 1e-13
    rng = np.random.default_rng(7)
    lam = np.repeat(np.arange(n // 4, dtype=float), 4)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    a = (q * lam) @ q.conj().T
    a = 0.5 * (a + a.conj().T)
    ev, z = ctx2.eigh_dense(a, "fp64", 8)
    assert np.abs(ev - lam).max() < 1e-11 * n
    assert np.abs(z.conj().T @ z - np.eye(8)).max() < 1e-10
    assert np.abs(a @ z - z * ev[None, :8]).max() < 1e-10 * n


def test_two_stage_bse_c2_against_fixture(ctx2):
    g = np.load(GOLD / "c2_honeycomb_magnetic_30.npz", allow_pickle=False)
    s = bse_amd.honeycomb_magnetic(30)
    print("c2 N=3600 two-stage fp64, 16 vectors", flush=True)
    ctx2.set_system(s).build("fp64").eigh(16)
    E = ctx2.eigenvalues()
    scale = np.abs(g["E"]).max()
    assert np.abs(E - g["E"]).max() < 1e-10 * scale
    f2 = np.abs(ctx2.oscillator()) ** 2
    assert cluster_f2_error(g["E"], f2, g["f2"], 1e-8 * scale) < 1e-9
    eps = ctx2.spectrum(g["omegas"], float(g["eta"]))
    assert np.linalg.norm(eps.imag - g["eps"].imag) < 1e-10 * np.linalg.norm(g["eps"].imag)
    Z = ctx2.eigenvectors(16)
    assert cluster_overlap_error(Z, g["Z"], g["E"], 1e-8 * scale, ncol=16) < 1e-10
    assert np.abs(Z.conj().T @ Z - np.eye(16)).max() < 1e-10


def test_two_stage_matches_one_stage_mixed(ctx1, ctx2):
    s = bse_amd.random_multiorbital(10)
    out = []
    for c in (ctx1, ctx2):
        c.set_system(s).build("bf16x2").eigh(16, refine=True)
        rn, rounds, status = c.refine_info(16)
        assert status == 0
        out.append((c.eigenvalues(), c.oscillator()))
    (E1, f1), (E2, f2) = out
    scale = np.abs(E1).max()
    assert np.abs(E1[:16] - E2[:16]).max() < 1e-10 * scale
    assert np.abs(E1 - E2).max() < 1e-6 * scale


def test_two_stage_structured_fp32(ctx2):
    """Zero panels in FP32 storage: the CholeskyQR2 path flags them and the
    TSQR fallback runs inside the look-ahead's side stream."""
    n = 300
    i = np.arange(n)
    lap = np.zeros((n, n), dtype=np.complex128)
    lap[i, i] = 2.0
    lap[i[:-1], i[:-1] + 1] = -1.0
    lap[i[:-1] + 1, i[:-1]] = -1.0
    ev, z = ctx2.eigh_dense(lap, "fp32", 4)
    ref = 2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    assert np.abs(ev - ref).max() < 1e-5 * 4
    assert np.abs(z.conj().T @ z - np.eye(4)).max() < 1e-10
    dg = np.diag(np.linspace(-3, 5, n)).astype(np.complex128)
    ev, _ = ctx2.eigh_dense(dg, "fp32", 3)
    assert np.abs(ev - np.linspace(-3, 5, n)).max() < 1e-6


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_two_stage_lookahead_is_exact(ctx2, prec):
    """The look-ahead splits each trailing update into the next panel's tile
    column (second stream) and the other tiles: the same arithmetic, so the
    results equal those of the single-stream order bit for bit."""
    a = _herm(1100, 5)
    old = os.environ.get("BSE_BAND_LOOKAHEAD")
    os.environ["BSE_BAND_LOOKAHEAD"] = "0"
    try:
        ev0, z0 = ctx2.eigh_dense(a, prec, 8)
    finally:
        if old is None:
            del os.environ["BSE_BAND_LOOKAHEAD"]
        else:
            os.environ["BSE_BAND_LOOKAHEAD"] = old
    ev1, z1 = ctx2.eigh_dense(a, prec, 8)
    assert np.array_equal(ev0, ev1)
    assert np.array_equal(z0, z1)


@pytest.mark.parametrize("n", [65, 333, 1000, 2049])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_two_stage_grouped_q1_matches_panel_order(ctx2, monkeypatch, n, prec):
    """Z = Q1 X four band panels per pass (BSE_BT16_GROUPED=1: FP64 matrix
    cores, the T of each group of four formed on the main stream) against the
    default one panel and 16 columns per pass (columns 16.. on the main
    stream beside the second stream's first 16): the same eigenvalues bit for bit
    (Q1 does not touch them) and the same vectors to FP64 rounding.  Band
    panels: 1 (n = 65), 10, 30 and 63 (n = 2049: the last group three panels,
    the last panel one reflector)."""
    a = _herm(n, 11)
    k = min(20, n)       # 20: the default form splits the columns 16 + 4 over two streams
    print(f"grouped Q1 vs panels n={n} {prec}", flush=True)
    monkeypatch.setenv("BSE_BT16_GROUPED", "1")
    ev_g, z_g = ctx2.eigh_dense(a, prec, k)
    monkeypatch.delenv("BSE_BT16_GROUPED")
    ev_p, z_p = ctx2.eigh_dense(a, prec, k)
    assert np.array_equal(ev_g, ev_p)
    assert np.abs(z_g - z_p).max() < 1e-12
    r = a @ z_g - z_g * ev_g[None, :k]
    assert np.abs(r).max() < (1e-10 if prec == "fp64" else 1e-4) * np.abs(ev_g).max()
