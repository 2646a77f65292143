"""The full BSE beyond Tamm-Dancoff on the MI355X against the oracle.

bse_build_full assembles the coupling block B next to the TDA block A;
bse_eigh_full solves [[A, B], [-conj B, -conj A]] by the structure-preserving
Cholesky reduction.  Checked against the oracle:
* B entry by entry (1e-11 relative; complex symmetric);
* the N positive energies (1e-10 relative), |f|^2 per degenerate cluster,
  the resonant + anti-resonant spectrum (relative L2 1e-9);
* eigenvectors through H x = E x with the oracle's H and x^H s3 x = 1;
* the c2 system (honeycomb_magnetic(30), N = 3600, order 7200) against the
  committed fixture;
* an unstable system (A not positive definite) is refused with
  BSE_ERR_NUMERIC instead of returning complex energies.
"""
import pathlib

import numpy as np
import pytest

import bse_amd
from bse_amd import _abi
import bse_oracle as O
from parity_util import cluster_f2_error

pytestmark = pytest.mark.gpu
GOLD = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def ctx():
    c = _abi.Context(0)
    yield c
    c.close()


SMALL = [lambda: bse_amd.honeycomb_plain(6), lambda: bse_amd.honeycomb_magnetic(4),
         lambda: bse_amd.honeycomb_magnetic(3).with_params(nv=1, nc=2), lambda: bse_amd.cubic_3d(3)]
IDS = ["plain6", "magnetic4", "magnetic3_1v2c", "cubic3d"]


@pytest.mark.parametrize("make", SMALL, ids=IDS)
def test_coupling_block_matches_oracle(ctx, make):
    s = make()
    print(f"coupling block {s.name} N={s.n_bse}", flush=True)
    p = O.problem_from_system(s)
    ref = O.assemble_coupling(p)
    ctx.set_system(s).build_full()
    B = ctx.coupling()
    scale = np.abs(ref).max()
    assert np.abs(B - ref).max() < 1e-11 * scale
    assert np.abs(B - B.T).max() < 1e-12 * scale


@pytest.mark.parametrize("make", SMALL, ids=IDS)
def test_full_bse_matches_oracle(ctx, make):
    s = make()
    print(f"full BSE {s.name} N={s.n_bse}", flush=True)
    p = O.problem_from_system(s)
    sol = O.solve_full(p)
    ctx.set_system(s).build_full().eigh_full(8)
    E = ctx.eigenvalues()
    scale = np.abs(sol["E"]).max()
    assert np.abs(E - sol["E"]).max() < 1e-10 * scale
    f2 = np.abs(ctx.oscillator()) ** 2
    assert cluster_f2_error(sol["E"], f2, np.abs(sol["f"]) ** 2, 1e-8 * scale) < 1e-9
    eps = ctx.spectrum(p.omegas, p.eta)
    assert np.linalg.norm(eps - sol["eps"]) < 1e-9 * np.linalg.norm(sol["eps"] - 1.0)
    X = ctx.eigenvectors(8)
    N = s.n_bse
    s3 = np.concatenate([np.ones(N), -np.ones(N)])
    assert np.abs(sol["H"] @ X - X * E[None, :8]).max() < 1e-10 * scale
    G = X.conj().T @ (s3[:, None] * X)
    assert np.abs(G - np.eye(8)).max() < 1e-9


def test_full_bse_all_vectors_and_tda_unchanged(ctx):
    s = bse_amd.honeycomb_plain(5)
    p = O.problem_from_system(s)
    sol = O.solve_full(p)
    ctx.set_system(s).build_full().eigh_full(-1)
    X = ctx.eigenvectors(s.n_bse)
    E = ctx.eigenvalues()
    assert np.abs(sol["H"] @ X - X * E[None, :]).max() < 1e-10 * np.abs(E).max()
    # the TDA path on the same context is unaffected
    ctx.build("fp64").eigh(-1)
    tda = O.solve(p)
    assert np.abs(ctx.eigenvalues() - tda["E"]).max() < 1e-10 * np.abs(tda["E"]).max()


def test_full_bse_c2_against_fixture(ctx):
    g = np.load(GOLD / "c2_full_honeycomb_magnetic_30.npz", allow_pickle=False)
    s = bse_amd.honeycomb_magnetic(30)
    print("full BSE c2 N=3600 (order 7200)", flush=True)
    ctx.set_system(s).build_full().eigh_full(16)
    E = ctx.eigenvalues()
    scale = np.abs(g["E"]).max()
    assert np.abs(E - g["E"]).max() < 1e-10 * scale
    f2 = np.abs(ctx.oscillator()) ** 2
    assert cluster_f2_error(g["E"], f2, g["f2"], 1e-8 * scale) < 1e-9
    eps = ctx.spectrum(g["omegas"], float(g["eta"]))
    assert np.linalg.norm(eps - g["eps"]) < 1e-9 * np.linalg.norm(g["eps"] - 1.0)


def test_unstable_system_refused(ctx):
    s = bse_amd.random_multiorbital(2)
    ctx.set_system(s).build_full()
    with pytest.raises(_abi.BSEError, match="positive definite"):
        ctx.eigh_full(4)
