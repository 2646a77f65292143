"""Pins of the oracle's full-BSE (beyond Tamm-Dancoff) restatement.

The coupling block, the general eigensolve with the augmentation (overlap)
form of the spectrum, and the structure-preserving Cholesky solve are checked
against each other and against what the mathematics fixes:

* the factorised coupling block equals a loop-literal brute force of the
  reference's add_coupling_part structure (2929-3040), and is complex
  symmetric;
* the full matrix's eigenvalues are real and come in +-E pairs, the
  eigenvectors are pseudo-orthonormal (X^H s3 X = 1), and with B = 0 the
  positive branch is the TDA spectrum;
* with no interaction (direct and exchange scales 0) the energies are the
  transition energies and the spectrum is the closed-form independent-
  particle one with its anti-resonant term;
* the augmentation form over 2N eigenpairs, the sign-weighted positive-branch
  sum and dense resolvent solves give the same spectrum;
* the structure-preserving solve (Cholesky, Hermitian W = L^H s3 L) equals the
  general eigensolve and, for real A and B, the product form (A-B)(A+B);
* f-sum rule: 2 sum_{E>0} E |f|^2 = rho~^H H s3 rho~.
"""
from dataclasses import replace

import numpy as np
import pytest

import bse_amd
import bse_oracle as O
import bse_bruteforce as BF

SYSTEMS = [lambda: bse_amd.honeycomb_plain(3), lambda: bse_amd.honeycomb_magnetic(2),
           lambda: bse_amd.random_multiorbital(2)]


@pytest.mark.parametrize("make", SYSTEMS)
def test_coupling_block_equals_loop_bruteforce(make):
    p = O.problem_from_system(make())
    B = O.assemble_coupling(p)
    Bb = BF.assemble_coupling_bruteforce(p)
    assert np.abs(B - Bb).max() < 1e-13 * np.abs(B).max()
    assert np.abs(B - B.T).max() < 1e-13 * np.abs(B).max()


@pytest.mark.parametrize("make", [lambda: bse_amd.honeycomb_plain(4), lambda: bse_amd.cubic_3d(3)])
def test_full_spectrum_paired_real_pseudo_orthonormal(make):
    p = O.problem_from_system(make())
    sol = O.solve_full(p)
    N = p.n_bse
    E2 = sol["E_all"]
    assert sol["max_imag"] < 1e-9 * np.abs(E2).max()
    assert np.abs(E2[:N][::-1] + E2[N:]).max() < 1e-10 * np.abs(E2).max()
    s3 = np.concatenate([np.ones(N), -np.ones(N)])
    X = sol["X"]
    G = X.conj().T @ (s3[:, None] * X)
    # pseudo-orthonormal between distinct eigenvalues (a general eigensolver
    # returns an arbitrary basis inside a degenerate cluster)
    E = sol["E"]
    distinct = np.abs(E[:, None] - E[None, :]) > 1e-6 * np.abs(E).max()
    assert np.abs(np.diag(G) - 1).max() < 1e-10
    assert np.abs(G[distinct]).max() < 1e-8
    assert np.all(sol["E"] > 0)


def test_zero_coupling_gives_tda():
    p = O.problem_from_system(bse_amd.honeycomb_magnetic(3))
    A = O.assemble(p)
    N = p.n_bse
    H = np.block([[A, np.zeros_like(A)], [np.zeros_like(A), -A.conj()]])
    E2, _, _ = O.eig_full(H)
    assert np.abs(E2[N:] - np.linalg.eigvalsh(A)).max() < 1e-12 * np.abs(E2).max()


def test_no_interaction_limit_closed_form():
    s = bse_amd.random_multiorbital(2).with_params(exchange_scale=0.0, direct_scale=0.0)
    p = O.problem_from_system(s)
    sol = O.solve_full(p)
    de = np.sort(O.transition_energies(p))
    assert np.abs(sol["E"] - de).max() < 1e-12 * de.max()
    rho = O.dipoles(p)
    dE = O.transition_energies(p)
    pref = O.C_EV_A / (p.volume * p.nk)
    w = p.omegas[:, None]
    ref = 1.0 - pref * ((np.abs(rho) ** 2)[None, :] * (1.0 / (w - dE[None, :] + 1j * p.eta)
                                                       - 1.0 / (w + dE[None, :] + 1j * p.eta))).sum(axis=1)
    assert np.abs(sol["eps"] - ref).max() < 1e-10 * np.abs(ref - 1).max()


@pytest.mark.parametrize("make", [lambda: bse_amd.honeycomb_plain(4), lambda: bse_amd.honeycomb_magnetic(2)])
def test_spectrum_forms_agree(make):
    p = O.problem_from_system(make())
    sol = O.solve_full(p)
    om = p.omegas[::97]
    pref = O.C_EV_A / (p.volume * p.nk)
    e_res = O.spectrum_resolvent(sol["H"], sol["rho_t"], om, p.eta, pref)
    e_aug = O.solve_full(p, omegas=om)["eps"]
    e_sgn = O.dielectric_full(sol["E"], sol["f"], om, p.eta, pref)
    scale = np.abs(e_res - 1).max()
    assert np.abs(e_aug - e_res).max() < 1e-10 * scale
    assert np.abs(e_sgn - e_res).max() < 1e-10 * scale


@pytest.mark.parametrize("make", [lambda: bse_amd.honeycomb_plain(4), lambda: bse_amd.honeycomb_magnetic(3)])
def test_structured_solve_equals_general_eig(make):
    p = O.problem_from_system(make())
    sol = O.solve_full(p)
    E, X, f = O.solve_full_structured(sol["A"], sol["B"], sol["rho_t"])
    scale = np.abs(sol["E"]).max()
    assert np.abs(E - sol["E"]).max() < 1e-11 * scale
    f2 = np.abs(f) ** 2
    for i, j in _clusters(E, 1e-8 * scale):
        assert abs(f2[i:j].sum() - (np.abs(sol["f"][i:j]) ** 2).sum()) < 1e-9 * np.sum(f2)
    N = p.n_bse
    H = sol["H"]
    assert np.abs(H @ X - X * E[None, :]).max() < 1e-10 * scale


def test_structured_solve_real_product_form():
    rng = np.random.default_rng(11)
    n = 40
    a = rng.normal(size=(n, n))
    A = a @ a.T / n + 3.0 * np.eye(n)
    b = rng.normal(size=(n, n))
    B = 0.3 * (b + b.T) / 2
    E, X, f = O.solve_full_structured(A.astype(complex), B.astype(complex), np.ones(2 * n, complex))
    ref = np.sqrt(np.sort(np.linalg.eigvals((A - B) @ (A + B)).real))
    assert np.abs(E - ref).max() < 1e-12 * ref.max()


def test_unstable_system_is_refused():
    """Strong binding on a tiny grid drives eigenvalues of A below zero; the
    full matrix then has complex eigenvalues and the oracle refuses it."""
    p = O.problem_from_system(bse_amd.random_multiorbital(2))
    with pytest.raises(ValueError):
        O.solve_full(p)


def test_full_f_sum_rule():
    p = O.problem_from_system(bse_amd.honeycomb_magnetic(3))
    sol = O.solve_full(p)
    N = p.n_bse
    s3 = np.concatenate([np.ones(N), -np.ones(N)])
    lhs = 2.0 * np.sum(sol["E"] * np.abs(sol["f"]) ** 2)
    rhs = np.real(sol["rho_t"].conj() @ (sol["H"] @ (s3 * sol["rho_t"])))
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)


def _clusters(E, tol):
    out, i = [], 0
    while i < len(E):
        j = i + 1
        while j < len(E) and E[j] - E[j - 1] < tol:
            j += 1
        out.append((i, j))
        i = j
    return out
