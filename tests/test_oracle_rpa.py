"""Pins of the oracle's static RPA screening (Dielectric_Function::pull_values,
bse.version13.cpp:2214-2303; W build 2575-2584).

* The point-orbital factorisation chi0 = phases^H Pi phases and the Dyson
  solve equal a loop-literal restatement of the reference's sums over
  transitions and G (bse_bruteforce.rpa_w_bruteforce) on several momenta.
* W_GG' is Hermitian; eps = 1 - V chi0 and eps^-1 = 1 + V chi are inverses.
* With chi0 scaled to 0 the table is the bare-potential table (W = v).
* First order in chi0: dW/dscale at 0 equals V chi0 V.
* The static chi0 is negative semidefinite (every transition energy > 0), so
  RPA screening lowers the head: W_00 < v_0 at Q != 0.
* The BSE with RPA screening goes through assemble / solve, and RPA binding
  energies lie between the unscreened and IPA limits.
"""
import numpy as np
import pytest

import bse_amd
import bse_oracle as O
import bse_bruteforce as BF

SYS = [lambda: bse_amd.honeycomb_plain(3).with_params(screening="rpa"),
       lambda: bse_amd.honeycomb_magnetic(2).with_params(screening="rpa"),
       lambda: bse_amd.random_multiorbital(2).with_params(screening="rpa", rpa_nv=2, rpa_nc=3)]


@pytest.mark.parametrize("make", SYS)
@pytest.mark.parametrize("d", [(0, 0, 0), (1, 0, 0), (-1, 1, 0)])
def test_rpa_factorised_equals_bruteforce(make, d):
    p = O.problem_from_system(make())
    W, _ = O.rpa_w_matrix(p, d)
    Wb = BF.rpa_w_bruteforce(p, d)
    assert np.abs(W - Wb).max() < 1e-11 * np.abs(W).max()
    assert np.abs(W - W.conj().T).max() < 1e-12 * np.abs(W).max()


def test_rpa_dielectric_inverse_and_sign():
    p = O.problem_from_system(bse_amd.honeycomb_plain(4).with_params(screening="rpa"))
    d = (1, 0, 0)
    W, qg = O.rpa_w_matrix(p, d)
    qn = np.linalg.norm(qg, axis=0)
    vg = O.v_model(p, qn)
    ph = np.exp(1j * (p.tau @ qg))
    chi0 = ph.conj().T @ O.rpa_pi(p, d) @ ph
    eps = np.eye(vg.size) - vg[:, None] * chi0
    epsinv = W / vg[None, :]
    assert np.abs(eps @ epsinv - np.eye(vg.size)).max() < 1e-12
    assert np.linalg.eigvalsh(0.5 * (chi0 + chi0.conj().T)).max() < 1e-12 * np.abs(chi0).max()
    g0 = qg.shape[1] // 2
    assert W[g0, g0].real < vg[g0]


@pytest.mark.parametrize("make", SYS[1:])
def test_rpa_reciprocity_without_time_reversal(make):
    """W_ij(-Q) = W_ji(Q) (chi0 reciprocal), which keeps the BSE matrix
    Hermitian, also for the magnetic and the complex random models; with
    time reversal the two kinds of transitions contribute equally."""
    p = O.problem_from_system(make())
    wt = O.rpa_w_table(p)
    w1, w2 = 2 * p.n1 - 2, 2 * p.n2 - 2
    flip = wt[::-1, ::-1, ::-1]
    assert np.abs(flip - np.swapaxes(wt, -1, -2)).max() < 1e-11 * np.abs(wt).max()
    H = O.assemble(p)
    assert np.abs(H - H.conj().T).max() < 1e-12 * np.abs(H).max()


def test_rpa_time_reversal_two_kinds_equal():
    p = O.problem_from_system(bse_amd.honeycomb_plain(4).with_params(screening="rpa"))
    from dataclasses import replace
    d = (1, 2, 0)
    full = O.rpa_pi(p, d)
    # with time reversal the first kind alone, doubled, is the reference's form
    kc = p.kpoints()
    e, v = O.tb_states(p, kc)
    ki, kj, kl = O._kidx(p)
    kq = (((ki - d[0]) % p.n1) * p.n2 + ((kj - d[1]) % p.n2)) * p.n3 + ((kl - d[2]) % p.n3)
    vidx, cidx = O.rpa_window(p)
    P1 = np.einsum("kic,kiv->kcvi", v[:, :, cidx].conj(), v[kq][:, :, vidx])
    f1 = -2.0 / ((e[:, cidx] + p.scissor)[:, :, None] - e[kq][:, vidx][:, None, :])
    ref = (p.f_s / (p.cell * p.nk)) * np.einsum("kcv,kcvi,kcvj->ij", f1, P1.conj(), P1)
    assert np.abs(full - ref).max() < 1e-12 * np.abs(ref).max()


def test_rpa_zero_screening_is_bare_table():
    p = O.problem_from_system(bse_amd.honeycomb_magnetic(3).with_params(screening="rpa"))
    t0 = O.rpa_w_table(p, scale=0.0)
    pb = O.problem_from_system(bse_amd.honeycomb_magnetic(3).with_params(r0=0.0, eps_env=1.0))
    tb = O.w_table(pb)
    assert np.abs(t0 - tb).max() < 1e-12 * np.abs(tb).max()


def test_rpa_first_order():
    p = O.problem_from_system(bse_amd.honeycomb_plain(3).with_params(screening="rpa"))
    d = (1, 1, 0)
    h = 1e-4
    Wp, qg = O.rpa_w_matrix(p, d, scale=h)
    Wm, _ = O.rpa_w_matrix(p, d, scale=-h)
    dW = (Wp - Wm) / (2 * h)
    vg = O.v_model(p, np.linalg.norm(qg, axis=0))
    ph = np.exp(1j * (p.tau @ qg))
    chi0 = ph.conj().T @ O.rpa_pi(p, d) @ ph
    ref = vg[:, None] * chi0 * vg[None, :]
    assert np.abs(dW - ref).max() < 1e-6 * np.abs(ref).max()


def test_rpa_bse_binding_between_limits():
    base = bse_amd.honeycomb_plain(6)
    e_rpa = O.solve(O.problem_from_system(base.with_params(screening="rpa")), omegas=None)["E"][0]
    e_bare = O.solve(O.problem_from_system(base.with_params(r0=0.0, eps_env=1.0)), omegas=None)["E"][0]
    e_ipa = np.sort(O.transition_energies(O.problem_from_system(base)))[0]
    assert e_bare < e_rpa < e_ipa
