"""Pins of the FP64 oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, and cannot be built
here (Armadillo is absent), so the oracle is pinned by what the mathematics
fixes: closed forms, invariants, special cases that reduce to textbook results,
a loop-literal brute force that follows the reference's own loop structure,
and finite differences for the Hellmann-Feynman gradients.  Each test is
chosen so that a plausible mistake (dropped term, wrong sign or index,
transposed operand, missing conjugate) fails.
"""
from dataclasses import replace

import numpy as np
import pytest

import bse_amd
import bse_oracle as O
import bse_bruteforce as B


def prob(s):
    return O.problem_from_system(s)


# ---------------------------------------------------------------- geometry
def test_reciprocal_identity():
    for lat in (bse_amd.tight_binding.honeycomb_lattice(), bse_amd.lattice.CrystalLattice(np.diag([3.0, 3.0, 20.0]))):
        b = O.reciprocal(lat.bravais)
        assert np.allclose(lat.bravais.T @ b, 2 * np.pi * np.eye(3), atol=1e-12)
        assert np.allclose(b, lat.reciprocal)


def test_glist_symmetric_and_centred():
    lat = bse_amd.tight_binding.honeycomb_lattice()
    g = O.glist2d(lat.bravais, 6.0)
    assert g.shape[1] == 25
    assert np.allclose(g[:, 12], 0.0)
    # G -> -G maps the list onto itself (needed for W_ij(-Q) = W_ij(Q)^*)
    assert np.allclose(g[:, ::-1], -g)
    assert np.allclose(g, bse_amd.lattice.g_list(lat, 6.0))


def test_kgrid_matches_package():
    s = bse_amd.honeycomb_plain(5)
    assert np.allclose(O.kgrid(s.model.lattice.bravais, 5, 5), s.kpoints())


# ---------------------------------------------------------------- TB states
def test_honeycomb_closed_form_bands():
    """eps_pm(k) = +-sqrt((D/2)^2 + t^2 |1 + e^{-ik.a1} + e^{-ik.a2}|^2)."""
    s = bse_amd.honeycomb_plain(9)
    p = prob(s)
    k = O.kgrid(p.bravais, 9, 9)
    e, v = O.tb_states(p, k)
    a1, a2 = p.bravais[:, 0], p.bravais[:, 1]
    f = 1 + np.exp(-1j * (a1 @ k)) + np.exp(-1j * (a2 @ k))
    ep = np.sqrt(1.5 ** 2 + 2.7 ** 2 * np.abs(f) ** 2)
    assert np.allclose(e[:, 1], ep, atol=1e-12)
    assert np.allclose(e[:, 0], -ep, atol=1e-12)
    # gap at K = (1/3, 2/3) is Delta
    kK = p.bravais @ np.zeros(3)
    b = O.reciprocal(p.bravais)
    kK = (b[:, 0] / 3 + 2 * b[:, 1] / 3)[:, None]
    eK, _ = O.tb_states(p, kK)
    assert abs(eK[0, 1] - eK[0, 0] - 3.0) < 1e-12


def test_band_window_order_and_scissor():
    s = bse_amd.random_multiorbital(3).with_params(scissor=0.7)
    p = prob(s)
    eps_c, C, eps_v, V = O.band_window(p)
    p0 = replace(p, scissor=0.0)
    eps_c0, _, eps_v0, _ = O.band_window(p0)
    assert np.allclose(eps_c - eps_c0, 0.7)
    assert np.allclose(eps_v, eps_v0)
    # v = 0 is the highest valence band, c = 0 the lowest conduction band
    assert np.all(np.diff(eps_v, axis=1) <= 1e-12)
    assert np.all(np.diff(eps_c, axis=1) >= -1e-12)
    # eigen-residuals of the selected states
    k = O.kgrid(p.bravais, 3, 3)
    hk = O.hk(p, k)
    r = np.einsum("kij,kcj->kci", hk, C) - eps_c0[:, :, None] * C
    assert np.abs(r).max() < 1e-12


# ---------------------------------------------------------------- assembly
@pytest.mark.parametrize("make", [lambda: bse_amd.honeycomb_plain(3),
                                  lambda: bse_amd.honeycomb_magnetic(3),
                                  lambda: bse_amd.random_multiorbital(2)])
def test_factorised_assembly_equals_loop_bruteforce(make):
    p = prob(make())
    assert np.abs(O.assemble(p) - B.assemble_bruteforce(p)).max() < 1e-12


def test_hermitian():
    H = O.assemble(prob(bse_amd.random_multiorbital(4)))
    assert np.abs(H - H.conj().T).max() < 1e-13


def test_ipa_limit():
    """With W = V = 0 the eigenvalues are the transition energies exactly."""
    s = bse_amd.random_multiorbital(4).with_params(direct_scale=0.0, exchange_scale=0.0)
    p = prob(s)
    E, _ = O.eigh(O.assemble(p))
    assert np.allclose(E, np.sort(O.transition_energies(p)), atol=1e-13)


def test_exchange_invariant_to_constant_shift():
    """sum_i conj(c_ck,i) c_vk,i = 0 up to O(q): adding const to V_ij changes nothing."""
    p = prob(bse_amd.random_multiorbital(3))
    bands = O.band_window(p)
    vx = O.v_exchange(p)
    H1 = O.assemble(p, bands, vx=vx)
    H2 = O.assemble(p, bands, vx=vx + 5.0)
    # P^H 1 = O(q) (valence at k - q), so the shift enters at the 1e-5 level only
    assert np.abs(H1 - H2).max() < 1e-6 * np.abs(H1).max()
    P = O.exchange_factors(p, bands[1], bands[3])
    assert np.linalg.matrix_rank(P @ vx @ P.conj().T, tol=1e-10) <= p.norb


def test_gauge_invariance_of_oscillators():
    """Random per-k phases on the TB states: H -> U^H H U, |f_m|^2 unchanged."""
    p = prob(bse_amd.random_multiorbital(3))
    eps_c, C, eps_v, V = O.band_window(p)
    rng = np.random.default_rng(1)
    ph_c = np.exp(1j * rng.uniform(0, 2 * np.pi, size=C.shape[:2]))
    ph_v = np.exp(1j * rng.uniform(0, 2 * np.pi, size=V.shape[:2]))
    b2 = (eps_c, C * ph_c[:, :, None], eps_v, V * ph_v[:, :, None])
    H1 = O.assemble(p, (eps_c, C, eps_v, V))
    H2 = O.assemble(p, b2)
    E1, Z1 = O.eigh(H1)
    E2, Z2 = O.eigh(H2)
    assert np.allclose(E1, E2, atol=1e-12)
    f1 = np.abs(O.oscillator(Z1, O.dipoles(p, (eps_c, C, eps_v, V)))) ** 2
    f2 = np.abs(O.oscillator(Z2, O.dipoles(p, b2))) ** 2
    assert np.allclose(f1, f2, atol=1e-12 * f1.max())


def test_spinor_without_rashba_decouples_spin_flip_exchange():
    """lambda_R = 0: spin-flip transitions carry no exchange (P = 0) and the
    spectrum is the union of the spin-conserving and spin-flip blocks."""
    s = bse_amd.honeycomb_magnetic(3)
    s = replace(s, model=bse_amd.tight_binding.honeycomb_spinful(rashba=0.0))
    p = prob(s)
    eps_c, C, eps_v, V = O.band_window(p)
    P = O.exchange_factors(p, C, V).reshape(p.nc, p.nv, p.nk, p.norb)
    H = O.assemble(p)
    flip = []
    for c in range(p.nc):
        for v in range(p.nv):
            spin_c = np.abs(C[:, c, 0::2]).sum(axis=1) > 0.5
            spin_v = np.abs(V[:, v, 0::2]).sum(axis=1) > 0.5
            fl = spin_c != spin_v
            assert np.abs(P[c, v][fl]).max(initial=0) < 1e-12
            flip.append(fl)
    flip = np.concatenate(flip)
    # no coupling between spin-conserving and spin-flip transitions
    assert np.abs(H[np.ix_(flip, ~flip)]).max() < 1e-12
    E = np.linalg.eigvalsh(H)
    Eu = np.sort(np.concatenate([np.linalg.eigvalsh(H[np.ix_(flip, flip)]),
                                 np.linalg.eigvalsh(H[np.ix_(~flip, ~flip)])]))
    assert np.allclose(E, Eu, atol=1e-12)


def test_mini_bz_average_by_quadrature():
    area, nk, r0, eps = 5.41266, 144, 10.0, 1.0
    da = (2 * np.pi) ** 2 / (area * nk)
    rad = np.sqrt(da / np.pi)
    q = np.linspace(0, rad, 200001)[1:]
    integ = np.trapezoid(2 * np.pi * q * O.keldysh(q, r0, eps), q) / da
    assert abs(O.keldysh_q0_average(area, nk, r0, eps) / integ - 1) < 1e-4


# ---------------------------------------------------------------- eig / spectrum
def test_sum_rule():
    p = prob(bse_amd.random_multiorbital(4))
    sol = O.solve(p)
    assert abs((np.abs(sol["f"]) ** 2).sum() / (np.abs(sol["rho0"]) ** 2).sum() - 1) < 1e-12


def test_ipa_spectrum_closed_form():
    s = bse_amd.honeycomb_plain(6).with_params(direct_scale=0.0, exchange_scale=0.0)
    p = prob(s)
    sol = O.solve(p)
    de = O.transition_energies(p)
    rho = O.dipoles(p)
    pref = O.C_EV_A / (p.volume * p.nk)
    w = p.omegas[:, None]
    im = pref * (np.abs(rho) ** 2 * p.eta / ((w - de) ** 2 + p.eta ** 2)).sum(axis=1)
    assert np.allclose(sol["eps"].imag, im, rtol=1e-10, atol=1e-14)


def test_lorentzian_area():
    """Each pole contributes pref*|f|^2*pi to the integral of Im eps."""
    p = prob(bse_amd.honeycomb_plain(4))
    E = np.array([2.0])
    f = np.array([1.0 + 0j])
    w = np.linspace(-200, 200, 2000001)
    eps = O.dielectric(E, f, w, 0.05, 1.0)
    assert abs(np.trapezoid(eps.imag, w) / np.pi - 1) < 1e-3


# ---------------------------------------------------------------- gradients
PARAMS = ("hopping_scale", "scissor", "r0", "eps_env", "exchange_field")


def test_scissor_gradient_is_one():
    p = prob(bse_amd.random_multiorbital(3))
    g, _ = O.hf_energy_gradients(p, n_exc=6, params=("scissor",))
    assert np.allclose(g[:, 0], 1.0, atol=1e-12)


def test_hf_gradients_match_central_differences():
    p = prob(bse_amd.random_multiorbital(3))
    g, E0 = O.hf_energy_gradients(p, n_exc=6, params=PARAMS)
    base = dict(hopping_scale=1.0, scissor=p.scissor, r0=p.r0, eps_env=p.eps_env, exchange_field=0.0)
    h = 1e-4
    for t, name in enumerate(PARAMS):
        Ep = np.linalg.eigvalsh(O.assemble(O.with_param(p, name, base[name] + h)))[:6]
        Em = np.linalg.eigvalsh(O.assemble(O.with_param(p, name, base[name] - h)))[:6]
        fd = (Ep - Em) / (2 * h)
        assert np.allclose(g[:, t], fd, rtol=1e-4, atol=1e-7), (name, g[:, t], fd)


def test_spectrum_gradient_matches_central_differences():
    s = bse_amd.random_multiorbital(2)
    s = s.with_params(n_omega=256)
    p = prob(s)
    rng = np.random.default_rng(5)
    base_eps = O.solve(p)["eps"].imag
    target = base_eps * (1 + 0.1 * rng.normal(size=base_eps.size))
    g, L = O.spectrum_l2_gradient(p, target, params=PARAMS)
    base = dict(hopping_scale=1.0, scissor=p.scissor, r0=p.r0, eps_env=p.eps_env, exchange_field=0.0)
    # The misfit of eta = 50 meV Lorentzians has large third derivatives: the
    # O(h^2) truncation error of a central difference at h = 1e-4 is ~2e-4
    # relative for the hopping scale, so the check uses h = 1e-5.
    h = 1e-5
    for t, name in enumerate(PARAMS):
        Lp, _ = O.spectrum_l2(O.with_param(p, name, base[name] + h), target, p.omegas)
        Lm, _ = O.spectrum_l2(O.with_param(p, name, base[name] - h), target, p.omegas)
        fd = (Lp - Lm) / (2 * h)
        assert abs(g[t] - fd) <= 1e-4 * abs(fd) + 1e-12 * abs(L), (name, g[t], fd)


# ---------------------------------------------------------------- 3D systems
def test_3d_factorised_assembly_equals_loop_bruteforce():
    """The 3D k grid, G-list and C/(q^2 + radius) potentials through the
    factorised W_ij(Q) table agree with the loop-literal sums."""
    s = bse_amd.cubic_3d(2).with_params(radius=0.05)
    p = prob(s)
    assert np.abs(O.assemble(p) - B.assemble_bruteforce(p)).max() < 1e-12


def test_3d_glist_centred_and_symmetric():
    s = bse_amd.cubic_3d(2)
    g = O.glist(prob(s))
    assert g.shape[1] == 27 and np.allclose(g[:, 13], 0.0) and np.allclose(g[:, ::-1], -g)
    assert np.allclose(g, s.gvectors())


@pytest.mark.parametrize("radius", [0.0, 0.3])
def test_3d_mini_bz_average_by_quadrature(radius):
    """Mean of C/(eps (q^2 + r)) over the ball of volume (2 pi)^3/(Omega Nk),
    by radial quadrature."""
    vol, nk, eps = 27.0, 64, 11.6
    vs = (2 * np.pi) ** 3 / (vol * nk)
    R = (3 * vs / (4 * np.pi)) ** (1 / 3)
    q = np.linspace(0.0, R, 200001)
    # q^2/(q^2 + r) with its limit at q = 0 (1 for r = 0, else 0)
    frac = np.where(q > 0, q ** 2 / np.where(q > 0, q ** 2 + radius, 1.0), 1.0 if radius == 0 else 0.0)
    f = 4 * np.pi * O.C_EV_A / eps * frac
    num = np.sum(0.5 * (f[1:] + f[:-1]) * np.diff(q)) / vs
    assert abs(O.coulomb3d_q0_average(vol, nk, radius, eps) / num - 1) < 1e-8


def test_3d_ipa_limit_and_hermitian():
    s = bse_amd.cubic_3d(3).with_params(direct_scale=0.0, exchange_scale=0.0)
    p = prob(s)
    H = O.assemble(p)
    assert np.abs(H - H.conj().T).max() == 0.0
    assert np.allclose(np.sort(np.linalg.eigvalsh(H)), np.sort(O.transition_energies(p)), atol=1e-13)
    H1 = O.assemble(prob(bse_amd.cubic_3d(3)))
    assert np.abs(H1 - H1.conj().T).max() < 1e-13
