"""Real-space Wannier route on the CPU: xsf files, grid layout, the oracle's
pair integrals and the tables built from them.

Pins of the oracle (oracle/wannier_oracle.py):
* the vectorised integrals equal a loop-literal restatement of the
  reference's loops (bse.version13.cpp:1513-1532, 1549-1568) on small grids;
* for normalised Gaussians they reach the closed-form integrals
  (A = exp(-|a-b|^2/4s^2) exp(iK.m - s^2 K^2/4), D = ... m, S = ...);
* the Bloch sum M_ij(k; G) = sum_R e^{-ik.R} A_ij(R, G) equals a brute-force
  integral of conj(psi_a,k) e^{iG.r} psi_b,k over a periodic supercell,
  psi_k(r) = sum_R e^{ik.R} sum_i c_i w_i(r - R) (the convention of
  H(k) = sum_R H(R) e^{ik.R}); the wrong sign of the phase fails it;
* D is the K-derivative of A at K = 0 (A(R, K) = S(R) + i K.D(R) + O(K^2)),
  so the dipole of the position table equals the finite-difference slope of
  the pair density built from M.
"""
import numpy as np
import pytest

import bse_amd
from bse_amd import wannier as W
import bse_oracle as O
import wannier_oracle as WO


def _random_grid(rng, nw=2, nspin=2, n_sc=(3, 2, 2), p=(2, 3, 2)):
    axes = np.array([[2.0, 0.3, 0.0], [0.1, 1.7, 0.2], [0.0, -0.2, 2.4]]).T
    g = W.WannierGrid(nw=nw, nspin=nspin, n_sc=n_sc, p=p, origin=np.array([-0.7, 0.4, -1.1]), axes=axes,
                      values=rng.normal(size=(nspin * nw, int(np.prod(n_sc)), int(np.prod(p)))))
    return g


def _honeycomb_gaussians(p=(12, 12, 24), n_sc=(5, 5, 1), sigma=0.45, nspin=1, c=6.0):
    s = bse_amd.honeycomb_plain(5)
    m = s.model
    centres = np.asarray(m.tau, dtype=float)
    if nspin == 2:
        centres = centres[::2]
    return s, W.gaussian_grid(np.asarray(m.lattice.bravais), centres, n_sc=n_sc, p=p, sigma=sigma,
                              nspin=nspin, c_extent=c)


def test_xsf_round_trip_and_header(tmp_path):
    rng = np.random.default_rng(1)
    g = _random_grid(rng)
    W.write_wannier_grid(g, "wan", directory=tmp_path)
    names = sorted(x.name for x in tmp_path.iterdir())
    assert names == ["wandown_00001.xsf", "wandown_00002.xsf", "wanup_00001.xsf", "wanup_00002.xsf"]
    # the reference's reader skips 22 lines and then reads the grid sizes
    lines = (tmp_path / "wanup_00001.xsf").read_text().split("\n")
    assert [int(x) for x in lines[W.XSF_HEADER_LINES].split()] == [6, 6, 4]
    h = W.read_wannier_grid("wan", 2, g.n_sc, spinful=True, directory=tmp_path, normalize=False)
    assert h.p == g.p and h.n_sc == g.n_sc
    assert np.allclose(h.origin, g.origin) and np.allclose(h.axes, g.axes)
    assert np.allclose(h.values, g.values, rtol=1e-9, atol=1e-12)


def test_xsf_names_single_spin():
    assert W.xsf_name("mos2", 0) == "mos2_00001.xsf"
    assert W.xsf_name("mos2", 11, spin=1, spinful=True) == "mos2down_00012.xsf"


def test_layout_maps_grid_points_to_positions(tmp_path):
    """A linear function of position written on the global grid (x fastest)
    reads back as the same function of WannierGrid.positions()."""
    rng = np.random.default_rng(2)
    g = _random_grid(rng, nw=1, nspin=1, n_sc=(2, 3, 2), p=(3, 2, 4))
    span = g.axes * np.asarray(g.n_sc)[None, :]
    n = [g.n_sc[a] * g.p[a] for a in range(3)]
    idx = np.stack(np.meshgrid(*[np.arange(n[a]) / n[a] for a in range(3)], indexing="ij"), -1)
    pos = g.origin + idx @ span.T                                        # [x, y, z, 3]
    f = lambda r: 1.0 + r[..., 0] - 2.0 * r[..., 1] + 0.5 * r[..., 2]
    W.write_xsf(tmp_path / "lin_00001.xsf", f(pos), g.origin, span)
    h = W.read_wannier_grid("lin", 1, g.n_sc, directory=tmp_path, normalize=False)
    assert np.allclose(h.values[0], f(h.positions()), atol=1e-8)


def test_normalise_as_the_reference():
    rng = np.random.default_rng(3)
    g = _random_grid(rng)
    W.normalise(g)
    assert np.allclose(g.weight * np.einsum("fcp,fcp->f", g.values, g.values), 1.0)
    assert np.isclose(g.weight, abs(np.linalg.det(g.axes)) / g.ppc)


def test_oracle_vectorised_equals_loops():
    rng = np.random.default_rng(4)
    g = _random_grid(rng)
    rc = np.array([[0, 0, 0], [1, 0, 0], [-1, 1, 0], [2, -1, 1], [0, 0, -1], [3, 0, 0]])
    kv = rng.normal(size=(3, 3))
    args = (g.values, g.nw, g.nspin, g.n_sc, g.p, g.origin, g.axes, rc, kv)
    A, D, S = WO.integrals(*args)
    A2, D2, S2 = WO.integrals_loops(*args)
    assert np.allclose(A, A2, rtol=1e-12, atol=1e-12)
    assert np.allclose(D, D2, rtol=1e-12, atol=1e-12)
    assert np.allclose(S, S2, rtol=1e-12, atol=1e-12)
    assert np.all(A[:, :, :, 5] == 0) and np.all(S[:, :, :, 5] == 0)     # translation beyond the supercell


def test_oracle_gaussian_closed_form():
    _, g = _honeycomb_gaussians()
    rc = W.translations((1, 1, 0))
    Rc = rc @ g.axes.T
    kv = np.array([[0.0, 0.0, 0.0], [0.7, -0.3, 0.0], [1.5, 0.9, 0.4]])
    A, D, S = WO.integrals(g.values, g.nw, 1, g.n_sc, g.p, g.origin, g.axes, rc, kv)
    tau = np.array([[0.0, 0.0, 0.0], [1.25, 0.72168784, 0.0]])
    for i in range(2):
        for j in range(2):
            for r in range(len(rc)):
                a, d, s = WO.gaussian_pair(tau[i] + Rc[r], tau[j], 0.45, kv)
                assert np.allclose(A[0, i, j, r], a, atol=2e-9)
                assert np.allclose(D[0, i, j, r], d, atol=2e-9)
                assert abs(S[0, i, j, r] - s) < 2e-9


def test_point_orbital_table_equals_point_route():
    """D_ij(R) = delta_ij delta_R0 tau_i reproduces the point-orbital dipole."""
    s = bse_amd.honeycomb_magnetic(3)
    p = O.problem_from_system(s)
    tau = np.asarray(s.model.tau, dtype=float)
    rc = W.translations((1, 1, 0))
    Rc = rc @ np.asarray(s.model.lattice.bravais).T
    D = np.zeros((len(rc), p.norb, p.norb, 3))
    D[len(rc) // 2] = np.einsum("ij,ia->ija", np.eye(p.norb), tau)
    assert np.all(rc[len(rc) // 2] == 0)
    pt = O.problem_from_system(s.with_params(dipole_r=Rc, dipole_d=D))
    bands = O.band_window(p)
    assert np.allclose(O.dipoles(pt, bands), O.dipoles(p, bands), rtol=1e-13, atol=1e-14)


def _bloch_brute(g, kc, a, b, G):
    """f sum_r conj(psi_a) e^{iG.r} psi_b / Ncell on the periodic supercell,
    psi_c(r) = sum_R e^{ik.R} sum_i c_i w_i(r - R), translated functions
    wrapped periodically (k on the supercell's grid, G a reciprocal vector)."""
    n0, n1, n2 = g.n_sc
    x = g.positions()                                                # (ncell, ppc, 3)
    w = g.values.reshape(g.nw, n0, n1, n2, g.ppc)
    psi_a = np.zeros((n0, n1, n2, g.ppc), dtype=complex)
    psi_b = np.zeros_like(psi_a)
    for r0 in range(n0):
        for r1 in range(n1):
            R = g.axes @ np.array([r0, r1, 0.0])
            ph = np.exp(1j * kc @ R)
            sh = np.roll(np.roll(w, r0, axis=1), r1, axis=2)               # w_i(r - R), periodic
            psi_a += ph * np.einsum("i,icdep->cdep", a, sh)
            psi_b += ph * np.einsum("i,icdep->cdep", b, sh)
    e = np.exp(1j * x @ G).reshape(n0, n1, n2, g.ppc)
    return g.weight * np.sum(psi_a.conj() * e * psi_b) / (n0 * n1 * n2)


def test_m_table_bloch_sum_equals_brute_force():
    s, g = _honeycomb_gaussians(p=(10, 10, 20))
    rng = np.random.default_rng(5)
    rc = W.translations((2, 2, 0))
    bv = np.asarray(s.model.lattice.bravais)
    b = 2 * np.pi * np.linalg.inv(bv).T                                # reciprocal vectors as columns
    Gs = np.array([np.zeros(3), b[:, 0], b[:, 0] - b[:, 1]])
    A, _, _ = WO.integrals(g.values, g.nw, 1, g.n_sc, g.p, g.origin, g.axes, rc, Gs)
    for m1, m2 in [(1, 0), (1, 2), (3, 1)]:
        kc = b @ np.array([m1 / 5, m2 / 5, 0.0])
        M = W.m_table(A[0], rc, g.axes, kc[None, :])[0]                    # (nw, nw, nG)
        Mwrong = W.m_table(A[0], rc, g.axes, -kc[None, :])[0]
        ca = rng.normal(size=2) + 1j * rng.normal(size=2)
        cb = rng.normal(size=2) + 1j * rng.normal(size=2)
        for q in range(len(Gs)):
            ref = _bloch_brute(g, kc, ca, cb, Gs[q])
            got = ca.conj() @ M[:, :, q] @ cb
            assert abs(got - ref) < 1e-9 * max(1.0, abs(ref)), (m1, m2, q, got, ref)
            assert abs(ca.conj() @ Mwrong[:, :, q] @ cb - ref) > 1e-3


def test_position_table_dipole_is_slope_of_pair_density():
    """rho = i sum conj(c) Dq c equals d/de of conj(c) M(k; e qhat) c at e = 0."""
    s, g = _honeycomb_gaussians(p=(10, 10, 20))
    rc = W.translations((2, 2, 0))
    eps = 1e-4
    qh = np.array([1.0, 0.0, 0.0])
    A, D, _ = WO.integrals(g.values, g.nw, 1, g.n_sc, g.p, g.origin, g.axes, rc, np.stack([eps * qh, -eps * qh]))
    # D is the slope of A
    assert np.allclose((A[..., 0] - A[..., 1]) / (2j * eps), D @ qh, atol=1e-8)
    Rc = rc @ g.axes.T
    st = s.with_params(dipole_r=Rc, dipole_d=np.ascontiguousarray(np.transpose(D[0], (2, 0, 1, 3))))
    p = O.problem_from_system(st)
    eps_c, C, eps_v, V = bands = O.band_window(p)
    rho = O.dipoles(p, bands).reshape(p.nc, p.nv, p.nk)
    kc = p.kpoints().T
    M = W.m_table(A[0], rc, g.axes, kc)                                   # (nk, nw, nw, 2)
    rp = np.einsum("kci,kij,kvj->cvk", C.conj(), M[..., 0], V)
    rm = np.einsum("kci,kij,kvj->cvk", C.conj(), M[..., 1], V)
    assert np.allclose(rho, (rp - rm) / (2 * eps), atol=1e-7 * np.abs(rho).max())
    # and the table route differs from the point route for these overlapping orbitals
    p0 = O.problem_from_system(s)
    assert np.abs(np.abs(rho.ravel()) - np.abs(O.dipoles(p0, bands))).max() > 1e-3


def test_translations_box():
    r = W.translations((1, 2, 0))
    assert r.shape == (15, 3) and np.all(r[7] == 0)
