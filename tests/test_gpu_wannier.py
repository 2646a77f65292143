"""Real-space Wannier route on the MI355X: bse_wannier_integrals against the
oracle (oracle/wannier_oracle.py) and the closed form of Gaussian pairs, and
the BSE with the Wannier position-table dipole against the oracle.

Tolerances: integrals 1e-12 x max|value| (FP64 sums in another order);
Gaussian closed form 2e-9 (quadrature error of the 0.2 A grid); BSE dipoles,
eigenvalues and spectrum as the FP64 parity tests (1e-12 / 1e-10).
"""
import numpy as np
import pytest

import bse_amd
from bse_amd import _abi
from bse_amd import wannier as W
import bse_oracle as O
import wannier_oracle as WO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _abi.Context(0)
    yield c
    c.close()


def _grid_random(seed=4, nw=3, nspin=2, n_sc=(3, 2, 2), p=(4, 3, 5)):
    rng = np.random.default_rng(seed)
    axes = np.array([[2.0, 0.3, 0.0], [0.1, 1.7, 0.2], [0.0, -0.2, 2.4]]).T
    return W.WannierGrid(nw=nw, nspin=nspin, n_sc=n_sc, p=p, origin=np.array([-0.7, 0.4, -1.1]), axes=axes,
                         values=rng.normal(size=(nspin * nw, int(np.prod(n_sc)), int(np.prod(p)))))


def _gaussians(s, nspin=1, p=(12, 12, 24), sigma=0.45):
    centres = np.asarray(s.model.tau, dtype=float)
    if nspin == 2:
        centres = centres[::2]
    return W.gaussian_grid(np.asarray(s.model.lattice.bravais), centres, n_sc=(5, 5, 1), p=p, sigma=sigma,
                           nspin=nspin, c_extent=6.0)


@pytest.mark.parametrize("nk", [0, 1, 8, 11])
def test_integrals_match_oracle(ctx, nk):
    print(f"wannier integrals random grid nk={nk}", flush=True)
    g = _grid_random()
    rc = np.array([[0, 0, 0], [1, 0, 0], [-1, 1, 0], [2, -1, 1], [0, 0, -1], [3, 0, 0], [-2, -1, -1]])
    kv = np.random.default_rng(7).normal(size=(nk, 3))
    A, D, S = ctx.wannier_integrals(g, rc, kv if nk else None)
    Ar, Dr, Sr = WO.integrals(g.values, g.nw, g.nspin, g.n_sc, g.p, g.origin, g.axes, rc,
                              kv if nk else np.zeros((0, 3)))
    sc = np.abs(Sr).max()
    assert np.abs(S - Sr).max() < 1e-12 * sc
    assert np.abs(D - Dr).max() < 1e-12 * np.abs(Dr).max()
    if nk:
        assert A.shape == Ar.shape
        assert np.abs(A - Ar).max() < 1e-12 * np.abs(Ar).max()
    assert np.all(S[:, :, :, 5] == 0) and np.all(D[:, :, :, 5] == 0)


def test_integrals_gaussian_closed_form(ctx):
    s = bse_amd.honeycomb_plain(5)
    g = _gaussians(s)
    rc = W.translations((1, 1, 0))
    Rc = rc @ g.axes.T
    kv = np.array([[0.0, 0.0, 0.0], [0.7, -0.3, 0.0], [1.5, 0.9, 0.4]])
    A, D, S = ctx.wannier_integrals(g, rc, kv)
    tau = np.asarray(s.model.tau, dtype=float)
    for i in range(2):
        for j in range(2):
            for r in range(len(rc)):
                a, d, o = WO.gaussian_pair(tau[i] + Rc[r], tau[j], 0.45, kv)
                assert np.allclose(A[0, i, j, r], a, atol=2e-9)
                assert np.allclose(D[0, i, j, r], d, atol=2e-9)
                assert abs(S[0, i, j, r] - o) < 2e-9


def _with_table(ctx, s, nspin=1):
    g = _gaussians(s, nspin=nspin)
    Rc, D = W.position_table(ctx, g, W.translations((2, 2, 0)))
    return s.with_params(dipole_r=Rc, dipole_d=D)


@pytest.mark.parametrize("which", ["plain", "magnetic"])
def test_bse_with_position_table_matches_oracle(ctx, which):
    s = bse_amd.honeycomb_plain(6) if which == "plain" else bse_amd.honeycomb_magnetic(4)
    st = _with_table(ctx, s, nspin=1 if which == "plain" else 2)
    print(f"bse wannier-dipole case {st.name}", flush=True)
    p = O.problem_from_system(st)
    sol = O.solve(p)
    ctx.set_system(st).build("fp64")
    rho = ctx.dipoles()
    assert np.abs(np.abs(rho) - np.abs(sol["rho0"])).max() < 1e-12 * np.abs(sol["rho0"]).max()
    ctx.eigh(-1)
    E = ctx.eigenvalues()
    scale = np.abs(sol["E"]).max()
    assert np.abs(E - sol["E"]).max() < 1e-10 * scale
    eps = ctx.spectrum(p.omegas, p.eta)
    rel = np.linalg.norm(eps.imag - sol["eps"].imag) / np.linalg.norm(sol["eps"].imag)
    assert rel < 1e-10
    # the overlapping Gaussians change the dipoles away from the point-orbital ones
    rho_pt = O.dipoles(O.problem_from_system(s), sol["bands"])
    assert np.abs(np.abs(rho) - np.abs(rho_pt)).max() > 1e-3 * np.abs(rho_pt).max()


def test_point_orbital_table_equals_point_route_on_gpu(ctx):
    s = bse_amd.random_multiorbital(4)
    rc = W.translations((1, 1, 0))
    Rc = rc @ np.asarray(s.model.lattice.bravais).T
    norb = s.model.norb
    D = np.zeros((len(rc), norb, norb, 3))
    D[len(rc) // 2] = np.einsum("ij,ia->ija", np.eye(norb), np.asarray(s.model.tau, dtype=float))
    ctx.set_system(s).build("fp64")
    rho = ctx.dipoles()
    ctx.set_system(s.with_params(dipole_r=Rc, dipole_d=D)).build("fp64")
    rho_t = ctx.dipoles()
    assert np.abs(rho_t - rho).max() < 1e-13 * np.abs(rho).max()


def test_refusals(ctx):
    g = W.WannierGrid(nw=1, nspin=1, n_sc=(1, 1, 1), p=(257, 1, 1), origin=np.zeros(3), axes=np.eye(3),
                      values=np.ones((1, 1, 257)))
    with pytest.raises(_abi.BSEError):
        ctx.wannier_integrals(g, np.zeros((1, 3), dtype=np.int32))
    st = _with_table(ctx, bse_amd.honeycomb_plain(4))
    with pytest.raises(_abi.BSEError):
        ctx.batch_prepare([st], precision="fp64", n_lowest=4, refine=False)
    p = O.problem_from_system(st)
    ctx.set_system(st).build("fp64").eigh(-1)
    with pytest.raises(_abi.BSEError):
        ctx.grad_spectrum(p.omegas, p.eta, np.zeros_like(p.omegas))
