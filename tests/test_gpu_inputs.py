"""Input paths of the drop-in boundary on the MI355X.

* 3D systems (3D k grid, 3D G-list, C/(q^2 + radius) potentials, prefactor
  1/(Omega Nk)) against the oracle, on the VALU (1v+1c) and the MFMA (2v+2c)
  assembly and in the matrix-free product.
* k-lists in the caller's order (bse_set_problem_klist): every reference-order
  output is the grid result with its k rows permuted; non-grids are refused.
* states at caller-given k (bse_tb_states_at) against the oracle's H(k) eig.
* a model read from collinear-spin hr.dat / centres.xyz files goes through
  the whole GPU path.
"""
import numpy as np
import pytest

import bse_amd
from bse_amd import _abi
from bse_amd import tight_binding as tb
import bse_oracle as O
from parity_util import cluster_overlap_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _abi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("make", [lambda: bse_amd.cubic_3d(4), lambda: bse_amd.random_cubic_3d(3)],
                         ids=["cubic_1v1c", "random_cubic_2v2c"])
def test_3d_system_matches_oracle(ctx, make):
    s = make()
    print(f"3D case {s.name} N={s.n_bse}", flush=True)
    p = O.problem_from_system(s)
    sol = O.solve(p)
    ctx.set_system(s).build("fp64")
    H = ctx.matrix()
    scale = np.abs(sol["H"]).max()
    assert np.abs(H - sol["H"]).max() < 1e-11 * scale
    rng = np.random.default_rng(3)
    X = rng.normal(size=(s.n_bse, 5)) + 1j * rng.normal(size=(s.n_bse, 5))
    assert np.abs(ctx.apply(X) - sol["H"] @ X).max() < 1e-11 * scale * np.sqrt(s.n_bse)
    ctx.eigh(-1)
    E = ctx.eigenvalues()
    assert np.abs(E - sol["E"]).max() < 1e-10 * np.abs(sol["E"]).max()
    eps = ctx.spectrum(p.omegas, p.eta)
    assert np.linalg.norm(eps.imag - sol["eps"].imag) < 1e-10 * np.linalg.norm(sol["eps"].imag)


def test_3d_mixed_refined(ctx):
    s = bse_amd.random_cubic_3d(3)
    ref = np.linalg.eigvalsh(O.assemble(O.problem_from_system(s)))
    ctx.set_system(s).build("fp32").eigh(8, refine=True)
    assert np.abs(ctx.eigenvalues()[:8] - ref[:8]).max() < 1e-10 * np.abs(ref).max()


def _permuted_case():
    s = bse_amd.honeycomb_magnetic(4)          # 2v+2c: MFMA assembly
    kc = s.grid.crystal()
    rng = np.random.default_rng(11)
    perm = rng.permutation(kc.shape[1])
    klist = kc[:, perm] + rng.integers(-1, 2, size=(3, kc.shape[1])) * np.array([[1], [1], [0]])
    return s, klist, perm


def test_klist_order_permutes_every_output(ctx):
    s, klist, perm = _permuted_case()
    nk, B = s.nk, s.nv * s.nc
    ctx.set_system(s).build("fp64")
    H0 = ctx.matrix()
    rho0 = ctx.dipoles()
    eps0, coef0 = ctx.tb_states()
    ctx.eigh(-1)
    E0 = ctx.eigenvalues()
    Z0 = ctx.eigenvectors(s.n_bse)
    ctx.set_system_klist(s, klist).build("fp64")
    # reference row r = cv * Nk + f holds grid row cv * Nk + perm[f]
    rows = (np.arange(B)[:, None] * nk + perm[None, :]).ravel()
    H1 = ctx.matrix()
    assert np.abs(H1 - H0[np.ix_(rows, rows)]).max() < 1e-13 * np.abs(H0).max()
    assert np.abs(ctx.dipoles() - rho0[rows]).max() < 1e-14
    eps1, coef1 = ctx.tb_states()
    assert np.array_equal(eps1, eps0[perm]) and np.array_equal(coef1, coef0[perm])
    X = np.random.default_rng(2).normal(size=(s.n_bse, 3)) + 0j
    assert np.abs(ctx.apply(X) - H1 @ X).max() < 1e-11 * np.abs(H0).max() * np.sqrt(s.n_bse)
    ctx.eigh(-1)
    assert np.abs(ctx.eigenvalues() - E0).max() < 1e-12 * np.abs(E0).max()
    Z1 = ctx.eigenvectors(s.n_bse)
    # the same matrix up to the row permutation: equal subspaces per cluster
    assert cluster_overlap_error(Z1, Z0[rows], E0, 1e-8 * np.abs(E0).max()) < 1e-10


def test_klist_refuses_non_grid(ctx):
    s, klist, _ = _permuted_case()
    bad = klist[:, :-1]
    with pytest.raises(_abi.BSEError, match="-6"):
        ctx.set_system_klist(s, bad)


def test_tb_states_at_given_k(ctx):
    m = tb.random_square(8, seed=5)
    rng = np.random.default_rng(4)
    k = rng.normal(size=(3, 37))
    k[2] = 0.0
    e, c = ctx.tb_states_at(k, model=m)
    ek, vk = np.linalg.eigh(m.hk(k))
    assert np.abs(e - ek).max() < 1e-12 * np.abs(ek).max()
    pa = np.einsum("kib,kjb->kbij", np.transpose(c, (0, 2, 1)), np.transpose(c, (0, 2, 1)).conj())
    pb = np.einsum("kib,kjb->kbij", vk, vk.conj())
    assert np.abs(pa - pb).max() < 1e-10
    # subset window: valence top-down, conduction bottom-up + scissor
    e2, c2 = ctx.tb_states_at(k, n_occ=4, nv=2, nc=3, scissor=0.25, model=m)
    assert np.allclose(e2[:, :2], ek[:, [3, 2]], atol=1e-12)
    assert np.allclose(e2[:, 2:], ek[:, [4, 5, 6]] + 0.25, atol=1e-12)


def test_collinear_files_through_gpu(ctx, tmp_path):
    m = tb.honeycomb_spinful(rashba=0.0, exchange=0.1)
    tb.write_hr_dat_collinear(tmp_path / "hr.dat", m)
    tb.write_centres_xyz_collinear(tmp_path / "centres.xyz", m)
    mf = tb.model_from_files(tmp_path / "hr.dat", tmp_path / "centres.xyz", m.lattice, collinear=True)
    s = bse_amd.honeycomb_magnetic(5).with_params(model=mf)
    ref = O.solve(O.problem_from_system(bse_amd.honeycomb_magnetic(5).with_params(model=m)))
    ctx.set_system(s).build("fp64").eigh(-1)
    E = ctx.eigenvalues()
    assert np.abs(E - ref["E"]).max() < 1e-10 * np.abs(ref["E"]).max()
