"""Parity of the HIP path against the FP64 oracle (runs on the MI355X).

Every comparison is on gauge-invariant quantities where per-k eigenvector
phases could differ (band energies, projectors, the assembled matrix once the
TB gauge is fixed identically on both sides, eigenvalues, |f_m|^2 summed over
degenerate clusters, the spectrum).

Tolerances: FP64 path 1e-10 relative (eigenvalues, spectrum) and 1e-12 x
max|H| (matrix entries); FP32-storage path 0.5 meV on the lowest 16 excitons.
"""
import numpy as np
import pytest

import bse_amd
from bse_amd import _abi
import bse_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _abi.Context(0)
    yield c
    c.close()


def small_systems():
    return [
        bse_amd.honeycomb_plain(6),
        bse_amd.honeycomb_magnetic(4),
        bse_amd.random_multiorbital(4),
    ]


def _projector_err(a, b):
    """max | a a^H - b b^H | over k and slots (gauge invariant)."""
    pa = np.einsum("ksi,ksj->ksij", a, a.conj())
    pb = np.einsum("ksi,ksj->ksij", b, b.conj())
    return np.abs(pa - pb).max()


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_tb_states_match_oracle(ctx, idx):
    s = small_systems()[idx]
    print(f"tb parity case {s.name}", flush=True)
    ctx.set_system(s)
    eps, coef = ctx.tb_states()
    p = O.problem_from_system(s)
    eps_c, C, eps_v, V = O.band_window(p)
    ref_eps = np.concatenate([eps_v, eps_c], axis=1)
    ref_coef = np.concatenate([V, C], axis=1)
    assert np.abs(eps - ref_eps).max() < 1e-12 * max(1.0, np.abs(ref_eps).max())
    assert _projector_err(coef, ref_coef) < 1e-10


def _oracle_H(s):
    p = O.problem_from_system(s)
    return O.assemble(p), p


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_assembly_fp64_matches_oracle(ctx, idx):
    s = small_systems()[idx]
    print(f"assembly parity case {s.name}", flush=True)
    ctx.set_system(s).build("fp64")
    H = ctx.matrix()
    Href, p = _oracle_H(s)
    # gauges agree (largest component real positive on both sides) except at
    # exact modulus ties; compare gauge-invariant eigenvalues as well
    err = np.abs(H - Href).max() / np.abs(Href).max()
    ev_g = np.linalg.eigvalsh(H)
    ev_o = np.linalg.eigvalsh(Href)
    assert np.abs(ev_g - ev_o).max() < 1e-12 * np.abs(ev_o).max()
    assert err < 1e-11
    assert np.abs(H - H.conj().T).max() < 1e-13 * np.abs(Href).max()


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x2", 1e-4)])
@pytest.mark.parametrize("idx", [1, 2])
def test_assembly_reduced_precision(ctx, prec, tol, idx):
    """FP32 MFMA and compensated BF16 (hi+lo) MFMA assembly, FP32 storage."""
    s = small_systems()[idx]
    ctx.set_system(s).build(prec)
    H = ctx.matrix().astype(np.complex128)
    Href, _ = _oracle_H(s)
    err = np.abs(H - Href).max() / np.abs(Href).max()
    print(f"{prec} {s.name}: max rel entry error {err:.2e}")
    assert err < tol


def test_dipoles_match_oracle(ctx):
    s = small_systems()[2]
    ctx.set_system(s).build("fp64")
    rho = ctx.dipoles()
    p = O.problem_from_system(s)
    ref = O.dipoles(p)
    # |rho_r| is gauge invariant
    assert np.abs(np.abs(rho) - np.abs(ref)).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 63, 64, 65, 100, 257])
def test_dense_eigh_random_hermitian(ctx, n):
    print(f"dense eigh n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = 0.5 * (a + a.conj().T)
    ev, z = ctx.eigh_dense(a, "fp64", -1)
    ref = np.linalg.eigvalsh(a)
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(ev - ref).max() < 1e-12 * scale * max(1, n) ** 0.5
    res = np.abs(a @ z - z * ev[None, :]).max()
    orth = np.abs(z.conj().T @ z - np.eye(n)).max()
    assert res < 1e-11 * scale * max(1, n) ** 0.5
    assert orth < 1e-10


def test_dense_eigh_degenerate(ctx):
    """Exactly degenerate spectrum: Q diag(1,1,1,2,2,3,...) Q^H."""
    rng = np.random.default_rng(7)
    n = 96
    lam = np.repeat(np.arange(n // 4, dtype=float), 4)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    a = (q * lam) @ q.conj().T
    a = 0.5 * (a + a.conj().T)
    ev, z = ctx.eigh_dense(a, "fp64", -1)
    assert np.abs(ev - lam).max() < 1e-12 * n
    assert np.abs(z.conj().T @ z - np.eye(n)).max() < 1e-10
    assert np.abs(a @ z - z * ev).max() < 1e-11 * n


def test_dense_eigh_diagonal(ctx):
    d = np.array([3.0, -1.0, 2.0, 2.0, 0.5, -1.0, 7.0])
    ev, z = ctx.eigh_dense(np.diag(d).astype(complex), "fp64", -1)
    assert np.array_equal(ev, np.sort(d))
    assert np.abs(np.abs(z) - np.abs(z) ** 2).max() < 1e-15   # permutation matrix


def test_dense_eigh_fp32_storage(ctx):
    rng = np.random.default_rng(3)
    n = 300
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = 0.5 * (a + a.conj().T)
    ev, z = ctx.eigh_dense(a, "fp32", 16)
    ref = np.linalg.eigvalsh(a)
    assert np.abs(ev - ref).max() < 1e-5 * np.abs(ref).max()
    assert z.shape == (n, 16)
    # the 16 vectors come through the thin back-transform: residual at the
    # FP32 rounding of the stored matrix, orthonormal to FP64 reflector accuracy
    scale = np.abs(ref).max()
    assert np.abs(a @ z - z * ev[None, :16]).max() < 1e-4 * scale
    assert np.abs(z.conj().T @ z - np.eye(16)).max() < 1e-10


@pytest.mark.parametrize("k", [1, 3, 5, 16])
def test_dense_eigh_few_vectors(ctx, k):
    """k <= 16 vectors use the thin back-transform (three kernels per panel,
    any k): residual and orthonormality at FP64 accuracy, order 333 (not a
    multiple of the 64-row blocks)."""
    print(f"dense eigh n=333 k={k}", flush=True)
    rng = np.random.default_rng(333 + k)
    n = 333
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = 0.5 * (a + a.conj().T)
    ev, z = ctx.eigh_dense(a, "fp64", k)
    ref = np.linalg.eigvalsh(a)
    scale = np.abs(ref).max()
    assert z.shape == (n, k)
    assert np.abs(ev - ref).max() < 1e-12 * scale * n ** 0.5
    assert np.abs(a @ z - z * ev[None, :k]).max() < 1e-11 * scale * n ** 0.5
    assert np.abs(z.conj().T @ z - np.eye(k)).max() < 1e-10


@pytest.mark.parametrize("n", [4096, 8192])
def test_dense_eigh_laplacian_closed_form(ctx, n):
    """Orders 4096 and 8192 run the bisection with 4 and 2 lanes per
    eigenvalue: the 1-D Laplacian tridiag(-1, 2, -1) has the eigenvalues
    2 - 2 cos(k pi / (n + 1)), k = 1..n."""
    print(f"dense eigh laplacian n={n}", flush=True)
    a = np.zeros((n, n), dtype=np.complex128, order="F")
    i = np.arange(n)
    a[i, i] = 2.0
    a[i[:-1], i[:-1] + 1] = -1.0
    a[i[:-1] + 1, i[:-1]] = -1.0
    ev, _ = ctx.eigh_dense(a, "fp64", 1)
    ref = 2.0 - 2.0 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    assert np.abs(ev - ref).max() < 1e-12 * 4.0 * n ** 0.5


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_bse_eig_and_spectrum_fp64(ctx, idx):
    s = small_systems()[idx]
    print(f"eig/spectrum parity case {s.name}", flush=True)
    p = O.problem_from_system(s)
    sol = O.solve(p)
    ctx.set_system(s).build("fp64").eigh(-1)
    E = ctx.eigenvalues()
    scale = np.abs(sol["E"]).max()
    assert np.abs(E - sol["E"]).max() < 1e-10 * scale
    Z = ctx.eigenvectors(s.n_bse)
    H = sol["H"]
    assert np.abs(H @ Z - Z * E[None, :]).max() < 1e-10 * scale
    assert np.abs(Z.conj().T @ Z - np.eye(s.n_bse)).max() < 1e-9
    f = ctx.oscillator()
    # |f_m|^2 summed over clusters of (near-)degenerate excitons
    _cluster_f2_check(E, np.abs(f) ** 2, np.abs(sol["f"]) ** 2, 1e-8 * scale)
    eps = ctx.spectrum(p.omegas, p.eta)
    rel = np.linalg.norm(eps.imag - sol["eps"].imag) / np.linalg.norm(sol["eps"].imag)
    assert rel < 1e-10
    rel_re = np.linalg.norm(eps.real - sol["eps"].real) / np.linalg.norm(sol["eps"].real)
    assert rel_re < 1e-10


def _cluster_f2_check(E, f2, f2ref, tol):
    i = 0
    n = len(E)
    tot = f2ref.sum()
    while i < n:
        j = i + 1
        while j < n and E[j] - E[j - 1] < tol:
            j += 1
        assert abs(f2[i:j].sum() - f2ref[i:j].sum()) <= 1e-9 * tot + 1e-12
        i = j


def test_sum_rule_on_gpu(ctx):
    s = bse_amd.random_multiorbital(5)
    ctx.set_system(s).build("fp64").eigh(0)
    f = ctx.oscillator()
    rho = ctx.dipoles()
    assert abs((np.abs(f) ** 2).sum() / (np.abs(rho) ** 2).sum() - 1.0) < 1e-11


def test_mixed_fp32_lowest16(ctx):
    s = bse_amd.random_multiorbital(8)
    p = O.problem_from_system(s)
    ref = np.linalg.eigvalsh(O.assemble(p))
    ctx.set_system(s).build("fp32").eigh(16)
    E = ctx.eigenvalues()
    assert np.abs(E[:16] - ref[:16]).max() < 0.5e-3


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_matrix_free_apply(ctx, idx):
    s = small_systems()[idx]
    ctx.set_system(s).build("fp32")
    H, _ = _oracle_H(s)
    rng = np.random.default_rng(idx)
    X = rng.normal(size=(s.n_bse, 37)) + 1j * rng.normal(size=(s.n_bse, 37))
    # the TB gauges agree on both sides, so H X is comparable entry by entry
    Y = ctx.apply(X)
    assert np.abs(Y - H @ X).max() < 1e-11 * np.abs(H).max() * np.sqrt(s.n_bse)


def test_refinement_reaches_fp64(ctx):
    s = bse_amd.random_multiorbital(8)
    p = O.problem_from_system(s)
    H = O.assemble(p)
    ref = np.linalg.eigvalsh(H)
    ctx.set_system(s).build("fp32").eigh(16, refine=True)
    E = ctx.eigenvalues()
    scale = np.abs(ref).max()
    assert np.abs(E[:16] - ref[:16]).max() < 1e-10 * scale
    Z = ctx.eigenvectors(16)
    assert np.abs(H @ Z - Z * E[None, :16]).max() < 1e-9 * scale
    assert np.abs(Z.conj().T @ Z - np.eye(16)).max() < 1e-10


PARAMS = ("hopping_scale", "scissor", "r0", "eps_env", "exchange_field")


def test_hf_gradients_match_oracle(ctx):
    s = bse_amd.random_multiorbital(4)
    p = O.problem_from_system(s)
    g_ref, E_ref = O.hf_energy_gradients(p, n_exc=8, params=PARAMS)
    ctx.set_system(s).build("fp64").eigh(8)
    g = ctx.grad_energies(8)
    assert np.allclose(ctx.eigenvalues()[:8], E_ref, atol=1e-10)
    assert np.allclose(g, g_ref, rtol=1e-8, atol=1e-10), (g, g_ref)


def test_hf_gradients_match_gpu_finite_differences(ctx):
    from bse_amd import autograd as ag
    base = bse_amd.random_multiorbital(5)
    th0 = ag.default_theta(base)
    E0, g = ag.energies_and_gradients(ctx, base, th0, 8)
    h = 1e-4
    for t in range(5):
        tp, tm = th0.copy(), th0.copy()
        tp[t] += h
        tm[t] -= h
        Ep, _ = ag.energies_and_gradients(ctx, base, tp, 8)
        Em, _ = ag.energies_and_gradients(ctx, base, tm, 8)
        fd = (Ep - Em) / (2 * h)
        assert np.allclose(g[:, t], fd, rtol=1e-4, atol=1e-7), (PARAMS[t], g[:, t], fd)


def test_torch_autograd_function(ctx):
    import torch
    from bse_amd import autograd as ag
    base = bse_amd.random_multiorbital(4)
    th = torch.tensor(ag.default_theta(base), dtype=torch.float64, requires_grad=True)
    E = ag.exciton_energies(th, base, n_exc=8, ctx=ctx)
    E.sum().backward()
    _, g = ag.energies_and_gradients(ctx, base, ag.default_theta(base), 8)
    assert np.allclose(th.grad.numpy(), g.sum(axis=0), rtol=1e-12, atol=1e-14)


def _misfit_target(s, seed=5):
    p = O.problem_from_system(s)
    rng = np.random.default_rng(seed)
    base_eps = O.solve(p)["eps"].imag
    return p, base_eps * (1 + 0.1 * rng.normal(size=base_eps.size))


def test_spectrum_gradient_matches_oracle(ctx):
    s = bse_amd.random_multiorbital(3).with_params(n_omega=256)
    p, target = _misfit_target(s)
    g_ref, L_ref = O.spectrum_l2_gradient(p, target, params=PARAMS)
    ctx.set_system(s).build("fp64").eigh(-1)
    L, g = ctx.grad_spectrum(p.omegas, p.eta, target)
    assert abs(L - L_ref) <= 1e-10 * abs(L_ref)
    assert np.allclose(g, g_ref, rtol=1e-8, atol=1e-10 * np.abs(g_ref).max()), (g, g_ref)


def test_spectrum_gradient_matches_gpu_finite_differences(ctx):
    from bse_amd import autograd as ag
    base = bse_amd.random_multiorbital(3).with_params(n_omega=256)
    _, target = _misfit_target(base, seed=9)
    th0 = ag.default_theta(base)
    L0, g = ag.spectrum_misfit_and_gradient(ctx, base, th0, target)
    # eta = 50 meV Lorentzians: central differences need h = 1e-5 to keep the
    # O(h^2) truncation error below 1e-4 relative
    h = 1e-5
    for t in range(5):
        tp, tm = th0.copy(), th0.copy()
        tp[t] += h
        tm[t] -= h
        Lp, _ = ag.spectrum_misfit_and_gradient(ctx, base, tp, target)
        Lm, _ = ag.spectrum_misfit_and_gradient(ctx, base, tm, target)
        fd = (Lp - Lm) / (2 * h)
        assert abs(g[t] - fd) <= 1e-4 * abs(fd) + 1e-10 * abs(L0), (PARAMS[t], g[t], fd)


def test_torch_spectrum_misfit(ctx):
    import torch
    from bse_amd import autograd as ag
    base = bse_amd.random_multiorbital(2).with_params(n_omega=128)
    _, target = _misfit_target(base, seed=2)
    th = torch.tensor(ag.default_theta(base), dtype=torch.float64, requires_grad=True)
    L = ag.spectrum_misfit(th, base, target, ctx=ctx)
    L.backward()
    _, g = ag.spectrum_misfit_and_gradient(ctx, base, ag.default_theta(base), target)
    assert np.allclose(th.grad.numpy(), g, rtol=1e-12, atol=1e-14)
