"""Static RPA screening on the MI355X against the oracle.

The screened table W_ij(Q) comes from the TB states of the screening window
(Dielectric_Function::pull_values, bse.version13.cpp:2214-2303, order 1 at
omega = 0; W = eps^-1 v): the assembled FP64 matrix must equal the oracle's
(1e-11 relative, which checks every table entry through the direct term),
the energies and spectrum likewise; the full BSE runs on the RPA table; the
gradients (which hold the model W fixed) are refused.
"""
import numpy as np
import pytest

import bse_amd
from bse_amd import _abi
import bse_oracle as O
from parity_util import cluster_f2_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _abi.Context(0)
    yield c
    c.close()


CASES = [
    (lambda: bse_amd.honeycomb_plain(6).with_params(screening="rpa"), "plain6"),
    (lambda: bse_amd.honeycomb_magnetic(4).with_params(screening="rpa"), "magnetic4"),
    (lambda: bse_amd.random_multiorbital(3).with_params(screening="rpa", rpa_nv=2, rpa_nc=3), "random3_window"),
    (lambda: bse_amd.cubic_3d(3).with_params(screening="rpa"), "cubic3d"),
]


@pytest.mark.parametrize("make,name", CASES, ids=[c[1] for c in CASES])
def test_rpa_matrix_and_energies_match_oracle(ctx, make, name):
    s = make()
    print(f"RPA {name} N={s.n_bse}", flush=True)
    p = O.problem_from_system(s)
    ref = O.solve(p)
    ctx.set_system(s).build("fp64")
    H = ctx.matrix()
    scale = np.abs(ref["H"]).max()
    assert np.abs(H - ref["H"]).max() < 1e-11 * scale
    ctx.eigh(-1)
    E = ctx.eigenvalues()
    assert np.abs(E - ref["E"]).max() < 1e-10 * np.abs(ref["E"]).max()
    f2 = np.abs(ctx.oscillator()) ** 2
    assert cluster_f2_error(ref["E"], f2, np.abs(ref["f"]) ** 2, 1e-8 * np.abs(ref["E"]).max()) < 1e-9
    eps = ctx.spectrum(p.omegas, p.eta)
    assert np.linalg.norm(eps - ref["eps"]) < 1e-9 * np.linalg.norm(ref["eps"] - 1.0)


# On these small grids the unscreened Q = 0 head overbinds (the lowest TDA
# energies are negative at the models' own scissor); a larger scissor makes
# [[A, B], [conj B, conj A]] positive definite so the full BSE is defined.
FULL = [(lambda: bse_amd.honeycomb_plain(5), 4.0), (lambda: bse_amd.cubic_3d(3), 1.0)]


@pytest.mark.parametrize("make,extra", FULL, ids=["plain5", "cubic3d"])
def test_rpa_full_bse_matches_oracle(ctx, make, extra):
    s0 = make()
    s = s0.with_params(screening="rpa", scissor=s0.scissor + extra)
    print(f"RPA full BSE {s.name} scissor {s.scissor}", flush=True)
    p = O.problem_from_system(s)
    sol = O.solve_full(p)
    ctx.set_system(s).build_full().eigh_full(4)
    E = ctx.eigenvalues()
    assert np.abs(E - sol["E"]).max() < 1e-10 * np.abs(sol["E"]).max()


def test_rpa_mixed_precision_refined(ctx):
    s = bse_amd.honeycomb_magnetic(6).with_params(screening="rpa")
    ref = O.solve(O.problem_from_system(s), omegas=None)
    ctx.set_system(s).build("fp32").eigh(8, refine=True)
    E = ctx.eigenvalues()
    assert np.abs(E[:8] - ref["E"][:8]).max() < 1e-10 * np.abs(ref["E"]).max()


def test_rpa_gradients_refused(ctx):
    s = bse_amd.honeycomb_plain(4).with_params(screening="rpa")
    ctx.set_system(s).build("fp64").eigh(-1)
    with pytest.raises(_abi.BSEError, match="model W"):
        ctx.grad_energies(2)
