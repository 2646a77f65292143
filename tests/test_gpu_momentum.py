"""Finite exciton momentum and band paths on the MI355X against the oracle.

* A finite exciton momentum Q (|Q| > 1e-4 1/A; adding_momentum,
  bse.version13.cpp:2549-2567): valence states at k - Q and the exchange over
  v(|Q + G|), G != 0.  The assembled matrix (1e-11 relative) and the
  energies (1e-10) match the oracle; the full BSE refuses a finite Q.
* Band energies along a path file (Hamiltonian_TB::pull_bands, 823-875)
  through bse_tb_states_at match the oracle's H(k) eigenvalues (1e-12),
  for every band and for a band window with scissor.
"""
import numpy as np
import pytest

import bse_amd
from bse_amd import _abi
from bse_amd import lattice as L
import bse_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = _abi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("make,q", [
    (lambda: bse_amd.honeycomb_magnetic(4), (0.07, -0.03, 0.0)),
    (lambda: bse_amd.honeycomb_plain(6), (0.2, 0.1, 0.0)),
    (lambda: bse_amd.random_multiorbital(3), (0.05, 0.15, 0.0)),
    (lambda: bse_amd.cubic_3d(3), (0.1, 0.05, -0.08)),
], ids=["magnetic4", "plain6", "random3", "cubic3d"])
def test_finite_momentum_matches_oracle(ctx, make, q):
    s = make().with_params(q_opt=q)
    print(f"finite Q {s.name} q={q}", flush=True)
    p = O.problem_from_system(s)
    ref = O.solve(p, omegas=None)
    ctx.set_system(s).build("fp64")
    H = ctx.matrix()
    scale = np.abs(ref["H"]).max()
    assert np.abs(H - ref["H"]).max() < 1e-11 * scale
    ctx.eigh(-1)
    E = ctx.eigenvalues()
    assert np.abs(E - ref["E"]).max() < 1e-10 * np.abs(ref["E"]).max()


def test_full_bse_refuses_finite_momentum(ctx):
    s = bse_amd.honeycomb_plain(4).with_params(q_opt=(0.1, 0.0, 0.0))
    ctx.set_system(s)
    with pytest.raises(_abi.BSEError, match="optical limit"):
        ctx.build_full()


def test_band_path_matches_oracle(ctx, tmp_path):
    s = bse_amd.honeycomb_magnetic(4)
    (tmp_path / "path.dat").write_text("0 0 0 0.5 0 0 7\n0.5 0 0 0.3333333333333333 0.3333333333333333 0 5\n"
                                       "0.3333333333333333 0.3333333333333333 0 0 0 0 6\n")
    segs = L.read_k_path(tmp_path / "path.dat")
    ctx.set_system(s)
    k, e_all = ctx.bands(segs)
    p = O.problem_from_system(s)
    ref = np.linalg.eigvalsh(O.hk(p, k))
    assert e_all.shape == (18, s.norb)
    assert np.abs(e_all - ref).max() < 1e-12 * np.abs(ref).max()
    k2, e_win = ctx.bands(segs, n_occ=s.n_occ, nv=2, nc=2, scissor=0.4)
    assert np.allclose(k2, k)
    win = np.concatenate([ref[:, s.n_occ - 1::-1][:, :2], ref[:, s.n_occ:s.n_occ + 2] + 0.4], axis=1)
    assert np.abs(e_win - win).max() < 1e-12 * np.abs(win).max()
    L.write_bands(tmp_path / "bands.dat", e_win)
    assert np.allclose(L.read_bands(tmp_path / "bands.dat"), e_win, atol=1e-10)
