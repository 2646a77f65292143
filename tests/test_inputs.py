"""Input formats of the reference (CPU): k-lists, collinear hr.dat, 3D models.

* k-list files (K_points::push_k_points_list_values, bse.version13.cpp:183-242)
  are accepted when they form a complete regular grid in any order, with any
  reciprocal-lattice translation of each point and one common offset; other
  lists are rejected (the C ABI's bse_set_problem_klist applies the same rule
  and returns BSE_ERR_UNSUPPORTED).
* collinear-spin hr.dat / centres.xyz (two stacked blocks, 494-552, 589-601)
  read into the spinor model with orbital 2 a + sigma.
"""
import numpy as np
import pytest

import bse_amd
from bse_amd import lattice as L
from bse_amd import tight_binding as tb


def test_klist_any_order_and_translation(tmp_path):
    g = L.KGrid(4, 6, (0.5, 0.0))
    kc = g.crystal()
    rng = np.random.default_rng(0)
    perm = rng.permutation(kc.shape[1])
    shifted = kc[:, perm] + rng.integers(-2, 3, size=(3, kc.shape[1])) * np.array([[1], [1], [0]])
    path = tmp_path / "klist"
    np.savetxt(path, shifted.T)
    k = L.read_k_list(path)
    grid, order = L.KGrid.from_crystal(k)
    assert (grid.n1, grid.n2, grid.n3) == (4, 6, 1)
    assert np.allclose(grid.shift, (0.5, 0.0))
    assert np.array_equal(order, perm)


def test_klist_number_and_shift_as_the_reference(tmp_path):
    path = tmp_path / "klist"
    np.savetxt(path, L.KGrid(3, 3).crystal().T)
    with open(path, "a") as f:
        f.write("\n\n")                       # trailing blank rows are not points
    k = L.read_k_list(path, number_k_points=9, shift=(1.0 / 6, 0.0, 0.0))
    grid, order = L.KGrid.from_crystal(k)
    assert np.allclose(grid.shift, (0.5, 0.0)) and np.array_equal(order, np.arange(9))


@pytest.mark.parametrize("bad", ["missing", "uneven", "duplicate", "kz"])
def test_klist_rejects_non_grids(bad):
    kc = L.KGrid(3, 4).crystal()
    if bad == "missing":
        kc = kc[:, :-1]
    elif bad == "uneven":
        kc = kc.copy()
        kc[0, kc[0] > 0.5] += 0.05
    elif bad == "duplicate":
        kc = kc.copy()
        kc[:, 1] = kc[:, 0]
    else:
        kc = kc.copy()
        kc[2] = 0.25
    with pytest.raises(ValueError):
        L.KGrid.from_crystal(kc)


def test_klist_3d():
    g = L.KGrid(2, 3, (0.0, 0.5, 0.25), 4)
    kc = g.crystal()
    perm = np.random.default_rng(1).permutation(kc.shape[1])
    grid, order = L.KGrid.from_crystal(kc[:, perm], dimension=3)
    assert (grid.n1, grid.n2, grid.n3) == (2, 3, 4)
    assert np.allclose(grid.shift, (0.0, 0.5, 0.25))
    assert np.array_equal(order, perm)


def test_collinear_hr_dat_roundtrip(tmp_path):
    m = tb.honeycomb_spinful(rashba=0.0, exchange=0.13)
    tb.write_hr_dat_collinear(tmp_path / "hr.dat", m)
    tb.write_centres_xyz_collinear(tmp_path / "centres.xyz", m)
    m2 = tb.model_from_files(tmp_path / "hr.dat", tmp_path / "centres.xyz", m.lattice, collinear=True)
    assert m2.spinful and m2.norb == 4
    assert np.array_equal(m2.r_cryst, m.r_cryst)
    assert np.abs(m2.h_r - m.h_r).max() < 1e-13
    assert np.abs(m2.tau - m.tau).max() < 1e-11
    # the up block of the file is what a spin-unpolarised reader sees first
    r, h = tb.read_hr_dat(tmp_path / "hr.dat")
    assert np.abs(h - m.h_r[:, 0::2, 0::2]).max() < 1e-13


def test_collinear_format_refuses_spin_mixing(tmp_path):
    with pytest.raises(ValueError):
        tb.write_hr_dat_collinear(tmp_path / "hr.dat", tb.honeycomb_spinful(rashba=0.02))


def test_collinear_reader_checks_blocks(tmp_path):
    m = tb.honeycomb_two_band()
    tb.write_hr_dat(tmp_path / "one.dat", m)
    with pytest.raises(ValueError):
        tb.read_hr_dat_collinear(tmp_path / "one.dat")


def test_cubic_model_hermitian():
    m = tb.cubic_two_band()
    assert m.hermiticity_error() < 1e-15
    s = bse_amd.cubic_3d(3)
    assert s.n_bse == 27 and s.gvectors().shape == (3, 27)
