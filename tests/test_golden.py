"""The oracle still reproduces the committed fixtures (CPU).

The fixtures in tests/golden/ are written by tests/golden/make_golden.py from
the oracle; the GPU parity tests compare against them.  Recomputing them here
pins the fixtures to the oracle (and the oracle to its own earlier output).
"""
import pathlib

import numpy as np
import pytest

import bse_amd
import bse_oracle as O
from parity_util import cluster_overlap_error, cluster_f2_error

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def _check(name, s, nz):
    g = np.load(GOLD / f"{name}.npz", allow_pickle=False)
    p = O.problem_from_system(s)
    sol = O.solve(p)
    scale = np.abs(sol["E"]).max()
    assert int(g["n"]) == s.n_bse
    assert np.abs(sol["E"] - g["E"]).max() < 1e-12 * scale
    assert cluster_f2_error(sol["E"], np.abs(sol["f"]) ** 2, g["f2"], 1e-8 * scale) < 1e-10
    assert np.linalg.norm(sol["eps"] - g["eps"]) < 1e-11 * np.linalg.norm(g["eps"])
    Z = sol["Z"] if nz is None else sol["Z"][:, :nz]
    assert cluster_overlap_error(Z, g["Z"], sol["E"], 1e-8 * scale, ncol=Z.shape[1]) < 1e-11


def test_c1_fixture():
    _check("c1_honeycomb_plain_12", bse_amd.honeycomb_plain(12), None)


@pytest.mark.slow
def test_c2_fixture():
    _check("c2_honeycomb_magnetic_30", bse_amd.honeycomb_magnetic(30), 16)
