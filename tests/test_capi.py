"""The library called from plain C (tests/capi/bse_c_driver.c, built by build()).

The driver reads the reference's own input files (Wannier90 hr.dat, one
block or two collinear blocks, and the centres .xyz) plus a key/value file of
the constants the reference's main() sets in code, runs set_problem -> build
-> eigh -> oscillator -> spectrum through the C ABI and writes text.  No
Python and no torch sit between the caller and the library.

CPU tests: the binary exists, links the in-tree library and refuses bad
parameter files before it touches a GPU.  GPU tests: the c1 system
(honeycomb_plain(12), N = 144) written to files and run through the driver
must reproduce the oracle fixture (every eigenvalue to 1e-10 relative, |f|^2
per degenerate cluster, the spectrum, all eigenvectors as cluster
subspaces); a collinear-spin model read from two-block files must match the
oracle; the FP32 + refinement path on c1 gives the lowest 16 to 1e-10.
"""
import pathlib
import subprocess

import numpy as np
import pytest

import bse_amd
from bse_amd import tight_binding as tb
import bse_oracle as O
from parity_util import cluster_f2_error, cluster_overlap_error

ROOT = pathlib.Path(__file__).resolve().parents[1]
DRIVER = ROOT / "tests" / "capi" / "bse_c_driver"
GOLD = ROOT / "tests" / "golden"


def _write_inputs(tmp, s, precision=0, n_vectors=-1, refine=0, omegas=None, eta=0.05, collinear=False):
    m = s.model
    if collinear:
        tb.write_hr_dat_collinear(tmp / "hr.dat", m)
        tb.write_centres_xyz_collinear(tmp / "centres.xyz", m)
    else:
        tb.write_hr_dat(tmp / "hr.dat", m)
        tb.write_centres_xyz(tmp / "centres.xyz", m)
    a = np.asarray(m.lattice.bravais, dtype=float)            # columns a_i
    sh = list(s.grid.shift)[:2] + [float(getattr(s.grid, "shift3", 0.0))]
    lines = [
        "# written by tests/test_capi.py",
        "hr_dat hr.dat",
        f"hr_blocks {2 if collinear else 1}",
        "centres centres.xyz",
        "centres_atoms 0",
        "lattice " + " ".join(f"{x:.17g}" for x in a.T.ravel()),
        f"grid {s.grid.n1} {s.grid.n2} {int(getattr(s.grid, 'n3', 1))}",
        "shift " + " ".join(f"{x:.17g}" for x in sh),
        f"bands {s.n_occ} {s.nv} {s.nc}",
        f"scissor {s.scissor:.17g}",
        "q " + " ".join(f"{x:.17g}" for x in s.q_opt),
        f"screening {s.r0:.17g} {s.eps_env:.17g}",
        f"g_cut {s.g_cut:.17g}",
        f"spin_factor {s.spin_factor:.17g}",
        f"scales {s.exchange_scale:.17g} {s.direct_scale:.17g}",
        "dipole_dir 1 0 0",
        f"dimension {int(getattr(s, 'dimension', 2))}",
        f"radius {float(getattr(s, 'radius', 0.0)):.17g}",
        f"screening_model {getattr(s, 'screening', 'model')}",
        f"rpa_window {int(getattr(s, 'rpa_nv', 0))} {int(getattr(s, 'rpa_nc', 0))}",
        f"precision {precision}",
        f"n_vectors {n_vectors}",
        f"refine {refine}",
        f"eta {eta:.17g}",
    ]
    if omegas is not None:
        lines.append(f"omega {omegas[0]:.17g} {omegas[-1]:.17g} {len(omegas)}")
    (tmp / "params.txt").write_text("\n".join(lines) + "\n")
    return tmp / "params.txt"


def _run(params, out, timeout=120):
    r = subprocess.run([str(DRIVER), str(params), str(out)], capture_output=True, text=True, timeout=timeout)
    return r


def _parse(out):
    E, f2, om, eps, z = [], [], [], [], []
    nz = None
    with open(out) as f:
        n = int(f.readline().split()[1])
        for line in f:
            t = line.split()
            if t[0] == "E":
                E.append(float(t[1]))
                f2.append(float(t[2]))
            elif t[0] == "eps":
                om.append(float(t[1]))
                eps.append(float(t[2]) + 1j * float(t[3]))
            elif t[0] == "Z":
                nz = int(t[1])
            else:
                z.append(float(t[0]) + 1j * float(t[1]))
    Z = np.array(z).reshape(nz, n).T if nz else np.zeros((n, 0))
    return n, np.array(E), np.array(f2), np.array(om), np.array(eps), Z


def test_driver_built_and_linked():
    assert DRIVER.exists(), "build() compiles tests/capi/bse_c_driver"
    r = subprocess.run(["readelf", "-d", str(DRIVER)], capture_output=True, text=True, check=True)
    assert "libbse_mi355x.so" in r.stdout


def test_driver_rejects_bad_params(tmp_path):
    p = tmp_path / "params.txt"
    p.write_text("hr_dat hr.dat\nbogus_key 1\n")
    r = _run(p, tmp_path / "out.txt", timeout=60)
    assert r.returncode == 1 and "unknown key" in r.stderr
    p.write_text("grid 4 4\n")
    r = _run(p, tmp_path / "out.txt", timeout=60)
    assert r.returncode == 1 and "bad line" in r.stderr


def test_driver_rejects_missing_files(tmp_path):
    p = tmp_path / "params.txt"
    p.write_text("hr_dat nothere.dat\ncentres nothere.xyz\ngrid 4 4 1\n")
    r = _run(p, tmp_path / "out.txt", timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr


@pytest.mark.gpu
def test_c_driver_c1_against_fixture(tmp_path):
    g = np.load(GOLD / "c1_honeycomb_plain_12.npz", allow_pickle=False)
    s = bse_amd.honeycomb_plain(12)
    params = _write_inputs(tmp_path, s, omegas=g["omegas"], eta=float(g["eta"]))
    om = np.linspace(g["omegas"][0], g["omegas"][-1], len(g["omegas"]))
    assert np.abs(om - g["omegas"]).max() < 1e-12
    print("C driver c1 N=144 fp64 all vectors", flush=True)
    r = _run(params, tmp_path / "out.txt")
    assert r.returncode == 0, r.stderr
    n, E, f2, omr, eps, Z = _parse(tmp_path / "out.txt")
    assert n == int(g["n"]) == 144
    scale = np.abs(g["E"]).max()
    assert np.abs(E - g["E"]).max() < 1e-10 * scale
    assert cluster_f2_error(g["E"], f2, g["f2"], 1e-8 * scale) < 1e-9
    assert np.linalg.norm(eps - g["eps"]) < 1e-9 * np.linalg.norm(g["eps"])
    assert cluster_overlap_error(Z, g["Z"], g["E"], 1e-8 * scale, ncol=n) < 1e-9
    assert np.abs(Z.conj().T @ Z - np.eye(n)).max() < 1e-10


@pytest.mark.gpu
def test_c_driver_c1_fp32_refined_lowest16(tmp_path):
    g = np.load(GOLD / "c1_honeycomb_plain_12.npz", allow_pickle=False)
    params = _write_inputs(tmp_path, bse_amd.honeycomb_plain(12), precision=1, n_vectors=16, refine=1)
    print("C driver c1 fp32 + refine, 16 vectors", flush=True)
    r = _run(params, tmp_path / "out.txt")
    assert r.returncode == 0, r.stderr
    n, E, f2, _, _, Z = _parse(tmp_path / "out.txt")
    scale = np.abs(g["E"]).max()
    assert np.abs(E[:16] - g["E"][:16]).max() < 1e-10 * scale
    assert cluster_overlap_error(Z, g["Z"][:, :16], g["E"], 1e-8 * scale, ncol=16) < 1e-8


@pytest.mark.gpu
def test_c_driver_rpa_screening(tmp_path):
    s = bse_amd.honeycomb_plain(5).with_params(screening="rpa")
    params = _write_inputs(tmp_path, s)
    print("C driver honeycomb 5x5 with RPA screening", flush=True)
    r = _run(params, tmp_path / "out.txt")
    assert r.returncode == 0, r.stderr
    n, E, f2, _, _, _ = _parse(tmp_path / "out.txt")
    ref = O.solve(O.problem_from_system(s), omegas=None)
    assert np.abs(E - ref["E"]).max() < 1e-10 * np.abs(ref["E"]).max()


@pytest.mark.gpu
def test_c_driver_collinear_files(tmp_path):
    m = tb.honeycomb_spinful(rashba=0.0, exchange=0.1)
    s = bse_amd.honeycomb_magnetic(5).with_params(model=m)
    params = _write_inputs(tmp_path, s, collinear=True)
    print("C driver collinear honeycomb 5x5", flush=True)
    r = _run(params, tmp_path / "out.txt")
    assert r.returncode == 0, r.stderr
    n, E, f2, _, _, _ = _parse(tmp_path / "out.txt")
    ref = O.solve(O.problem_from_system(s))
    scale = np.abs(ref["E"]).max()
    assert n == s.n_bse
    assert np.abs(E - ref["E"]).max() < 1e-10 * scale
    assert cluster_f2_error(ref["E"], f2, np.abs(ref["f"]) ** 2, 1e-8 * scale) < 1e-8
