"""Write the oracle fixtures used by the GPU parity tests (run from the repo root).

    python tests/golden/make_golden.py

c1: honeycomb_plain(12), N = 144 -- every eigenvalue, |f_m|^2, the spectrum
    (2048 points, eta = 50 meV) and all eigenvectors (reference order).
c2: honeycomb_magnetic(30), N = 3600 -- every eigenvalue, |f_m|^2, the
    spectrum and the lowest 16 eigenvectors.
c2_full: the same system beyond Tamm-Dancoff (coupling block included) --
    the N positive energies, |f|^2 and the full spectrum, from the oracle's
    structure-preserving solve (Cholesky of [[A, B], [conj B, conj A]] and
    numpy eigh of L^H s3 L; the general-eig route agrees with it on the
    small systems of tests/test_oracle_full_bse.py, and order 7200 is too
    large for it here).

The fixtures are outputs of the FP64 oracle (numpy complex128, LAPACK zheevd),
stored as .npz without pickles.  tests/test_golden.py checks that the oracle
still reproduces them.
"""
import importlib.util
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
PKG = ROOT / "bethe-salpeter-equation_amd"
sys.path.insert(0, str(ROOT / "oracle"))


def _pkg():
    spec = importlib.util.spec_from_file_location("bse_amd", PKG / "__init__.py",
                                                  submodule_search_locations=[str(PKG)])
    m = importlib.util.module_from_spec(spec)
    sys.modules["bse_amd"] = m
    spec.loader.exec_module(m)
    return m


CASES = {
    "c1_honeycomb_plain_12": (lambda m: m.honeycomb_plain(12), None),
    "c2_honeycomb_magnetic_30": (lambda m: m.honeycomb_magnetic(30), 16),
}


def solve_case(m, name):
    import bse_oracle as O
    make, nz = CASES[name]
    s = make(m)
    p = O.problem_from_system(s)
    sol = O.solve(p)
    Z = sol["Z"] if nz is None else sol["Z"][:, :nz]
    return dict(E=sol["E"], f2=np.abs(sol["f"]) ** 2, eps=sol["eps"], Z=np.ascontiguousarray(Z),
                omegas=p.omegas, eta=np.float64(p.eta), n=np.int64(s.n_bse))


def solve_full_case(m, make):
    import bse_oracle as O
    s = make(m)
    p = O.problem_from_system(s)
    bands = O.band_window(p)
    A = O.assemble(p, bands)
    B = O.assemble_coupling(p, bands)
    rho_t = O.full_dipoles(p, bands)
    E, _, f = O.solve_full_structured(A, B, rho_t)
    eps = O.dielectric_full(E, f, p.omegas, p.eta, O.C_EV_A / (p.volume * p.nk))
    return dict(E=E, f2=np.abs(f) ** 2, eps=eps, omegas=p.omegas, eta=np.float64(p.eta), n=np.int64(s.n_bse))


FULL_CASES = {"c2_full_honeycomb_magnetic_30": lambda m: m.honeycomb_magnetic(30)}


if __name__ == "__main__":
    m = _pkg()
    out = pathlib.Path(__file__).resolve().parent
    only = sys.argv[1:]
    for name in CASES:
        if only and name not in only:
            continue
        d = solve_case(m, name)
        np.savez_compressed(out / f"{name}.npz", **d)
        print(name, "N =", int(d["n"]), "E0 =", d["E"][0])
    for name, make in FULL_CASES.items():
        if only and name not in only:
            continue
        d = solve_full_case(m, make)
        np.savez_compressed(out / f"{name}.npz", **d)
        print(name, "N =", int(d["n"]), "E0 =", d["E"][0])
