"""Output formats of the reference and the oracle's reference-quirks mode (CPU).

* The number formatting of the spectrum writer is checked against what
  g++'s std::ostream prints for the same doubles and std::complex values.
* The reference-quirks matrix (exchange erased by the direct row's '=',
  direct block conjugated) has the eigenvalues of the intended matrix with
  the exchange switched off; the difference to the intended spectrum is the
  exchange shift it documents.
"""
import shutil
import subprocess

import numpy as np
import pytest

import bse_amd
from bse_amd import formats as F
import bse_oracle as O

VALUES = [0.0, 1.0, -1.0, 0.1, 1e-7, 123456789.0, 3.14159265358979, -2.5e-5, 1e20, 2.0 / 3.0,
          100000.0, 1000000.0, 0.000123456789, 9.9999995, -0.0]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_number_format_matches_ostream(tmp_path):
    src = tmp_path / "fmt.cpp"
    vals = ", ".join(repr(v) for v in VALUES)
    src.write_text(
        "#include <complex>\n#include <iostream>\n"
        f"int main() {{ double v[] = {{{vals}}}; int n = sizeof(v) / sizeof(v[0]);\n"
        "  for (int i = 0; i < n; ++i) std::cout << v[i] << \"\\n\";\n"
        "  for (int i = 0; i + 1 < n; ++i) std::cout << std::complex<double>(v[i], v[i + 1]) << \"\\n\";\n"
        "  return 0; }\n")
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = [F.fmt_real(v) for v in VALUES]
    want += [F.fmt_complex(complex(VALUES[i], VALUES[i + 1])) for i in range(len(VALUES) - 1)]
    assert got[:len(want)] == want


def test_spectrum_file_round_trip(tmp_path):
    s = bse_amd.honeycomb_plain(6)
    p = O.problem_from_system(s)
    sol = O.solve(p)
    path = tmp_path / "spectrum.data"
    F.write_spectrum(path, p.omegas, sol["eps"])
    lines = path.read_text().split("\n")
    assert lines[0] == "### omega xx xy xz yx yy yz zx zy zz"
    assert len(lines) == len(p.omegas) + 2 and lines[-1] == ""
    w = F.fmt_complex(complex(p.omegas[5]))
    assert lines[6].startswith(f"{w} {F.fmt_complex(sol['eps'][5])} {w} (0,0) {w} (0,0) ")
    om, ex, last = F.read_spectrum(path)
    assert np.allclose(om.real, p.omegas, rtol=1e-5, atol=1e-12) and np.all(om.imag == 0)
    assert np.allclose(ex, sol["eps"], rtol=1e-5, atol=1e-6 * np.abs(sol["eps"]).max())
    assert np.allclose(last, sol["eps"].imag / 3, rtol=1e-5, atol=1e-6 * np.abs(sol["eps"]).max())


def test_spectrum_rejects_mismatched_lengths(tmp_path):
    with pytest.raises(ValueError):
        F.write_spectrum(tmp_path / "x", np.zeros(3), np.zeros(4))


def test_reference_quirks_mode():
    s = bse_amd.honeycomb_plain(6)
    p = O.problem_from_system(s)
    bands = O.band_window(p)
    Hq = O.assemble_reference_quirks(p, bands)
    H0 = O.assemble(O.problem_from_system(s.with_params(exchange_scale=0.0)), bands)
    H = O.assemble(p, bands)
    assert np.allclose(Hq, H0.conj(), atol=1e-14)
    Eq, E0, E = np.linalg.eigvalsh(Hq), np.linalg.eigvalsh(H0), np.linalg.eigvalsh(H)
    assert np.allclose(Eq, E0, atol=1e-12)
    # the erased exchange: the intended spectrum lies higher (repulsive exchange)
    assert E.sum() - Eq.sum() > 1e-3
    with pytest.raises(ValueError):
        O.assemble_reference_quirks(O.problem_from_system(bse_amd.honeycomb_magnetic(3)))
