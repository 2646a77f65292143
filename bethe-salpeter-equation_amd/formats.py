"""Output files in the reference's text formats.

* ``write_spectrum``: the macroscopic dielectric function file of
  ``Excitonic_Hamiltonian::pull_macroscopic_bse_dielectric_function``
  (bse.version13.cpp:3643-3656).  Header ``### omega xx xy xz yx yy yz zx zy
  zz``; then one line per frequency holding, for each diagonal component,
  the frequency and the component as ``std::complex`` prints them
  (``(re,im)``), and finally Im(average)/3.  The reference computes only
  eps_xx (its direction loops run over one value, 3574-3575), so eps_yy and
  eps_zz print as ``(0,0)`` and the average is eps_xx.
* Numbers follow ``std::ostream``'s default formatting: 6 significant
  digits, fixed or exponent notation as ``%g`` chooses.
"""
from __future__ import annotations

import re
import numpy as np

SPECTRUM_HEADER = "### omega xx xy xz yx yy yz zx zy zz"


def fmt_real(x: float) -> str:
    """A double as ``operator<<`` prints it by default (``%g``, precision 6)."""
    return "%g" % float(x)


def fmt_complex(z: complex) -> str:
    """A ``std::complex<double>`` as ``operator<<`` prints it: ``(re,im)``."""
    z = complex(z)
    return f"({fmt_real(z.real)},{fmt_real(z.imag)})"


def spectrum_lines(omegas, eps_xx):
    """The lines of the reference's spectrum file (without newlines)."""
    om = np.asarray(omegas)
    ex = np.asarray(eps_xx, dtype=np.complex128)
    if om.shape != ex.shape:
        raise ValueError("one eps_xx value per frequency")
    out = [SPECTRUM_HEADER]
    for w, e in zip(om, ex):
        wz = fmt_complex(complex(w))
        comps = [e, 0.0, 0.0]
        line = "".join(f"{wz} {fmt_complex(c)} " for c in comps)
        out.append(line + fmt_real(e.imag / 3.0))
    return out


def write_spectrum(path, omegas, eps_xx):
    """Write eps_xx(omega) (omegas in eV, real or complex) in the reference's
    spectrum format."""
    with open(path, "w") as f:
        f.write("\n".join(spectrum_lines(omegas, eps_xx)) + "\n")


_CX = re.compile(r"\(([^,()]+),([^,()]+)\)")


def read_spectrum(path):
    """(omegas complex, eps_xx complex, last column) from a spectrum file in
    the reference's format."""
    om, ex, last = [], [], []
    with open(path) as f:
        for line in f:
            if not line.strip() or line.startswith("###"):
                continue
            cx = _CX.findall(line)
            if len(cx) != 6:
                raise ValueError(f"{path}: expected 6 complex fields, got {len(cx)}")
            om.append(complex(float(cx[0][0]), float(cx[0][1])))
            ex.append(complex(float(cx[1][0]), float(cx[1][1])))
            last.append(float(line.rsplit(")", 1)[1]))
    return np.array(om), np.array(ex), np.array(last)
