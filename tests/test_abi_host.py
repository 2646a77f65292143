"""C ABI and host-logic tests that need no GPU."""
import ctypes
import pathlib
import re
import subprocess

import numpy as np
import pytest

import bse_amd
from bse_amd import _abi, tight_binding as tb

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "bse_mi355x.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bse_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_abi.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 17
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_abi.ABI_SYMBOLS) <= set(names)


def test_abi_version():
    lib = _abi.load_library()
    assert lib.bse_abi_version() == 3


def test_create_without_gpu_fails_cleanly():
    """On a host with no visible GPU the context constructor reports an error
    code instead of crashing (the product has no CPU fallback)."""
    lib = _abi.load_library()
    import os
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU device node is present")
    h = ctypes.c_void_p()
    assert lib.bse_create(ctypes.byref(h), 0) != 0
    with pytest.raises(_abi.BSEError):
        _abi.Context(0)


def test_struct_layout_matches_c(tmp_path):
    """ctypes mirrors of bse_tb_model / bse_system match the C compiler's layout."""
    src = tmp_path / "layout.c"
    src.write_text(f'''#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
int main(void) {{
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(bse_tb_model), sizeof(bse_system),
         offsetof(bse_system, n_occ), offsetof(bse_system, scissor),
         offsetof(bse_system, r0), offsetof(bse_system, dipole_dir),
         offsetof(bse_system, dimension), offsetof(bse_system, n3),
         offsetof(bse_system, shift3), offsetof(bse_system, radius),
         offsetof(bse_system, screening), offsetof(bse_system, rpa_nc));
  return 0;
}}''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                            check=True).stdout.split()]
    S = _abi.SystemC
    assert vals == [ctypes.sizeof(_abi.TBModelC), ctypes.sizeof(S), S.n_occ.offset,
                    S.scissor.offset, S.r0.offset, S.dipole_dir.offset, S.dimension.offset,
                    S.n3.offset, S.shift3.offset, S.radius.offset, S.screening.offset,
                    S.rpa_nc.offset]


def test_hr_dat_roundtrip(tmp_path):
    m = tb.random_square(8, seed=3)
    tb.write_hr_dat(tmp_path / "m_hr.dat", m)
    r, h = tb.read_hr_dat(tmp_path / "m_hr.dat")
    order = np.lexsort(m.r_cryst.T[::-1])
    order2 = np.lexsort(r.T[::-1])
    assert np.array_equal(r[order2], m.r_cryst[order])
    assert np.allclose(h[order2], m.h_r[order], atol=1e-13)
    tb.write_centres_xyz(tmp_path / "c.xyz", m)
    assert np.allclose(tb.read_centres_xyz(tmp_path / "c.xyz", 8), m.tau)
    tb.write_bravais(tmp_path / "b.dat", m.lattice)
    assert np.allclose(tb.read_bravais(tmp_path / "b.dat").bravais, m.lattice.bravais)


def test_hr_dat_size_mismatch_raises(tmp_path):
    m = tb.honeycomb_two_band()
    p = tmp_path / "bad_hr.dat"
    tb.write_hr_dat(p, m)
    lines = p.read_text().splitlines()
    p.write_text("\n".join(lines[:-3]) + "\n")
    with pytest.raises(ValueError):
        tb.read_hr_dat(p)


def test_models_hermitian_and_sizes():
    for s, n in [(bse_amd.honeycomb_plain(12), 144), (bse_amd.honeycomb_magnetic(30), 3600),
                 (bse_amd.random_multiorbital(36), 20736)]:
        assert s.n_bse == n
        assert s.model.hermiticity_error() < 1e-15
    sweep = bse_amd.honeycomb_sweep()
    assert len(sweep) == 64
    nks = [x.nk for x in sweep]
    assert min(nks) >= 144 and max(nks) <= 1600
    # seeded: the same list twice
    assert [x.name for x in bse_amd.honeycomb_sweep()] == [x.name for x in sweep]


def test_random_model_seeded():
    a = tb.random_square(8, seed=20736)
    b = tb.random_square(8, seed=20736)
    c = tb.random_square(8, seed=1)
    assert np.array_equal(a.h_r, b.h_r) and np.array_equal(a.tau, b.tau)
    assert not np.array_equal(a.h_r, c.h_r)
