"""MI355X-native excitonic hot path of the tight-binding Bethe-Salpeter solver.

Host-side mirror of the reference's seams (bse.version13.cpp):
``Hamiltonian_TB.pull_ks_states_subset`` (778), ``Excitonic_Hamiltonian.
extract_hbse_and_rcv`` (3041), ``common_diagonalization`` (3120),
``pull_excitonic_oscillator_force`` (3486) and the dielectric-function sum
(3525), all executed by hand-written gfx950 kernels behind the C ABI declared in
``include/bse_mi355x.h`` (``libbse_mi355x.so``, loaded through ctypes).

The directory name is not a Python identifier; load it with
:func:`load_package` from a file path, e.g. ``importlib`` with
``submodule_search_locations`` (see ``__graft_entry__.py``).
"""
from . import lattice, tight_binding, system, wannier, formats  # noqa: F401
from .system import (BSESystem, honeycomb_plain, honeycomb_magnetic,  # noqa: F401
                     random_multiorbital, honeycomb_sweep, cubic_3d, random_cubic_3d, GRAD_PARAMS)
