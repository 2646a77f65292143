"""One excitonic problem: TB model + k-grid + band window + screening + spectrum.

The fields keep the reference's names where it has them (bse.version13.cpp main,
3686-3848): ``number_valence_bands``/``number_conduction_bands`` (nv, nc),
``scissor_operator`` (816), the optical momentum q = minval * x (3562, 3574),
``cutoff_g_points_list`` (3726).  Screening is the Rytova-Keldysh model and the
exchange uses the bare 2D Coulomb potential (the reference's empirical
eps^-1 = 0.086206897, 2698, and its 2D Coulomb with |k|^4, 1303, are replaced).

Band selection is by index: with ``n_occ`` occupied bands, valence v = 0 is band
n_occ-1 (highest valence, 812-813) and conduction c = 0 is band n_occ (lowest
conduction, 815-816), the reference's looking_from_fermi=1 order for an insulator.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
import numpy as np

from .lattice import CONVERSION_PARAMETER, MINVAL, KGrid, g_list
from . import tight_binding as tb


@dataclass
class BSESystem:
    model: tb.TBModel
    grid: KGrid
    n_occ: int
    nv: int
    nc: int
    r0: float = 10.0                # Keldysh screening length, Angstrom
    eps_env: float = 1.0            # environment dielectric constant
    g_cut: float = 6.0              # G-list cutoff, 1/Angstrom
    scissor: float = 0.0            # eV, added to conduction energies
    exchange_scale: float = 1.0     # multiplies the exchange term
    direct_scale: float = 1.0       # multiplies the direct (screened) term
    q_opt: tuple = (MINVAL, 0.0, 0.0)
    eta: float = 0.05               # Lorentzian broadening of the spectrum, eV
    n_omega: int = 2048
    omega_max: float = 6.0
    name: str = "system"
    meta: dict = field(default_factory=dict)
    dimension: int = 2              # 2: Keldysh W, 2D v; 3: C/(eps (q^2+radius)), C/(q^2+radius)
    radius: float = 0.0             # 3D Coulomb regularisation (1281-1285), 1/Angstrom^2
    screening: str = "model"        # "model" (r0, eps_env) or "rpa": static RPA from the TB states
    rpa_nv: int = 0                 # RPA screening window: highest occupied bands (0: all)
    rpa_nc: int = 0                 # lowest empty bands (0: all)
    # optical dipole from a Wannier position table (wannier.position_table):
    # Cartesian lattice vectors (nR, 3) and D_ij(R) = int w_i(r - R) r w_j(r)
    # (nR, norb, norb, 3); None: point orbitals (tau_i)
    dipole_r: np.ndarray = None
    dipole_d: np.ndarray = None

    # ---- derived sizes -------------------------------------------------
    @property
    def nk(self) -> int:
        return self.grid.nk

    @property
    def norb(self) -> int:
        return self.model.norb

    @property
    def n_bse(self) -> int:
        """N = Nk * nv * nc (bse.version13.cpp:2496)."""
        return self.nk * self.nv * self.nc

    @property
    def spin_factor(self) -> float:
        """f_s = 2 - spinorial (2801): 2 spin-degenerate singlet, 1 spinor."""
        return 1.0 if self.model.spinful else 2.0

    @property
    def area(self) -> float:
        return self.model.lattice.area

    @property
    def volume(self) -> float:
        return self.model.lattice.volume

    @property
    def cell(self) -> float:
        """Measure of the kernel prefactor 1/(cell Nk): area (2D) or volume (3D)."""
        return self.volume if self.dimension == 3 else self.area

    def kpoints(self) -> np.ndarray:
        return self.grid.cartesian(self.model.lattice)

    def gvectors(self) -> np.ndarray:
        return g_list(self.model.lattice, self.g_cut, dimension=self.dimension)

    def omegas(self) -> np.ndarray:
        """omega_s = s * omega_max / n_omega, endpoint excluded (3828-3829)."""
        return np.arange(self.n_omega) * (self.omega_max / self.n_omega)

    @property
    def spectrum_prefactor(self) -> float:
        """C / (Omega * Nk) of the dielectric function (3532)."""
        return CONVERSION_PARAMETER / (self.volume * self.nk)

    def with_grid(self, n1, n2=None) -> "BSESystem":
        return replace(self, grid=KGrid(n1, n1 if n2 is None else n2))

    def with_params(self, **kw) -> "BSESystem":
        return replace(self, **kw)


# ---------------------------------------------------------------------------
# Workload presets
# ---------------------------------------------------------------------------

def honeycomb_plain(n: int = 12, gap: float = 3.0, r0: float = 10.0,
                    eps_env: float = 1.0) -> BSESystem:
    """Two-band honeycomb (a = 2.5 A, t = -2.7 eV, gap 3 eV), 1v+1c, n x n grid.

    At n = 12: N = 144.  Keldysh r0 = 10 A, eps_env = 1; G_cut = 6 /A gives the
    5 x 5 G-list (|b| = 2.902 /A).
    """
    return BSESystem(tb.honeycomb_two_band(gap=gap), KGrid(n, n), n_occ=1, nv=1, nc=1,
                     r0=r0, eps_env=eps_env, g_cut=6.0, name=f"honeycomb_plain_{n}x{n}")


def honeycomb_magnetic(n: int = 30) -> BSESystem:
    """Spinful honeycomb, exchange 0.1 eV along z, Rashba 0.02 eV, 2v+2c.

    At n = 30: N = 3600, f_s = 1.
    """
    return BSESystem(tb.honeycomb_spinful(), KGrid(n, n), n_occ=2, nv=2, nc=2,
                     r0=10.0, eps_env=1.0, g_cut=6.0, name=f"honeycomb_magnetic_{n}x{n}")


def random_multiorbital(n: int = 36, seed: int = 20736) -> BSESystem:
    """Random 8-orbital square lattice (a = 3 A), NN+NNN complex Gaussian hoppings
    sigma = 0.5 eV, on-site U(-2, 2) eV; bands 0-3 valence, 4-7 conduction;
    4v+4c, Keldysh r0 = 15 A.  At n = 36: N = 20736.  G_cut = 4.5 /A (5 x 5)."""
    return BSESystem(tb.random_square(8, seed=seed), KGrid(n, n), n_occ=4, nv=4, nc=4,
                     r0=15.0, eps_env=1.0, g_cut=4.5, name=f"random_multiorbital_{n}x{n}")


def cubic_3d(n: int = 4, gap: float = 4.0, t: float = -1.0, eps_env: float = 11.6) -> BSESystem:
    """3D two-band cubic model (CsCl arrangement, a = 3 A), 1v+1c, n x n x n grid.

    The direct term uses the reference's 3D potential C/(q^2 + radius) screened
    by a static eps (the reference's empirical eps^-1 = 1/11.6, 2698), the
    exchange its bare form; G_cut = 2.2 /A gives the 3 x 3 x 3 G-list.
    At n = 4: N = 64.
    """
    return BSESystem(tb.cubic_two_band(gap=gap, t=t), KGrid(n, n, (0.0, 0.0, 0.0), n), n_occ=1,
                     nv=1, nc=1, r0=0.0, eps_env=eps_env, g_cut=2.2, dimension=3,
                     name=f"cubic_3d_{n}x{n}x{n}")


def random_cubic_3d(n: int = 3, seed: int = 3, radius: float = 0.1) -> BSESystem:
    """Random 4-orbital simple-cubic model, 2v+2c, n x n x n grid, 3D potentials
    with radius 0.1 /A^2 and eps = 4; G_cut 2.2 /A (3 x 3 x 3).  N = 4 n^3."""
    return BSESystem(tb.random_cubic(4, seed=seed), KGrid(n, n, (0.0, 0.0, 0.0), n), n_occ=2,
                     nv=2, nc=2, r0=0.0, eps_env=4.0, g_cut=2.2, dimension=3, radius=radius,
                     name=f"random_cubic_3d_{n}x{n}x{n}")


def honeycomb_sweep(n_systems: int = 64, seed: int = 4, n_min: int = 12,
                    n_max: int = 40) -> list:
    """Batched sweep of two-band honeycomb systems with uneven k-grids.

    Draw order from ``Generator(PCG64(seed))`` per system: n1, n2 = integers in
    [n_min, n_max] (inclusive), gap ~ U(1.5, 4), r0 ~ U(5, 30), eps_env ~ U(1, 5).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for s in range(n_systems):
        n1, n2 = (int(x) for x in rng.integers(n_min, n_max + 1, size=2))
        gap = float(rng.uniform(1.5, 4.0))
        r0 = float(rng.uniform(5.0, 30.0))
        eps = float(rng.uniform(1.0, 5.0))
        sysm = BSESystem(tb.honeycomb_two_band(gap=gap), KGrid(n1, n2), n_occ=1, nv=1,
                         nc=1, r0=r0, eps_env=eps, g_cut=6.0,
                         name=f"sweep{s}_{n1}x{n2}")
        out.append(sysm)
    return out


# Parameters differentiated by the Hellmann-Feynman kernel, in this order.
GRAD_PARAMS = ("hopping_scale", "scissor", "r0", "eps_env", "exchange_field")
