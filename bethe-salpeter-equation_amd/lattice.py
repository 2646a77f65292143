"""Geometry inputs of the excitonic path: Bravais lattice, k-grid and G-list.

Mirrors the reference's three geometry classes:

* ``Crystal_Lattice``  (bse.version13.cpp:85-124): file row i is a_i, stored as
  column i; volume = |det|; b_i = 2*pi*(a_j x a_k)/volume.
* ``K_points``         (bse.version13.cpp:183-242): k-list in crystal coordinates
  plus a shift, converted to Cartesian with the reciprocal vectors.
* ``G_points``         (bse.version13.cpp:319-385): G = i*b1 + j*b2 (+ k*b3) for
  |i| <= int(G_cut/|b_i|), enumerated lexicographically, so that index Ng//2 is
  G = 0 (the reference reads G0 at ``int(number_g_points_list/2)``, 2811).  The
  reference's 2D branch is broken (count initialised to 0, wrong index, fill
  starting at 1: 356, 361, 368); this module builds the intended 2D list.

Units: lengths in Angstrom, k and G in 1/Angstrom.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import numpy as np

# e^2/epsilon_0 in eV*Angstrom (bse.version13.cpp:32).
CONVERSION_PARAMETER = (1.602176634 * 1000.0) / 8.8541878176
# Small optical momentum magnitude (bse.version13.cpp:28, used as q at 3574-3579).
MINVAL = 1.0e-5


@dataclass
class CrystalLattice:
    """Bravais lattice with columns a1, a2, a3 (Angstrom)."""

    bravais: np.ndarray  # (3, 3), column i = a_i

    def __post_init__(self):
        self.bravais = np.asarray(self.bravais, dtype=np.float64).reshape(3, 3)

    @classmethod
    def from_rows(cls, rows):
        """Build from a 3x3 array whose row i is a_i (the file layout, 93-96)."""
        return cls(np.asarray(rows, dtype=np.float64).reshape(3, 3).T.copy())

    @property
    def volume(self) -> float:
        return float(abs(np.linalg.det(self.bravais)))

    @property
    def area(self) -> float:
        """In-plane cell area |a1 x a2| used by the 2D kernel prefactor."""
        return float(np.linalg.norm(np.cross(self.bravais[:, 0], self.bravais[:, 1])))

    @property
    def reciprocal(self) -> np.ndarray:
        a0, a1, a2 = (self.bravais[:, i] for i in range(3))
        b = np.stack([np.cross(a1, a2), np.cross(a2, a0), np.cross(a0, a1)], axis=1)
        return (2.0 * np.pi / self.volume) * b


@dataclass
class KGrid:
    """Monkhorst-Pack style grid k = ((i + s1)/n1) b1 + ((j + s2)/n2) b2.

    k index = i * n2 + j (i along b1 slowest).  ``shift`` is in crystal units of
    one grid step, so (0, 0) is Gamma-centred.  Differences k_a - k_b are then
    integer combinations (di/n1) b1 + (dj/n2) b2, which the W table exploits.
    """

    n1: int
    n2: int
    shift: tuple = (0.0, 0.0)

    @property
    def nk(self) -> int:
        return self.n1 * self.n2

    def crystal(self) -> np.ndarray:
        i, j = np.meshgrid(np.arange(self.n1), np.arange(self.n2), indexing="ij")
        c = np.zeros((3, self.nk))
        c[0] = (i.ravel() + self.shift[0]) / self.n1
        c[1] = (j.ravel() + self.shift[1]) / self.n2
        return c

    def cartesian(self, lattice: CrystalLattice) -> np.ndarray:
        """3 x nk Cartesian k-points (bse.version13.cpp:205-216)."""
        return lattice.reciprocal @ self.crystal()

    def index_pairs(self):
        """(i, j) integer coordinates of each k index."""
        i, j = np.meshgrid(np.arange(self.n1), np.arange(self.n2), indexing="ij")
        return i.ravel(), j.ravel()


def g_list(lattice: CrystalLattice, g_cut: float, dimension: int = 2) -> np.ndarray:
    """G-list of the reference's parallelepiped form, G = 0 at index Ng//2.

    3D follows bse.version13.cpp:334-351; 2D is the intended version of 352-379
    (directions b1, b2 only).  Returns 3 x Ng.
    """
    b = lattice.reciprocal
    dirs = 3 if dimension == 3 else 2
    n = [int(g_cut / np.linalg.norm(b[:, d])) for d in range(dirs)]
    ranges = [np.arange(-m, m + 1) for m in n]
    if dirs == 3:
        ii, jj, kk = np.meshgrid(*ranges, indexing="ij")
        coeff = np.stack([ii.ravel(), jj.ravel(), kk.ravel()])
    else:
        ii, jj = np.meshgrid(*ranges, indexing="ij")
        coeff = np.stack([ii.ravel(), jj.ravel(), np.zeros(ii.size, dtype=int)])
    g = b @ coeff.astype(np.float64)
    assert np.allclose(g[:, g.shape[1] // 2], 0.0)
    return g


@dataclass
class GList:
    """Cached G-list plus the integer extents used to build it."""

    vectors: np.ndarray = field(repr=False)

    @property
    def ng(self) -> int:
        return self.vectors.shape[1]

    @property
    def g0(self) -> int:
        return self.ng // 2
