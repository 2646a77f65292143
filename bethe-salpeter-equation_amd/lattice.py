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
    """Monkhorst-Pack style grid k = ((i + s1)/n1) b1 + ((j + s2)/n2) b2
    (+ ((l + s3)/n3) b3 for 3D systems).

    k index = (i * n2 + j) * n3 + l (i along b1 slowest).  ``shift`` is in
    crystal units of one grid step, so zeros are Gamma-centred.  Differences
    k_a - k_b are then integer combinations (di/n1) b1 + (dj/n2) b2 (+ (dl/n3)
    b3), which the W table exploits.
    """

    n1: int
    n2: int
    shift: tuple = (0.0, 0.0)
    n3: int = 1

    @property
    def nk(self) -> int:
        return self.n1 * self.n2 * self.n3

    @property
    def shift3(self) -> float:
        return float(self.shift[2]) if len(self.shift) > 2 else 0.0

    def crystal(self) -> np.ndarray:
        i, j, l = np.meshgrid(np.arange(self.n1), np.arange(self.n2), np.arange(self.n3),
                              indexing="ij")
        c = np.zeros((3, self.nk))
        c[0] = (i.ravel() + self.shift[0]) / self.n1
        c[1] = (j.ravel() + self.shift[1]) / self.n2
        if self.n3 > 1 or self.shift3 != 0.0:
            c[2] = (l.ravel() + self.shift3) / self.n3
        return c

    def cartesian(self, lattice: CrystalLattice) -> np.ndarray:
        """3 x nk Cartesian k-points (bse.version13.cpp:205-216)."""
        return lattice.reciprocal @ self.crystal()

    def index_pairs(self):
        """(i, j) integer coordinates of each k index (2D grids)."""
        i, j = np.meshgrid(np.arange(self.n1), np.arange(self.n2), indexing="ij")
        return i.ravel(), j.ravel()

    @classmethod
    def from_crystal(cls, kc, dimension: int = 2, tol: float = 1e-7):
        """(grid, order) for a k-list in crystal coordinates (3 x nk) that is a
        complete regular grid with one common offset, in any order and with
        any reciprocal-lattice translation of each point.  order[f] is the
        grid index of list entry f.  Raises ValueError for anything else (the
        rule bse_set_problem_klist applies on the device side)."""
        kc = np.asarray(kc, dtype=np.float64).reshape(3, -1)
        nk = kc.shape[1]
        fr = kc - np.floor(kc)
        fr[fr > 1.0 - tol] = 0.0
        n, shift, idx = [1, 1, 1], [0.0, 0.0, 0.0], []
        for a in range(3):
            if a == 2 and dimension == 2:
                if np.any(fr[2] > tol):
                    raise ValueError("2D k-list with a component along b3")
                idx.append(np.zeros(nk, dtype=np.int64))
                continue
            u = np.unique(np.round(fr[a] / tol) * tol)
            vals = [u[0]]
            for v in u[1:]:
                if v - vals[-1] > 2 * tol:
                    vals.append(v)
            na = len(vals)
            if not np.allclose(np.array(vals) - vals[0], np.arange(na) / na, atol=10 * tol):
                raise ValueError("k-list coordinates are not equally spaced along an axis")
            t = (fr[a] - vals[0]) * na
            j = np.rint(t).astype(np.int64)
            if np.abs(t - j).max() > 1e-5 or j.min() < 0 or j.max() >= na:
                raise ValueError("k-list point off the grid")
            n[a], shift[a] = na, vals[0] * na
            idx.append(j)
        if n[0] * n[1] * n[2] != nk:
            raise ValueError("k-list is not a complete grid")
        order = (idx[0] * n[1] + idx[1]) * n[2] + idx[2]
        if np.unique(order).size != nk:
            raise ValueError("k-list repeats a grid point")
        sh = (shift[0], shift[1]) if dimension == 2 else tuple(shift)
        return cls(n[0], n[1], sh, n[2]), order


def read_k_list(path, number_k_points=None, shift=(0.0, 0.0, 0.0)):
    """k-list file of K_points::push_k_points_list_values (183-242): lines of
    ``k1 k2 k3``, reading stops after ``number_k_points`` lines (all if None);
    ``shift`` is added to every point.  Returns 3 x nk (crystal coordinates
    when the file holds them)."""
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) < 3:
                continue
            rows.append([float(x) for x in parts[:3]])
            if number_k_points is not None and len(rows) >= number_k_points:
                break
    k = np.array(rows, dtype=np.float64).T
    return k + np.asarray(shift, dtype=np.float64)[:, None]


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


def read_k_path(path, number_segments=None):
    """Band-path file of Hamiltonian_TB::pull_bands (823-875): per segment the
    start k, the end k and the number of points, ``k0x k0y k0z k1x k1y k1z n``
    (whitespace-separated, line breaks free); reading stops after
    ``number_segments`` segments (all if None) or at the end of the numbers.
    Returns a list of (k0 (3,), k1 (3,), n)."""
    with open(path) as f:
        vals = f.read().split()
    segs, i = [], 0
    while i + 7 <= len(vals):
        if number_segments is not None and len(segs) >= number_segments:
            break
        k0 = np.array([float(x) for x in vals[i:i + 3]])
        k1 = np.array([float(x) for x in vals[i + 3:i + 6]])
        n = int(float(vals[i + 6]))
        if n < 2:
            raise ValueError(f"{path}: a segment needs at least 2 points")
        segs.append((k0, k1, n))
        i += 7
    return segs


def band_path_points(segments, lattice: CrystalLattice | None = None, crystal: bool = True) -> np.ndarray:
    """Points of the band path (pull_bands 846-860): segment (k0, k1, n) gives
    k_j = k0 (1 - j/(n-1)) + k1 j/(n-1), j = 0 .. n-1, every segment in turn
    (end points repeated as in the reference).  crystal: the file's k are in
    reciprocal-lattice coordinates and are converted with the reciprocal
    vectors of ``lattice`` (the reference's crystal_coordinates = 1 with its
    primitive_vectors).  Returns 3 x nk Cartesian (1/Angstrom)."""
    pts = []
    for k0, k1, n in segments:
        t = np.arange(n) / (n - 1)
        pts.append(np.outer(k0, 1.0 - t) + np.outer(k1, t))
    k = np.concatenate(pts, axis=1)
    if crystal:
        if lattice is None:
            raise ValueError("crystal coordinates need the lattice")
        k = lattice.reciprocal @ k
    return k


def write_bands(path, energies) -> None:
    """Bands file of pull_bands (864-868): one line per k point, the band
    energies (eV) separated by spaces."""
    e = np.asarray(energies, dtype=np.float64)
    with open(path, "w") as f:
        for row in e:
            f.write(" ".join(f"{x:.12g}" for x in row) + " \n")


def read_bands(path) -> np.ndarray:
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                rows.append([float(x) for x in line.split()])
    return np.array(rows)
