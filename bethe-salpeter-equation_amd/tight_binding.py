"""Tight-binding models: Wannier90-style H(R) cubes, orbital centres, generators.

The reference reads ``hr.dat`` + ``centres.xyz`` (``Hamiltonian_TB`` ctor,
bse.version13.cpp:460-608) and forms H_lm(k) = sum_R w_R H_lm(R) e^{i k.R} with
Cartesian R (554-564, 689-700).  Two details are kept on purpose:

* the phase carries R only, never the orbital centre ("lattice gauge");
* the reference *multiplies* by the ndegen weights (676-677, 696-697) where
  Wannier90 divides.  :func:`read_hr_dat` divides (the Wannier90 meaning) and
  the generators here write ndegen = 1, so both conventions agree on them.

Orbital index conventions for spinful models: i = 2*site + sigma (sigma = 0 up,
1 down); every spin component sits on its site's centre.  The reference's
collinear mode stacks two full hr blocks instead (494, 548-551); a spin-diagonal
spinor model reproduces it.

Model generators for the five workloads use ``numpy.random.Generator(PCG64(seed))``
with a fixed draw order documented on each function, so every run rebuilds the
same Hamiltonian.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
import numpy as np

from .lattice import CrystalLattice


@dataclass
class TBModel:
    lattice: CrystalLattice
    r_cryst: np.ndarray            # (nr, 3) integer lattice coordinates of R
    h_r: np.ndarray                # (nr, norb, norb) complex128, H_lm(R)
    tau: np.ndarray                # (norb, 3) orbital centres, Angstrom
    spinful: bool = False
    name: str = "tb"
    meta: dict = field(default_factory=dict)

    @property
    def norb(self) -> int:
        return self.h_r.shape[1]

    @property
    def nr(self) -> int:
        return self.h_r.shape[0]

    @property
    def r_cart(self) -> np.ndarray:
        """(nr, 3) Cartesian R (bse.version13.cpp:554-564)."""
        return (self.lattice.bravais @ self.r_cryst.T.astype(np.float64)).T

    def hermiticity_error(self) -> float:
        """max |H(R) - H(-R)^dagger| over R; zero for a valid model."""
        idx = {tuple(r): n for n, r in enumerate(self.r_cryst.tolist())}
        err = 0.0
        for n, r in enumerate(self.r_cryst.tolist()):
            m = idx.get(tuple(-x for x in r))
            if m is None:
                return float("inf")
            err = max(err, float(np.max(np.abs(self.h_r[n] - self.h_r[m].conj().T))))
        return err

    def hk(self, kcart: np.ndarray) -> np.ndarray:
        """H(k) for a 3 x nk array of Cartesian k (host helper; nk x norb x norb)."""
        ph = np.exp(1j * (self.r_cart @ np.asarray(kcart)))  # (nr, nk)
        return np.einsum("rk,rlm->klm", ph, self.h_r)

    def scaled(self, hopping_scale: float = 1.0, exchange_field: float = 0.0,
               onsite_shift: np.ndarray | None = None) -> "TBModel":
        """Copy with every R != 0 block scaled and an on-site sigma_z field added.

        ``exchange_field`` adds h * diag(+1, -1, +1, -1, ...) to H(0): orbitals are
        read as (site, spin) pairs, i = 2*site + sigma.
        """
        h = self.h_r.copy()
        zero = np.all(self.r_cryst == 0, axis=1)
        h[~zero] *= hopping_scale
        i0 = int(np.nonzero(zero)[0][0])
        sz = np.where(np.arange(self.norb) % 2 == 0, 1.0, -1.0)
        h[i0] += np.diag(exchange_field * sz)
        if onsite_shift is not None:
            h[i0] += np.diag(onsite_shift)
        return replace(self, h_r=h)


def _complete_hermitian(blocks: dict, norb: int):
    """Add H(-R) = H(R)^dagger for every given R != 0; returns r list and cube."""
    full = {}
    for r, h in blocks.items():
        full[r] = np.asarray(h, dtype=np.complex128)
        if any(r):
            mr = tuple(-x for x in r)
            if mr in blocks:
                raise ValueError("give only one of R and -R")
            full[mr] = full[r].conj().T.copy()
    rs = sorted(full.keys())
    return np.array(rs, dtype=np.int64), np.stack([full[r] for r in rs])


def honeycomb_lattice(a: float = 2.5, c: float = 20.0) -> CrystalLattice:
    a1 = np.array([a, 0.0, 0.0])
    a2 = np.array([0.5 * a, 0.5 * np.sqrt(3.0) * a, 0.0])
    a3 = np.array([0.0, 0.0, c])
    return CrystalLattice(np.stack([a1, a2, a3], axis=1))


def honeycomb_two_band(gap: float = 3.0, t: float = -2.7, a: float = 2.5) -> TBModel:
    """Two-band honeycomb: on-site +gap/2 (A), -gap/2 (B); NN hopping t.

    tau_A = 0, tau_B = (a1 + a2)/3; H_AB(R) = t for R in {0, -a1, -a2}, so the
    three A->B bond vectors tau_B + R - tau_A all have length a/sqrt(3).
    """
    lat = honeycomb_lattice(a)
    tau = np.zeros((2, 3))
    tau[1] = (lat.bravais[:, 0] + lat.bravais[:, 1]) / 3.0
    blocks = {(0, 0, 0): np.array([[gap / 2, t], [t, -gap / 2]], dtype=complex)}
    for r in [(-1, 0, 0), (0, -1, 0)]:
        h = np.zeros((2, 2), dtype=complex)
        h[0, 1] = t
        blocks[r] = h
    rs, h = _complete_hermitian(blocks, 2)
    return TBModel(lat, rs, h, tau, spinful=False, name="honeycomb2",
                   meta=dict(gap=gap, t=t, a=a))


_SX = np.array([[0, 1], [1, 0]], dtype=complex)
_SY = np.array([[0, -1j], [1j, 0]], dtype=complex)
_SZ = np.array([[1, 0], [0, -1]], dtype=complex)


def honeycomb_spinful(gap: float = 3.0, t: float = -2.7, exchange: float = 0.1,
                      rashba: float = 0.02, a: float = 2.5) -> TBModel:
    """Spinful honeycomb with exchange field h*sigma_z and Kane-Mele Rashba term.

    Orbitals (A up, A down, B up, B down).  NN block for a hop from site i to site j
    with unit bond vector d: t*1 + i*lambda_R*(sigma x d)_z = t + i*lambda_R*(sx*dy - sy*dx).
    """
    lat = honeycomb_lattice(a)
    pos = [np.zeros(3), (lat.bravais[:, 0] + lat.bravais[:, 1]) / 3.0]
    tau = np.array([pos[0], pos[0], pos[1], pos[1]])
    h0 = np.zeros((4, 4), dtype=complex)
    h0[0:2, 0:2] = gap / 2 * np.eye(2) + exchange * _SZ
    h0[2:4, 2:4] = -gap / 2 * np.eye(2) + exchange * _SZ
    blocks = {}

    def ab_block(r):
        d = pos[1] + lat.bravais @ np.array(r, dtype=float) - pos[0]
        d = d / np.linalg.norm(d)
        return t * np.eye(2) + 1j * rashba * (_SX * d[1] - _SY * d[0])

    h0[0:2, 2:4] = ab_block((0, 0, 0))
    h0[2:4, 0:2] = h0[0:2, 2:4].conj().T
    blocks[(0, 0, 0)] = h0
    for r in [(-1, 0, 0), (0, -1, 0)]:
        h = np.zeros((4, 4), dtype=complex)
        h[0:2, 2:4] = ab_block(r)
        blocks[r] = h
    rs, h = _complete_hermitian(blocks, 4)
    return TBModel(lat, rs, h, tau, spinful=True, name="honeycomb_spinful",
                   meta=dict(gap=gap, t=t, exchange=exchange, rashba=rashba, a=a))


def random_square(norb: int = 8, a: float = 3.0, sigma: float = 0.5,
                  seed: int = 20736, onsite_range: float = 2.0) -> TBModel:
    """Random multi-orbital square-lattice model (seeded PCG64).

    Draw order from ``Generator(PCG64(seed))``:
      1. tau crystal coordinates: uniform(size=(norb, 2)) in the cell;
      2. on-site energies: uniform(-onsite_range, onsite_range, size=norb);
      3. for R in (a1, a2, a1+a2, a1-a2): Re = normal(0, sigma/sqrt2, (norb, norb)),
         then Im the same; H(-R) = H(R)^dagger.
    Intra-cell off-diagonal elements are zero.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    lat = CrystalLattice(np.diag([a, a, 20.0]))
    u = rng.uniform(size=(norb, 2))
    tau = np.zeros((norb, 3))
    tau[:, 0] = u[:, 0] * a
    tau[:, 1] = u[:, 1] * a
    onsite = rng.uniform(-onsite_range, onsite_range, size=norb)
    blocks = {(0, 0, 0): np.diag(onsite).astype(complex)}
    s = sigma / np.sqrt(2.0)
    for r in [(1, 0, 0), (0, 1, 0), (1, 1, 0), (1, -1, 0)]:
        re = rng.normal(0.0, s, size=(norb, norb))
        im = rng.normal(0.0, s, size=(norb, norb))
        blocks[r] = re + 1j * im
    rs, h = _complete_hermitian(blocks, norb)
    return TBModel(lat, rs, h, tau, spinful=False, name=f"random_square{norb}",
                   meta=dict(seed=seed, sigma=sigma, a=a))


def cubic_two_band(gap: float = 4.0, t: float = -1.0, a: float = 3.0) -> TBModel:
    """3D two-band model on a simple cubic lattice (CsCl arrangement).

    tau_A = 0, tau_B = (a/2, a/2, a/2); on-site +gap/2 (A), -gap/2 (B); hopping
    t from A to each of its 8 nearest B sites, i.e. H_AB(R) = t for the eight
    R in {0, -1}^3 (lattice units).
    """
    lat = CrystalLattice(np.diag([a, a, a]))
    tau = np.zeros((2, 3))
    tau[1] = 0.5 * a
    blocks = {(0, 0, 0): np.array([[gap / 2, t], [t, -gap / 2]], dtype=complex)}
    for r in [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, -1, 0), (-1, 0, -1), (0, -1, -1),
              (-1, -1, -1)]:
        h = np.zeros((2, 2), dtype=complex)
        h[0, 1] = t
        blocks[r] = h
    rs, h = _complete_hermitian(blocks, 2)
    return TBModel(lat, rs, h, tau, spinful=False, name="cubic2",
                   meta=dict(gap=gap, t=t, a=a))


def random_cubic(norb: int = 4, a: float = 3.0, sigma: float = 0.5, seed: int = 3,
                 onsite_range: float = 2.0) -> TBModel:
    """Random multi-orbital simple-cubic model (seeded PCG64), the 3D analogue
    of random_square: tau uniform in the cell, on-site U(-2, 2) eV, complex
    Gaussian hoppings to the six nearest cells (R = a1, a2, a3 drawn, -R by
    Hermiticity).  Draw order: tau (norb x 3), on-site (norb), then Re and Im
    (norb x norb) for each R."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lat = CrystalLattice(np.diag([a, a, a]))
    tau = rng.uniform(size=(norb, 3)) * a
    onsite = rng.uniform(-onsite_range, onsite_range, size=norb)
    blocks = {(0, 0, 0): np.diag(onsite).astype(complex)}
    sd = sigma / np.sqrt(2.0)
    for r in [(1, 0, 0), (0, 1, 0), (0, 0, 1)]:
        re = rng.normal(0.0, sd, size=(norb, norb))
        im = rng.normal(0.0, sd, size=(norb, norb))
        blocks[r] = re + 1j * im
    rs, h = _complete_hermitian(blocks, norb)
    return TBModel(lat, rs, h, tau, spinful=False, name=f"random_cubic{norb}",
                   meta=dict(seed=seed, sigma=sigma, a=a))


# ----------------------------------------------------------------------------
# Reference file formats (Appendix B of the blueprint; bse.version13.cpp lines)
# ----------------------------------------------------------------------------

def write_bravais(path, lattice: CrystalLattice):
    """9 numbers, file row i = a_i (read as column i at 93-96)."""
    with open(path, "w") as f:
        for i in range(3):
            f.write(" ".join(f"{x:.12f}" for x in lattice.bravais[:, i]) + "\n")


def read_bravais(path) -> CrystalLattice:
    vals = np.loadtxt(path, dtype=np.float64).reshape(3, 3)
    return CrystalLattice.from_rows(vals)


def write_hr_dat(path, model: TBModel, header: str = "written by bse_amd"):
    """Wannier90 hr.dat: header, num_wann, nrpts, ndegen (15 per line), then
    ``R1 R2 R3 l m Re Im`` with l fastest inside each R (read at 531-545)."""
    with open(path, "w") as f:
        f.write(_hr_text(model, header))


def _read_hr_block(f):
    """One hr block from an open file (header line, num_wann, nrpts, ndegen,
    then nrpts*num_wann^2 lines), or None at the end of the file."""
    head = f.readline()
    if not head:
        return None
    line = f.readline()
    while line and not line.strip():
        line = f.readline()
    if not line:
        return None
    norb = int(line.split()[0])
    nr = int(f.readline().split()[0])
    deg = []
    while len(deg) < nr:
        deg.extend(int(x) for x in f.readline().split())
    deg = np.array(deg[:nr], dtype=np.float64)
    rows = []
    while len(rows) < nr * norb * norb:
        line = f.readline()
        if not line:
            break
        parts = line.split()
        if parts:
            rows.append([float(x) for x in parts[:7]])
    data = np.array(rows, dtype=np.float64).reshape(-1, 7)
    if data.shape[0] != nr * norb * norb:
        raise ValueError(f"hr.dat: expected {nr * norb * norb} lines, got {data.shape[0]}")
    r = data[:, 0:3].astype(np.int64).reshape(nr, norb * norb, 3)[:, 0, :]
    l = data[:, 3].astype(int) - 1
    m = data[:, 4].astype(int) - 1
    h = np.zeros((nr, norb, norb), dtype=np.complex128)
    rid = np.repeat(np.arange(nr), norb * norb)
    h[rid, l, m] = data[:, 5] + 1j * data[:, 6]
    h /= deg[:, None, None]
    return r, h


def read_hr_dat(path):
    """Parse hr.dat; returns (r_cryst (nr,3) int, h_r (nr,norb,norb) complex).

    H(R) is divided by ndegen (Wannier90's meaning; the reference multiplies,
    676-677).  A header/size mismatch raises instead of printing and going on
    (the reference prints ``ERROR!!!!!!!`` and continues, 499-501).
    """
    with open(path) as f:
        blk = _read_hr_block(f)
    if blk is None:
        raise ValueError("hr.dat: empty file")
    return blk


def read_hr_dat_collinear(path):
    """Collinear-spin hr.dat: two complete blocks, spin up then spin down
    (Hamiltonian_TB ctor, 494-552, spinorial_calculation = 1), each with the
    same num_wann and R list.  Returns (r_cryst, h_r) of the equivalent spinor
    model, orbital i = 2 a + sigma: H_{2a+s, 2b+s'}(R) = delta_ss' H_s(R)_ab."""
    with open(path) as f:
        up = _read_hr_block(f)
        dn = _read_hr_block(f)
    if up is None or dn is None:
        raise ValueError("collinear hr.dat needs two blocks (up, down)")
    (r_u, h_u), (r_d, h_d) = up, dn
    if h_u.shape != h_d.shape or not np.array_equal(r_u, r_d):
        raise ValueError("collinear hr.dat: the two spin blocks differ in size or R list")
    nr, nw, _ = h_u.shape
    h = np.zeros((nr, 2 * nw, 2 * nw), dtype=np.complex128)
    h[:, 0::2, 0::2] = h_u
    h[:, 1::2, 1::2] = h_d
    return r_u, h


def write_hr_dat_collinear(path, model: TBModel, header: str = "written by bse_amd"):
    """Two stacked hr blocks (up, then down) of a spin-diagonal spinor model."""
    h = model.h_r
    if np.abs(h[:, 0::2, 1::2]).max(initial=0.0) > 0 or np.abs(h[:, 1::2, 0::2]).max(initial=0.0) > 0:
        raise ValueError("the model mixes spins; the collinear format cannot hold it")
    from dataclasses import replace as _replace
    with open(path, "w") as f:
        for sig in (0, 1):
            blk = _replace(model, h_r=h[:, sig::2, sig::2].copy(), tau=model.tau[sig::2])
            f.write(_hr_text(blk, header + (" up" if sig == 0 else " down")))


def _hr_text(model: TBModel, header: str) -> str:
    norb, nr = model.norb, model.nr
    out = [header + "\n", f"{norb:12d}\n{nr:12d}\n"]
    for i in range(0, nr, 15):
        out.append("".join(f"{1:5d}" for _ in range(min(15, nr - i))) + "\n")
    for n in range(nr):
        r = model.r_cryst[n]
        for m in range(norb):
            for l in range(norb):
                h = model.h_r[n, l, m]
                out.append(f"{r[0]:5d}{r[1]:5d}{r[2]:5d}{l + 1:5d}{m + 1:5d}"
                           f"{h.real:22.14e}{h.imag:22.14e}\n")
    return "".join(out)


def write_centres_xyz(path, model: TBModel, atoms=None):
    """centres.xyz: count line, comment line, ``X x y z`` per orbital, then atoms."""
    atoms = [] if atoms is None else atoms
    with open(path, "w") as f:
        f.write(f"{model.norb + len(atoms)}\n written by bse_amd\n")
        for t in model.tau:
            f.write(f"X {t[0]:.12f} {t[1]:.12f} {t[2]:.12f}\n")
        for sym, p in atoms:
            f.write(f"{sym} {p[0]:.12f} {p[1]:.12f} {p[2]:.12f}\n")


def read_centres_xyz(path, norb: int, number_atoms: int = 0, spin_blocks: int = 1) -> np.ndarray:
    """Wannier centres (589-601): count line, comment line, ``X x y z`` per
    Wannier function, then number_atoms atom lines that are skipped.  With
    spin_blocks = 2 (collinear spin) two such blocks follow each other and the
    result is in spinor order, row 2 a + sigma."""
    blocks = []
    with open(path) as f:
        for _ in range(spin_blocks):
            f.readline()
            f.readline()
            blocks.append(np.array([[float(x) for x in f.readline().split()[1:4]]
                                    for _ in range(norb)]))
            for _ in range(number_atoms):
                f.readline()
    if spin_blocks == 1:
        return blocks[0]
    tau = np.zeros((2 * norb, 3))
    tau[0::2] = blocks[0]
    tau[1::2] = blocks[1]
    return tau


def model_from_files(hr_path, centres_path, lattice: CrystalLattice, collinear: bool = False,
                     number_atoms: int = 0, name: str = "from_files") -> TBModel:
    """TB model from the reference's inputs: hr.dat (one block, or two for
    collinear spin) and centres.xyz (one or two blocks), R in lattice units."""
    if collinear:
        r, h = read_hr_dat_collinear(hr_path)
        tau = read_centres_xyz(centres_path, h.shape[1] // 2, number_atoms, spin_blocks=2)
    else:
        r, h = read_hr_dat(hr_path)
        tau = read_centres_xyz(centres_path, h.shape[1], number_atoms)
    return TBModel(lattice, r, h, tau, spinful=collinear, name=name)


def write_centres_xyz_collinear(path, model: TBModel):
    """Two centres blocks (up, then down) of a spinor model."""
    with open(path, "w") as f:
        for sig in (0, 1):
            t = model.tau[sig::2]
            f.write(f"{len(t)}\n written by bse_amd\n")
            for x in t:
                f.write(f"X {x[0]:.12f} {x[1]:.12f} {x[2]:.12f}\n")


def write_k_list(path, crystal_k: np.ndarray):
    """Nk lines of ``k1 k2 k3`` in crystal coordinates (read at 190-216)."""
    np.savetxt(path, np.asarray(crystal_k).T, fmt="%.15f")
