"""Real-space Wannier ingredients: xsf grids, their pair integrals and the
pair-density tables built from them (the reference's real-space route,
bse.version13.cpp: ``Real_space_wannier`` 908-1212, ``Dipole_Elements::
function_building_real_space_wannier_dipole_ij`` 1503-1533, ``_small_q``
1535-1570, ``function_building_M_k1k2_ij`` 1572-1919).

Grid layout.  A supercell of ``n_sc[a]`` unit cells along each lattice vector
a_a carries ``p[a]`` points per cell; point (cell c, sub-point s) lies at
``origin + sum_a (c_a + s_a / p_a) a_a``.  ``values[f, cell, point]`` with
cell = (c0 n1 + c1) n2 + c2 and point = (s0 p1 + s1) p2 + s2 is the
reference's ``real_space_wannier_functions_list`` (rows = sub-points,
columns = (spin, function, cell), 1134), function f = spin * nw + w.

The pair integrals run on the GPU (``Context.wannier_integrals``); this
module reads and writes the files, normalises the functions as the reference
does, and contracts the integrals into the tables the BSE uses:

* ``m_table``: M_ij(k; K) = sum_R e^{-i k.R} A_ij(R, K), the pair density
  rho_nm(k, K) = sum_ij conj(c_nk,i) M_ij(k; K) c_m,k-q,j of Bloch states
  |nk> = N^-1/2 sum_R sum_i c_nk,i e^{ik.R} |w_iR> (the convention of
  H(k) = sum_R H(R) e^{ik.R}).  With w_i concentrated at tau_i it reduces to
  the point-orbital M_ij = delta_ij e^{iK.tau_i} (1459-1501).
* ``position_table``: D_ij(R) = int w_i(r - R) r w_j(r), the small-q dipole
  of the real-space route (the BSE's ``dipole_d``); with point orbitals it is
  delta_ij delta_R0 tau_i.

Two differences from the reference, both deliberate: the phase of the
diagonal-k M multiplies a Cartesian vector (the reference adds the integer
cell index ``origin_unitcell`` to a Cartesian translation, 1597, 1636), and
the k dependence of M is kept (the reference's diagonal-k M uses the exciton
momentum only).
"""
from __future__ import annotations

from dataclasses import dataclass
import os
import numpy as np

XSF_HEADER_LINES = 22      # lines the reference skips before the grid sizes (1116-1118)


@dataclass
class WannierGrid:
    nw: int                  # functions per spin block
    nspin: int               # 1, or 2 (collinear: up block, then down block)
    n_sc: tuple              # unit cells of the supercell along a1, a2, a3
    p: tuple                 # grid points per unit cell along a1, a2, a3
    origin: np.ndarray       # Cartesian position of grid point 0
    axes: np.ndarray         # (3, 3), column a = unit-cell vector a_a
    values: np.ndarray       # (nspin * nw, ncell, ppc)

    @property
    def ncell(self) -> int:
        return int(np.prod(self.n_sc))

    @property
    def ppc(self) -> int:
        return int(np.prod(self.p))

    @property
    def cell_volume(self) -> float:
        return float(abs(np.linalg.det(self.axes)))

    @property
    def weight(self) -> float:
        """f = Omega_supercell / N_points (1506)."""
        return self.ncell * self.cell_volume / (self.ncell * self.ppc)

    def positions(self) -> np.ndarray:
        """(ncell, ppc, 3) Cartesian positions of the grid points."""
        n0, n1, n2 = self.n_sc
        p0, p1, p2 = self.p
        c = np.stack(np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij"), -1).reshape(-1, 3)
        s = np.stack(np.meshgrid(np.arange(p0) / p0, np.arange(p1) / p1, np.arange(p2) / p2, indexing="ij"),
                     -1).reshape(-1, 3)
        u = c[:, None, :] + s[None, :, :]
        return self.origin[None, None, :] + u @ self.axes.T

    def home_cell(self) -> tuple:
        """The supercell cell whose corner lies nearest the Cartesian origin
        (origin_unitcell, 1190-1210)."""
        best, arg = None, (0, 0, 0)
        for i in range(self.n_sc[0]):
            for j in range(self.n_sc[1]):
                for k in range(self.n_sc[2]):
                    d = np.linalg.norm(self.origin + self.axes @ np.array([i, j, k], dtype=float))
                    if best is None or d < best:
                        best, arg = d, (i, j, k)
        return arg


# ----------------------------------------------------------------------------
# xsf files (one function per file, Real_space_wannier ctor 1081-1163)
# ----------------------------------------------------------------------------

def read_xsf(path):
    """(n[3], origin, spanning vectors (3, 3) as columns, values (n0*n1*n2,)
    with x fastest) of the 3D data grid in an xsf file.  The reference skips
    22 header lines (1116-1118); this reader finds the BEGIN_DATAGRID_3D line,
    which gives the same numbers on those files and on other header lengths."""
    with open(path) as f:
        lines = f.read().split("\n")
    start = None
    for i, ln in enumerate(lines):
        if ln.strip().upper().startswith("BEGIN_DATAGRID_3D"):
            start = i + 1
            break
    if start is None:
        start = XSF_HEADER_LINES
    toks = " ".join(lines[start:]).split()
    n = [int(float(t)) for t in toks[:3]]
    origin = np.array([float(t) for t in toks[3:6]])
    span = np.array([float(t) for t in toks[6:15]]).reshape(3, 3).T      # columns
    total = n[0] * n[1] * n[2]
    vals = np.array([float(t) for t in toks[15:15 + total]])
    if vals.size != total:
        raise ValueError(f"{path}: {vals.size} grid values, expected {total}")
    return n, origin, span, vals


def write_xsf(path, values_xyz: np.ndarray, origin, span, comment="written by bse_amd wannier.write_xsf"):
    """Write one 3D grid (values_xyz[x, y, z]) with exactly 22 header lines, so
    the reference's reader (skip 22 lines, 1116) reads it too; span: the
    supercell's spanning vectors as columns."""
    nx, ny, nz = values_xyz.shape
    span = np.asarray(span, dtype=float)
    hdr = [f" # {comment}", " #", " #", " #", "", "CRYSTAL", "PRIMVEC"]
    hdr += ["  " + " ".join(f"{x:.10f}" for x in span[:, a]) for a in range(3)]
    hdr += ["CONVEC"]
    hdr += ["  " + " ".join(f"{x:.10f}" for x in span[:, a]) for a in range(3)]
    hdr += ["PRIMCOORD", "1 1", "X " + " ".join(f"{x:.10f}" for x in origin), "", "",
            "BEGIN_BLOCK_DATAGRID_3D", "3D_field", "BEGIN_DATAGRID_3D_UNKNOWN"]
    assert len(hdr) == XSF_HEADER_LINES
    flat = np.transpose(values_xyz, (2, 1, 0)).ravel()      # x fastest
    with open(path, "w") as f:
        f.write("\n".join(hdr) + "\n")
        f.write(f"{nx} {ny} {nz}\n")
        f.write(" ".join(f"{x:.10f}" for x in origin) + "\n")
        for a in range(3):
            f.write(" ".join(f"{x:.10f}" for x in span[:, a]) + "\n")
        for i in range(0, flat.size, 6):
            f.write(" ".join(f"{x:.10e}" for x in flat[i:i + 6]) + "\n")
        f.write("END_DATAGRID_3D\nEND_BLOCK_DATAGRID_3D\n")


def xsf_name(seedname, w, spin=0, spinful=False):
    """seedname_00001.xsf, or seednameup_/seednamedown_ for the two collinear
    blocks (1100-1108)."""
    tag = "_" if not spinful else ("up_" if spin == 0 else "down_")
    return f"{seedname}{tag}{w + 1:05d}.xsf"


def _grid_to_cells(vals, n_sc, p):
    """Grid values (x fastest) -> (ncell, ppc) in the reference's layout."""
    n = [n_sc[a] * p[a] for a in range(3)]
    v = vals.reshape(n[2], n[1], n[0])                      # [z][y][x]
    v = v.reshape(n_sc[2], p[2], n_sc[1], p[1], n_sc[0], p[0])
    # -> [c0][c1][c2][s0][s1][s2]
    v = np.transpose(v, (4, 2, 0, 5, 3, 1))
    return v.reshape(int(np.prod(n_sc)), int(np.prod(p)))


def _cells_to_grid(vals, n_sc, p):
    """Inverse of _grid_to_cells: (ncell, ppc) -> values_xyz[x, y, z]."""
    v = vals.reshape(n_sc[0], n_sc[1], n_sc[2], p[0], p[1], p[2])
    v = np.transpose(v, (0, 3, 1, 4, 2, 5))
    return v.reshape(n_sc[0] * p[0], n_sc[1] * p[1], n_sc[2] * p[2])


def normalise(grid: WannierGrid) -> np.ndarray:
    """Scale every function to f sum |w|^2 = 1 (1141-1163); returns the norms
    before scaling."""
    norms = grid.weight * np.einsum("fcp,fcp->f", grid.values, grid.values)
    grid.values /= np.sqrt(norms)[:, None, None]
    return norms


def read_wannier_grid(seedname, nw, n_sc, spinful=False, directory=".", normalize=True) -> WannierGrid:
    """The reference's Real_space_wannier: nw (x 2 collinear spin blocks) xsf
    files, the supercell's spanning vectors divided by n_sc (1184-1187),
    each function normalised over the supercell."""
    nspin = 2 if spinful else 1
    funcs, n0, org, span = [], None, None, None
    for s in range(nspin):
        for w in range(nw):
            n, origin, sp, vals = read_xsf(os.path.join(directory, xsf_name(seedname, w, s, spinful)))
            if n0 is None:
                n0, org, span = n, origin, sp
            elif n != n0:
                raise ValueError("xsf grids of different sizes")
            funcs.append(vals)
    p = tuple(n0[a] // n_sc[a] for a in range(3))
    if any(p[a] * n_sc[a] != n0[a] for a in range(3)):
        raise ValueError(f"grid {n0} is not n_sc x points per cell")
    vals = np.stack([_grid_to_cells(v, n_sc, p) for v in funcs])
    g = WannierGrid(nw=nw, nspin=nspin, n_sc=tuple(n_sc), p=p, origin=np.asarray(org, dtype=float),
                    axes=span / np.asarray(n_sc, dtype=float)[None, :], values=vals)
    if normalize:
        normalise(g)
    return g


def write_wannier_grid(grid: WannierGrid, seedname, directory="."):
    """One xsf file per function, in the reference's names."""
    spinful = grid.nspin == 2
    span = grid.axes * np.asarray(grid.n_sc, dtype=float)[None, :]
    for s in range(grid.nspin):
        for w in range(grid.nw):
            v = _cells_to_grid(grid.values[s * grid.nw + w], grid.n_sc, grid.p)
            write_xsf(os.path.join(directory, xsf_name(seedname, w, s, spinful)), v, grid.origin, span)


def gaussian_grid(lattice_bravais, centres, n_sc=(3, 3, 1), p=(16, 16, 16), sigma=0.5, nspin=1,
                  c_extent=None) -> WannierGrid:
    """Normalised Gaussians exp(-|r - tau_i|^2 / (2 sigma^2)) on a supercell
    whose middle cell is the home cell (its corner at the Cartesian origin);
    centres (nw, 3) Cartesian.  c_extent (optional): length of a3 for 2D
    lattices whose third vector is a vacuum spacing (the grid then spans
    [-c/2, c/2) along a3 with one cell)."""
    axes = np.array(lattice_bravais, dtype=float).reshape(3, 3)
    if c_extent is not None:
        axes = axes.copy()
        axes[:, 2] = axes[:, 2] / np.linalg.norm(axes[:, 2]) * c_extent
    home = np.array([n // 2 for n in n_sc], dtype=float)
    origin = -axes @ home
    if c_extent is not None and n_sc[2] == 1:
        origin = origin - 0.5 * axes[:, 2]
    g = WannierGrid(nw=len(centres), nspin=nspin, n_sc=tuple(n_sc), p=tuple(p), origin=origin, axes=axes,
                    values=np.zeros((nspin * len(centres), int(np.prod(n_sc)), int(np.prod(p)))))
    x = g.positions()
    for s in range(nspin):
        for i, tau in enumerate(np.asarray(centres, dtype=float)):
            d2 = np.sum((x - tau[None, None, :]) ** 2, axis=-1)
            g.values[s * g.nw + i] = np.exp(-0.5 * d2 / sigma ** 2)
    normalise(g)
    return g


# ----------------------------------------------------------------------------
# Tables for the BSE (M_k1k2 1572-1919 in the diagonal-k, small-q form)
# ----------------------------------------------------------------------------

def translations(nmax=(1, 1, 0)) -> np.ndarray:
    """Cell translations R = (i, j, k), |i| <= nmax[0], ... (the reference's
    number_primitive_cells_integration box centred at 0, 1441-1449)."""
    r = [(i, j, k) for i in range(-nmax[0], nmax[0] + 1) for j in range(-nmax[1], nmax[1] + 1)
         for k in range(-nmax[2], nmax[2] + 1)]
    return np.array(r, dtype=np.int32)


def position_table(ctx, grid: WannierGrid, r_cells, spinor_order=True):
    """(R Cartesian (nR, 3), D (nR, norb, norb, 3)) for the BSE's Wannier
    dipole, from the GPU integrals; norb = nw (one spin block) or 2 nw with
    orbital i = 2 a + sigma (the collinear model's order) and no spin-flip
    elements."""
    r_cells = np.asarray(r_cells, dtype=np.int32).reshape(-1, 3)
    _, D, _ = ctx.wannier_integrals(grid, r_cells)
    nw, ns, nr = grid.nw, grid.nspin, r_cells.shape[0]
    Rc = r_cells.astype(float) @ grid.axes.T
    if ns == 1:
        return Rc, np.ascontiguousarray(np.transpose(D[0], (2, 0, 1, 3)))
    out = np.zeros((nr, 2 * nw, 2 * nw, 3))
    for s in range(2):
        blk = np.transpose(D[s], (2, 0, 1, 3))
        if spinor_order:
            out[:, s::2, s::2, :] = blk
        else:
            out[:, s * nw:(s + 1) * nw, s * nw:(s + 1) * nw, :] = blk
    return Rc, out


def m_table(A: np.ndarray, r_cells, axes, kpts) -> np.ndarray:
    """M_ij(k; K) = sum_R e^{-i k.R} A_ij(R, K) for one spin block:
    A (nw, nw, nR, nK) -> (nk, nw, nw, nK)."""
    Rc = np.asarray(r_cells, dtype=float).reshape(-1, 3) @ np.asarray(axes, dtype=float).T
    ph = np.exp(-1j * np.asarray(kpts, dtype=float).reshape(-1, 3) @ Rc.T)     # (nk, nR)
    return np.einsum("kr,ijrq->kijq", ph, A)
