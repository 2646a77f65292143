"""CPU oracle of the real-space Wannier integrals -- TEST INFRASTRUCTURE ONLY.

Imported by ``tests/`` as the checker of ``bse_wannier_integrals``; never by
the product path.  Restates, in numpy FP64, the reference's grid sums
(bse.version13.cpp):

  A_ij(R, K) = f sum_r w_i(r - R) e^{iK.r} w_j(r)   function_building_real_space_
                                                    wannier_dipole_ij 1503-1533
  D_ij(R)    = f sum_r w_i(r - R) r w_j(r)          ..._small_q 1535-1570
  S_ij(R)    = f sum_r w_i(r - R) w_j(r)            orthogonality test 1165-1180

f = Omega_supercell / N_points (1506); the translated function takes its grid
value from cell c - R, zero outside the supercell (indexingi/j/k 1441-1456).

Pins (tests/test_wannier.py): the vectorised form equals a loop-literal
restatement of the reference's loops on small grids, and for Gaussian
functions it approaches the closed form below as the grid is refined.
"""
from __future__ import annotations

import numpy as np


def _cells(n_sc):
    n0, n1, n2 = n_sc
    return np.stack(np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij"), -1).reshape(-1, 3)


def _positions(n_sc, p, origin, axes):
    c = _cells(n_sc).astype(float)
    s = np.stack(np.meshgrid(np.arange(p[0]) / p[0], np.arange(p[1]) / p[1], np.arange(p[2]) / p[2],
                             indexing="ij"), -1).reshape(-1, 3)
    u = c[:, None, :] + s[None, :, :]
    return np.asarray(origin, dtype=float)[None, None, :] + u @ np.asarray(axes, dtype=float).T


def integrals(values, nw, nspin, n_sc, p, origin, axes, r_cells, kvec):
    """Vectorised A (nspin, nw, nw, nR, nK), D (..., 3), S (nspin, nw, nw, nR)."""
    axes = np.asarray(axes, dtype=float)
    ncell, ppc = int(np.prod(n_sc)), int(np.prod(p))
    f = abs(np.linalg.det(axes)) / ppc
    x = _positions(n_sc, p, origin, axes)                         # (ncell, ppc, 3)
    kv = np.asarray(kvec, dtype=float).reshape(-1, 3)
    ph = np.exp(1j * np.einsum("cpa,qa->cpq", x, kv))             # (ncell, ppc, nK)
    cells = _cells(n_sc)
    r_cells = np.asarray(r_cells, dtype=int).reshape(-1, 3)
    nr = r_cells.shape[0]
    A = np.zeros((nspin, nw, nw, nr, kv.shape[0]), dtype=np.complex128)
    D = np.zeros((nspin, nw, nw, nr, 3))
    S = np.zeros((nspin, nw, nw, nr))
    w = np.asarray(values, dtype=float).reshape(nspin * nw, ncell, ppc)
    for r, R in enumerate(r_cells):
        t = cells - R[None, :]
        ok = np.all((t >= 0) & (t < np.asarray(n_sc)[None, :]), axis=1)
        tidx = (t[:, 0] * n_sc[1] + t[:, 1]) * n_sc[2] + t[:, 2]
        for s in range(nspin):
            wi = np.zeros((nw, ncell, ppc))
            wi[:, ok, :] = w[s * nw:(s + 1) * nw][:, tidx[ok], :]
            wj = w[s * nw:(s + 1) * nw]
            prod = np.einsum("icp,jcp->ijcp", wi, wj)
            S[s, :, :, r] = f * prod.sum(axis=(2, 3))
            D[s, :, :, r, :] = f * np.einsum("ijcp,cpa->ija", prod, x)
            A[s, :, :, r, :] = f * np.einsum("ijcp,cpq->ijq", prod, ph)
    return A, D, S


def integrals_loops(values, nw, nspin, n_sc, p, origin, axes, r_cells, kvec):
    """Loop-literal restatement of the reference's integral loops (1513-1532,
    1549-1568): translations outer, then spin, supercell cells with the
    translated cell index, then the points of a cell.  Small grids only."""
    axes = np.asarray(axes, dtype=float)
    origin = np.asarray(origin, dtype=float)
    ppc = p[0] * p[1] * p[2]
    f = abs(np.linalg.det(axes)) / ppc
    w = np.asarray(values, dtype=float).reshape(nspin * nw, n_sc[0] * n_sc[1] * n_sc[2], ppc)
    kv = np.asarray(kvec, dtype=float).reshape(-1, 3)
    r_cells = np.asarray(r_cells, dtype=int).reshape(-1, 3)
    nr = r_cells.shape[0]
    A = np.zeros((nspin, nw, nw, nr, kv.shape[0]), dtype=np.complex128)
    D = np.zeros((nspin, nw, nw, nr, 3))
    S = np.zeros((nspin, nw, nw, nr))
    for w1 in range(nw):
        for w2 in range(nw):
            for r in range(nr):
                R = r_cells[r]
                for spin in range(nspin):
                    for i2 in range(n_sc[0]):
                        for j2 in range(n_sc[1]):
                            for k2 in range(n_sc[2]):
                                ii, jj, kk = i2 - R[0], j2 - R[1], k2 - R[2]      # indexingi(i1, i2)
                                if not (0 <= ii < n_sc[0] and 0 <= jj < n_sc[1] and 0 <= kk < n_sc[2]):
                                    continue
                                c_t = (ii * n_sc[1] + jj) * n_sc[2] + kk
                                c_2 = (i2 * n_sc[1] + j2) * n_sc[2] + k2
                                for s2 in range(p[0]):
                                    for t2 in range(p[1]):
                                        for l2 in range(p[2]):
                                            pt = (s2 * p[1] + t2) * p[2] + l2
                                            pos = (origin + (i2 + s2 / p[0]) * axes[:, 0]
                                                   + (j2 + t2 / p[1]) * axes[:, 1]
                                                   + (k2 + l2 / p[2]) * axes[:, 2])
                                            prod = (w[spin * nw + w1, c_t, pt] * w[spin * nw + w2, c_2, pt])
                                            S[spin, w1, w2, r] += f * prod
                                            D[spin, w1, w2, r] += f * prod * pos
                                            for q in range(kv.shape[0]):
                                                ex = float(kv[q] @ pos)
                                                A[spin, w1, w2, r, q] += f * prod * complex(np.cos(ex), np.sin(ex))
    return A, D, S


def gaussian_pair(a, b, sigma, kvec):
    """Closed form for normalised 3D Gaussians g_x(r) = (pi sigma^2)^(-3/4)
    exp(-|r - x|^2 / (2 sigma^2)): int g_a g_b e^{iK.r} dr =
    exp(-|a - b|^2 / (4 sigma^2)) exp(iK.m - sigma^2 |K|^2 / 4), m = (a + b)/2;
    the moment int g_a g_b r dr = exp(-|a - b|^2 / (4 sigma^2)) m.
    Returns (A (nK,), D (3,), S)."""
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    m = 0.5 * (a + b)
    ov = np.exp(-np.sum((a - b) ** 2) / (4.0 * sigma ** 2))
    kv = np.asarray(kvec, dtype=float).reshape(-1, 3)
    A = ov * np.exp(1j * kv @ m - 0.25 * sigma ** 2 * np.sum(kv ** 2, axis=1))
    return A, ov * m, ov
