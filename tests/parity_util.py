"""Gauge-invariant comparisons shared by the GPU parity tests.

Eigenvectors of (near-)degenerate excitons are defined only up to a unitary
mixing inside their cluster, and every TB state carries an arbitrary phase,
so eigenvectors are compared as subspaces: for each cluster of eigenvalues
closer than ``tol`` the singular values of Z_gpu[:, C]^H Z_ref[:, C] are the
cosines of the principal angles; 1 - min(sigma) is the cluster's overlap
error.  The TB gauge is fixed identically on both sides (largest component
real positive), so the subspaces are directly comparable.
"""
import numpy as np


def clusters(E, tol):
    """Index ranges [i, j) of eigenvalues whose neighbours are closer than tol."""
    out, i, n = [], 0, len(E)
    while i < n:
        j = i + 1
        while j < n and E[j] - E[j - 1] < tol:
            j += 1
        out.append((i, j))
        i = j
    return out


def cluster_overlap_error(Z, Zref, E, tol, ncol=None):
    """max over clusters fully inside the first ncol columns of 1 - min sigma."""
    ncol = Z.shape[1] if ncol is None else ncol
    worst = 0.0
    for i, j in clusters(E[:Zref.shape[1]], tol):
        if j > ncol:
            break
        s = np.linalg.svd(Z[:, i:j].conj().T @ Zref[:, i:j], compute_uv=False)
        worst = max(worst, 1.0 - float(s.min()))
    return worst


def cluster_f2_error(E, f2, f2ref, tol):
    """max over clusters of |sum f2 - sum f2ref|, relative to sum f2ref."""
    tot = float(np.sum(f2ref))
    worst = 0.0
    for i, j in clusters(E, tol):
        worst = max(worst, abs(float(f2[i:j].sum() - f2ref[i:j].sum())) / tot)
    return worst
