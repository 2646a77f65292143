"""CPU FP64 oracle for the excitonic hot path -- TEST INFRASTRUCTURE ONLY.

This module is the checker: it is imported by ``tests/``, by the smoke check
and by the CPU-baseline leg of the benchmark, never by the product path.  It
restates, in numpy complex128, the reference's algorithm (bse.version13.cpp)
with point-orbital pair densities:

  geometry ........ Crystal_Lattice 85-124, K_points 183-242, G_points 319-385
  potentials ...... Coulomb_Potential::pull 1273-1309 (3D C/(q^2 + radius));
                    the 2D Rytova-Keldysh / bare 2D models replace its 2D branch
  TB states ....... FFT 647-705, pull_ks_states 706-777, _subset 778-821
                    (numpy.linalg.eigh = LAPACK zheevd stands in for
                    Armadillo eig_sym; ascending, unit columns)
  point-orbital M . function_building_exponential_factor 1459-1501
  W / v tables .... pull_coulomb_potentials 2547-2728 (model W, q=0 average)
  assembly ........ pull_resonant_part_and_rcv 2767-2928 in textbook orientation
                    (the reference stores the complex conjugate of the direct
                    term through .t() at 2878), with the exchange kept after
                    the direct rows are written (the reference's row
                    assignment at 2876 overwrites it) and in its Hermitian form
  eig ............. common_diagonalization 3318-3345 (zheevd, sort, normalise)
  oscillator ...... pull_excitonic_oscillator_force 3486-3502
  spectrum ........ pull_macroscopic_bse_dielectric_function 3525-3657

Index convention of the BSE basis: r = c*nv*Nk + v*Nk + k (k fastest), as in
the energies_diff / rho layout (2004, 2022) and the row index of 2876.

Pins (see tests/test_oracle_*.py): closed-form honeycomb bands, IPA limit,
Hermiticity, exchange invariance under a constant V shift, gauge invariance of
|f_m|^2, trace identity, sum rule, brute-force agreement with the reference's
loop structure (``bse_bruteforce``), mpmath-free small-N exact cases, finite
differences for the Hellmann-Feynman gradients.  No golden vectors exist in the
reference (it ships no tests or data), so nothing else can pin it.
"""
from __future__ import annotations

from dataclasses import dataclass
import numpy as np

C_EV_A = (1.602176634 * 1000.0) / 8.8541878176   # e^2/eps0 in eV*A (line 32)
MINVAL = 1.0e-5                                   # line 28


# ----------------------------------------------------------------------------
# Geometry (restated; lines 85-124, 183-242, 319-385)
# ----------------------------------------------------------------------------

def reciprocal(bravais):
    a = np.asarray(bravais, dtype=np.float64)
    vol = abs(np.linalg.det(a))
    b = np.stack([np.cross(a[:, 1], a[:, 2]), np.cross(a[:, 2], a[:, 0]),
                  np.cross(a[:, 0], a[:, 1])], axis=1)
    return 2.0 * np.pi / vol * b


def kgrid(bravais, n1, n2, shift=(0.0, 0.0), n3=1, shift3=0.0):
    """Cartesian k (3 x nk), k = ((i+s1)/n1) b1 + ((j+s2)/n2) b2 (+ ((l+s3)/n3) b3),
    k index = (i*n2 + j)*n3 + l (K_points 183-242 with the grid written out)."""
    b = reciprocal(bravais)
    i, j, l = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    c3 = (l.ravel() + shift3) / n3 if n3 > 1 or shift3 != 0.0 else np.zeros(l.size)
    c = np.stack([(i.ravel() + shift[0]) / n1, (j.ravel() + shift[1]) / n2, c3])
    return b @ c


def glist2d(bravais, g_cut):
    """Parallelepiped 2D G-list, G = 0 at index Ng//2 (intended form of 352-379)."""
    b = reciprocal(bravais)
    n = [int(g_cut / np.linalg.norm(b[:, d])) for d in range(2)]
    out = []
    for i in range(-n[0], n[0] + 1):
        for j in range(-n[1], n[1] + 1):
            out.append(i * b[:, 0] + j * b[:, 1])
    return np.array(out).T


def glist3d(bravais, g_cut):
    """Parallelepiped 3D G-list G = i b1 + j b2 + l b3, |i| <= int(G_cut/|b_i|),
    lexicographic, G = 0 at index Ng//2 (334-351)."""
    b = reciprocal(bravais)
    n = [int(g_cut / np.linalg.norm(b[:, d])) for d in range(3)]
    out = []
    for i in range(-n[0], n[0] + 1):
        for j in range(-n[1], n[1] + 1):
            for l in range(-n[2], n[2] + 1):
                out.append(i * b[:, 0] + j * b[:, 1] + l * b[:, 2])
    return np.array(out).T


def glist(p):
    return glist3d(p.bravais, p.g_cut) if p.dimension == 3 else glist2d(p.bravais, p.g_cut)


# ----------------------------------------------------------------------------
# Inputs in plain arrays
# ----------------------------------------------------------------------------

@dataclass
class Problem:
    bravais: np.ndarray      # 3x3 columns a_i
    r_cart: np.ndarray       # (nr, 3)
    h_r: np.ndarray          # (nr, norb, norb)
    tau: np.ndarray          # (norb, 3)
    n1: int
    n2: int
    n_occ: int
    nv: int
    nc: int
    r0: float
    eps_env: float
    g_cut: float
    scissor: float = 0.0
    f_s: float = 2.0
    exchange_scale: float = 1.0
    direct_scale: float = 1.0
    q: tuple = (MINVAL, 0.0, 0.0)
    eta: float = 0.05
    omegas: np.ndarray = None
    dimension: int = 2       # 2: Keldysh W and 2D v; 3: C/(eps (q^2 + radius)), C/(q^2 + radius)
    n3: int = 1
    shift: tuple = (0.0, 0.0, 0.0)
    radius: float = 0.0
    screening: str = "model"  # "model": W of the dimension's model; "rpa": static RPA from the TB states
    rpa_nv: int = 0          # screening window (0: every occupied / every empty band)
    rpa_nc: int = 0
    dipole_r: np.ndarray = None   # Wannier position table: Cartesian R (nR, 3) ...
    dipole_d: np.ndarray = None   # ... and D_ij(R) = int w_i(r - R) r w_j(r) (nR, norb, norb, 3)

    @property
    def nk(self):
        return self.n1 * self.n2 * self.n3

    @property
    def cell(self):
        """Measure of the kernel prefactor 1/(cell Nk): the in-plane area for the
        2D potentials, the volume for the 3D ones (2846)."""
        return self.volume if self.dimension == 3 else self.area

    def kpoints(self):
        return kgrid(self.bravais, self.n1, self.n2, self.shift[:2], self.n3, self.shift[2])

    @property
    def norb(self):
        return self.h_r.shape[1]

    @property
    def n_bse(self):
        return self.nk * self.nv * self.nc

    @property
    def area(self):
        a = np.asarray(self.bravais)
        return float(np.linalg.norm(np.cross(a[:, 0], a[:, 1])))

    @property
    def volume(self):
        return float(abs(np.linalg.det(np.asarray(self.bravais))))


def problem_from_system(s) -> Problem:
    """Read plain arrays off a host-side system description (attributes only)."""
    m = s.model
    g = s.grid
    sh = tuple(g.shift) + (0.0,) * (3 - len(g.shift))
    return Problem(bravais=np.array(m.lattice.bravais), r_cart=np.array(m.r_cart),
                   h_r=np.array(m.h_r), tau=np.array(m.tau), n1=g.n1, n2=g.n2,
                   n_occ=s.n_occ, nv=s.nv, nc=s.nc, r0=s.r0, eps_env=s.eps_env,
                   g_cut=s.g_cut, scissor=s.scissor, f_s=s.spin_factor,
                   exchange_scale=s.exchange_scale, direct_scale=s.direct_scale,
                   q=tuple(s.q_opt), eta=s.eta, omegas=s.omegas(),
                   dimension=getattr(s, "dimension", 2), n3=getattr(g, "n3", 1), shift=sh,
                   radius=getattr(s, "radius", 0.0), screening=getattr(s, "screening", "model"),
                   rpa_nv=getattr(s, "rpa_nv", 0), rpa_nc=getattr(s, "rpa_nc", 0),
                   dipole_r=getattr(s, "dipole_r", None), dipole_d=getattr(s, "dipole_d", None))


# ----------------------------------------------------------------------------
# Single-particle states (FFT 647-705; pull_ks_states 706-777; _subset 778-821)
# ----------------------------------------------------------------------------

def fix_gauge(vecs):
    """Make the largest-modulus component of each column real and positive.

    The reference keeps LAPACK's arbitrary phases; every quantity compared is
    gauge invariant, this only makes outputs deterministic.  vecs: (..., n, m).
    """
    idx = np.argmax(np.abs(vecs), axis=-2)
    piv = np.take_along_axis(vecs, idx[..., None, :], axis=-2)
    ph = piv / np.abs(piv)
    return vecs / ph


def hk(p: Problem, kcart):
    ph = np.exp(1j * (p.r_cart @ kcart))             # (nr, nk)
    return np.einsum("rk,rlm->klm", ph, p.h_r)


def tb_states(p: Problem, kcart):
    """Ascending eigenpairs of H(k) (zheevd), unit columns, fixed gauge."""
    e, v = np.linalg.eigh(hk(p, kcart))
    return e, fix_gauge(v)


def band_window(p: Problem, kcart=None):
    """Conduction at k and valence at k - q (1964-1966, 2836).

    Returns eps_c (nk, nc) incl. scissor, C (nk, nc, norb), eps_v (nk, nv),
    V (nk, nv, norb); v = 0 is the highest valence band (812-813), c = 0 the
    lowest conduction band (815-816).
    """
    if kcart is None:
        kcart = p.kpoints()
    q = np.asarray(p.q, dtype=np.float64)[:, None]
    ek, vk = tb_states(p, kcart)
    eq, vq = tb_states(p, kcart - q)
    cidx = p.n_occ + np.arange(p.nc)
    vidx = p.n_occ - 1 - np.arange(p.nv)
    eps_c = ek[:, cidx] + p.scissor
    C = np.transpose(vk[:, :, cidx], (0, 2, 1))
    eps_v = eq[:, vidx]
    V = np.transpose(vq[:, :, vidx], (0, 2, 1))
    return eps_c, C, eps_v, V


# ----------------------------------------------------------------------------
# Potentials and tables
# ----------------------------------------------------------------------------

def keldysh(qn, r0, eps_env):
    """Rytova-Keldysh W(Q) = (C/2)/(eps Q (1 + r0 Q)), eV*A^2."""
    return 0.5 * C_EV_A / (eps_env * qn * (1.0 + r0 * qn))


def keldysh_q0_average(area, nk, r0, eps_env):
    """Mean of W over the disk pi R^2 = (2 pi)^2/(A Nk) around Q = 0."""
    da = (2.0 * np.pi) ** 2 / (area * nk)
    rad = np.sqrt(da / np.pi)
    if r0 == 0.0:
        return (np.pi * C_EV_A / eps_env) * rad / da
    return (np.pi * C_EV_A / eps_env) * np.log1p(r0 * rad) / (r0 * da)


def bare2d(gn):
    return 0.5 * C_EV_A / gn


def coulomb3d(q2, radius, eps_env=1.0):
    """C / (eps (|q|^2 + radius)), the reference's 3D potential (1281-1285)
    screened by a static dielectric constant (eV A^3)."""
    return C_EV_A / (eps_env * (q2 + radius))


def coulomb3d_q0_average(volume, nk, radius, eps_env):
    """Mean of C/(eps (q^2 + radius)) over the ball of volume (2 pi)^3/(Omega Nk)."""
    vs = (2.0 * np.pi) ** 3 / (volume * nk)
    rad = (3.0 * vs / (4.0 * np.pi)) ** (1.0 / 3.0)
    integ = rad - np.sqrt(radius) * np.arctan(rad / np.sqrt(radius)) if radius > 0 else rad
    return 4.0 * np.pi * C_EV_A * integ / (eps_env * vs)


def w_model(p, qn, r0, eps_env):
    """Screened W(|q|) of the problem's dimension (q != 0)."""
    if p.dimension == 3:
        return coulomb3d(qn * qn, p.radius, eps_env)
    return keldysh(qn, r0, eps_env)


def w_q0_average(p, r0, eps_env):
    if p.dimension == 3:
        return coulomb3d_q0_average(p.volume, p.nk, p.radius, eps_env)
    return keldysh_q0_average(p.area, p.nk, r0, eps_env)


def v_model(p, gn):
    """Bare exchange potential v(|G|), G != 0."""
    if p.dimension == 3:
        return coulomb3d(gn * gn, p.radius)
    return bare2d(gn)


def w_table(p: Problem, r0=None, eps_env=None):
    """W_ij(Q) on every grid difference Q = (d1/n1) b1 + (d2/n2) b2 + (d3/n3) b3.

    Returns (2n1-1, 2n2-1, 2n3-1, norb, norb) indexed [d1 + n1-1, d2 + n2-1,
    d3 + n3-1] (n3 = 1 in 2D).  W_ij(Q) = sum_G W(|Q+G|) e^{i(Q+G).(tau_i -
    tau_j)}; the single term with Q + G = 0 takes the mini-BZ average.
    """
    r0 = p.r0 if r0 is None else r0
    eps_env = p.eps_env if eps_env is None else eps_env
    b = reciprocal(p.bravais)
    g = glist(p)
    d1 = np.arange(-(p.n1 - 1), p.n1)
    d2 = np.arange(-(p.n2 - 1), p.n2)
    d3 = np.arange(-(p.n3 - 1), p.n3)
    dt = p.tau[:, None, :] - p.tau[None, :, :]          # (norb, norb, 3)
    wbar = w_q0_average(p, r0, eps_env)
    out = np.zeros((d1.size, d2.size, d3.size, p.norb, p.norb), dtype=np.complex128)
    for a, x in enumerate(d1):
        for c, y in enumerate(d2):
            for e, z in enumerate(d3):
                qv = (x / p.n1) * b[:, 0] + (y / p.n2) * b[:, 1] + (z / p.n3) * b[:, 2]
                qg = qv[:, None] + g                        # (3, ng)
                qn = np.linalg.norm(qg, axis=0)
                zero = qn < 1e-12
                wq = np.where(zero, wbar, w_model(p, np.where(zero, 1.0, qn), r0, eps_env))
                ph = np.exp(1j * np.einsum("xg,ijx->gij", qg, dt))
                out[a, c, e] = np.einsum("g,gij->ij", wq, ph)
    return out


def bare_q0_average(p: Problem):
    """Mini-BZ average of the bare potential around q = 0 (the 2D bare v is
    the Keldysh form at r0 = 0, eps = 1; in 3D eps = 1)."""
    if p.dimension == 3:
        return coulomb3d_q0_average(p.volume, p.nk, p.radius, 1.0)
    return keldysh_q0_average(p.area, p.nk, 0.0, 1.0)


def rpa_window(p: Problem):
    """Screening bands: the rpa_nv highest occupied and rpa_nc lowest empty
    (0: all of them)."""
    nvs = p.n_occ if p.rpa_nv <= 0 else min(p.rpa_nv, p.n_occ)
    ncs = (p.norb - p.n_occ) if p.rpa_nc <= 0 else min(p.rpa_nc, p.norb - p.n_occ)
    vidx = p.n_occ - 1 - np.arange(nvs)
    cidx = p.n_occ + np.arange(ncs)
    return vidx, cidx


def rpa_pi(p: Problem, d, states=None):
    """Point-orbital factor of the static chi0 at the grid difference d:

      Pi_ij(Q) = (f_s/(cell Nk)) sum_{k,c,v} [ f1 conj(P1_i) P1_j + f2 conj(P2_i) P2_j ],
      P1_i = conj(c_{c,k,i}) c_{v,k-Q,i},  f1 = -1/(e_c(k) - e_v(k-Q)),
      P2_i = conj(c_{v,k,i}) c_{c,k-Q,i},  f2 = -1/(e_c(k-Q) - e_v(k)),

    so that chi0_GG'(Q) = sum_ij e^{-i(Q+G).tau_i} Pi_ij e^{i(Q+G').tau_j}.
    The reference (Dielectric_Function::pull_values 2214-2255) keeps only the
    first kind of transition, with the static factor 1/(w - dE + i eta) -
    1/(w + dE - i eta) = -2/dE at w = 0, eta -> 0: the two kinds are equal
    under time reversal, and summing them separately keeps chi0 reciprocal
    (chi0_GG'(-Q) = chi0_{-G',-G}(Q)) and the BSE matrix Hermitian without it.
    The state at k - Q is the grid point k - Q folded into the grid (the TB
    coefficients are periodic in k)."""
    kc = p.kpoints()
    e, v = tb_states(p, kc) if states is None else states
    ki, kj, kl = _kidx(p)
    kq = (((ki - d[0]) % p.n1) * p.n2 + ((kj - d[1]) % p.n2)) * p.n3 + ((kl - d[2]) % p.n3)
    vidx, cidx = rpa_window(p)
    ec, ev = e[:, cidx] + p.scissor, e[:, vidx]
    P1 = np.einsum("kic,kiv->kcvi", v[:, :, cidx].conj(), v[kq][:, :, vidx])
    f1 = -1.0 / (ec[:, :, None] - ev[kq][:, None, :])
    P2 = np.einsum("kiv,kic->kcvi", v[:, :, vidx].conj(), v[kq][:, :, cidx])
    f2 = -1.0 / (ec[kq][:, :, None] - ev[:, None, :])
    pi = (np.einsum("kcv,kcvi,kcvj->ij", f1, P1.conj(), P1)
          + np.einsum("kcv,kcvi,kcvj->ij", f2, P2.conj(), P2))
    return (p.f_s / (p.cell * p.nk)) * pi


def rpa_w_matrix(p: Problem, d, states=None, scale=1.0):
    """Static RPA W_GG'(Q) = eps^-1_GG' v(|Q+G'|) = V + V chi V with
    chi = (1 - chi0 V)^-1 chi0 (the Dyson solve of pull_values 2261-2303,
    order_approximation = 1), V = diag v(|Q+G|); the Q + G = 0 entry of V is
    the mini-BZ average of the bare potential (the head and wings of chi0
    vanish there, the states at k and k - 0 being orthogonal).  scale
    multiplies chi0 (1: RPA; 0: the bare potential).  Returns (W, Q+G (3, Ng))."""
    b = reciprocal(p.bravais)
    qv = (d[0] / p.n1) * b[:, 0] + (d[1] / p.n2) * b[:, 1] + (d[2] / p.n3) * b[:, 2]
    qg = qv[:, None] + glist(p)
    qn = np.linalg.norm(qg, axis=0)
    zero = qn < 1e-12
    vg = np.where(zero, bare_q0_average(p), v_model(p, np.where(zero, 1.0, qn)))
    ph = np.exp(1j * (p.tau @ qg))                     # (norb, Ng): e^{i(Q+G).tau_i}
    chi0 = scale * (ph.conj().T @ rpa_pi(p, d, states) @ ph)
    chi = np.linalg.solve(np.eye(vg.size) - chi0 * vg[None, :], chi0)
    W = np.diag(vg) + vg[:, None] * chi * vg[None, :]
    return W, qg


def rpa_w_table(p: Problem, scale=1.0):
    """W_ij(Q) = sum_GG' e^{i(Q+G).tau_i} W_GG'(Q) e^{-i(Q+G').tau_j} on every
    grid difference (the layout of w_table); the reference forms
    v_coulomb_gg(G, G', k-k') from eps^-1 and the potential (2575-2584)."""
    states = tb_states(p, p.kpoints())
    d1 = np.arange(-(p.n1 - 1), p.n1)
    d2 = np.arange(-(p.n2 - 1), p.n2)
    d3 = np.arange(-(p.n3 - 1), p.n3)
    out = np.zeros((d1.size, d2.size, d3.size, p.norb, p.norb), dtype=np.complex128)
    for a, x in enumerate(d1):
        for c, y in enumerate(d2):
            for e, z in enumerate(d3):
                W, qg = rpa_w_matrix(p, (x, y, z), states, scale)
                ph = np.exp(1j * (p.tau @ qg))
                out[a, c, e] = ph @ W @ ph.conj().T
    return out


def interaction_table(p: Problem):
    """The direct term's W_ij(Q): the model W, or the static RPA."""
    return rpa_w_table(p) if p.screening == "rpa" else w_table(p)


OPTICAL_Q = 1e-4   # exciton momenta up to this length are the optical limit q -> 0


def exciton_momentum(p: Problem):
    """The exchange's Q: 0 in the optical limit |q| <= OPTICAL_Q (the
    reference's minval = 1e-5 lies inside it), else the exciton momentum q."""
    q = np.asarray(p.q, dtype=np.float64)
    return np.zeros(3) if np.linalg.norm(q) <= OPTICAL_Q else q


def v_exchange(p: Problem):
    """V_ij = sum_{G != 0} v(|Q + G|) e^{i(Q + G).(tau_i - tau_j)} with the bare
    v: v_coulomb_g = pull(excitonic_momentum + G) (2567) with G0 excluded
    (2811-2816); Q = exciton_momentum(p)."""
    g = glist(p)
    g = np.delete(g, g.shape[1] // 2, axis=1)             # exclude G0 (2811-2816)
    g = g + exciton_momentum(p)[:, None]
    gn = np.linalg.norm(g, axis=0)
    dt = p.tau[:, None, :] - p.tau[None, :, :]
    ph = np.exp(1j * np.einsum("xg,ijx->gij", g, dt))
    return np.einsum("g,gij->ij", v_model(p, gn), ph)


# ----------------------------------------------------------------------------
# Assembly (2767-2928), reference index order r = c*nv*Nk + v*Nk + k
# ----------------------------------------------------------------------------

def _kidx(p: Problem):
    i, j, l = np.meshgrid(np.arange(p.n1), np.arange(p.n2), np.arange(p.n3), indexing="ij")
    return i.ravel(), j.ravel(), l.ravel()


def direct_term(p: Problem, Cl, Cr, Vl, Vr, wt):
    """D[(c,v,k),(c',v',k')] = -(1/(A Nk)) sum_ij X_{cc',i} W_ij(k-k') Y_{vv',j}.

    X_{cc',i} = conj(Cl[k,c,i]) Cr[k',c',i];  Y_{vv',j} = Vl[k,v,j] conj(Vr[k',v',j]).
    The four coefficient arrays are separate arguments so that derivatives can
    substitute one factor at a time.  Returns (nc,nv,nk,nc,nv,nk).
    """
    ki, kj, kl = _kidx(p)
    nk, nc, nv = p.nk, p.nc, p.nv
    alpha = -p.direct_scale / (p.cell * nk)
    out = np.empty((nc, nv, nk, nc, nv, nk), dtype=np.complex128)
    for k in range(nk):
        wk = wt[ki[k] - ki + p.n1 - 1, kj[k] - kj + p.n2 - 1, kl[k] - kl + p.n3 - 1]  # (nk', i, j)
        X = np.einsum("ci,kdi->kcdi", Cl[k].conj(), Cr)              # (k', c, c', i)
        A = np.einsum("kcdi,kij->kcdj", X, wk)                       # (k', c, c', j)
        Y = np.einsum("vj,kwj->kvwj", Vl[k], Vr.conj())              # (k', v, v', j)
        blk = np.einsum("kcdj,kvwj->cvdwk", A, Y)                    # (c, v, c', v', k')
        out[:, :, k] = alpha * blk
    return out


def exchange_factors(p: Problem, C, V):
    """P[(c,v,k), i] = conj(C[k,c,i]) V[k,v,i] (N x norb)."""
    P = np.einsum("kci,kvi->cvki", C.conj(), V)
    return P.reshape(p.n_bse, p.norb)


def assemble(p: Problem, bands=None, wt=None, vx=None):
    """Full TDA matrix H = diag(eps_c(k) - eps_v(k-q)) + D + X (reference order)."""
    eps_c, C, eps_v, V = band_window(p) if bands is None else bands
    wt = interaction_table(p) if wt is None else wt
    vx = v_exchange(p) if vx is None else vx
    N = p.n_bse
    H = direct_term(p, C, C, V, V, wt).reshape(N, N)
    P = exchange_factors(p, C, V)
    H += (p.exchange_scale * p.f_s / (p.cell * p.nk)) * (P @ vx @ P.conj().T)
    de = (eps_c[:, :, None] - eps_v[:, None, :])                     # (k, c, v)
    H[np.diag_indices(N)] += np.transpose(de, (1, 2, 0)).ravel()
    return H


def assemble_reference_quirks(p: Problem, bands=None, wt=None):
    """The matrix the reference's build stores for a spin-less system, with
    two of its defects kept (the reference-quirks mode; the product path
    implements the intended equations only):

    * the direct row is written with '=' over the whole row after the
      exchange was added there (2818-2825 vs 2876-2878): the exchange is
      lost, whatever exchange_scale says;
    * the direct row goes through Armadillo's conjugating .t() (2878): the
      stored block is conj(K^d) (same eigenvalues, conjugated eigenvectors).

    The collinear-spin form of the first defect (only the spin-flip exchange
    blocks survive) belongs to the reference's spin-block layout, which the
    spinor formulation here does not have; spinful problems are refused.
    """
    if p.f_s != 2.0:
        raise ValueError("reference-quirks mode covers spin-less problems (f_s = 2) only")
    eps_c, C, eps_v, V = band_window(p) if bands is None else bands
    wt = interaction_table(p) if wt is None else wt
    N = p.n_bse
    H = np.conj(direct_term(p, C, C, V, V, wt).reshape(N, N))
    de = (eps_c[:, :, None] - eps_v[:, None, :])
    H[np.diag_indices(N)] += np.transpose(de, (1, 2, 0)).ravel()
    return H


def transition_energies(p: Problem, bands=None):
    eps_c, C, eps_v, V = band_window(p) if bands is None else bands
    return np.transpose(eps_c[:, :, None] - eps_v[:, None, :], (1, 2, 0)).ravel()


def dipoles(p: Problem, bands=None, direction=(1.0, 0.0, 0.0)):
    """rho_r(G0) = i sum_i conj(c_{ck,i}) (tau_i . qhat) c_{v,k-q,i}.

    Point-orbital small-q limit of the reference's dipole route (1535-1570,
    1614-1633): M_ii = i tau_i . qhat, independent of G.  With a Wannier
    position table (p.dipole_r, p.dipole_d) the small-q M of the real-space
    route: rho = i sum_ij conj(c_{ck,i}) Dq_ij(k) c_{v,k-q,j}, Dq_ij(k) =
    sum_R e^{-ik.R} qhat.D_ij(R) (the Bloch sum of <w_iR| r |w_j0>).
    """
    eps_c, C, eps_v, V = band_window(p) if bands is None else bands
    qh = np.asarray(direction, dtype=np.float64)
    qh = qh / np.linalg.norm(qh)
    if p.dipole_d is not None:
        Rc = np.asarray(p.dipole_r, dtype=float).reshape(-1, 3)
        Dr = np.asarray(p.dipole_d, dtype=float) @ qh                     # (nR, norb, norb)
        ph = np.exp(-1j * p.kpoints().T @ Rc.T)                            # (nk, nR)
        Dq = np.einsum("kr,rij->kij", ph, Dr)
        rho = 1j * np.einsum("kci,kij,kvj->cvk", C.conj(), Dq, V)
        return rho.ravel()
    x = p.tau @ qh
    rho = 1j * np.einsum("kci,i,kvi->cvk", C.conj(), x, V)
    return rho.ravel()


# ----------------------------------------------------------------------------
# Eig, oscillator strengths, spectrum (3318-3345, 3486-3502, 3621-3634)
# ----------------------------------------------------------------------------

def eigh(H):
    """Ascending eigenvalues, unit eigenvectors (LAPACK zheevd via numpy)."""
    e, z = np.linalg.eigh(H)
    return e, z


def oscillator(Z, rho0):
    """f_m = sum_r conj(rho_r) Z_rm (3490-3493)."""
    return Z.T @ rho0.conj()


def dielectric(E, f, omegas, eta, prefactor):
    """eps_xx(w) = 1 - pref * sum_m |f_m|^2 / (w - E_m + i eta) (3621-3634)."""
    w = np.asarray(omegas)[:, None]
    s = (np.abs(f) ** 2)[None, :] / (w - E[None, :] + 1j * eta)
    return 1.0 - prefactor * s.sum(axis=1)


def solve(p: Problem, omegas=None):
    """Whole path; returns dict with H, E, Z, f, eps."""
    bands = band_window(p)
    H = assemble(p, bands)
    E, Z = eigh(H)
    rho0 = dipoles(p, bands)
    f = oscillator(Z, rho0)
    om = p.omegas if omegas is None else omegas
    eps = dielectric(E, f, om, p.eta, C_EV_A / (p.volume * p.nk)) if om is not None else None
    return dict(H=H, E=E, Z=Z, f=f, rho0=rho0, eps=eps, bands=bands)


# ----------------------------------------------------------------------------
# Hellmann-Feynman reference gradients (restated analytically)
# ----------------------------------------------------------------------------

def _tb_derivative(p: Problem, kcart, dh_of_k):
    """First-order band derivatives: d eps_n = c_n^H dH c_n,
    dc_n = sum_{m != n} c_m (c_m^H dH c_n)/(eps_n - eps_m) (parallel transport)."""
    e, v = tb_states(p, kcart)
    dh = dh_of_k(kcart)                                   # (nk, norb, norb)
    m = np.einsum("kin,kij,kjm->knm", v.conj(), dh, v)    # c_n^H dH c_m
    de = np.real(np.einsum("knn->kn", m))
    gap = e[:, None, :] - e[:, :, None]                   # [k, m, n] = e_n - e_m
    with np.errstate(divide="ignore", invalid="ignore"):
        coef = np.where(np.abs(gap) > 0, m / gap, 0.0)    # coefficient of c_m in dc_n
    idx = np.arange(p.norb)
    coef[:, idx, idx] = 0.0
    dv = np.einsum("kim,kmn->kin", v, coef)
    return e, v, de, dv


def _window(p, e, v, which):
    if which == "c":
        idx = p.n_occ + np.arange(p.nc)
    else:
        idx = p.n_occ - 1 - np.arange(p.nv)
    return e[:, idx], np.transpose(v[:, :, idx], (0, 2, 1))


def dH_dtheta(p: Problem, name: str):
    """Exact derivative of the assembled matrix (reference order) w.r.t. one
    parameter of the gradient set (hopping_scale, scissor, r0, eps_env,
    exchange_field).  Uses the product rule on the four coefficient factors."""
    N = p.n_bse
    if p.screening == "rpa":
        raise ValueError("the gradient restatement holds W fixed: model screening only")
    if name == "scissor":
        return np.eye(N, dtype=np.complex128)
    bands = band_window(p)
    eps_c, C, eps_v, V = bands
    if name in ("r0", "eps_env"):
        h = 1e-6 * (p.r0 if name == "r0" else p.eps_env)
        kw_p = {name: getattr(p, name) + h}
        kw_m = {name: getattr(p, name) - h}
        # W depends analytically on r0/eps; difference of tables at +-h is
        # exact to O(h^2) and only enters the (linear) direct term.
        dwt = (w_table(p, **kw_p) - w_table(p, **kw_m)) / (2 * h)
        return direct_term(p, C, C, V, V, dwt).reshape(N, N)
    kc = p.kpoints()
    q = np.asarray(p.q, dtype=np.float64)[:, None]
    zero = np.all(np.abs(p.r_cart) < 1e-12, axis=1)
    if name == "hopping_scale":
        def dh_of_k(kk):
            ph = np.exp(1j * (p.r_cart[~zero] @ kk))
            return np.einsum("rk,rlm->klm", ph, p.h_r[~zero])
    elif name == "exchange_field":
        sz = np.where(np.arange(p.norb) % 2 == 0, 1.0, -1.0)

        def dh_of_k(kk):
            return np.broadcast_to(np.diag(sz).astype(complex), (kk.shape[1], p.norb, p.norb))
    else:
        raise ValueError(name)
    ek, vk, dek, dvk = _tb_derivative(p, kc, dh_of_k)
    eq, vq, deq, dvq = _tb_derivative(p, kc - q, dh_of_k)
    dec, dC = _window(p, dek, dvk, "c")
    dev, dV = _window(p, deq, dvq, "v")
    wt = w_table(p)
    vx = v_exchange(p)
    dD = (direct_term(p, dC, C, V, V, wt) + direct_term(p, C, dC, V, V, wt)
          + direct_term(p, C, C, dV, V, wt) + direct_term(p, C, C, V, dV, wt)).reshape(N, N)
    P = exchange_factors(p, C, V)
    dP = exchange_factors(p, dC, V) + exchange_factors(p, C, dV)
    bx = p.exchange_scale * p.f_s / (p.cell * p.nk)
    dX = bx * (dP @ vx @ P.conj().T + P @ vx @ dP.conj().T)
    dde = np.transpose(dec[:, :, None] - dev[:, None, :], (1, 2, 0)).ravel()
    out = dD + dX
    out[np.diag_indices(N)] += dde
    return out


def hf_energy_gradients(p: Problem, n_exc=8, params=("hopping_scale", "scissor", "r0",
                                                     "eps_env", "exchange_field")):
    """dE_m/dtheta = A_m^H (dH/dtheta) A_m for the lowest n_exc excitons."""
    sol = solve(p, omegas=None)
    Z = sol["Z"][:, :n_exc]
    g = np.zeros((n_exc, len(params)))
    for t, name in enumerate(params):
        dH = dH_dtheta(p, name)
        g[:, t] = np.real(np.einsum("rm,rs,sm->m", Z.conj(), dH, Z))
    return g, sol["E"][:n_exc]


def with_param(p: Problem, name: str, value: float) -> Problem:
    """Copy of the problem with one gradient parameter set (for finite differences)."""
    from dataclasses import replace
    if name in ("r0", "eps_env", "scissor"):
        return replace(p, **{name: value})
    zero = np.all(np.abs(p.r_cart) < 1e-12, axis=1)
    h = p.h_r.copy()
    if name == "hopping_scale":
        base = p.h_r.copy()
        h[~zero] = base[~zero] * value
        return replace(p, h_r=h)
    if name == "exchange_field":
        sz = np.where(np.arange(p.norb) % 2 == 0, 1.0, -1.0)
        i0 = int(np.nonzero(zero)[0][0])
        h[i0] = h[i0] + np.diag(value * sz)
        return replace(p, h_r=h)
    raise ValueError(name)


def spectrum_l2(p: Problem, target, omegas=None):
    """L = sum_w (Im eps(w) - target(w))^2."""
    sol = solve(p, omegas)
    return float(np.sum((np.imag(sol["eps"]) - target) ** 2)), sol


def spectrum_l2_gradient(p: Problem, target, params=("hopping_scale", "scissor", "r0",
                                                     "eps_env", "exchange_field")):
    """dL/dtheta via the resolvent: Im eps = -pref Im(rho^H G(z) rho),
    G = (z - H)^-1 = A diag(1/(z - E)) A^H.  Dense restatement for small N."""
    om = p.omegas
    L, sol = spectrum_l2(p, target, om)
    E, Z, rho = sol["E"], sol["Z"], sol["rho0"]
    pref = C_EV_A / (p.volume * p.nk)
    r = np.imag(sol["eps"]) - target                        # (nw,)
    z = om + 1j * p.eta
    g = 1.0 / (z[:, None] - E[None, :])                     # (nw, N)
    f = Z.T @ rho.conj()                                    # f_m = rho^H A_m
    # K_mn = f_m conj(f_n) sum_w r_w g_m g_n
    K = (f[:, None] * f.conj()[None, :]) * np.einsum("w,wm,wn->mn", r, g, g)
    out = np.zeros(len(params))
    for t, name in enumerate(params):
        dH = dH_dtheta(p, name)
        M = Z.conj().T @ dH @ Z
        term = np.sum(K * M)
        val = term
        if name in ("hopping_scale", "exchange_field"):
            drho = _drho(p, name)
            # d/dtheta of rho^H G rho from rho: drho^H G rho + rho^H G drho
            Gr = Z @ (g.T * f.conj()[:, None])              # (N, nw): G(z_w) rho
            rG = (f[:, None] * g.T)                         # (N_m, nw) coefficients
            t1 = np.einsum("rw,r->w", Gr, drho.conj())      # drho^H G rho
            t2 = np.einsum("mw,rm,r->w", rG, Z.conj(), drho)  # rho^H G drho
            val = term + np.sum(r * (t1 + t2))
        out[t] = -2.0 * pref * np.imag(val)
    return out, L


def _drho(p: Problem, name):
    """Derivative of the dipole vector from the TB coefficient derivatives."""
    kc = p.kpoints()
    q = np.asarray(p.q, dtype=np.float64)[:, None]
    zero = np.all(np.abs(p.r_cart) < 1e-12, axis=1)
    if name == "hopping_scale":
        def dh_of_k(kk):
            ph = np.exp(1j * (p.r_cart[~zero] @ kk))
            return np.einsum("rk,rlm->klm", ph, p.h_r[~zero])
    else:
        sz = np.where(np.arange(p.norb) % 2 == 0, 1.0, -1.0)

        def dh_of_k(kk):
            return np.broadcast_to(np.diag(sz).astype(complex), (kk.shape[1], p.norb, p.norb))
    ek, vk, dek, dvk = _tb_derivative(p, kc, dh_of_k)
    eq, vq, deq, dvq = _tb_derivative(p, kc - q, dh_of_k)
    _, C = _window(p, ek, vk, "c")
    _, V = _window(p, eq, vq, "v")
    _, dC = _window(p, dek, dvk, "c")
    _, dV = _window(p, deq, dvq, "v")
    x = p.tau[:, 0]
    d = 1j * (np.einsum("kci,i,kvi->cvk", dC.conj(), x, V)
              + np.einsum("kci,i,kvi->cvk", C.conj(), x, dV))
    return d.ravel()


# ----------------------------------------------------------------------------
# Beyond Tamm-Dancoff: coupling block, full BSE, spectrum
# (add_coupling_part 2929-3040; eig_gen 3326; pull_excitonic_oscillator_force
# 3494-3500; pull_augmentation_matrix 3504-3522; non-TDA sum 3607-3620)
# ----------------------------------------------------------------------------

def coupling_direct_term(p: Problem, C, V, wt):
    """Bd[(c,v,k),(c',v',k')] = -(1/(A Nk)) sum_ij X'_{cv',i} W_ij(k-k') Y'_{vc',j}

    X'_{cv',i} = conj(C[k,c,i]) V[k',v',i];  Y'_{vc',j} = V[k,v,j] conj(C[k',c',j]):
    the direct term with the primed conduction and valence states exchanged
    (the temporary_matrix3 / temporary_vector3 contraction of 2986-3027 with
    point-orbital pair densities).  Returns (nc,nv,nk,nc,nv,nk)."""
    ki, kj, kl = _kidx(p)
    nk, nc, nv = p.nk, p.nc, p.nv
    alpha = -p.direct_scale / (p.cell * nk)
    out = np.empty((nc, nv, nk, nc, nv, nk), dtype=np.complex128)
    for k in range(nk):
        wk = wt[ki[k] - ki + p.n1 - 1, kj[k] - kj + p.n2 - 1, kl[k] - kl + p.n3 - 1]  # (k', i, j)
        X = np.einsum("ci,kwi->kcwi", C[k].conj(), V)               # (k', c, v', i)
        A = np.einsum("kcwi,kij->kcwj", X, wk)                       # (k', c, v', j)
        Y = np.einsum("vj,kdj->kvdj", V[k], C.conj())                # (k', v, c', j)
        blk = np.einsum("kcwj,kvdj->cvdwk", A, Y)                    # (c, v, c', v', k')
        out[:, :, k] = alpha * blk
    return out


def assemble_coupling(p: Problem, bands=None, wt=None, vx=None):
    """Coupling block B = Bd + (f_s/(A Nk)) P V P^T (reference order).

    The exchange half is the reference's factor_v * v(G) * conj(rho_vc) rho_cv
    sum over G != G0 (2937-2967) with point-orbital pair densities, which for
    the Hermitian form of the resonant exchange (P V P^H) is P V P^T: both
    blocks are complex symmetric, B = B^T."""
    if np.any(exciton_momentum(p) != 0):
        raise ValueError("the coupling block is formed in the optical limit |q| <= OPTICAL_Q")
    eps_c, C, eps_v, V = band_window(p) if bands is None else bands
    wt = interaction_table(p) if wt is None else wt
    vx = v_exchange(p) if vx is None else vx
    N = p.n_bse
    B = coupling_direct_term(p, C, V, wt).reshape(N, N)
    P = exchange_factors(p, C, V)
    B += (p.exchange_scale * p.f_s / (p.cell * p.nk)) * (P @ vx @ P.T)
    return B


def full_hamiltonian(p: Problem, bands=None):
    """H = [[A, B], [-conj(B), -conj(A)]] (2N x 2N): the resonant block, the
    coupling block and the anti-resonant rows filled as -conj of the transpose
    (the "Conjugating HBSE" loop of 3034-3037 with B = B^T)."""
    bands = band_window(p) if bands is None else bands
    A = assemble(p, bands)
    B = assemble_coupling(p, bands)
    H = np.block([[A, B], [-B.conj(), -A.conj()]])
    return A, B, H


def full_dipoles(p: Problem, bands=None, direction=(1.0, 0.0, 0.0)):
    """rho~ = (rho_cv; rho_vc): the anti-resonant half holds the v <- c pair
    densities (rho_q_diagk_cv rows N..2N, 2951), rho_vc = -conj(rho_cv) in the
    point-orbital dipole limit (i tau.qhat is anti-Hermitian)."""
    rho = dipoles(p, bands, direction)
    return np.concatenate([rho, -rho.conj()])


def eig_full(H):
    """Eigenpairs of the full BSE matrix through a general eigensolver (the
    reference's eig_gen, 3326), sorted by real part (3334-3345).  Returns
    (E, X) with E (2N,) real parts, X unit columns, and the largest imaginary
    part as a diagnostic."""
    e, x = np.linalg.eig(H)
    order = np.argsort(e.real, kind="stable")
    e, x = e[order], x[:, order]
    x = x / np.linalg.norm(x, axis=0)[None, :]
    return e.real, x, float(np.abs(e.imag).max())


def spectrum_augmented(E, X, rho_t, omegas, eta, prefactor):
    """eps(w) = 1 - pref sum_{l,l'} (rho~^H X_l) S^-1_{ll'} (X_l'^H s3 rho~) / (w - E_l + i eta)
    over all 2N eigenpairs, S = X^H X the overlap ("augmentation") matrix of
    the non-orthogonal eigenvectors (pull_augmentation_matrix 3504-3522).  This
    is the resolvent rho~^H (w - H)^-1 s3 rho~ written in the eigenbasis of H,
    which the reference's sum at 3607-3620 approximates with the N-vector
    overlap and without the metric s3."""
    n2 = X.shape[0]
    s3 = np.concatenate([np.ones(n2 // 2), -np.ones(n2 // 2)])
    S = X.conj().T @ X
    left = X.T @ rho_t.conj()                          # rho~^H X_l
    right = np.linalg.solve(S, X.conj().T @ (s3 * rho_t))
    w = np.asarray(omegas)[:, None]
    s = (left * right)[None, :] / (w - E[None, :] + 1j * eta)
    return 1.0 - prefactor * s.sum(axis=1)


def spectrum_resolvent(H, rho_t, omegas, eta, prefactor):
    """eps(w) = 1 - pref rho~^H (w + i eta - H)^-1 s3 rho~ by dense solves
    (independent of any eigendecomposition; small N only)."""
    n2 = H.shape[0]
    s3 = np.concatenate([np.ones(n2 // 2), -np.ones(n2 // 2)])
    out = np.empty(len(omegas), dtype=np.complex128)
    for t, w in enumerate(omegas):
        y = np.linalg.solve((w + 1j * eta) * np.eye(n2) - H, s3 * rho_t)
        out[t] = 1.0 - prefactor * (rho_t.conj() @ y)
    return out


def dielectric_full(E, f, omegas, eta, prefactor):
    """eps(w) = 1 - pref sum_{l>0} |f_l|^2 [1/(w - E_l + i eta) - 1/(w + E_l + i eta)]
    over the positive branch, f_l = rho~^H X_l with X_l^H s3 X_l = 1 (the
    negative branch pairs with (conj y; conj x) and conj f_l)."""
    w = np.asarray(omegas)[:, None]
    a = (np.abs(f) ** 2)[None, :]
    s = a / (w - E[None, :] + 1j * eta) - a / (w + E[None, :] + 1j * eta)
    return 1.0 - prefactor * s.sum(axis=1)


def solve_full_structured(A, B, rho_t):
    """Structure-preserving solve of [[A, B], [-conj B, -conj A]] for a
    positive definite M = [[A, B], [conj B, conj A]] (the stability condition
    of the BSE): H = s3 M, M = L L^H (Cholesky), W = L^H s3 L Hermitian with
    the eigenvalues of H; positive branch E_l, y_l; X_l = sqrt(E_l) L^-H y_l
    (X^H s3 X = 1); f_l = rho~^H X_l = sqrt(E_l) (L^-1 rho~)^H y_l
    (Shao, da Jornada, Yang, Deslippe & Lin 2016, cited at 3371)."""
    N = A.shape[0]
    M = np.block([[A, B], [B.conj(), A.conj()]])
    L = np.linalg.cholesky(M)
    s3 = np.concatenate([np.ones(N), -np.ones(N)])
    W = L.conj().T @ (s3[:, None] * L)
    W = 0.5 * (W + W.conj().T)
    lam, Y = np.linalg.eigh(W)
    pos = lam > 0
    E, Yp = lam[pos], Y[:, pos]
    z = np.linalg.solve(L, rho_t)
    f = np.sqrt(E) * (Yp.T @ z.conj())
    X = np.linalg.solve(L.conj().T, Yp) * np.sqrt(E)[None, :]
    return E, X, f


def solve_full(p: Problem, omegas=None):
    """Full BSE (no Tamm-Dancoff): positive excitation energies E (N,),
    eigenvectors X (2N x N, X^H s3 X = 1), f = rho~^H X, and the spectrum by the
    augmentation form over the general eigensolver's 2N eigenpairs."""
    bands = band_window(p)
    A, B, H = full_hamiltonian(p, bands)
    rho_t = full_dipoles(p, bands)
    E2, X2, im = eig_full(H)
    N = p.n_bse
    Ep, Xp = E2[N:], X2[:, N:]
    s3 = np.concatenate([np.ones(N), -np.ones(N)])
    nrm = np.real(np.einsum("rl,r,rl->l", Xp.conj(), s3, Xp))
    if im > 1e-9 * np.abs(E2).max() or np.any(nrm <= 0):
        raise ValueError("unstable full BSE: [[A, B], [conj B, conj A]] is not positive definite")
    Xn = Xp / np.sqrt(nrm)[None, :]
    f = Xn.T @ rho_t.conj()
    om = p.omegas if omegas is None else omegas
    pref = C_EV_A / (p.volume * p.nk)
    eps = spectrum_augmented(E2, X2, rho_t, om, p.eta, pref) if om is not None else None
    return dict(A=A, B=B, H=H, E=Ep, X=Xn, f=f, rho_t=rho_t, eps=eps, E_all=E2, max_imag=im)
