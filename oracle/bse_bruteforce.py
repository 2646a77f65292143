"""Loop-literal brute force of the BSE assembly -- TEST INFRASTRUCTURE ONLY.

Follows the reference's own loop structure instead of the factorised W_ij(Q)
form used by ``bse_oracle``, so that agreement between the two pins the
algebra (index order, conjugations, signs, the k - k' + G argument of W):

* pair densities with point-orbital M (function_building_exponential_factor,
  off-diagonal branch, bse.version13.cpp:1484-1499, contracted as in
  pull_values 2103-2117):
    rho_cc(c2,c1,k2,k1,G) = sum_i conj(c_{c2,k2,i}) e^{i(G+k2-k1).tau_i} c_{c1,k1,i}
    rho_vv(v2,v1,k2,k1,G) = same with valence states taken at k - q (2836)
* diagonal model W_GG'(k2-k1) = delta_GG' W(|k2-k1+G|) (2697-2728) with the
  mini-BZ average replacing the single Q + G = 0 term, over the 2D or 3D
  G-list (319-385) and k grid;
* the direct row of 2850-2874 (temporary_vector2), kept in its textbook
  orientation (the reference stores its conjugate through .t(), 2878);
* the exchange of 2796-2830 in its Hermitian form sum_{G != G0} v rho_r rho_s^*.

Pure Python loops: use only for N up to a few tens.
"""
from __future__ import annotations

import numpy as np

from bse_oracle import (C_EV_A, Problem, exciton_momentum, glist, tb_states, w_model, w_q0_average, v_model)


def assemble_bruteforce(p: Problem):
    kc = p.kpoints()
    q = np.asarray(p.q, dtype=np.float64)[:, None]
    ek, vk = tb_states(p, kc)
    eq, vq = tb_states(p, kc - q)
    nk, nc, nv = p.nk, p.nc, p.nv
    cidx = [p.n_occ + c for c in range(nc)]
    vidx = [p.n_occ - 1 - v for v in range(nv)]
    g = glist(p)
    ng = g.shape[1]
    g0 = ng // 2
    tau = p.tau
    wbar = w_q0_average(p, p.r0, p.eps_env)

    def rho(left, right, kl, kr, G, vecs):
        ph = np.exp(1j * (tau @ (g[:, G] + kc[:, kl] - kc[:, kr])))
        return np.sum(vecs[kl][:, left].conj() * ph * vecs[kr][:, right])

    def wmodel(k2, k1, G):
        qv = kc[:, k2] - kc[:, k1] + g[:, G]
        qn = np.linalg.norm(qv)
        return wbar if qn < 1e-12 else w_model(p, qn, p.r0, p.eps_env)

    N = p.n_bse
    H = np.zeros((N, N), dtype=np.complex128)
    factor_w = -p.direct_scale / (p.cell * nk)

    def ridx(c, v, k):
        return c * nv * nk + v * nk + k

    for c1 in range(nc):
        for v1 in range(nv):
            for k1 in range(nk):
                r = ridx(c1, v1, k1)
                for c2 in range(nc):
                    for k2 in range(nk):
                        tmp = np.zeros(ng, dtype=np.complex128)
                        for s in range(ng):
                            # W is diagonal in G, so the g'-sum has one term
                            tmp[s] = np.conj(rho(cidx[c2], cidx[c1], k2, k1, s, vk)) * wmodel(k2, k1, s)
                        for v2 in range(nv):
                            acc = 0.0j
                            for s in range(ng):
                                acc += factor_w * tmp[s] * rho(vidx[v2], vidx[v1], k2, k1, s, vq)
                            H[r, ridx(c2, v2, k2)] = acc

    # exchange, Hermitian form, G0 excluded, pair densities at Q + G
    Q = exciton_momentum(p)
    rho_cv = np.zeros((N, ng), dtype=np.complex128)
    for c in range(nc):
        for v in range(nv):
            for k in range(nk):
                for s in range(ng):
                    ph = np.exp(1j * (tau @ (Q + g[:, s])))
                    rho_cv[ridx(c, v, k), s] = np.sum(vk[k][:, cidx[c]].conj() * ph * vq[k][:, vidx[v]])
    fx = p.exchange_scale * p.f_s / (p.cell * nk)
    for s in range(ng):
        if s == g0:
            continue
        vg = v_model(p, np.linalg.norm(Q + g[:, s]))
        H += fx * vg * np.outer(rho_cv[:, s], rho_cv[:, s].conj())

    for c in range(nc):
        for v in range(nv):
            for k in range(nk):
                r = ridx(c, v, k)
                H[r, r] += (ek[k, cidx[c]] + p.scissor) - eq[k, vidx[v]]
    return H


def assemble_coupling_bruteforce(p: Problem):
    """Coupling block B by the loop structure of add_coupling_part (2929-3040):

    * exchange: factor_v sum_{G != G0} v(G) rho_cv(r, G) rho_cv(s, -G)
      (the v(G) conj(rho_vc) rho_cv product of 2944-2967 with point-orbital
      pair densities, rho_vc(s, G) = conj(rho_cv(s, -G)));
    * direct: factor_w sum_G conj(rho(v' at k', c at k, G)) W(|k'-k+G|)
      rho(c' at k', v at k, G) -- temporary_matrix3 / temporary_vector3 of
      2986-3027, the resonant direct loop with the primed pair reversed.
    """
    kc = p.kpoints()
    q = np.asarray(p.q, dtype=np.float64)[:, None]
    ek, vk = tb_states(p, kc)
    eq, vq = tb_states(p, kc - q)
    nk, nc, nv = p.nk, p.nc, p.nv
    cidx = [p.n_occ + c for c in range(nc)]
    vidx = [p.n_occ - 1 - v for v in range(nv)]
    g = glist(p)
    ng = g.shape[1]
    g0 = ng // 2
    tau = p.tau
    wbar = w_q0_average(p, p.r0, p.eps_env)

    def rho(vl, left, vr, right, kl, kr, gvec):
        ph = np.exp(1j * (tau @ (gvec + kc[:, kl] - kc[:, kr])))
        return np.sum(vl[kl][:, left].conj() * ph * vr[kr][:, right])

    def wmodel(k2, k1, G):
        qv = kc[:, k2] - kc[:, k1] + g[:, G]
        qn = np.linalg.norm(qv)
        return wbar if qn < 1e-12 else w_model(p, qn, p.r0, p.eps_env)

    N = p.n_bse
    B = np.zeros((N, N), dtype=np.complex128)
    factor_w = -p.direct_scale / (p.cell * nk)

    def ridx(c, v, k):
        return c * nv * nk + v * nk + k

    for c1 in range(nc):
        for v1 in range(nv):
            for k1 in range(nk):
                r = ridx(c1, v1, k1)
                for c2 in range(nc):
                    for v2 in range(nv):
                        for k2 in range(nk):
                            acc = 0.0j
                            for s in range(ng):
                                left = np.conj(rho(vq, vidx[v2], vk, cidx[c1], k2, k1, g[:, s]))
                                right = rho(vk, cidx[c2], vq, vidx[v1], k2, k1, g[:, s])
                                acc += factor_w * left * wmodel(k2, k1, s) * right
                            B[r, ridx(c2, v2, k2)] = acc

    fx = p.exchange_scale * p.f_s / (p.cell * nk)
    for c1 in range(nc):
        for v1 in range(nv):
            for k1 in range(nk):
                r = ridx(c1, v1, k1)
                for c2 in range(nc):
                    for v2 in range(nv):
                        for k2 in range(nk):
                            acc = 0.0j
                            for s in range(ng):
                                if s == g0:
                                    continue
                                vg = v_model(p, np.linalg.norm(g[:, s]))
                                a = np.sum(vk[k1][:, cidx[c1]].conj() * np.exp(1j * (tau @ g[:, s]))
                                           * vq[k1][:, vidx[v1]])
                                b = np.sum(vk[k2][:, cidx[c2]].conj() * np.exp(-1j * (tau @ g[:, s]))
                                           * vq[k2][:, vidx[v2]])
                                acc += fx * vg * a * b
                            B[r, ridx(c2, v2, k2)] += acc
    return B


def rpa_w_bruteforce(p: Problem, d):
    """Static RPA W_GG'(Q) by the loop structure of Dielectric_Function::
    pull_values (2214-2303, order_approximation = 1, omega = 0, eta -> 0):

      chi0(i, j) = factor_chi sum_r conj(rho_r(i)) rho_r(j) factor_r   (2250-2283)
      tmp1(i, j) = delta_ij - v_j chi0(i, j);  chi = tmp1^-1 chi0      (2287-2293)
      eps^-1(i, j) = delta_ij + v_i chi(i, j)                           (2294-2300)

    with the transitions r of both kinds, c at k with v at k - Q and v at k
    with c at k - Q, each with its own -1/dE (the reference's -2/dE on the
    first kind is their sum under time reversal), rho_r(G) = sum_a
    conj(c_left,a) e^{i(Q+G).tau_a} c_right,a over the screening window,
    factor_chi = f_s/(cell Nk); then W_GG' = eps^-1_GG' v(|Q+G'|) (the Q + G = 0
    potential is the bare mini-BZ average)."""
    from bse_oracle import _kidx, bare_q0_average, reciprocal, rpa_window
    kc = p.kpoints()
    e, vecs = tb_states(p, kc)
    ki, kj, kl = _kidx(p)
    b = reciprocal(p.bravais)
    qv = (d[0] / p.n1) * b[:, 0] + (d[1] / p.n2) * b[:, 1] + (d[2] / p.n3) * b[:, 2]
    g = glist(p)
    ng = g.shape[1]
    vidx, cidx = rpa_window(p)
    vg = np.zeros(ng)
    for s in range(ng):
        qn = np.linalg.norm(qv + g[:, s])
        vg[s] = bare_q0_average(p) if qn < 1e-12 else v_model(p, qn)
    rows, fac = [], []
    for k in range(p.nk):
        kq = (((ki[k] - d[0]) % p.n1) * p.n2 + ((kj[k] - d[1]) % p.n2)) * p.n3 + ((kl[k] - d[2]) % p.n3)
        for c in cidx:
            for v in vidx:
                for left, kl_, right, kr_, de in (
                        (c, k, v, kq, e[k, c] + p.scissor - e[kq, v]),
                        (v, k, c, kq, e[kq, c] + p.scissor - e[k, v])):
                    fac.append(-1.0 / de)
                    rho = np.zeros(ng, dtype=np.complex128)
                    for s in range(ng):
                        ph = np.exp(1j * (p.tau @ (qv + g[:, s])))
                        rho[s] = np.sum(vecs[kl_][:, left].conj() * ph * vecs[kr_][:, right])
                    rows.append(rho)
    rho = np.array(rows)
    fac = np.array(fac)
    factor_chi = p.f_s / (p.cell * p.nk)
    chi0 = np.zeros((ng, ng), dtype=np.complex128)
    for i in range(ng):
        for j in range(ng):
            chi0[i, j] = factor_chi * np.sum(rho[:, i].conj() * rho[:, j] * fac)
    tmp1 = np.eye(ng) - chi0 * vg[None, :]
    chi = np.linalg.solve(tmp1, chi0)
    epsinv = np.eye(ng) + vg[:, None] * chi
    return epsinv * vg[None, :]
