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
