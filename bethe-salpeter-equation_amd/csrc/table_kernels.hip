// K3: interaction tables and per-transition factors.
//
// Replaces Excitonic_Hamiltonian::pull_coulomb_potentials
// (bse.version13.cpp:2547-2728) and the pair-density construction of
// Dipole_Elements::pull_values (1922-2158) with point-orbital M
// (function_building_exponential_factor, 1459-1501).  The reference tabulates
// W_GG'(k - k') for all Nk^2 k pairs; on a Monkhorst-Pack grid k - k' takes
// only (2 n1 - 1)(2 n2 - 1) values, so the orbital-resolved table
//     W_ij(Q) = sum_G W(|Q + G|) e^{i (Q + G).(tau_i - tau_j)}
// is built once per distinct Q (the single Q + G = 0 term takes the mini-BZ
// average).  The exchange needs V_ij = sum_{G != 0} v(|G|) e^{i G.(tau_i -
// tau_j)} (G0 excluded as at 2811-2816).
#include "bse_common.hpp"

// Screened interaction W(Q) and its parameter derivatives (mode 0 value,
// 1 d/dr0, 2 d/deps), C = e^2/eps0 in eV*Angstrom:
//   2D: Rytova-Keldysh W(Q) = (C/2) / (eps Q (1 + r0 Q))                 [eV A^2]
//   3D: W(Q) = C / (eps (Q^2 + radius)), the reference's regularised bare
//       Coulomb potential (Coulomb_Potential::pull, 1281-1285) screened by a
//       static dielectric constant as its empirical W = eps^-1 v (2698)   [eV A^3]
__device__ inline double screened_w(int dim, double q2, double r0, double eps, double radius,
                                    double cc, int mode) {
    if (dim == 3) {
        if (mode == 1) return 0.0;
        const double w = cc / (eps * (q2 + radius));
        return mode == 0 ? w : -w / eps;
    }
    const double qn = sqrt(q2), ch = 0.5 * cc;
    const double den = eps * qn * (1.0 + r0 * qn);
    if (mode == 0) return ch / den;
    if (mode == 1) return -ch / (eps * (1.0 + r0 * qn) * (1.0 + r0 * qn));  // d/dr0
    return -ch / (eps * den);                                             // d/deps
}

// Bare exchange potential v(G): 2D (C/2)/|G|, 3D C/(|G|^2 + radius).
__device__ inline double bare_v(int dim, double g2, double radius, double cc) {
    return dim == 3 ? cc / (g2 + radius) : 0.5 * cc / sqrt(g2);
}

struct TableGeom {
    int n1, n2, n3, dim;
    double b[9];         // reciprocal vectors b1 | b2 | b3 (Cartesian, 1/Angstrom)
    double r0, eps, radius, cc, wbar;
    int mode;
};

// Table entry (grid difference dd, i, j): dd enumerates (d1, d2, d3) in the
// mixed radix of ktable_index, d = digit - (n - 1).
__global__ __launch_bounds__(256) void k_w_table(TableGeom tg, int norb, int ng,
                                                 const double* __restrict__ gvec /* 3 x ng */,
                                                 const double* __restrict__ tau /* norb x 3 */,
                                                 double2* __restrict__ wtab)
{
    const int64_t total = ktable_size(tg.n1, tg.n2, tg.n3) * norb * norb;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int j = (int)(t % norb);
    const int i = (int)((t / norb) % norb);
    const int64_t dd = t / ((int64_t)norb * norb);
    const int w2 = 2 * tg.n2 - 1, w3 = 2 * tg.n3 - 1;
    const int d3i = (int)(dd % w3);
    const int d2i = (int)((dd / w3) % w2);
    const int d1i = (int)(dd / ((int64_t)w3 * w2));
    const double f1 = (double)(d1i - (tg.n1 - 1)) / tg.n1, f2 = (double)(d2i - (tg.n2 - 1)) / tg.n2,
                 f3 = (double)(d3i - (tg.n3 - 1)) / tg.n3;
    const double qx = f1 * tg.b[0] + f2 * tg.b[3] + f3 * tg.b[6];
    const double qy = f1 * tg.b[1] + f2 * tg.b[4] + f3 * tg.b[7];
    const double qz = f1 * tg.b[2] + f2 * tg.b[5] + f3 * tg.b[8];
    const double tx = tau[3 * i] - tau[3 * j], ty = tau[3 * i + 1] - tau[3 * j + 1],
                 tz = tau[3 * i + 2] - tau[3 * j + 2];
    double2 acc = cmk(0, 0);
    for (int g = 0; g < ng; ++g) {
        const double gx = qx + gvec[3 * g], gy = qy + gvec[3 * g + 1], gz = qz + gvec[3 * g + 2];
        const double q2 = gx * gx + gy * gy + gz * gz;
        const double w = (q2 < 1e-24) ? tg.wbar
                                      : screened_w(tg.dim, q2, tg.r0, tg.eps, tg.radius, tg.cc, tg.mode);
        double s, c;
        sincos(gx * tx + gy * ty + gz * tz, &s, &c);
        acc.x += w * c;
        acc.y += w * s;
    }
    wtab[t] = acc;
}

// V_ij (exchange), sum over G != G0 of v(G) e^{i G.(tau_i - tau_j)}.
// V_ij = sum_{G != G0} v(|Q + G|) e^{i (Q + G).(tau_i - tau_j)}: Q = 0 in the
// optical limit, else the exciton momentum (v_coulomb_g of
// pull_coulomb_potentials, 2567, with G0 excluded as at 2811-2816).
__global__ void k_v_exchange(int norb, int ng, int dim, const double* __restrict__ gvec,
                             const double* __restrict__ tau, double cc, double radius,
                             double qx, double qy, double qz, double2* __restrict__ vx)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= norb * norb) return;
    const int i = t / norb, j = t % norb;
    const double tx = tau[3 * i] - tau[3 * j], ty = tau[3 * i + 1] - tau[3 * j + 1],
                 tz = tau[3 * i + 2] - tau[3 * j + 2];
    double2 acc = cmk(0, 0);
    for (int g = 0; g < ng; ++g) {
        if (g == ng / 2) continue;
        const double gx = qx + gvec[3 * g], gy = qy + gvec[3 * g + 1], gz = qz + gvec[3 * g + 2];
        const double v = bare_v(dim, gx * gx + gy * gy + gz * gz, radius, cc);
        double s, c;
        sincos(gx * tx + gy * ty + gz * tz, &s, &c);
        acc.x += v * c;
        acc.y += v * s;
    }
    vx[t] = acc;
}

// Per-transition factors in the k-major internal order r' = k*B + c*nv + v:
//   P[r', i]  = conj(c_{ck,i}) c_{v,k-q,i}            (exchange factor)
//   PV[r', j] = sum_i P[r', i] V_ij
//   rho[r']   = i sum_i conj(c_{ck,i}) (tau_i . qhat) c_{v,k-q,i}
//   dE[r']    = eps_c(k) - eps_v(k - q)
// With dcoef != nullptr the derivative factors (product rule) are formed
// instead: dP = conj(dc_c) c_v + conj(c_c) dc_v, d rho likewise, d dE.
__global__ __launch_bounds__(256) void k_transitions(
    int nk, int nv, int nc, int norb, const double* __restrict__ eps,
    const double2* __restrict__ coef, const double* __restrict__ deps,
    const double2* __restrict__ dcoef, const double* __restrict__ tau, double qhx,
    double qhy, double qhz, const double2* __restrict__ vx, double2* __restrict__ P,
    double2* __restrict__ PV, double2* __restrict__ rho, double* __restrict__ dE,
    const double2* __restrict__ dq)
{
    const int B = nc * nv, ns = nv + nc;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (int64_t)nk * B) return;
    const int k = (int)(r / B), b = (int)(r % B);
    const int c = b / nv, v = b % nv;
    const double2* cc = coef + ((size_t)k * ns + nv + c) * norb;
    const double2* cv = coef + ((size_t)k * ns + v) * norb;
    double2 p[BSE_MAXORB];
    double2 rh = cmk(0, 0);
    for (int i = 0; i < norb; ++i) {
        double2 pi;
        if (dcoef == nullptr) {
            pi = cmulc(cc[i], cv[i]);
        } else {
            const double2* dcc = dcoef + ((size_t)k * ns + nv + c) * norb;
            const double2* dcv = dcoef + ((size_t)k * ns + v) * norb;
            pi = cadd(cmulc(dcc[i], cv[i]), cmulc(cc[i], dcv[i]));
        }
        p[i] = pi;
        P[r * norb + i] = pi;
        const double x = tau[3 * i] * qhx + tau[3 * i + 1] * qhy + tau[3 * i + 2] * qhz;
        rh = cadd(rh, cscale(pi, x));
    }
    if (dq) {
        // Wannier position matrix: rh = sum_ij conj(c_c,i) Dq_ij(k) c_v,j
        // (with derivatives: the product rule over the two coefficient sets)
        const double2* D = dq + (size_t)k * norb * norb;
        const double2* dcc = dcoef ? dcoef + ((size_t)k * ns + nv + c) * norb : nullptr;
        const double2* dcv = dcoef ? dcoef + ((size_t)k * ns + v) * norb : nullptr;
        rh = cmk(0, 0);
        for (int i = 0; i < norb; ++i) {
            double2 t = cmk(0, 0), td = cmk(0, 0);
            for (int j = 0; j < norb; ++j) {
                const double2 d = D[i * norb + j];
                cfma(t, d, dcoef ? dcv[j] : cv[j]);
                if (dcoef) cfma(td, d, cv[j]);
            }
            rh = cadd(rh, cmulc(cc[i], t));
            if (dcoef) rh = cadd(rh, cmulc(dcc[i], td));
        }
    }
    rho[r] = cmk(-rh.y, rh.x);   // i * rh
    for (int j = 0; j < norb; ++j) {
        double2 acc = cmk(0, 0);
        for (int i = 0; i < norb; ++i) cfma(acc, p[i], vx[i * norb + j]);
        PV[r * norb + j] = acc;
    }
    const double* e = (dcoef == nullptr) ? eps : deps;
    dE[r] = e[(size_t)k * ns + nv + c] - e[(size_t)k * ns + v];
}

// b: 9 doubles (b1 | b2 | b3); cc = e^2/eps0; the table has
// ktable_size(n1, n2, n3) * norb^2 entries.
extern "C" hipError_t bse_launch_w_table(hipStream_t st, int n1, int n2, int n3, int dim, int norb,
                                         int ng, const double* gvec, const double* b,
                                         const double* tau, double r0, double eps, double radius,
                                         double cc, double wbar, int mode, double2* wtab)
{
    TableGeom tg;
    tg.n1 = n1; tg.n2 = n2; tg.n3 = n3; tg.dim = dim;
    for (int i = 0; i < 9; ++i) tg.b[i] = b[i];
    tg.r0 = r0; tg.eps = eps; tg.radius = radius; tg.cc = cc; tg.wbar = wbar; tg.mode = mode;
    const int64_t total = ktable_size(n1, n2, n3) * norb * norb;
    hipLaunchKernelGGL(k_w_table, dim3((unsigned)ceil_div64(total, 256)), dim3(256), 0, st, tg, norb,
                       ng, gvec, tau, wtab);
    return hipGetLastError();
}

// q: the exciton momentum (3 doubles, host); |q| <= BSE_OPTICAL_Q is the
// optical limit q -> 0.
extern "C" hipError_t bse_launch_v_exchange(hipStream_t st, int norb, int ng, int dim,
                                            const double* gvec, const double* tau, double cc,
                                            double radius, const double* q, double2* vx)
{
    double qx = q[0], qy = q[1], qz = q[2];
    if (qx * qx + qy * qy + qz * qz <= BSE_OPTICAL_Q * BSE_OPTICAL_Q) qx = qy = qz = 0.0;
    hipLaunchKernelGGL(k_v_exchange, dim3(1), dim3(256), 0, st, norb, ng, dim, gvec, tau, cc, radius, qx, qy,
                       qz, vx);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_transitions(hipStream_t st, int nk, int nv, int nc, int norb,
                                             const double* eps, const double2* coef,
                                             const double* deps, const double2* dcoef,
                                             const double* tau, const double* qhat,
                                             const double2* vx, double2* P, double2* PV,
                                             double2* rho, double* dE, const double2* dq)
{
    const int64_t n = (int64_t)nk * nv * nc;
    hipLaunchKernelGGL(k_transitions, dim3((unsigned)ceil_div64(n, 256)), dim3(256), 0, st, nk,
                       nv, nc, norb, eps, coef, deps, dcoef, tau, qhat[0], qhat[1], qhat[2], vx,
                       P, PV, rho, dE, dq);
    return hipGetLastError();
}
