// K1 + K2: per-k tight-binding Hamiltonian and its FP64 Hermitian eigenpairs.
//
// Replaces Hamiltonian_TB::FFT (bse.version13.cpp:647-705) and
// pull_ks_states / pull_ks_states_subset (706-821): H_lm(k) = sum_R H_lm(R)
// e^{i k.R} (Cartesian R, no orbital centre in the phase), eigenpairs sorted
// ascending with unit columns, then the band window: valence slots v = 0..nv-1
// are bands n_occ-1-v at k - q (top down, 812-813), conduction slots are bands
// n_occ+c at k (bottom up, + scissor, 815-816).
//
// One thread per (k, half): half 0 diagonalises H(k) for the conduction
// states, half 1 diagonalises H(k - q) for the valence states.  The eig is a
// cyclic complex Jacobi method in FP64 (norb <= 16), which is accurate to
// working precision and has no branches that depend on LAPACK-style
// deflation.  Each eigenvector's largest-modulus component is made real and
// positive so outputs are deterministic (every downstream quantity is gauge
// invariant).
#include "bse_common.hpp"

#define TB_MAXORB BSE_MAXORB

// Derivative mode: dh_r != nullptr adds a first-order perturbation dH(k)
// (same Fourier form) and returns d eps and dc (parallel transport gauge).
__global__ __launch_bounds__(64) void k_tb_states(
    int nk, int norb, int nr, const double* __restrict__ rcart,
    const double2* __restrict__ hr, const double* __restrict__ kpts,
    double qx, double qy, double qz, int n_occ, int nv, int nc, double scissor,
    const double2* __restrict__ dhr,   // optional nr x norb x norb perturbation
    double* __restrict__ eps_out,      // nk x (nv+nc)
    double2* __restrict__ coef_out,    // nk x (nv+nc) x norb
    double* __restrict__ deps_out,     // optional nk x (nv+nc)
    double2* __restrict__ dcoef_out)   // optional nk x (nv+nc) x norb
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * nk) return;
    const int k = t >> 1, half = t & 1;
    const int n = norb;
    double kx = kpts[3 * k], ky = kpts[3 * k + 1], kz = kpts[3 * k + 2];
    if (half) { kx -= qx; ky -= qy; kz -= qz; }

    double2 A[TB_MAXORB * TB_MAXORB];
    double2 V[TB_MAXORB * TB_MAXORB];
    for (int i = 0; i < n * n; ++i) { A[i] = cmk(0, 0); V[i] = cmk(0, 0); }
    for (int i = 0; i < n; ++i) V[i * n + i] = cmk(1, 0);
    for (int r = 0; r < nr; ++r) {
        const double ph = kx * rcart[3 * r] + ky * rcart[3 * r + 1] + kz * rcart[3 * r + 2];
        double s, c;
        sincos(ph, &s, &c);
        const double2 e = cmk(c, s);
        const double2* h = hr + (size_t)r * n * n;
        for (int i = 0; i < n * n; ++i) cfma(A[i], h[i], e);   // A[l*n+m] = H_lm
    }
    // Enforce exact Hermiticity of the input (averaging the two triangles).
    for (int l = 0; l < n; ++l) {
        A[l * n + l].y = 0.0;
        for (int m = l + 1; m < n; ++m) {
            double2 a = A[l * n + m], b = A[m * n + l];
            double2 h = cmk(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
            A[l * n + m] = h;
            A[m * n + l] = cconj(h);
        }
    }
    // Cyclic Jacobi sweeps.
    double fro = 0.0;
    for (int i = 0; i < n * n; ++i) fro += cabs2(A[i]);
    for (int sweep = 0; sweep < 40; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += cabs2(A[p * n + q]);
        if (off <= 1e-34 * fro) break;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                const double2 apq = A[p * n + q];
                const double absa = sqrt(cabs2(apq));
                if (absa == 0.0) continue;
                const double app = A[p * n + p].x, aqq = A[q * n + q].x;
                const double2 u = cmk(apq.x / absa, -apq.y / absa);   // e^{-i phi}
                const double theta = (aqq - app) / (2.0 * absa);
                double tt = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                if (theta < 0.0) tt = -tt;
                const double c = 1.0 / sqrt(1.0 + tt * tt), s = tt * c;
                // columns: B = A J, J = [[c, s], [-s u, c u]]
                for (int i = 0; i < n; ++i) {
                    const double2 aip = A[i * n + p], aiq = A[i * n + q];
                    const double2 su = cscale(u, s), cu = cscale(u, c);
                    A[i * n + p] = csub(cscale(aip, c), cmul(aiq, su));
                    A[i * n + q] = cadd(cscale(aip, s), cmul(aiq, cu));
                    const double2 vip = V[i * n + p], viq = V[i * n + q];
                    V[i * n + p] = csub(cscale(vip, c), cmul(viq, su));
                    V[i * n + q] = cadd(cscale(vip, s), cmul(viq, cu));
                }
                // rows: A' = J^H B
                const double2 uc = cconj(u);
                for (int j = 0; j < n; ++j) {
                    const double2 bpj = A[p * n + j], bqj = A[q * n + j];
                    A[p * n + j] = csub(cscale(bpj, c), cmul(bqj, cscale(uc, s)));
                    A[q * n + j] = cadd(cscale(bpj, s), cmul(bqj, cscale(uc, c)));
                }
                A[p * n + q] = cmk(0, 0);
                A[q * n + p] = cmk(0, 0);
                A[p * n + p] = cmk(app - tt * absa, 0.0);
                A[q * n + q] = cmk(aqq + tt * absa, 0.0);
            }
        }
    }
    double ev[TB_MAXORB];
    int ord[TB_MAXORB];
    for (int i = 0; i < n; ++i) { ev[i] = A[i * n + i].x; ord[i] = i; }
    // insertion sort ascending (stable)
    for (int i = 1; i < n; ++i) {
        int o = ord[i];
        double x = ev[o];
        int j = i - 1;
        while (j >= 0 && ev[ord[j]] > x) { ord[j + 1] = ord[j]; --j; }
        ord[j + 1] = o;
    }
    const int ns = nv + nc;
    // Gather the needed bands, normalise, fix gauge.
    const int b_lo = half ? 0 : nv, b_hi = half ? nv : ns;
    for (int slot = b_lo; slot < b_hi; ++slot) {
        const int band = half ? (n_occ - 1 - slot) : (n_occ + (slot - nv));
        const int col = ord[band];
        double nrm = 0.0, best = -1.0;
        int ib = 0;
        for (int i = 0; i < n; ++i) {
            double a2 = cabs2(V[i * n + col]);
            nrm += a2;
            if (a2 > best) { best = a2; ib = i; }
        }
        const double2 piv = V[ib * n + col];
        const double pa = sqrt(cabs2(piv));
        const double inv = 1.0 / sqrt(nrm);
        const double2 g = cmk(piv.x / pa * inv, -piv.y / pa * inv);  // conj(piv)/|piv|/||v||
        double2* out = coef_out + ((size_t)k * ns + slot) * n;
        for (int i = 0; i < n; ++i) out[i] = cmul(V[i * n + col], g);
        eps_out[(size_t)k * ns + slot] = ev[col] + (half ? 0.0 : scissor);

        if (dhr != nullptr) {
            // dH(k) in the eigenbasis: dc_n = sum_{m != n} c_m <m|dH|n>/(e_n - e_m)
            double2 dH[TB_MAXORB * TB_MAXORB];
            for (int i = 0; i < n * n; ++i) dH[i] = cmk(0, 0);
            for (int r = 0; r < nr; ++r) {
                const double ph = kx * rcart[3 * r] + ky * rcart[3 * r + 1] + kz * rcart[3 * r + 2];
                double s, c;
                sincos(ph, &s, &c);
                const double2 e = cmk(c, s);
                const double2* h = dhr + (size_t)r * n * n;
                for (int i = 0; i < n * n; ++i) cfma(dH[i], h[i], e);
            }
            // w = dH c_n (with the gauge-fixed c_n)
            double2 cn[TB_MAXORB], w[TB_MAXORB];
            for (int i = 0; i < n; ++i) cn[i] = cmul(V[i * n + col], g);
            for (int l = 0; l < n; ++l) {
                double2 acc = cmk(0, 0);
                for (int m = 0; m < n; ++m) cfma(acc, dH[l * n + m], cn[m]);
                w[l] = acc;
            }
            double2 den = cmk(0, 0);
            for (int i = 0; i < n; ++i) cfmac(den, cn[i], w[i]);
            deps_out[(size_t)k * ns + slot] = den.x;
            double2 dc[TB_MAXORB];
            for (int i = 0; i < n; ++i) dc[i] = cmk(0, 0);
            for (int mb = 0; mb < n; ++mb) {
                const int cm = ord[mb];
                if (cm == col) continue;
                const double gap = ev[col] - ev[cm];
                if (gap == 0.0) continue;
                double2 proj = cmk(0, 0);                 // <m|dH|n>
                for (int i = 0; i < n; ++i) cfmac(proj, V[i * n + cm], w[i]);
                proj = cscale(proj, 1.0 / gap);
                for (int i = 0; i < n; ++i) cfma(dc[i], V[i * n + cm], proj);
            }
            double2* dout = dcoef_out + ((size_t)k * ns + slot) * n;
            for (int i = 0; i < n; ++i) dout[i] = dc[i];
        }
    }
}

extern "C" hipError_t bse_launch_tb_states(
    hipStream_t st, int nk, int norb, int nr, const double* rcart, const double2* hr,
    const double* kpts, const double* q, int n_occ, int nv, int nc, double scissor,
    const double2* dhr, double* eps, double2* coef, double* deps, double2* dcoef)
{
    const int threads = 64;
    const int blocks = (int)ceil_div64(2LL * nk, threads);
    hipLaunchKernelGGL(k_tb_states, dim3(blocks), dim3(threads), 0, st, nk, norb, nr, rcart,
                       hr, kpts, q[0], q[1], q[2], n_occ, nv, nc, scissor, dhr, eps, coef,
                       deps, dcoef);
    return hipGetLastError();
}
