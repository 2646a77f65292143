// K3, static RPA screening of the direct term (Dielectric_Function::
// pull_values, bse.version13.cpp:2214-2303, order_approximation = 1 at
// omega = 0; the W build of pull_coulomb_potentials, 2575-2584), in place
// of the model W, on every grid difference Q of the W_ij(Q) table.
//
// With point orbitals the pair density of a transition (left state at k,
// right state at k - Q) is rho(Q + G) = sum_i P_i e^{i (Q + G).tau_i},
// P_i = conj(c_{left,k,i}) c_{right,k-Q,i}, so the independent-particle
// response factors through an norb x norb matrix per Q:
//   Pi_ij(Q) = (f_s/(cell Nk)) sum_{k,c,v} [f1 conj(P1_i) P1_j + f2 conj(P2_i) P2_j]
//   P1: c at k, v at k - Q,  f1 = -1/(e_c(k) - e_v(k-Q));
//   P2: v at k, c at k - Q,  f2 = -1/(e_c(k-Q) - e_v(k));
//   chi0_GG'(Q) = sum_ij e^{-i(Q+G).tau_i} Pi_ij e^{i(Q+G').tau_j}
// (the static factors of 1/(w - dE + i eta) and -1/(w + dE - i eta) at w = 0,
// eta -> 0; the reference writes the first kind with -2/dE, the two kinds
// being equal under time reversal; summing them separately keeps chi0
// reciprocal and the BSE matrix Hermitian without it).  States at k - Q are
// the grid point k - Q folded into the grid, TB coefficients being periodic
// in k.  Then, per Q, in LDS:
//   chi = (1 - chi0 V)^-1 chi0 (Gaussian elimination, partial pivoting),
//   W = V + V chi V  (= eps^-1 V),  V = diag v(|Q + G|),
//   W_ij(Q) = sum_GG' e^{i(Q+G).tau_i} W_GG' e^{-i(Q+G').tau_j},
// the Q + G = 0 entry of V being the mini-BZ average of the bare potential
// (chi0's head and wings vanish there).  Every sum runs in a fixed order.
#include "bse_common.hpp"

#define RPA_MAXG 48     // G vectors held in LDS
#define RPA_MAXO 16     // orbitals
#define RPA_CHUNK 32    // transitions staged per step

struct RpaArgs {
    int n1, n2, n3, nk, norb, ng, nvs, ncs, dim;
    double b[9];          // reciprocal vectors b1 | b2 | b3
    double pref;          // f_s / (cell Nk)
    double vbar;          // mini-BZ average of the bare potential
    double radius, cc;
};

__device__ inline void rpa_decode(const RpaArgs& a, int64_t dd, int& d1, int& d2, int& d3) {
    const int w2 = 2 * a.n2 - 1, w3 = 2 * a.n3 - 1;
    d3 = (int)(dd % w3) - (a.n3 - 1);
    d2 = (int)((dd / w3) % w2) - (a.n2 - 1);
    d1 = (int)(dd / ((int64_t)w3 * w2)) - (a.n1 - 1);
}

__device__ inline int rpa_fold(int k, int d1, int d2, int d3, int n1, int n2, int n3) {
    const int l = k % n3, j = (k / n3) % n2, i = k / (n3 * n2);
    const int i2 = ((i - d1) % n1 + n1) % n1, j2 = ((j - d2) % n2 + n2) % n2, l2 = ((l - d3) % n3 + n3) % n3;
    return (i2 * n2 + j2) * n3 + l2;
}

// Pi(Q) for every grid difference; one workgroup per Q, one thread per (i, j).
// eps, coef: the screening window's states at the grid points (slots
// 0..nvs-1 valence top-down, nvs.. conduction bottom-up + scissor).
__global__ __launch_bounds__(256) void k_rpa_pi(RpaArgs a, const double* __restrict__ eps,
                                                const double2* __restrict__ coef, double2* __restrict__ Pi)
{
    __shared__ double2 sP[RPA_CHUNK][RPA_MAXO + 1];
    __shared__ double sf[RPA_CHUNK];
    int d1, d2, d3;
    rpa_decode(a, blockIdx.x, d1, d2, d3);
    const int no = a.norb, ns = a.nvs + a.ncs, ncv = a.ncs * a.nvs;
    const int tid = threadIdx.x;
    const int i = tid / no, j = tid % no;
    const bool own = tid < no * no;
    double2 acc = cmk(0, 0);
    const int64_t T = (int64_t)a.nk * ncv * 2;     // transition t = 2 (k ncv + cv) + kind
    for (int64_t t0 = 0; t0 < T; t0 += RPA_CHUNK) {
        for (int e = tid; e < RPA_CHUNK * no; e += 256) {
            const int tt = e / no, o = e % no;
            const int64_t t = t0 + tt;
            double2 p = cmk(0, 0);
            if (t < T) {
                const int64_t kcv = t >> 1;
                const int k = (int)(kcv / ncv), cv = (int)(kcv % ncv), c = cv / a.nvs, v = cv % a.nvs;
                const int kq = rpa_fold(k, d1, d2, d3, a.n1, a.n2, a.n3);
                p = (t & 1) ? cmulc(coef[((int64_t)k * ns + v) * no + o], coef[((int64_t)kq * ns + a.nvs + c) * no + o])
                            : cmulc(coef[((int64_t)k * ns + a.nvs + c) * no + o], coef[((int64_t)kq * ns + v) * no + o]);
            }
            sP[tt][o] = p;
        }
        for (int tt = tid; tt < RPA_CHUNK; tt += 256) {
            const int64_t t = t0 + tt;
            double f = 0.0;
            if (t < T) {
                const int64_t kcv = t >> 1;
                const int k = (int)(kcv / ncv), cv = (int)(kcv % ncv), c = cv / a.nvs, v = cv % a.nvs;
                const int kq = rpa_fold(k, d1, d2, d3, a.n1, a.n2, a.n3);
                f = (t & 1) ? -1.0 / (eps[(int64_t)kq * ns + a.nvs + c] - eps[(int64_t)k * ns + v])
                            : -1.0 / (eps[(int64_t)k * ns + a.nvs + c] - eps[(int64_t)kq * ns + v]);
            }
            sf[tt] = f;
        }
        __syncthreads();
        if (own) {
#pragma unroll 8
            for (int tt = 0; tt < RPA_CHUNK; ++tt) {
                const double2 x = cmulc(sP[tt][i], sP[tt][j]);
                acc.x += sf[tt] * x.x;
                acc.y += sf[tt] * x.y;
            }
        }
        __syncthreads();
    }
    if (own) Pi[(int64_t)blockIdx.x * no * no + tid] = cscale(acc, a.pref);
}

// W_ij(Q) from Pi(Q); one workgroup per Q.
__global__ __launch_bounds__(256) void k_rpa_w(RpaArgs a, const double* __restrict__ gvec,
                                               const double* __restrict__ tau, const double2* __restrict__ Pi,
                                               double2* __restrict__ wtab)
{
    __shared__ double2 sPh[RPA_MAXG][RPA_MAXO + 1];   // e^{i (Q+G).tau_i}
    __shared__ double2 sT[RPA_MAXO][RPA_MAXG + 1];    // Pi ph^T, then (W conj(ph))^T
    __shared__ double2 sA[RPA_MAXG][RPA_MAXG + 1];    // 1 - chi0 V, reduced in place
    __shared__ double2 sX[RPA_MAXG][RPA_MAXG + 1];    // chi0 -> chi -> W
    __shared__ double sv[RPA_MAXG];
    __shared__ int spiv;
    const int tid = threadIdx.x;
    const int no = a.norb, ng = a.ng;
    int d1, d2, d3;
    rpa_decode(a, blockIdx.x, d1, d2, d3);
    const double f1 = (double)d1 / a.n1, f2 = (double)d2 / a.n2, f3 = (double)d3 / a.n3;
    const double qx = f1 * a.b[0] + f2 * a.b[3] + f3 * a.b[6];
    const double qy = f1 * a.b[1] + f2 * a.b[4] + f3 * a.b[7];
    const double qz = f1 * a.b[2] + f2 * a.b[5] + f3 * a.b[8];
    // phases and the bare potential
    for (int e = tid; e < ng * no; e += 256) {
        const int g = e / no, i = e % no;
        const double gx = qx + gvec[3 * g], gy = qy + gvec[3 * g + 1], gz = qz + gvec[3 * g + 2];
        double s, c;
        sincos(gx * tau[3 * i] + gy * tau[3 * i + 1] + gz * tau[3 * i + 2], &s, &c);
        sPh[g][i] = cmk(c, s);
    }
    for (int g = tid; g < ng; g += 256) {
        const double gx = qx + gvec[3 * g], gy = qy + gvec[3 * g + 1], gz = qz + gvec[3 * g + 2];
        const double q2 = gx * gx + gy * gy + gz * gz;
        sv[g] = (q2 < 1e-24) ? a.vbar : (a.dim == 3 ? a.cc / (q2 + a.radius) : 0.5 * a.cc / sqrt(q2));
    }
    __syncthreads();
    // T[i][G'] = sum_j Pi_ij ph[G'][j]
    const double2* pq = Pi + (int64_t)blockIdx.x * no * no;
    for (int e = tid; e < no * ng; e += 256) {
        const int i = e / ng, g = e % ng;
        double2 s = cmk(0, 0);
        for (int j = 0; j < no; ++j) cfma(s, pq[i * no + j], sPh[g][j]);
        sT[i][g] = s;
    }
    __syncthreads();
    // chi0[G][G'] = sum_i conj(ph[G][i]) T[i][G'];  A = 1 - chi0 V
    for (int e = tid; e < ng * ng; e += 256) {
        const int g = e / ng, h = e % ng;
        double2 s = cmk(0, 0);
        for (int i = 0; i < no; ++i) cfmac(s, sPh[g][i], sT[i][h]);
        sX[g][h] = s;
        double2 x = cscale(s, -sv[h]);
        if (g == h) x.x += 1.0;
        sA[g][h] = x;
    }
    __syncthreads();
    // solve A chi = chi0: forward elimination with partial pivoting
    for (int k = 0; k < ng; ++k) {
        if (tid == 0) {
            int p = k;
            double best = cabs2(sA[k][k]);
            for (int r = k + 1; r < ng; ++r) {
                const double m2 = cabs2(sA[r][k]);
                if (m2 > best) { best = m2; p = r; }
            }
            spiv = p;
        }
        __syncthreads();
        const int p = spiv;
        if (p != k) {
            for (int c = tid; c < ng; c += 256) {
                const double2 ta = sA[k][c]; sA[k][c] = sA[p][c]; sA[p][c] = ta;
                const double2 tx = sX[k][c]; sX[k][c] = sX[p][c]; sX[p][c] = tx;
            }
        }
        __syncthreads();
        const double2 piv = sA[k][k];
        const int rows = ng - k - 1;
        for (int e = tid; e < rows * ng; e += 256) {
            const int r = k + 1 + e / ng, c = e % ng;
            const double2 l = cdiv(sA[r][k], piv);
            if (c > k) sA[r][c] = csub(sA[r][c], cmul(l, sA[k][c]));
            sX[r][c] = csub(sX[r][c], cmul(l, sX[k][c]));
        }
        __syncthreads();
    }
    // back substitution, one column of chi per thread
    for (int c = tid; c < ng; c += 256) {
        for (int k = ng - 1; k >= 0; --k) {
            double2 s = sX[k][c];
            for (int j = k + 1; j < ng; ++j) s = csub(s, cmul(sA[k][j], sX[j][c]));
            sX[k][c] = cdiv(s, sA[k][k]);
        }
    }
    __syncthreads();
    // W = V + V chi V
    for (int e = tid; e < ng * ng; e += 256) {
        const int g = e / ng, h = e % ng;
        double2 w = cscale(sX[g][h], sv[g] * sv[h]);
        if (g == h) w.x += sv[g];
        sX[g][h] = w;
    }
    __syncthreads();
    // U[G][j] = sum_G' W[G][G'] conj(ph[G'][j]) (into sT transposed), then
    // W_ij = sum_G ph[G][i] U[G][j]
    for (int e = tid; e < ng * no; e += 256) {
        const int g = e / no, j = e % no;
        double2 s = cmk(0, 0);
        for (int h = 0; h < ng; ++h) cfma(s, sX[g][h], cconj(sPh[h][j]));
        sA[g][j] = s;
    }
    __syncthreads();
    for (int e = tid; e < no * no; e += 256) {
        const int i = e / no, j = e % no;
        double2 s = cmk(0, 0);
        for (int g = 0; g < ng; ++g) cfma(s, sPh[g][i], sA[g][j]);
        wtab[(int64_t)blockIdx.x * no * no + e] = s;
    }
}

extern "C" int64_t bse_rpa_pi_elems(int n1, int n2, int n3, int norb) {
    return ktable_size(n1, n2, n3) * norb * norb;
}

extern "C" hipError_t bse_launch_rpa_table(hipStream_t st, int n1, int n2, int n3, int norb, int ng, int dim,
                                           const double* recip, int nvs, int ncs, double pref, double vbar,
                                           double radius, double cc, const double* eps, const double2* coef,
                                           const double* gvec, const double* tau, double2* Pi, double2* wtab)
{
    if (norb > RPA_MAXO || ng > RPA_MAXG || nvs < 1 || ncs < 1) return hipErrorInvalidValue;
    RpaArgs a;
    a.n1 = n1; a.n2 = n2; a.n3 = n3; a.nk = n1 * n2 * n3; a.norb = norb; a.ng = ng; a.nvs = nvs; a.ncs = ncs;
    a.dim = dim;
    for (int i = 0; i < 9; ++i) a.b[i] = recip[i];
    a.pref = pref; a.vbar = vbar; a.radius = radius; a.cc = cc;
    const unsigned nq = (unsigned)ktable_size(n1, n2, n3);
    hipLaunchKernelGGL(k_rpa_pi, dim3(nq), dim3(256), 0, st, a, eps, coef, Pi);
    hipLaunchKernelGGL(k_rpa_w, dim3(nq), dim3(256), 0, st, a, gvec, tau, (const double2*)Pi, wtab);
    return hipGetLastError();
}
