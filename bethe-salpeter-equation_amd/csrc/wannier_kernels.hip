// Real-space Wannier ingredients (next row 3 of the blueprint): integrals of
// pairs of Wannier functions tabulated on the real-space grid of a supercell,
// the reference's Real_space_wannier (bse.version13.cpp:1081-1212) and
// Dipole_Elements::function_building_real_space_wannier_dipole_ij(_small_q)
// (1503-1570):
//   A_ij(R, K) = f sum_r w_i(r - R) e^{i K.r} w_j(r)     (1513-1532)
//   D_ij(R)_a  = f sum_r w_i(r - R) r_a w_j(r)           (1549-1568, the
//                                                          small-q moments)
//   S_ij(R)    = f sum_r w_i(r - R) w_j(r)               (overlaps, 1165-1180)
// with f = Omega_supercell / N_points (1506).  The translated function is
// the same grid value taken from cell c - R (zero when that cell is outside
// the supercell: indexingi/j/k, 1441-1456).  Both spin blocks of a
// collinear model are integrated separately (same-spin pairs only).
//
// Grid: n_sc[a] unit cells along axis a, p[a] points per cell; point
// (cell c, sub-point s) lies at x = origin + sum_a (c_a + s_a / p_a) axis_a.
// Function f's values: w[f * N + cell(c) * P + sub(s)], cell(c) =
// (c0 n1 + c1) n2 + c2, sub(s) = (s0 p1 + s1) p2 + s2, P = p0 p1 p2 (the
// reference's real_space_wannier_functions_list, rows = sub-points,
// columns = (spin, function, cell)).
#include "bse_internal.hpp"

#define WANN_KB 8           // K vectors per workgroup
#define WANN_AXMAX 256      // points along one supercell axis held in the phase tables

struct WannGeom {
    int nsc[3], p[3];
    int ncell, pcell;
    double origin[3], axis[9];   // axis column-major: axis_a = axis[3a .. 3a+2]
    double f;                    // Omega_supercell / N_points
};

// One workgroup per (spin, i, j, R) row and block of WANN_KB K vectors
// (blockIdx.y); the 256 threads stride the grid points.  Phases
// e^{iK.x} = e^{iK.origin} prod_a e^{i (K.axis_a) g_a / p_a} from per-axis
// tables in LDS (g_a the global point index along axis a).  Block y = 0
// also forms the moments and the overlap.
__global__ __launch_bounds__(256) void k_wann_int(WannGeom g, int nw, int nspin, int nR,
                                                  const int* __restrict__ Rc, int nK,
                                                  const double* __restrict__ Kv,
                                                  const double* __restrict__ w, double2* __restrict__ A,
                                                  double* __restrict__ D, double* __restrict__ S)
{
    __shared__ double2 tab[3][WANN_KB][WANN_AXMAX];
    __shared__ double red[32];
    const int row = blockIdx.x;                  // ((spin nw + i) nw + j) nR + r
    const int r = row % nR, ij = row / nR;
    const int j = ij % nw, i = (ij / nw) % nw, spin = ij / (nw * nw);
    const int k0 = blockIdx.y * WANN_KB;
    const int nk = min(WANN_KB, nK - k0);
    const int N = g.ncell * g.pcell;
    const int na[3] = {g.nsc[0] * g.p[0], g.nsc[1] * g.p[1], g.nsc[2] * g.p[2]};
    for (int e = threadIdx.x; e < 3 * WANN_KB * WANN_AXMAX; e += 256) {
        const int a = e / (WANN_KB * WANN_AXMAX), q = (e / WANN_AXMAX) % WANN_KB, m = e % WANN_AXMAX;
        double2 v = cmk(0, 0);
        if (q < nk && m < na[a]) {
            const double* K = Kv + 3 * (k0 + q);
            const double th = (K[0] * g.axis[3 * a] + K[1] * g.axis[3 * a + 1] + K[2] * g.axis[3 * a + 2]) *
                              ((double)m / g.p[a]);
            double sn, cs;
            sincos(th, &sn, &cs);
            v = cmk(cs, sn);
        }
        tab[a][q][m] = v;
    }
    __syncthreads();
    const int R0 = Rc[3 * r], R1 = Rc[3 * r + 1], R2 = Rc[3 * r + 2];
    const double* wi = w + (int64_t)(spin * nw + i) * N;
    const double* wj = w + (int64_t)(spin * nw + j) * N;
    double2 acc[WANN_KB];
#pragma unroll
    for (int q = 0; q < WANN_KB; ++q) acc[q] = cmk(0, 0);
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, s0 = 0.0;
    const bool moments = blockIdx.y == 0;
    for (int e = threadIdx.x; e < N; e += 256) {
        const int c = e / g.pcell, sp = e % g.pcell;
        const int c2 = c % g.nsc[2], c1 = (c / g.nsc[2]) % g.nsc[1], c0 = c / (g.nsc[1] * g.nsc[2]);
        const int t0 = c0 - R0, t1 = c1 - R1, t2 = c2 - R2;     // cell of the translated function
        if (t0 < 0 || t0 >= g.nsc[0] || t1 < 0 || t1 >= g.nsc[1] || t2 < 0 || t2 >= g.nsc[2]) continue;
        const double prod = wi[(int64_t)((t0 * g.nsc[1] + t1) * g.nsc[2] + t2) * g.pcell + sp] * wj[e];
        if (prod == 0.0) continue;
        const int s2 = sp % g.p[2], s1 = (sp / g.p[2]) % g.p[1], s0i = sp / (g.p[1] * g.p[2]);
        const int g0 = c0 * g.p[0] + s0i, g1 = c1 * g.p[1] + s1, g2 = c2 * g.p[2] + s2;
        if (moments) {
            const double u0 = (double)c0 + (double)s0i / g.p[0], u1 = (double)c1 + (double)s1 / g.p[1],
                         u2 = (double)c2 + (double)s2 / g.p[2];
            m0 += prod * (g.origin[0] + u0 * g.axis[0] + u1 * g.axis[3] + u2 * g.axis[6]);
            m1 += prod * (g.origin[1] + u0 * g.axis[1] + u1 * g.axis[4] + u2 * g.axis[7]);
            m2 += prod * (g.origin[2] + u0 * g.axis[2] + u1 * g.axis[5] + u2 * g.axis[8]);
            s0 += prod;
        }
#pragma unroll
        for (int q = 0; q < WANN_KB; ++q) {
            const double2 ph = cmul(cmul(tab[0][q][g0], tab[1][q][g1]), tab[2][q][g2]);
            acc[q] = cadd(acc[q], cscale(ph, prod));
        }
    }
    // block sums in a fixed order; the phase e^{iK.origin} and f applied once
    for (int q = 0; q < nk; ++q) {
        const double re = block_sum(acc[q].x, red);
        const double im = block_sum(acc[q].y, red);
        if (threadIdx.x == 0) {
            const double* K = Kv + 3 * (k0 + q);
            double sn, cs;
            sincos(K[0] * g.origin[0] + K[1] * g.origin[1] + K[2] * g.origin[2], &sn, &cs);
            A[(int64_t)row * nK + k0 + q] = cscale(cmul(cmk(re, im), cmk(cs, sn)), g.f);
        }
    }
    if (moments) {
        const double a0 = block_sum(m0, red), a1 = block_sum(m1, red), a2 = block_sum(m2, red);
        const double as = block_sum(s0, red);
        if (threadIdx.x == 0) {
            D[3 * (int64_t)row] = g.f * a0;
            D[3 * (int64_t)row + 1] = g.f * a1;
            D[3 * (int64_t)row + 2] = g.f * a2;
            S[row] = g.f * as;
        }
    }
}

// Position table projected on the dipole direction at every k of the grid:
// Dq_ij(k) = sum_R e^{-i k.R} qhat.D_ij(R), the Bloch sum of the Wannier
// position matrix (point orbitals: delta_ij qhat.tau_i).  One thread per
// (k, i, j).  kc: nk x 3 Cartesian k; Rcart: nR x 3 Cartesian R.
__global__ __launch_bounds__(256) void k_dipole_table(int nk, int norb, int nR, const double* __restrict__ kc,
                                                      const double* __restrict__ Rcart,
                                                      const double* __restrict__ Dr, double qx, double qy,
                                                      double qz, double2* __restrict__ Dq)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)nk * norb * norb) return;
    const int k = (int)(e / (norb * norb)), ij = (int)(e % (norb * norb));
    double2 s = cmk(0, 0);
    for (int r = 0; r < nR; ++r) {
        const double* d = Dr + ((int64_t)r * norb * norb + ij) * 3;
        const double x = qx * d[0] + qy * d[1] + qz * d[2];
        const double ph = -(kc[3 * k] * Rcart[3 * r] + kc[3 * k + 1] * Rcart[3 * r + 1] + kc[3 * k + 2] * Rcart[3 * r + 2]);
        double sn, cs;
        sincos(ph, &sn, &cs);
        s = cadd(s, cmk(cs * x, sn * x));
    }
    Dq[e] = s;
}

extern "C" hipError_t bse_launch_dipole_table(hipStream_t st, int nk, int norb, int nR, const double* kc,
                                              const double* Rcart, const double* Dr, const double* qhat,
                                              double2* Dq)
{
    const int64_t n = (int64_t)nk * norb * norb;
    hipLaunchKernelGGL(k_dipole_table, dim3((unsigned)ceil_div64(n, 256)), dim3(256), 0, st, nk, norb, nR, kc, Rcart,
                       Dr, qhat[0], qhat[1], qhat[2], Dq);
    return hipGetLastError();
}

// geom: nsc[3], p[3] (ints), origin[3], axis[9] (doubles); outputs sized
// nspin nw^2 nR (x nK complex, x 3, x 1).
extern "C" hipError_t bse_launch_wannier_integrals(hipStream_t st, const int* nsc, const int* p,
                                                   const double* origin, const double* axis, double cell_volume,
                                                   int nw, int nspin, int nR, const int* d_Rc, int nK,
                                                   const double* d_K, const double* d_w, double2* d_A,
                                                   double* d_D, double* d_S)
{
    WannGeom g;
    for (int a = 0; a < 3; ++a) {
        g.nsc[a] = nsc[a];
        g.p[a] = p[a];
        g.origin[a] = origin[a];
        if (nsc[a] < 1 || p[a] < 1 || nsc[a] * p[a] > WANN_AXMAX) return hipErrorInvalidValue;
    }
    for (int e = 0; e < 9; ++e) g.axis[e] = axis[e];
    g.ncell = nsc[0] * nsc[1] * nsc[2];
    g.pcell = p[0] * p[1] * p[2];
    g.f = (double)g.ncell * cell_volume / ((double)g.ncell * g.pcell);
    const int nrow = nspin * nw * nw * nR;
    const int nkb = (int)std::max<int64_t>(1, ceil_div64(nK, WANN_KB));
    hipLaunchKernelGGL(k_wann_int, dim3((unsigned)nrow, (unsigned)nkb), dim3(256), 0, st, g, nw, nspin, nR, d_Rc,
                       nK, d_K, d_w, d_A, d_D, d_S);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// C ABI: the integrals for host arrays (buffers from the context's pool).
// ---------------------------------------------------------------------------
extern "C" int bse_wannier_integrals(bse_ctx* ctx, const bse_wannier_grid* gr, int n_r, const int32_t* r_cells,
                                     int n_k, const double* kvec, double* A, double* D, double* S)
{
    if (!ctx) return BSE_ERR_ARG;
    if (!gr || !gr->w || n_r < 1 || !r_cells || n_k < 0 || (n_k > 0 && (!kvec || !A)) || !D || !S)
        return fail(ctx, BSE_ERR_ARG, "null or empty argument");
    if (gr->nw < 1 || (gr->nspin != 1 && gr->nspin != 2)) return fail(ctx, BSE_ERR_ARG, "nw >= 1, nspin 1 or 2");
    for (int a = 0; a < 3; ++a) {
        if (gr->n_sc[a] < 1 || gr->p[a] < 1) return fail(ctx, BSE_ERR_ARG, "grid sizes must be >= 1");
        if (gr->n_sc[a] * gr->p[a] > WANN_AXMAX)
            return fail(ctx, BSE_ERR_UNSUPPORTED, "at most 256 grid points along a supercell axis");
    }
    const double* ax = gr->axis;
    const double vol = fabs(ax[0] * (ax[4] * ax[8] - ax[5] * ax[7]) - ax[3] * (ax[1] * ax[8] - ax[2] * ax[7]) +
                            ax[6] * (ax[1] * ax[5] - ax[2] * ax[4]));
    if (!(vol > 0.0)) return fail(ctx, BSE_ERR_ARG, "degenerate unit-cell axes");
    (void)hipSetDevice(ctx->device);
    const int64_t npts = (int64_t)gr->n_sc[0] * gr->n_sc[1] * gr->n_sc[2] * gr->p[0] * gr->p[1] * gr->p[2];
    const int64_t nfun = (int64_t)gr->nspin * gr->nw;
    const int64_t rows = nfun * gr->nw * n_r;
    double* d_w = BUF(double, S_WGRID, nfun * npts);
    int* d_r = BUF(int, S_WR, 3 * (size_t)n_r);
    double* d_k = BUF(double, S_WK, 3 * (size_t)std::max(n_k, 1));
    double2* d_a = BUF(double2, S_WA, rows * std::max(n_k, 1));
    double* d_d = BUF(double, S_WD, 3 * rows);
    double* d_s = BUF(double, S_WS, rows);
    hipStream_t st = ctx->st;
    HCK(hipMemcpyAsync(d_w, gr->w, sizeof(double) * nfun * npts, hipMemcpyHostToDevice, st));
    HCK(hipMemcpyAsync(d_r, r_cells, sizeof(int) * 3 * n_r, hipMemcpyHostToDevice, st));
    if (n_k > 0) HCK(hipMemcpyAsync(d_k, kvec, sizeof(double) * 3 * n_k, hipMemcpyHostToDevice, st));
    HCK(bse_launch_wannier_integrals(st, gr->n_sc, gr->p, gr->origin, gr->axis, vol, gr->nw, gr->nspin, n_r, d_r,
                                     n_k, d_k, d_w, d_a, d_d, d_s));
    if (n_k > 0) HCK(hipMemcpyAsync(A, d_a, sizeof(double2) * rows * n_k, hipMemcpyDeviceToHost, st));
    HCK(hipMemcpyAsync(D, d_d, sizeof(double) * 3 * rows, hipMemcpyDeviceToHost, st));
    HCK(hipMemcpyAsync(S, d_s, sizeof(double) * rows, hipMemcpyDeviceToHost, st));
    HCK(hipStreamSynchronize(st));
    return BSE_OK;
}
