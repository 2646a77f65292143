// K7, all eigenvectors: back-transform Z <- Q Z with the one-stage
// reduction's panels, Q = H_0 H_1 ... = prod_p (I - V_p T_p V_p^H), for the
// FP64 path that forms every eigenvector (the eig_sym / zheevd drop-in,
// bse.version13.cpp:3330).
//
// One launch per panel p (last panel first).  The eigenvector matrix is cut
// into strips of BT_CS columns, one workgroup per strip, which applies the
// whole panel to its strip:
//   phase 1  W1 = V^H Z_strip (nbp x BT_CS), rows streamed through LDS in
//            chunks of BT_R1, four row groups of threads, each thread a
//            4 x 4 block of W1 in FP64 registers, the row groups summed in a
//            fixed order;
//   phase 2  W2 = T W1 (T upper triangular), in LDS;
//   phase 3  Z_strip -= V W2, rows in chunks of 2 BT_R1 (V staged in LDS,
//            each thread 2 rows x 4 columns of Z in registers).
// The strip is read twice and written once per panel, V once per phase from
// L2.  Work per panel 16 m nbp ncols real flops (m = n - j0).  Every sum runs
// in a fixed order, so the result is bitwise reproducible.
#include "bse_common.hpp"
#include <algorithm>

#define BT_CS 32       // columns per strip
#define BT_NB 32       // reflectors per panel (at most)
#define BT_R1 32       // rows per phase-1 chunk (phase 3: 2 BT_R1)

__global__ __launch_bounds__(256) void k_bt_panel(int m, int nbp, const double2* __restrict__ V, int64_t ldv,
                                                  const double2* __restrict__ T, int ldt,
                                                  double2* __restrict__ Z, int64_t ldz, int ncols)
{
    __shared__ double2 sA[2 * BT_R1][BT_NB + 1];   // phase 1: rows 0..31 V, 32..63 Z; phase 3: V rows
    __shared__ double2 sW[BT_NB][BT_CS + 1];       // W1 (and the row-group partials)
    __shared__ double2 sW2[BT_NB][BT_CS + 1];      // W2 = T W1
    const int tid = threadIdx.x;
    const int c0 = blockIdx.x * BT_CS;
    const int nc = min(BT_CS, ncols - c0);
    double2* Zs = Z + (int64_t)c0 * ldz;

    // ---- phase 1: W1 = V^H Z_strip.  The next chunk's loads are issued into
    // registers before the current chunk is multiplied.
    const int rg = tid >> 6, blk = tid & 63, lb = blk >> 3, cb = blk & 7;
    double2 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = cmk(0, 0);
    constexpr int P1 = BT_R1 * BT_NB / 256;       // elements per thread and chunk
    double2 pv[P1], pz[P1];
    auto load1 = [&](int r0) {
#pragma unroll
        for (int u = 0; u < P1; ++u) {
            const int e = tid + 256 * u, rr = e & (BT_R1 - 1), l = e / BT_R1, r = r0 + rr;
            pv[u] = (r < m && l < nbp) ? V[r + (int64_t)l * ldv] : cmk(0, 0);
            pz[u] = (r < m && l < nc) ? Zs[r + (int64_t)l * ldz] : cmk(0, 0);
        }
    };
    load1(0);
    for (int r0 = 0; r0 < m; r0 += BT_R1) {
#pragma unroll
        for (int u = 0; u < P1; ++u) {
            const int e = tid + 256 * u, rr = e & (BT_R1 - 1), l = e / BT_R1;
            sA[rr][l] = pv[u];
            sA[BT_R1 + rr][l] = pz[u];
        }
        __syncthreads();
        if (r0 + BT_R1 < m) load1(r0 + BT_R1);
#pragma unroll
        for (int t = 0; t < BT_R1 / 4; ++t) {
            const int rr = rg + 4 * t;
            double2 va[4], zb[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) va[a] = sA[rr][lb * 4 + a];
#pragma unroll
            for (int b = 0; b < 4; ++b) zb[b] = sA[BT_R1 + rr][cb * 4 + b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) cfmac(acc[a][b], va[a], zb[b]);
        }
        __syncthreads();
    }
    // row groups 1..3 add into group 0, in order
    for (int q = 1; q < 4; ++q) {
        if (rg == q) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) sW[lb * 4 + a][cb * 4 + b] = acc[a][b];
        }
        __syncthreads();
        if (rg == 0) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = cadd(acc[a][b], sW[lb * 4 + a][cb * 4 + b]);
        }
        __syncthreads();
    }
    if (rg == 0) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) sW[lb * 4 + a][cb * 4 + b] = acc[a][b];
    }
    __syncthreads();

    // ---- phase 2: W2 = T W1 (T upper triangular, nbp x nbp, ld ldt)
    for (int e = tid; e < BT_NB * BT_CS; e += 256) {
        const int l = e % BT_NB, c = e / BT_NB;
        double2 s = cmk(0, 0);
        if (l < nbp)
            for (int k = l; k < nbp; ++k) cfma(s, T[l + (int64_t)k * ldt], sW[k][c]);
        sW2[l][c] = s;
    }
    __syncthreads();

    // ---- phase 3: Z_strip -= V W2, the next chunk's V and Z in flight
    // while the current one is updated
    const int cq = tid >> 5, rb = tid & 31;      // 8 column groups of 4, 32 row slots
    constexpr int P3 = 2 * BT_R1 * BT_NB / 256;
    double2 qv[P3], z[2][4];
    auto load3 = [&](int r0) {
#pragma unroll
        for (int u = 0; u < P3; ++u) {
            const int e = tid + 256 * u, rr = e & (2 * BT_R1 - 1), l = e / (2 * BT_R1), r = r0 + rr;
            qv[u] = (r < m && l < nbp) ? V[r + (int64_t)l * ldv] : cmk(0, 0);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int r = r0 + rb + BT_R1 * a, c = cq * 4 + b;
                z[a][b] = (r < m && c < nc) ? Zs[r + (int64_t)c * ldz] : cmk(0, 0);
            }
    };
    load3(0);
    for (int r0 = 0; r0 < m; r0 += 2 * BT_R1) {
#pragma unroll
        for (int u = 0; u < P3; ++u) {
            const int e = tid + 256 * u, rr = e & (2 * BT_R1 - 1), l = e / (2 * BT_R1);
            sA[rr][l] = qv[u];
        }
        double2 zc[2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) zc[a][b] = z[a][b];
        __syncthreads();
        if (r0 + 2 * BT_R1 < m) load3(r0 + 2 * BT_R1);
        for (int l = 0; l < nbp; ++l) {
            const double2 v0 = sA[rb][l], v1 = sA[rb + BT_R1][l];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double2 w = sW2[l][cq * 4 + b];
                zc[0][b] = csub(zc[0][b], cmul(v0, w));
                zc[1][b] = csub(zc[1][b], cmul(v1, w));
            }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int r = r0 + rb + BT_R1 * a, c = cq * 4 + b;
                if (r < m && c < nc) Zs[r + (int64_t)c * ldz] = zc[a][b];
            }
        __syncthreads();
    }
}

// Z (n x ncols, ld ldz) <- Q Z for the panels of the packed reflector store
// (panel p: rows j0 = p nb .., nbp = min(nb, n - 1 - j0) reflectors at
// Vst + voff[p], ld n - j0; T_p at Tall + p nb nb, ld nb), last panel first.
extern "C" hipError_t bse_apply_q_strips(hipStream_t st, int n, int nb, int npan, const double2* Vst,
                                         const int64_t* voff, const double2* Tall, double2* Z, int64_t ldz,
                                         int ncols)
{
    if (nb > BT_NB || ncols <= 0) return nb > BT_NB ? hipErrorInvalidValue : hipSuccess;
    const unsigned g = (unsigned)ceil_div64(ncols, BT_CS);
    for (int p = npan - 1; p >= 0; --p) {
        const int j0 = p * nb, m = n - j0, nbp = std::min(nb, n - 1 - j0);
        if (nbp <= 0) continue;
        hipLaunchKernelGGL(k_bt_panel, dim3(g), dim3(256), 0, st, m, nbp, Vst + voff[p], (int64_t)m,
                           Tall + (size_t)p * nb * nb, nb, Z + j0, ldz, ncols);
    }
    return hipGetLastError();
}
