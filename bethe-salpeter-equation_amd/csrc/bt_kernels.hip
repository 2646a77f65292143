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

// ---------------------------------------------------------------------------
// K7 on FP64 matrix cores: the panels are applied four at a time.  A group
// of four consecutive panels (128 reflectors, rows j0 = 128 g ..) is one
// compact-WY block Q_g = Q_4g Q_4g+1 Q_4g+2 Q_4g+3 = I - V T V^H with
//   T = [[T_0:q, -T_0:q V_0:q^H V_q T_q], [0, T_q]]    (q = 1, 2, 3),
// so Z is read twice and written once per four panels instead of per panel.
// Per group (last group first):
//   k_btg_w1   W1_s = V^H Z over row chunk s (split K, partials in FP64),
//   k_btg_w2   W2 = -T sum_s W1_s (partials summed in chunk order),
//   k_btg_upd  Z += V W2.
// Once per call, for every group: k_btg_gram (the six 32 x 32 blocks
// V_i^H V_q, i < q) and k_btg_tbig (T of the group, 128 x 128).
// Products are complex FP64 built from four v_mfma_f64_16x16x4 per K step;
// every operand is read from global memory straight into the MFMA layout
// (16 lanes of one operand register read 16 consecutive rows of a
// column-major matrix, or 16 consecutive columns of a row-major one).  All
// sums run in a fixed order: the result is bitwise reproducible.
// Column k of a group is reflector l = k mod 32 of panel q = k / 32; it is
// zero above group row 32 q and, for a short last panel, beyond its count.
// ---------------------------------------------------------------------------
typedef double btg_d4 __attribute__((ext_vector_type(4)));
#define BTG_P 4                   // panels per group
#define BTG_K (BTG_P * BT_NB)     // reflectors per group
#ifndef BTG_KB
#define BTG_KB 4                  // K steps per prefetched block
#endif
#ifndef BTG_WG
#define BTG_WG 4096               // target workgroups of a full-height V^H Z launch
#endif

__device__ __forceinline__ btg_d4 btg_mfma(double a, double b, btg_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// c += a b for MT x NT complex 16 x 16 tiles (one K step of 4); a[i] is the
// lane's A element of row tile i, b[j] its B element of column tile j.
template <int MT, int NT>
__device__ __forceinline__ void btg_cmma(btg_d4 (&cr)[MT][NT], btg_d4 (&ci)[MT][NT], const double2 (&a)[MT],
                                         const double2 (&b)[NT]) {
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) cr[i][j] = btg_mfma(a[i].x, b[j].x, cr[i][j]);
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) ci[i][j] = btg_mfma(a[i].x, b[j].y, ci[i][j]);
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) cr[i][j] = btg_mfma(-a[i].y, b[j].y, cr[i][j]);
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) ci[i][j] = btg_mfma(a[i].y, b[j].x, ci[i][j]);
}

template <int MT, int NT>
__device__ __forceinline__ void btg_zero(btg_d4 (&cr)[MT][NT], btg_d4 (&ci)[MT][NT]) {
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
            cr[i][j] = btg_d4{0.0, 0.0, 0.0, 0.0};
            ci[i][j] = btg_d4{0.0, 0.0, 0.0, 0.0};
        }
}

// Panel p starts at row rb + 32 p and holds min(32, n - co - (rb + 32 p))
// reflectors: rb = 0, co = 1 for the one-stage reduction (reflector j acts
// below row j), rb = 32, co = 0 for the band reduction's panels (panel p
// factors rows 32 (p + 1) ..).
struct BtgGroup {
    int p0, gq, j0, m, co;
    __device__ BtgGroup(int n, int npan, int g, int rb, int co_) {
        p0 = g * BTG_P;
        gq = min(BTG_P, npan - p0);
        j0 = rb + p0 * BT_NB;
        m = n - j0;
        co = co_;
    }
    // reflectors of panel q of the group
    __device__ int nbp(int n, int q) const { return min(BT_NB, n - co - j0 - BT_NB * q); }
};

// Reduction over group rows r in [rlo, rhi) of conj(Va[r][a]) Vb[r][b] or
// conj(Va[r][a]) Z[r][b] for one wave's 32 x 32 output (2 x 2 tiles).  The
// lane's operand pointers pa[i], pb[j] address row 0 of its column (valid
// flags va, vb).  Inside a block of 4 BTG_KB rows, K step t of lane quad lq
// reads row rb + lq BTG_KB + t, so a lane reads BTG_KB consecutive rows.
__device__ __forceinline__ void btg_rows(btg_d4 (&cr)[2][2], btg_d4 (&ci)[2][2], const double2* const (&pa)[2],
                                         const bool (&va)[2], const double2* const (&pb)[2], const bool (&vb)[2],
                                         int rlo, int rhi, int lq) {
    double2 a[BTG_KB][2], b[BTG_KB][2];
    auto load = [&](int rb, double2 (&A)[BTG_KB][2], double2 (&B)[BTG_KB][2]) {
#pragma unroll
        for (int t = 0; t < BTG_KB; ++t) {
            const int r = rb + lq * BTG_KB + t;
            const bool in = r < rhi;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const double2 v = (in && va[i]) ? pa[i][r] : cmk(0, 0);
                A[t][i] = cmk(v.x, -v.y);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) B[t][j] = (in && vb[j]) ? pb[j][r] : cmk(0, 0);
        }
    };
    if (rlo >= rhi) return;
    load(rlo, a, b);
    for (int rb = rlo; rb < rhi; rb += 4 * BTG_KB) {
        double2 an[BTG_KB][2], bn[BTG_KB][2];
        const bool more = rb + 4 * BTG_KB < rhi;
        if (more) load(rb + 4 * BTG_KB, an, bn);
#pragma unroll
        for (int t = 0; t < BTG_KB; ++t) btg_cmma<2, 2>(cr, ci, a[t], b[t]);
        if (more) {
#pragma unroll
            for (int t = 0; t < BTG_KB; ++t)
#pragma unroll
                for (int i = 0; i < 2; ++i) { a[t][i] = an[t][i]; b[t][i] = bn[t][i]; }
        }
    }
}

// Gram blocks V_i^H V_q (32 x 32, column-major) of every group: grid
// (6 pairs, groups); the four waves split the rows, summed in wave order.
__global__ __launch_bounds__(256) void k_btg_gram(int n, int npan, int rb, int co, const double2* __restrict__ Vst,
                                                  const int64_t* __restrict__ voff, double2* __restrict__ Gram)
{
    __shared__ double2 sR[4][BT_NB * BT_NB];
    const int pi[6] = {0, 0, 1, 0, 1, 2}, pq[6] = {1, 2, 2, 3, 3, 3};
    const int pair = blockIdx.x, g = blockIdx.y;
    const BtgGroup G(n, npan, g, rb, co);
    const int i = pi[pair], q = pq[pair];
    if (q >= G.gq) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
    const int nbi = G.nbp(n, i), nbq = G.nbp(n, q);
    const int64_t ldi = G.m - BT_NB * i, ldq = G.m - BT_NB * q;
    // rows of V_q start at group row 32 q; rows are split in blocks of 4 BTG_KB
    const int r0 = BT_NB * q, R = G.m - r0;
    const int per = (int)ceil_div64(ceil_div64(R, 4), 4 * BTG_KB) * 4 * BTG_KB;
    const int rlo = r0 + w * per, rhi = min(G.m, rlo + per);
    // operand pointers address group row 0 of their column
    const double2* pa[2];
    const double2* pb[2];
    bool va[2], vb[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int a = 16 * t + li;
        va[t] = a < nbi;
        vb[t] = a < nbq;
        pa[t] = Vst + voff[G.p0 + i] + (int64_t)a * ldi - BT_NB * i;
        pb[t] = Vst + voff[G.p0 + q] + (int64_t)a * ldq - BT_NB * q;
    }
    btg_d4 cr[2][2], ci[2][2];
    btg_zero<2, 2>(cr, ci);
    btg_rows(cr, ci, pa, va, pb, vb, rlo, rhi, lq);
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
                const int a = 16 * ti + mfma16_row_f64(lane, reg), b = 16 * tj + li;
                sR[w][a + BT_NB * b] = cmk(cr[ti][tj][reg], ci[ti][tj][reg]);
            }
    __syncthreads();
    double2* out = Gram + ((int64_t)g * 6 + pair) * BT_NB * BT_NB;
    for (int e = threadIdx.x; e < BT_NB * BT_NB; e += 256)
        out[e] = cadd(cadd(cadd(sR[0][e], sR[1][e]), sR[2][e]), sR[3][e]);
}

// T of every group (128 x 128, column-major, zero below the diagonal and for
// missing reflectors): one workgroup per group.
__global__ __launch_bounds__(256) void k_btg_tbig(int n, int npan, int rb, int co, const double2* __restrict__ Tall,
                                                  const double2* __restrict__ Gram, double2* __restrict__ Tbig)
{
    __shared__ double2 sS[3 * BT_NB][BT_NB + 1];
    __shared__ double2 sY[3 * BT_NB][BT_NB + 1];
    __shared__ double2 sT[BT_NB][BT_NB + 1];
    const int g = blockIdx.x, tid = threadIdx.x;
    const BtgGroup G(n, npan, g, rb, co);
    double2* T = Tbig + (int64_t)g * BTG_K * BTG_K;
    for (int e = tid; e < BTG_K * BTG_K; e += 256) {
        const int k = e % BTG_K, kp = e / BTG_K, q = k / BT_NB;
        double2 v = cmk(0, 0);
        if (q == kp / BT_NB && q < G.gq) {
            const int nb = G.nbp(n, q), a = k % BT_NB, b = kp % BT_NB;
            if (a < nb && b < nb && a <= b) v = Tall[(int64_t)(G.p0 + q) * BT_NB * BT_NB + a + BT_NB * b];
        }
        T[e] = v;
    }
    const int pidx[4] = {0, 0, 1, 3};   // first pair index of column block q
    for (int q = 1; q < G.gq; ++q) {
        __syncthreads();
        const int nb = G.nbp(n, q), rows = BT_NB * q;
        for (int e = tid; e < BT_NB * BT_NB; e += 256) {
            const int a = e % BT_NB, b = e / BT_NB;
            sT[a][b] = (a < nb && b < nb && a <= b) ? Tall[(int64_t)(G.p0 + q) * BT_NB * BT_NB + a + BT_NB * b]
                                                    : cmk(0, 0);
        }
        for (int e = tid; e < rows * BT_NB; e += 256) {
            const int r = e % rows, b = e / rows, i = r / BT_NB, a = r % BT_NB;
            sS[r][b] = Gram[((int64_t)g * 6 + pidx[q] + i) * BT_NB * BT_NB + a + BT_NB * b];
        }
        __syncthreads();
        // Y = S_0:q,q T_q
        for (int e = tid; e < rows * BT_NB; e += 256) {
            const int r = e % rows, b = e / rows;
            double2 s = cmk(0, 0);
            for (int k = 0; k <= b; ++k) cfma(s, sS[r][k], sT[k][b]);
            sY[r][b] = s;
        }
        __syncthreads();
        // T_0:q,q = -T_0:q,0:q Y (T_0:q,0:q upper triangular, written above)
        for (int e = tid; e < rows * BT_NB; e += 256) {
            const int r = e % rows, b = e / rows;
            double2 s = cmk(0, 0);
            for (int k = r; k < rows; ++k) cfma(s, T[r + (int64_t)BTG_K * k], sY[k][b]);
            T[r + (int64_t)BTG_K * (rows + b)] = cmk(-s.x, -s.y);
        }
    }
}

// W1p[s][k][c] = sum over group rows r of chunk s of conj(V[r][k]) Z[j0 + r][c]
// (ld ncols).  Grid (column tiles of 32, chunks); wave w: panel w of the group.
__global__ __launch_bounds__(256) void k_btg_w1(int n, int npan, int rb, int co, int g, const double2* __restrict__ Vst,
                                                const int64_t* __restrict__ voff, const double2* __restrict__ Z,
                                                int64_t ldz, int ncols, int rchunk, double2* __restrict__ W1p)
{
    const BtgGroup G(n, npan, g, rb, co);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
    const int c0 = blockIdx.x * 32, s = blockIdx.y;
    const int q = w;
    btg_d4 cr[2][2], ci[2][2];
    btg_zero<2, 2>(cr, ci);
    if (q < G.gq) {
        const int nb = G.nbp(n, q);
        const int64_t ldq = G.m - BT_NB * q;
        const double2* pa[2];
        const double2* pb[2];
        bool va[2], vb[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int a = 16 * t + li, c = c0 + 16 * t + li;
            va[t] = a < nb;
            vb[t] = c < ncols;
            pa[t] = Vst + voff[G.p0 + q] + (int64_t)a * ldq - BT_NB * q;
            pb[t] = Z + G.j0 + (int64_t)min(c, ncols - 1) * ldz;
        }
        const int rlo = max(s * rchunk, BT_NB * q), rhi = min(G.m, (s + 1) * rchunk);
        btg_rows(cr, ci, pa, va, pb, vb, rlo, rhi, lq);
    }
    double2* out = W1p + (int64_t)s * BTG_K * ncols;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
                const int k = BT_NB * w + 16 * ti + mfma16_row_f64(lane, reg), c = c0 + 16 * tj + li;
                if (c < ncols) out[(int64_t)k * ncols + c] = cmk(cr[ti][tj][reg], ci[ti][tj][reg]);
            }
}

// W2[k][c] = -sum_k' T[k][k'] sum_s W1p[s][k'][c]: grid column tiles of 32;
// the partials are summed into LDS, wave w forms rows 32 w .. of W2.
__global__ __launch_bounds__(256) void k_btg_w2(int nsplit, const double2* __restrict__ T,
                                                const double2* __restrict__ W1p, int ncols, double2* __restrict__ W2)
{
    __shared__ double2 sW[BTG_K][32];
    const int tid = threadIdx.x, c0 = blockIdx.x * 32;
    for (int e = tid; e < BTG_K * 32; e += 256) {
        const int k = e / 32, c = c0 + e % 32;
        double2 v = cmk(0, 0);
        if (c < ncols)
            for (int s = 0; s < nsplit; ++s) v = cadd(v, W1p[((int64_t)s * BTG_K + k) * ncols + c]);
        sW[k][e % 32] = v;
    }
    __syncthreads();
    const int w = tid >> 6, lane = tid & 63, li = lane & 15, lq = lane >> 4;
    btg_d4 cr[2][2], ci[2][2];
    btg_zero<2, 2>(cr, ci);
    for (int k0 = BT_NB * w; k0 < BTG_K; k0 += 4) {
        double2 a[2], b[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            a[t] = T[(BT_NB * w + 16 * t + li) + (int64_t)BTG_K * (k0 + lq)];
            b[t] = sW[k0 + lq][16 * t + li];
        }
        btg_cmma<2, 2>(cr, ci, a, b);
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
                const int k = BT_NB * w + 16 * ti + mfma16_row_f64(lane, reg), c = c0 + 16 * tj + li;
                if (c < ncols) W2[(int64_t)k * ncols + c] = cmk(-cr[ti][tj][reg], -ci[ti][tj][reg]);
            }
}

// Z[j0 + r][c] += sum_k V[r][k] W2[k][c], computed as the transposed product
// (output rows = columns c, output columns = rows r, so 16 lanes of one
// register hold 16 consecutive rows of Z).  Grid (column tiles of 64, row
// tiles of 64); wave w: columns 32 (w & 1) .., rows 32 (w >> 1) ..
__global__ __launch_bounds__(256) void k_btg_upd(int n, int npan, int rb, int co, int g, const double2* __restrict__ Vst,
                                                 const int64_t* __restrict__ voff, const double2* __restrict__ W2,
                                                 double2* __restrict__ Z, int64_t ldz, int ncols)
{
    const BtgGroup G(n, npan, g, rb, co);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lq = lane >> 4;
    const int c0 = blockIdx.x * 64 + 32 * (w & 1), r0 = blockIdx.y * 64 + 32 * (w >> 1);
    btg_d4 cr[2][2], ci[2][2];
    double2* Zg = Z + G.j0;
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
                const int c = c0 + 16 * ti + mfma16_row_f64(lane, reg), r = r0 + 16 * tj + li;
                const double2 v = (c < ncols && r < G.m) ? Zg[r + (int64_t)c * ldz] : cmk(0, 0);
                cr[ti][tj][reg] = v.x;
                ci[ti][tj][reg] = v.y;
            }
    // K step s: k = 4 s + lq, panel s / 8, reflector 4 (s % 8) + lq
    const int nsteps = G.gq * (BT_NB / 4);
    double2 a[BTG_KB][2], b[BTG_KB][2];
    auto load = [&](int s0, double2 (&A)[BTG_KB][2], double2 (&B)[BTG_KB][2]) {
        const int q = s0 / (BT_NB / 4);
        const int nb = G.nbp(n, q);
        const int64_t ldq = G.m - BT_NB * q;
        const double2* Vq = Vst + voff[G.p0 + q];
#pragma unroll
        for (int t = 0; t < BTG_KB; ++t) {
            const int s = s0 + t, k = 4 * s + lq, l = k % BT_NB;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int c = c0 + 16 * i + li;
                A[t][i] = c < ncols ? W2[(int64_t)k * ncols + c] : cmk(0, 0);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int r = r0 + 16 * j + li;
                B[t][j] = (l < nb && r >= BT_NB * q && r < G.m) ? Vq[(r - BT_NB * q) + (int64_t)l * ldq] : cmk(0, 0);
            }
        }
    };
    load(0, a, b);
    for (int s0 = 0; s0 < nsteps; s0 += BTG_KB) {
        double2 an[BTG_KB][2], bn[BTG_KB][2];
        const bool more = s0 + BTG_KB < nsteps;
        if (more) load(s0 + BTG_KB, an, bn);
#pragma unroll
        for (int t = 0; t < BTG_KB; ++t) btg_cmma<2, 2>(cr, ci, a[t], b[t]);
        if (more) {
#pragma unroll
            for (int t = 0; t < BTG_KB; ++t)
#pragma unroll
                for (int i = 0; i < 2; ++i) { a[t][i] = an[t][i]; b[t][i] = bn[t][i]; }
        }
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
                const int c = c0 + 16 * ti + mfma16_row_f64(lane, reg), r = r0 + 16 * tj + li;
                if (c < ncols && r < G.m) Zg[r + (int64_t)c * ldz] = cmk(cr[ti][tj][reg], ci[ti][tj][reg]);
            }
}

static int btg_nsplit(int ncols) {
    // about BTG_WG workgroups for the V^H Z partials of a full-height group,
    // at most 32 partials for the reduction to sum
    return (int)std::min<int64_t>(32, std::max<int64_t>(1, ceil_div64(BTG_WG, ceil_div64(ncols, 32))));
}

// Workspace (complex elements) of the grouped back-transform: the Gram blocks
// and T of every group, then W2 and the W1 partials for ncols columns.
extern "C" int64_t bse_apply_q_grouped_workspace(int n, int npan, int ncols)
{
    const int64_t ng = ceil_div64(npan, BTG_P);
    return ng * 6 * BT_NB * BT_NB + ng * BTG_K * BTG_K + (int64_t)(btg_nsplit(ncols) + 1) * BTG_K * ncols;
}

// T of every group of four panels into the workspace (independent of Z).
// Panels as for BtgGroup (rb, co); panel p at Vst + voff_d[p] with ld
// n - rb - 32 p, T_p at Tall + p 32 32.
extern "C" hipError_t bse_apply_q_grouped_prepare(hipStream_t st, int n, int npan, int rb, int co,
                                                  const double2* Vst, const int64_t* voff_d, const double2* Tall,
                                                  double2* ws)
{
    if (npan <= 0) return hipSuccess;
    const int ng = (int)ceil_div64(npan, BTG_P);
    double2* gram = ws;
    double2* tbig = gram + (int64_t)ng * 6 * BT_NB * BT_NB;
    hipLaunchKernelGGL(k_btg_gram, dim3(6, ng), dim3(256), 0, st, n, npan, rb, co, Vst, voff_d, gram);
    hipLaunchKernelGGL(k_btg_tbig, dim3(ng), dim3(256), 0, st, n, npan, rb, co, Tall, gram, tbig);
    return hipGetLastError();
}

// Z (n x ncols, ld ldz) <- Q Z, last group first, with the T blocks that
// bse_apply_q_grouped_prepare left in ws.
extern "C" hipError_t bse_apply_q_grouped_run(hipStream_t st, int n, int npan, int rb, int co, const double2* Vst,
                                              const int64_t* voff_d, double2* Z, int64_t ldz, int ncols,
                                              double2* ws, int64_t ws_el)
{
    if (ncols <= 0 || npan <= 0) return hipSuccess;
    if (ws_el < bse_apply_q_grouped_workspace(n, npan, ncols)) return hipErrorInvalidValue;
    const int ng = (int)ceil_div64(npan, BTG_P);
    double2* tbig = ws + (int64_t)ng * 6 * BT_NB * BT_NB;
    double2* w2 = tbig + (int64_t)ng * BTG_K * BTG_K;
    double2* w1p = w2 + (int64_t)BTG_K * ncols;
    const int ns = btg_nsplit(ncols);
    const unsigned c32 = (unsigned)ceil_div64(ncols, 32), c64 = (unsigned)ceil_div64(ncols, 64);
    for (int g = ng - 1; g >= 0; --g) {
        const int m = n - rb - g * BTG_K;
        // chunks of at least 64 rows, a multiple of 4 BTG_KB
        const int rchunk = (int)std::max<int64_t>(64, ceil_div64(ceil_div64(m, ns), 4 * BTG_KB) * 4 * BTG_KB);
        const int nsg = (int)ceil_div64(m, rchunk);
        hipLaunchKernelGGL(k_btg_w1, dim3(c32, nsg), dim3(256), 0, st, n, npan, rb, co, g, Vst, voff_d, Z, ldz,
                           ncols, rchunk, w1p);
        hipLaunchKernelGGL(k_btg_w2, dim3(c32), dim3(256), 0, st, nsg, tbig + (int64_t)g * BTG_K * BTG_K, w1p, ncols,
                           w2);
        hipLaunchKernelGGL(k_btg_upd, dim3(c64, (unsigned)ceil_div64(m, 64)), dim3(256), 0, st, n, npan, rb, co, g,
                           Vst, voff_d, w2, Z, ldz, ncols);
    }
    return hipGetLastError();
}

// Z (n x ncols, ld ldz) <- Q Z for the one-stage panels (nb = 32; panel p at
// Vst + voff[p] with ld n - 32 p, T_p at Tall + p 32 32), four panels per
// pass.  voff_d: the panel offsets in device memory.
extern "C" hipError_t bse_apply_q_grouped(hipStream_t st, int n, int nb, int npan, const double2* Vst,
                                          const int64_t* voff_d, const double2* Tall, double2* Z, int64_t ldz,
                                          int ncols, double2* ws, int64_t ws_el)
{
    if (nb != BT_NB || ncols <= 0 || npan <= 0) return nb != BT_NB ? hipErrorInvalidValue : hipSuccess;
    if (ws_el < bse_apply_q_grouped_workspace(n, npan, ncols)) return hipErrorInvalidValue;
    hipError_t e = bse_apply_q_grouped_prepare(st, n, npan, 0, 1, Vst, voff_d, Tall, ws);
    if (e != hipSuccess) return e;
    return bse_apply_q_grouped_run(st, n, npan, 0, 1, Vst, voff_d, Z, ldz, ncols, ws, ws_el);
}
