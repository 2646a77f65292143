// K5, two-stage form, first stage: dense Hermitian -> Hermitian band of
// half-bandwidth b = 32 by blocked Householder panels (the dense half of
// the two-stage reduction that replaces the one-stage hetrd when only a few
// eigenvectors are wanted; the matrix is the one Armadillo's eig_sym hands to
// LAPACK zheevd, bse.version13.cpp:3330).
//
// Panel p covers columns [p b, (p+1) b); with s = (p+1) b and m = n - s, its
// part below the band, P = A[s:, p b : s] (m x b), is factored P = Q [R; 0]
// and the trailing block A22 = A[s:, s:] becomes Q^H A22 Q.  Per panel:
//   gather      P (FP64, dense) and the diagonal block into band storage;
//   TSQR        leaf QRs of 64..127-row blocks in LDS, a tree of QRs of
//               stacked R factors (8 per node), then Q_thin = Q [I; 0] formed
//               back down the tree (Demmel, Grigori, Hoemmen & Langou 2012);
//   reconstruct the Householder form I - V T V^H of Q from Q_thin by an LU
//               factorisation of E - Q_thin D without pivoting, D a diagonal
//               of unit phases chosen during the elimination so that every
//               pivot has modulus >= 1 (Ballard, Demmel, Grigori, Jacquelin,
//               Nguyen & Solomonik 2014): V = L, T = U L1^{-H}, and the band
//               block R' = D^H R;
//   X = A22 V   Hermitian block product over the tiled lower storage;
//   Z = X T - V (T^H (V^H X T)) / 2;
//   her2k       A22 -= V Z^H + Z V^H (the one-stage trailing-update kernels).
// Every step is a parallel pass or a small one-workgroup kernel; there is no
// per-column device-wide synchronisation.  Panels, reflectors and all
// products are FP64; an FP32-stored matrix is read as FP32 and rounded once
// per update, as in the one-stage path.
//
// Band storage (FP64, used by the second stage): AB[c * ldab + d] = A(c + d, c)
// for 0 <= d <= 2b, so that the bulge chasing has room for its fill.
#include "bse_common.hpp"
#include <type_traits>

#define BR_B 32            // band half-width = panel width
#define TSQR_RB 64         // leaf block rows (the last leaf takes 64..127)
#define TSQR_G 8           // R factors stacked per tree node
#define TSQR_MAXROWS 256   // rows of one in-LDS QR: max(2 RB - 1, G b)

// ---------------------------------------------------------------------------
// Gather of panel p from the tiled lower storage: P (m x b, ld m, FP64) and
// the lower triangle of the diagonal block A[c0 : c0+b, c0 : c0+b] into the
// band store.  One thread per row of the panel (b columns each); the threads
// of the first block also copy the diagonal block.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_band_gather(int n, int c0, int b, const T* __restrict__ A,
                                                     double2* __restrict__ P, double2* __restrict__ AB,
                                                     int ldab)
{
    const int nt = tile_count(n);
    const int s = c0 + b, m = n - s;
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r < m) {
        for (int j = 0; j < b; ++j) P[r + (int64_t)j * m] = cplx<T>::load(A + tix(s + r, c0 + j, nt));
    }
    if (blockIdx.x == 0) {
        // diagonal block, lower triangle (rows and columns below n)
        for (int e = threadIdx.x; e < b * b; e += 256) {
            const int i = e % b, j = e / b;
            const int gr = c0 + i, gc = c0 + j;
            if (i >= j && gr < n) {
                double2 x = cplx<T>::load(A + tix(gr, gc, nt));
                if (i == j) x.y = 0.0;
                AB[(int64_t)gc * ldab + (i - j)] = x;
            }
        }
    }
}

// The trailing block that is already inside the band after the last panel:
// every (r, c) with c0 <= c <= r < n goes to the band store.
template <typename T>
__global__ __launch_bounds__(256) void k_band_gather_tail(int n, int c0, const T* __restrict__ A,
                                                          double2* __restrict__ AB, int ldab)
{
    const int nt = tile_count(n);
    const int w = n - c0;
    for (int e = blockIdx.x * 256 + threadIdx.x; e < w * w; e += gridDim.x * 256) {
        const int i = e % w, j = e / w;
        if (i < j) continue;
        double2 x = cplx<T>::load(A + tix(c0 + i, c0 + j, nt));
        if (i == j) x.y = 0.0;
        AB[(int64_t)(c0 + j) * ldab + (i - j)] = x;
    }
}

// ---------------------------------------------------------------------------
// Householder QR of a rows x b matrix held in LDS (column-major, ld lda), in
// place: R on and above the diagonal, the reflector vectors below it (their
// leading 1 implicit).  H_c = I - tau_c v v^H with H_c^H x = beta e_1, beta
// real (the LAPACK zlarfg convention); a column that is already zero below its
// diagonal with a real diagonal gets tau = 0.  256 threads.
// ---------------------------------------------------------------------------
__device__ void lds_house_qr(int rows, int b, double2* A, int lda, double2* tau, double* red,
                             double2* wbuf /* >= 8 * 32 */)
{
    const int tid = threadIdx.x;
    const int kmax = min(rows, b);
    for (int c = 0; c < kmax; ++c) {
        double ss = 0.0;
        for (int i = c + 1 + tid; i < rows; i += 256) ss += cabs2(A[i + c * lda]);
        ss = block_sum(ss, red);
        const double2 alpha = A[c + c * lda];
        double2 tc = cmk(0, 0), scal = cmk(0, 0);
        double beta = alpha.x;
        if (!(ss == 0.0 && alpha.y == 0.0)) {
            const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + ss);
            beta = (alpha.x >= 0.0) ? -nrm : nrm;
            tc = cmk((beta - alpha.x) / beta, -alpha.y / beta);
            scal = cdiv(cmk(1.0, 0.0), cmk(alpha.x - beta, alpha.y));
        }
        __syncthreads();   // alpha read by all before the column is rewritten
        for (int i = c + 1 + tid; i < rows; i += 256) A[i + c * lda] = cmul(A[i + c * lda], scal);
        if (tid == 0) {
            A[c + c * lda] = cmk(beta, 0.0);
            tau[c] = tc;
        }
        __syncthreads();
        // columns j > c:  A(c:, j) -= conj(tc) v (v^H A(c:, j)),  v(c) = 1;
        // the loops are unrolled over the at most TSQR_MAXROWS / 8 rows of a
        // thread so that its LDS reads are issued together
        constexpr int RT = TSQR_MAXROWS / 8;
        const int nc = b - c - 1;
        const int jj = tid & 31, g = tid >> 5;
        const bool act = jj < nc && (tc.x != 0.0 || tc.y != 0.0);
        double2 vr[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const int i = c + g + 8 * t;
            vr[t] = (act && i < rows) ? ((i == c) ? cmk(1.0, 0.0) : A[i + c * lda]) : cmk(0, 0);
        }
        double2 acc = cmk(0, 0);
        if (act) {
            const int j = c + 1 + jj;
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const int i = c + g + 8 * t;
                if (i < rows) cfmac(acc, vr[t], A[i + j * lda]);
            }
        }
        wbuf[g * 32 + jj] = acc;
        __syncthreads();
        if (act) {
            double2 w = wbuf[jj];
#pragma unroll
            for (int q = 1; q < 8; ++q) w = cadd(w, wbuf[q * 32 + jj]);
            const double2 f = cmul(cconj(tc), w);
            const int j = c + 1 + jj;
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const int i = c + g + 8 * t;
                if (i < rows) A[i + j * lda] = csub(A[i + j * lda], cmul(vr[t], f));
            }
        }
        __syncthreads();
    }
}

// Y <- (H_0 H_1 ... H_{k-1}) Y for the reflectors stored below the diagonal of
// V (rows x k, ld ldv, in LDS or global), Y (rows x nc, ld ldy) in LDS.
__device__ void lds_apply_q(int rows, int k, const double2* V, int ldv, const double2* tau,
                            double2* Y, int ldy, int nc, double2* wbuf /* >= 8 * 32 */)
{
    const int tid = threadIdx.x;
    const int jj = tid & 31, g = tid >> 5;
    // a thread's reflector entries (rows c + g + 8 t) are loaded into
    // registers at once per reflector: V may live in global memory
    constexpr int RT = TSQR_MAXROWS / 8;
    for (int c = k - 1; c >= 0; --c) {
        const double2 tc = tau[c];
        if (tc.x == 0.0 && tc.y == 0.0) continue;   // uniform: tau is shared
        double2 vr[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const int i = c + g + 8 * t;
            vr[t] = (i < rows) ? ((i == c) ? cmk(1.0, 0.0) : V[i + (int64_t)c * ldv]) : cmk(0, 0);
        }
        double2 acc = cmk(0, 0);
        if (jj < nc) {
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const int i = c + g + 8 * t;
                if (i < rows) cfmac(acc, vr[t], Y[i + jj * ldy]);
            }
        }
        wbuf[g * 32 + jj] = acc;
        __syncthreads();
        if (jj < nc) {
            double2 w = wbuf[jj];
#pragma unroll
            for (int q = 1; q < 8; ++q) w = cadd(w, wbuf[q * 32 + jj]);
            const double2 f = cmul(tc, w);
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const int i = c + g + 8 * t;
                if (i < rows) Y[i + jj * ldy] = csub(Y[i + jj * ldy], cmul(vr[t], f));
            }
        }
        __syncthreads();
    }
}

// Leaf rows of a TSQR of an m-row panel: nblk = max(1, m / RB) blocks of RB
// rows, the last one taking the remainder.
__host__ __device__ inline int tsqr_nblk(int m) { return m < TSQR_RB ? 1 : m / TSQR_RB; }
// LDS leading dimension of an in-LDS panel of `rows` rows: odd, so that the
// 32 columns a wavefront reads at one row fall in different banks.
__host__ __device__ inline int tsqr_ld(int rows) { return rows | 1; }
__host__ __device__ inline int tsqr_r0(int blk) { return blk * TSQR_RB; }
__host__ __device__ inline int tsqr_r1(int blk, int nblk, int m) { return blk == nblk - 1 ? m : (blk + 1) * TSQR_RB; }

// Leaf QRs: block blk of P (m x b, ld m) is factored in LDS; its reflector
// vectors overwrite the block below the diagonal, R (b x b, zero below the
// diagonal) goes to R + blk b b and tau to tau + blk b.
__global__ __launch_bounds__(256) void k_tsqr_leaf(int m, int b, double2* __restrict__ P,
                                                   double2* __restrict__ R, double2* __restrict__ tau,
                                                   const int* __restrict__ gate)
{
    extern __shared__ double2 sA[];
    __shared__ double red[32];
    __shared__ double2 wbuf[8 * 32], stau[BR_B];
    if (gate && *gate == 0) return;
    const int nblk = tsqr_nblk(m);
    const int blk = blockIdx.x;
    const int r0 = tsqr_r0(blk), rows = tsqr_r1(blk, nblk, m) - r0;
    const int ld = tsqr_ld(rows);
    for (int e = threadIdx.x; e < rows * b; e += 256) {
        const int i = e % rows, j = e / rows;
        sA[i + j * ld] = P[(r0 + i) + (int64_t)j * m];
    }
    if (threadIdx.x < BR_B) stau[threadIdx.x] = cmk(0, 0);
    __syncthreads();
    lds_house_qr(rows, b, sA, ld, stau, red, wbuf);
    for (int e = threadIdx.x; e < rows * b; e += 256) {
        const int i = e % rows, j = e / rows;
        if (i > j) P[(r0 + i) + (int64_t)j * m] = sA[i + j * ld];
    }
    for (int e = threadIdx.x; e < b * b; e += 256) {
        const int i = e % b, j = e / b;
        R[(int64_t)blk * b * b + e] = (i <= j && i < rows) ? sA[i + j * ld] : cmk(0, 0);
    }
    if (threadIdx.x < b) tau[(int64_t)blk * b + threadIdx.x] = stau[threadIdx.x];
}

// Tree node: group g stacks R factors g G .. g G + cnt - 1 (each b x b) of the
// level below, factors the (cnt b) x b stack; its reflectors go to
// Vn + g G b b (ld cnt b), tau to taun + g b, the new R to Rout + g b b.
__global__ __launch_bounds__(256) void k_tsqr_node(int nin, int b, const double2* __restrict__ Rin,
                                                   double2* __restrict__ Vn, double2* __restrict__ taun,
                                                   double2* __restrict__ Rout, const int* __restrict__ gate)
{
    extern __shared__ double2 sA[];
    __shared__ double red[32];
    __shared__ double2 wbuf[8 * 32], stau[BR_B];
    if (gate && *gate == 0) return;
    const int g = blockIdx.x;
    const int cnt = min(TSQR_G, nin - g * TSQR_G);
    const int rows = cnt * b;
    const int ld = tsqr_ld(rows);
    for (int e = threadIdx.x; e < rows * b; e += 256) {
        const int i = e % rows, j = e / rows;
        const int t = i / b, ii = i % b;
        sA[i + j * ld] = Rin[(int64_t)(g * TSQR_G + t) * b * b + ii + j * b];
    }
    if (threadIdx.x < BR_B) stau[threadIdx.x] = cmk(0, 0);
    __syncthreads();
    lds_house_qr(rows, b, sA, ld, stau, red, wbuf);
    double2* V = Vn + (int64_t)g * TSQR_G * b * b;
    for (int e = threadIdx.x; e < rows * b; e += 256) {
        const int i = e % rows, j = e / rows;
        V[e] = sA[i + j * ld];
    }
    for (int e = threadIdx.x; e < b * b; e += 256) {
        const int i = e % b, j = e / b;
        Rout[(int64_t)g * b * b + e] = (i <= j) ? sA[i + j * ld] : cmk(0, 0);
    }
    if (threadIdx.x < b) taun[(int64_t)g * b + threadIdx.x] = stau[threadIdx.x];
}

// Downward pass through a tree level: node g gets its b x k input X (the
// identity for the root, root != 0) and writes Q_node [X; 0] as cnt pieces of
// b x k into Xout + (g G + t) b k.
__global__ __launch_bounds__(256) void k_tsqr_node_q(int nin, int b, int k, const double2* __restrict__ Vn,
                                                     const double2* __restrict__ taun,
                                                     const double2* __restrict__ Xin, int root,
                                                     double2* __restrict__ Xout, const int* __restrict__ gate)
{
    extern __shared__ double2 sY[];
    __shared__ double2 wbuf[8 * 32], stau[BR_B];
    if (gate && *gate == 0) return;
    const int g = blockIdx.x;
    const int cnt = min(TSQR_G, nin - g * TSQR_G);
    const int rows = cnt * b;
    const int ld = tsqr_ld(rows);
    for (int e = threadIdx.x; e < rows * k; e += 256) {
        const int i = e % rows, j = e / rows;
        double2 x = cmk(0, 0);
        if (i < b) x = root ? cmk(i == j ? 1.0 : 0.0, 0.0) : Xin[(int64_t)g * b * k + i + j * b];
        sY[i + j * ld] = x;
    }
    if (threadIdx.x < b) stau[threadIdx.x] = taun[(int64_t)g * b + threadIdx.x];
    __syncthreads();
    lds_apply_q(rows, b, Vn + (int64_t)g * TSQR_G * b * b, rows, stau, sY, ld, k, wbuf);
    for (int e = threadIdx.x; e < rows * k; e += 256) {
        const int i = e % rows, j = e / rows;
        const int t = i / b, ii = i % b;
        Xout[(int64_t)(g * TSQR_G + t) * b * k + ii + j * b] = sY[i + j * ld];
    }
}

// Leaves of the downward pass: block blk forms Q_blk [X_blk; 0] (rows x k)
// into Q (m x k, ld m).  root != 0: a single leaf, X = I.
__global__ __launch_bounds__(256) void k_tsqr_leaf_q(int m, int b, int k, const double2* __restrict__ P,
                                                     const double2* __restrict__ tau,
                                                     const double2* __restrict__ Xin, int root,
                                                     double2* __restrict__ Q, const int* __restrict__ gate)
{
    extern __shared__ double2 sY[];
    __shared__ double2 wbuf[8 * 32], stau[BR_B];
    if (gate && *gate == 0) return;
    const int nblk = tsqr_nblk(m);
    const int blk = blockIdx.x;
    const int r0 = tsqr_r0(blk), rows = tsqr_r1(blk, nblk, m) - r0;
    const int kv = min(rows, b);
    const int ld = tsqr_ld(rows);
    for (int e = threadIdx.x; e < rows * k; e += 256) {
        const int i = e % rows, j = e / rows;
        double2 x = cmk(0, 0);
        if (i < b) x = root ? cmk(i == j ? 1.0 : 0.0, 0.0) : Xin[(int64_t)blk * b * k + i + j * b];
        sY[i + j * ld] = x;
    }
    if (threadIdx.x < b) stau[threadIdx.x] = tau[(int64_t)blk * b + threadIdx.x];
    __syncthreads();
    lds_apply_q(rows, kv, P + r0, m, stau, sY, ld, k, wbuf);
    for (int e = threadIdx.x; e < rows * k; e += 256) {
        const int i = e % rows, j = e / rows;
        Q[(r0 + i) + (int64_t)j * m] = sY[i + j * ld];
    }
}

// ---------------------------------------------------------------------------
// Householder form of Q_thin (m x k, k = min(m, b)), from its top k x k block
// W (in LDS, W[i][c]):
//   LU of E - Q D with d_j = -conj(w_j)/|w_j| from the running Schur entry
//   w_j, so u_jj = 1 + |w_j| >= 1; U(j, c) = -d_c W(j, c);
//   L1 (unit lower) -> V rows 0..k-1;  T = U L1^{-H} -> Tp (ld b);
//   Mlow = -D U^{-1} (for the rows below: L2 = Q2 Mlow) -> Mlow (ld b), and
//   left in LDS as Ui (U^{-1}) and d for the caller;
//   R' = D^H Rfin (k x b, Rfin ld b) -> the band store, rows s.. of columns c0..
// 256 threads; W is overwritten.
// ---------------------------------------------------------------------------
typedef double2 hh_sq[BR_B][BR_B + 1];

__device__ inline void lds_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
}

// X = R^{-1} for an upper triangular R of order k <= 32 in LDS, by one wave:
// rows from the bottom, lane c forms X(i, c) from the rows below it (all
// reads of a row issued together).  R(i, l) = A[i][l], or conj(A[l][i])
// when lower != 0 (R = L^H for a lower L).  Entries outside the triangle
// (and rows, columns >= k) are zero.
__device__ inline void lds_upper_inv(hh_sq& A, bool lower, hh_sq& X, int k)
{
    const int c = threadIdx.x & 63;
    const int cc = c & (BR_B - 1);
    if (c < BR_B) {
#pragma unroll
        for (int i = 0; i < BR_B; ++i) X[i][c] = cmk(0, 0);
    }
    lds_wave_sync();
    for (int i = k - 1; i >= 0; --i) {
        double2 s = cmk(i == c ? 1.0 : 0.0, 0.0);
#pragma unroll 8
        for (int l = 0; l < BR_B; ++l) {
            const double2 r = lower ? cconj(A[l][i]) : A[i][l];
            const double2 x = X[l][cc];
            if (l > i && l < k) s = csub(s, cmul(r, x));
        }
        const double2 d = lower ? cconj(A[i][i]) : A[i][i];
        if (c < k && c >= i) X[i][c] = cmul(s, crecip(d));
        lds_wave_sync();
    }
}

// X = L^{-1} for a unit lower triangular L of order k <= 32, by one wave:
// rows from the top.
__device__ inline void lds_unit_lower_inv(hh_sq& L, hh_sq& X, int k)
{
    const int c = threadIdx.x & 63;
    const int cc = c & (BR_B - 1);
    if (c < BR_B) {
#pragma unroll
        for (int i = 0; i < BR_B; ++i) X[i][c] = cmk(0, 0);
    }
    lds_wave_sync();
    for (int i = 0; i < k; ++i) {
        double2 s = cmk(i == c ? 1.0 : 0.0, 0.0);
#pragma unroll 8
        for (int l = 0; l < BR_B; ++l) {
            const double2 x = X[l][cc];
            if (l < i) s = csub(s, cmul(L[i][l], x));
        }
        if (c < k && c <= i) X[i][c] = s;
        lds_wave_sync();
    }
}
// The LU runs in place in W: the multipliers (L1 below the unit diagonal)
// replace W's strict lower part, U (diagonal u_jj, upper -d_c W(j, c)) its
// upper part, so W, Ui and Li are the only LDS matrices (the kernels that
// call this fit beside a running trailing update).
__device__ __attribute__((always_inline)) inline void hh_reconstruct_lds(
    int m, int b, int k, hh_sq& W, hh_sq& Ui, hh_sq& Li, double2* d, double2* ud,
    const double2* __restrict__ Rfin, double2* __restrict__ V, double2* __restrict__ Tp,
    double2* __restrict__ Mlow, double2* __restrict__ AB, int ldab, int c0)
{
    const int tid = threadIdx.x;
    for (int j = 0; j < k; ++j) {
        const double2 wj = W[j][j];
        const double aw = sqrt(cabs2(wj));
        const double2 dj = (aw > 0.0) ? cmk(-wj.x / aw, wj.y / aw) : cmk(1.0, 0.0);
        const double2 ujj = csub(cmk(1.0, 0.0), cmul(dj, wj));
        if (tid == 0) {
            d[j] = dj;
            ud[j] = ujj;
        }
        // multipliers of column j (below the diagonal), then the Schur
        // update of the later columns
        if (tid > j && tid < k) W[tid][j] = cdiv(cmul(cmk(-dj.x, -dj.y), W[tid][j]), ujj);
        __syncthreads();
        for (int e = tid; e < BR_B * BR_B; e += 256) {
            const int i = e % BR_B, c = e / BR_B;
            if (i > j && i < k && c > j && c < k) W[i][c] = csub(W[i][c], cmul(W[i][j], W[j][c]));
        }
        __syncthreads();
    }
    // U on and above the diagonal
    for (int e = tid; e < BR_B * BR_B; e += 256) {
        const int j = e % BR_B, c = e / BR_B;
        if (j < c && c < k) W[j][c] = cmul(cmk(-d[c].x, -d[c].y), W[j][c]);
        else if (j == c && c < k) W[j][j] = ud[j];
    }
    __syncthreads();
    // U^{-1} (upper, wave 0) and L1^{-1} (unit lower, wave 1), by rows
    if (tid < 64) lds_upper_inv(W, false, Ui, k);
    else if (tid < 128) lds_unit_lower_inv(W, Li, k);
    __syncthreads();
    for (int e = tid; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, c = e / BR_B;
        if (i < k && c < k) {
            // V rows 0..k-1 = L1
            V[i + (int64_t)c * m] = (i == c) ? cmk(1.0, 0.0) : (i > c ? W[i][c] : cmk(0, 0));
            // T = U L1^{-H}:  T(i, c) = sum_{l >= i} U(i, l) conj(Li(c, l))
            double2 t = cmk(0, 0);
            for (int l = i; l < k; ++l) cfma(t, W[i][l], cconj(Li[c][l]));
            Tp[i + c * b] = t;
            // Mlow = -D U^{-1}: (-Q2 D) U^{-1} = Q2 (-D U^{-1})
            if (Mlow) Mlow[i + c * b] = cmul(cmk(-d[i].x, -d[i].y), Ui[i][c]);
        } else if (i < b && c < b) {
            Tp[i + c * b] = cmk(0, 0);
            if (Mlow) Mlow[i + c * b] = cmk(0, 0);
        }
    }
    // R' = D^H Rfin (k x b): rows s .. s + k - 1 of columns c0 .. c0 + b - 1
    for (int e = tid; e < k * b; e += 256) {
        const int i = e % k, j = e / k;
        const double2 r = cmul(cconj(d[i]), Rfin[i + j * b]);
        const int gr = c0 + b + i, gc = c0 + j;
        AB[(int64_t)gc * ldab + (gr - gc)] = r;
    }
}

// gate (may be null): the kernel runs only when *gate != 0 (the CholeskyQR2
// path below sets it when a panel needs the TSQR fallback).
__device__ inline bool gate_closed(const int* gate) { return gate && *gate == 0; }

__global__ __launch_bounds__(256) void k_hh_reconstruct(int m, int b, int k, const double2* __restrict__ Q,
                                                        const double2* __restrict__ Rfin,
                                                        double2* __restrict__ V, double2* __restrict__ Tp,
                                                        double2* __restrict__ Mlow, double2* __restrict__ AB,
                                                        int ldab, int c0, const int* __restrict__ gate)
{
    __shared__ hh_sq W, Ui, Li;
    __shared__ double2 d[BR_B], ud[BR_B];
    if (gate_closed(gate)) return;
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, c = e / BR_B;
        W[i][c] = (i < k && c < k) ? Q[i + (int64_t)c * m] : cmk(0, 0);
    }
    __syncthreads();
    hh_reconstruct_lds(m, b, k, W, Ui, Li, d, ud, Rfin, V, Tp, Mlow, AB, ldab, c0);
}

// Y (rows k0 .. m-1, ld m) = X M (mode 0) or Y -= X M (mode 1), X (ld m) with
// kx columns, M (kx x ny, ld b) staged in LDS.  Workgroup: 64 rows x 4 groups
// of 8 output columns (a wave reads 64 consecutive rows of a column).
__global__ __launch_bounds__(256) void k_rows_small(int m, int k0, int kx, int ny, int b,
                                                    const double2* __restrict__ X,
                                                    const double2* __restrict__ M, double2* __restrict__ Y,
                                                    int mode, const int* __restrict__ gate)
{
    __shared__ double2 sM[BR_B][BR_B + 1];
    if (gate_closed(gate)) return;
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, c = e / BR_B;
        sM[i][c] = (i < kx && c < ny) ? M[i + c * b] : cmk(0, 0);
    }
    __syncthreads();
    const int r = k0 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int cg = (threadIdx.x >> 6) * 8;
    if (r >= m || cg >= ny) return;
    double2 x[BR_B];
#pragma unroll
    for (int l = 0; l < BR_B; ++l) x[l] = (l < kx) ? X[r + (int64_t)l * m] : cmk(0, 0);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int c = cg + u;
        if (c >= ny) break;
        double2 s = mode ? Y[r + (int64_t)c * m] : cmk(0, 0);
        double2 t = cmk(0, 0);
#pragma unroll
        for (int l = 0; l < BR_B; ++l) cfma(t, x[l], sM[l][c]);
        Y[r + (int64_t)c * m] = mode ? csub(s, t) : t;
    }
}

// ---------------------------------------------------------------------------
// CholeskyQR2 of the m x 32 panel (the fast path; Fukaya, Nakatsukasa,
// Yanagisawa & Yamamoto 2014): G1 = P^H P = R1^H R1, Q1 = P R1^{-1},
// G2 = Q1^H Q1 = R2^H R2, Q = Q1 R2^{-1}, R = R2 R1.  Every pass over the
// panel is row-parallel; the two Cholesky factorisations and the Householder
// reconstruction run in one workgroup each.  The panel goes to the TSQR
// (flag = 1) when a pivot of G1 falls below 1e-12 of its diagonal entry
// (cond(P) beyond ~1e6, where one more pass no longer restores orthogonality
// to rounding level), when a column is zero, or when G2 is not within 0.1 of
// the identity.
// ---------------------------------------------------------------------------
#define CQ_G 64           // workgroups of the row-parallel passes

// Partial Gram matrices G = P^H P over each workgroup's 64-row chunks
// (chunks g, g + grid, ...), partial g (b x b, column-major) -> Gp + g b b.
// PASS 0 reads the panel and clears the three flags (flag[0]: the TSQR is
// needed; flag[1], flag[2]: the first, second factorisation succeeded);
// PASS 1 reads Q1 = P R1^{-1} and runs only when flag[1] is set.
template <int PASS>
__global__ __launch_bounds__(256) void k_cq_gram(int m, const double2* __restrict__ P,
                                                 double2* __restrict__ Gp, int* __restrict__ flag)
{
    __shared__ double2 sP[64][BR_B + 1];
    const int tid = threadIdx.x;
    if (PASS == 0) {
        if (blockIdx.x == 0 && tid == 0) { flag[0] = 0; flag[1] = 0; flag[2] = 0; }
    } else {
        if (flag[1] == 0) return;
    }
    double2 g[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) g[u] = cmk(0, 0);
    const int nch = (int)ceil_div64(m, 64);
    for (int ch = blockIdx.x; ch < nch; ch += gridDim.x) {
        const int r0 = ch * 64, rows = min(64, m - r0);
        __syncthreads();
        for (int e = tid; e < 64 * BR_B; e += 256) {
            const int r = e & 63, c = e >> 6;
            sP[r][c] = (r < rows) ? P[(r0 + r) + (int64_t)c * m] : cmk(0, 0);
        }
        __syncthreads();
        const int i = tid & 31, j0 = tid >> 5;
#pragma unroll 4
        for (int r = 0; r < 64; ++r) {
            const double2 pi = sP[r][i];
#pragma unroll
            for (int u = 0; u < 4; ++u) cfmac(g[u], pi, sP[r][j0 + 8 * u]);
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) Gp[(int64_t)blockIdx.x * BR_B * BR_B + tid + 256 * u] = g[u];
}

// G = sum of the ng partial Gram matrices (fixed order), one element per
// thread over 4 workgroups, 16 loads in flight per batch.  Runs when flag[f]
// is set (f < 0: always).
__global__ __launch_bounds__(256) void k_cq_reduce(int ng, const double2* __restrict__ Gp,
                                                   double2* __restrict__ G, const int* __restrict__ flag, int f)
{
    if (f >= 0 && flag[f] == 0) return;
    const int e = blockIdx.x * 256 + threadIdx.x;        // < BR_B * BR_B
    double2 acc = cmk(0, 0);
    for (int g0 = 0; g0 < ng; g0 += 16) {
        double2 t[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) t[q] = (g0 + q < ng) ? Gp[(int64_t)(g0 + q) * BR_B * BR_B + e] : cmk(0, 0);
#pragma unroll
        for (int q = 0; q < 16; ++q) acc = cadd(acc, t[q]);
    }
    G[e] = acc;
}

// The Gram matrix G (Gp, summed by k_cq_reduce) into sG, then its Cholesky
// factor in place (lower triangle).
// pass 1 also checks |G - I| <= 0.1.  Returns false (uniformly) on failure.
__device__ bool cq_sum_chol(int ng, const double2* __restrict__ Gp, int pass, hh_sq& sG, double* diag0,
                            int* fl)
{
    const int tid = threadIdx.x;
    if (tid == 0) *fl = 0;
    // G summed by k_cq_reduce (ng = 1: Gp is G itself)
    (void)ng;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = tid + 256 * u, i = e % BR_B, j = e / BR_B;
        const double2 x = Gp[e];
        sG[i][j] = x;
        if (i == j) diag0[i] = x.x;
    }
    __syncthreads();
    if (pass == 1) {
        for (int e = tid; e < BR_B * BR_B; e += 256) {
            const int i = e % BR_B, j = e / BR_B;
            const double2 x = csub(sG[i][j], cmk(i == j ? 1.0 : 0.0, 0.0));
            if (!(fabs(x.x) <= 0.1 && fabs(x.y) <= 0.1)) *fl = 1;
        }
        __syncthreads();
    }
    // right-looking Cholesky, the trailing update spread over the workgroup
    for (int j = 0; j < BR_B; ++j) {
        if (*fl) break;                                   // uniform
        const double dj = sG[j][j].x;
        if (!(dj > 1e-12 * diag0[j] && diag0[j] > 0.0)) {
            __syncthreads();
            if (tid == 0) *fl = 1;
            __syncthreads();
            break;
        }
        const double ljj = sqrt(dj);
        __syncthreads();                                  // every thread has read sG[j][j]
        if (tid > j && tid < BR_B) sG[tid][j] = cscale(sG[tid][j], 1.0 / ljj);
        if (tid == j) sG[j][j] = cmk(ljj, 0.0);
        __syncthreads();
        for (int e = tid; e < BR_B * BR_B; e += 256) {
            const int i = e % BR_B, c = e / BR_B;
            if (c > j && i >= c) sG[i][c] = csub(sG[i][c], cmul(sG[i][j], cconj(sG[c][j])));
        }
        __syncthreads();
    }
    __syncthreads();
    return *fl == 0;
}

// Ri = R^{-1} for R = L^H (L lower in sL).
__device__ void cq_rinv(hh_sq& sL, hh_sq& Ri)
{
    if (threadIdx.x < 64) lds_upper_inv(sL, true, Ri, BR_B);
    __syncthreads();
}

// First factorisation: R1 (upper, ld b) and R1^{-1} to global.
__global__ __launch_bounds__(256) void k_cq_chol1(int ng, const double2* __restrict__ Gp,
                                                  double2* __restrict__ R1, double2* __restrict__ R1i,
                                                  int* __restrict__ flag, int* __restrict__ fail)
{
    __shared__ hh_sq sG, Ri;
    __shared__ double diag0[BR_B];
    __shared__ int fl;
    if (!cq_sum_chol(ng, Gp, 0, sG, diag0, &fl)) {
        if (threadIdx.x == 0) {
            flag[0] = 1;
            *fail = 1;
        }
        return;
    }
    cq_rinv(sG, Ri);
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, j = e / BR_B;
        R1[e] = (i <= j) ? cconj(sG[j][i]) : cmk(0, 0);
        R1i[e] = Ri[i][j];
    }
    if (threadIdx.x == 0) flag[1] = 1;
}

// Second factorisation and the Householder form: R = R2 R1 -> Rf; the top
// block W = Q1(0:32, :) R2^{-1}; hh_reconstruct_lds; Mp = R2^{-1} Mlow for
// the rows below (V = Q1 Mp there).  Nothing is written when it fails.
__global__ __launch_bounds__(256) void k_cq_chol2(int m, int b, int ng, const double2* __restrict__ Gp,
                                                  const double2* __restrict__ Q1, const double2* __restrict__ R1,
                                                  double2* __restrict__ Rf, double2* __restrict__ V,
                                                  double2* __restrict__ Tp, double2* __restrict__ Mp,
                                                  double2* __restrict__ AB, int ldab, int c0,
                                                  int* __restrict__ flag, int* __restrict__ fail)
{
    __shared__ hh_sq W, Ui, Li;
    __shared__ double2 d[BR_B], ud[BR_B];
    __shared__ double diag0[BR_B];
    __shared__ int fl;
    if (flag[1] == 0) return;
    // Cholesky of G2 in W, R2^{-1} in Li
    if (!cq_sum_chol(ng, Gp, 1, W, diag0, &fl)) {
        if (threadIdx.x == 0) {
            flag[0] = 1;
            *fail = 1;
        }
        return;
    }
    cq_rinv(W, Li);
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) Ui[e % BR_B][e / BR_B] = R1[e];
    __syncthreads();
    // R = R2 R1 (both upper; R2(i, l) = conj(W(l, i))) into Rf
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, j = e / BR_B;
        double2 s = cmk(0, 0);
        for (int l = i; l <= j; ++l) cfma(s, cconj(W[l][i]), Ui[l][j]);
        Rf[e] = (i <= j) ? s : cmk(0, 0);
    }
    __syncthreads();
    // the top 32 x 32 block of Q1 into Ui, then the top block of
    // Q = Q1 R2^{-1} into W; R2^{-1} is kept in registers (4 entries per
    // thread) while the reconstruction uses Li
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, j = e / BR_B;
        Ui[i][j] = Q1[i + (int64_t)j * m];
    }
    __syncthreads();
    double2 r2i[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = threadIdx.x + 256 * u, i = e % BR_B, j = e / BR_B;
        double2 t = cmk(0, 0);
        for (int l = 0; l <= j; ++l) cfma(t, Ui[i][l], Li[l][j]);
        W[i][j] = t;
        r2i[u] = Li[i][j];
    }
    __syncthreads();
    hh_reconstruct_lds(m, b, BR_B, W, Ui, Li, d, ud, Rf, V, Tp, nullptr, AB, ldab, c0);
    __syncthreads();
    // Mlow = -D U^{-1} into W, R2^{-1} back into Li, then Mp = R2^{-1} Mlow
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, j = e / BR_B;
        W[i][j] = cmul(cmk(-d[i].x, -d[i].y), Ui[i][j]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = threadIdx.x + 256 * u;
        Li[e % BR_B][e / BR_B] = r2i[u];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, j = e / BR_B;
        double2 s = cmk(0, 0);
        for (int l = i; l < BR_B; ++l) cfma(s, Li[i][l], W[l][j]);
        Mp[i + j * b] = s;
    }
    if (threadIdx.x == 0) flag[2] = 1;
}

// ---------------------------------------------------------------------------
// X = A22 V for the trailing block A22 = A[s:, s:] over its storage tiles
// (tile rows and columns tcs = s / 64 ..).  Workgroup (i, q) owns tile row
// I = tcs + i and the q-th of KSPLIT slices of the nT tiles that row i of the
// full Hermitian block touches: tiles (I, J) for J <= I (stored) and the
// conjugate transposes of the stored tiles (J, I) for J > I.  Its 64 x k
// partial goes to Xp + q (nT 64) k.  V is read as zero above row s, so the
// first tile row and column need no masks; rows of X above s are not used.
// Diagonal tiles contribute (L + L^H - diag): the stored strict upper half of
// a diagonal tile is zero (storage invariant).
// 16 x 16 threads: rows tx + 16 a (a < 4), columns ty + 16 c (c < 2).
// ---------------------------------------------------------------------------
#ifndef HEMM_KSPLIT
#define HEMM_KSPLIT 16
#endif
template <typename T>
__global__ __launch_bounds__(256) void k_band_hemm(int n, int s, int k, const T* __restrict__ A,
                                                   const double2* __restrict__ V, int64_t ldv,
                                                   double2* __restrict__ Xp)
{
    __shared__ double2 sA[8][65];     // sA[kk][row]: 8 contraction indices x 64 rows
    __shared__ double2 sV[8][BR_B];   // sV[kk][col]
    const int nt = tile_count(n), tcs = s >> 6, nT = nt - tcs;
    const int i = blockIdx.x, q = blockIdx.y;
    const int I = tcs + i;
    const int j0 = (int)((int64_t)nT * q / HEMM_KSPLIT), j1 = (int)((int64_t)nT * (q + 1) / HEMM_KSPLIT);
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double2 acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a) { acc[a][0] = cmk(0, 0); acc[a][1] = cmk(0, 0); }
    for (int jt = j0; jt < j1; ++jt) {
        const int J = tcs + jt;
        const bool tr = J > I;                       // use conj(A(J, I))^T
        const T* tile = A + tile_pos(tr ? J : I, tr ? I : J, nt) * (BSE_TILE * BSE_TILE);
        const bool diag = (J == I);
        for (int kc = 0; kc < 64; kc += 8) {
            // stage 8 contraction columns: A_IJ(row, kc + kk)
            for (int e = threadIdx.x; e < 8 * 64; e += 256) {
                int kk, row;
                double2 x;
                if (!tr) {
                    row = e & 63; kk = e >> 6;
                    x = cplx<T>::load(tile + row + (kc + kk) * 64);
                    if (diag) {
                        // (L + L^H - diag): above the diagonal take conj(L(col, row))
                        if (row < kc + kk) x = cconj(cplx<T>::load(tile + (kc + kk) + row * 64));
                        else if (row == kc + kk) x.y = 0.0;
                    }
                } else {
                    // A_IJ(row, col) = conj(A_JI(col, row)): read 8 consecutive col entries
                    kk = e & 7; row = e >> 3;
                    x = cconj(cplx<T>::load(tile + (kc + kk) + row * 64));
                }
                sA[kk][row] = x;
            }
            for (int e = threadIdx.x; e < 8 * BR_B; e += 256) {
                const int kk = e / BR_B, c = e % BR_B;
                const int64_t gr = (int64_t)J * 64 + kc + kk;   // global row of V
                sV[kk][c] = (c < k && gr >= s && gr < n) ? V[(gr - s) + (int64_t)c * ldv] : cmk(0, 0);
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) {
                const double2 v0 = sV[kk][ty], v1 = sV[kk][ty + 16];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const double2 x = sA[kk][tx + 16 * a];
                    cfma(acc[a][0], x, v0);
                    cfma(acc[a][1], x, v1);
                }
            }
            __syncthreads();
        }
    }
    double2* out = Xp + (int64_t)q * nT * 64 * BR_B + (int64_t)i * 64;
    const int64_t ldx = (int64_t)nT * 64;
    (void)ldx;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        out[(tx + 16 * a) + (int64_t)ty * nT * 64] = acc[a][0];
        out[(tx + 16 * a) + (int64_t)(ty + 16) * nT * 64] = acc[a][1];
    }
}

// ---------------------------------------------------------------------------
// The same product for FP32 storage on matrix cores (v_mfma_f32_16x16x4_f32,
// FP32 operands and accumulation; the KSPLIT partials are summed in FP64 by
// k_band_xt).  Workgroup (i, q) as in k_band_hemm; wave w owns rows 16 w ..
// 16 w + 15 of the 64-row block and all 32 output columns.  Lanes 16 h ..
// 16 h + 15 take complex k = 4 step + h, and the four real products
//   Re X = sum Re A Re V - Im A Im V,   Im X = sum Re A Im V + Im A Re V
// have accumulators of their own, so every operand read serves four MFMAs
// and no lane selects or negates.  The tile A_IJ (or A_JI^H) and the V block
// are staged in LDS; the next tile's loads are issued into registers before
// the current tile's products, and within the products each step's operands
// are read one step ahead.
// ---------------------------------------------------------------------------
// FP32 copy of the panel's V (m x k, ld m) for the FP32 block product, row
// major with 32 entries per row (zero beyond k): the 64 rows a tile needs
// are one contiguous 16 KB run.
__global__ __launch_bounds__(256) void k_v_rows_f32(int m, int k, const double2* __restrict__ V,
                                                    float2* __restrict__ Vr)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;     // row-major element
    if (e >= (int64_t)m * BR_B) return;
    const int r = (int)(e / BR_B), c = (int)(e % BR_B);
    const double2 v = (c < k) ? V[r + (int64_t)c * m] : cmk(0, 0);
    Vr[e] = make_float2((float)v.x, (float)v.y);
}

typedef float hemm_f4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_band_hemm32(int n, int s, int k, const float2* __restrict__ A,
                                                     const float2* __restrict__ V, int64_t ldv,
                                                     double2* __restrict__ Xp)     // V: FP32, row major (32 per row)
{
    // A tile (64 x 64) and V block (64 x 32) in LDS with XOR swizzles so
    // that the four 16-lane groups of an operand read hit disjoint banks:
    // A_IJ(r, kk) at sA[kk][r ^ 16 (kk & 1)] for a stored tile, and at
    // sA[r][kk ^ 2 (r & 15)] for a transposed one (J > I, read from A_JI);
    // V(kk, c) at sV[kk][c ^ 16 (kk & 1)].
    __shared__ float2 sA[64 * 64];
    __shared__ float2 sV[64 * BR_B];
    const int nt = tile_count(n), tcs = s >> 6, nT = nt - tcs;
    const int i = blockIdx.x, q = blockIdx.y;
    const int I = tcs + i;
    const int j0 = (int)((int64_t)nT * q / HEMM_KSPLIT), j1 = (int)((int64_t)nT * (q + 1) / HEMM_KSPLIT);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int h = lane >> 4, l16 = lane & 15;
    const int row = 16 * w + l16;
    // accumulators per column block c: Re A Re V, Im A Im V, Re A Im V,
    // Im A Re V (each MFMA takes four distinct complex k, one per 16-lane
    // group); per tile their FP32 partials (64 real products) go into FP64
    // sums X = (RR - II) + i (RI + IR)
    hemm_f4 acc[2][4];
    double dR[2][4], dI[2][4];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[c][t] = hemm_f4{0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 4; ++r) { dR[c][r] = 0.0; dI[c][r] = 0.0; }
    }
    float2 ra[16], rv[8];
    // loads of tile jt into registers: element e = tid + 256 u
    auto fetch = [&](int jt) {
        const int J = tcs + jt;
        const bool tr = J > I;
        const float2* tile = A + tile_pos(tr ? J : I, tr ? I : J, nt) * (BSE_TILE * BSE_TILE);
#pragma unroll
        for (int u = 0; u < 16; ++u) ra[u] = tile[tid + 256 * u];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = tid + 256 * u, kk = e >> 5;
            const int64_t gr = (int64_t)J * 64 + kk;
            rv[u] = (gr >= s && gr < n) ? V[(gr - s) * BR_B + (e & 31)] : make_float2(0.f, 0.f);
        }
    };
    if (j0 < j1) fetch(j0);
    const int swn = 16 * (h & 1);                 // kk & 1 = h & 1 for every kk this lane reads
    const int aoff = h * 64 + (row ^ swn);        // stored tile: + 256 st
    const int swt = 2 * (row & 15);               // transposed tile
    const int voff0 = h * BR_B + (l16 ^ swn), voff1 = h * BR_B + ((16 + l16) ^ swn);   // + 128 st
    for (int jt = j0; jt < j1; ++jt) {
        const int J = tcs + jt;
        const bool tr = J > I, diag = (J == I);
        __syncthreads();    // the previous tile's products are done with sA, sV
        if (!tr) {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = tid + 256 * u, r = e & 63, kk = e >> 6;
                float2 x = ra[u];
                if (diag && r == kk) x.y = 0.0f;
                if (!(diag && r < kk)) sA[kk * 64 + (r ^ (16 * (kk & 1)))] = x;
                // on a diagonal tile the stored element (r, kk), r > kk, also
                // gives A(kk, r) = conj(A(r, kk))
                if (diag && r > kk) sA[r * 64 + (kk ^ (16 * (r & 1)))] = make_float2(x.x, -x.y);
            }
        } else {
            // element e = kk + 64 r of A_JI: A_IJ(r, kk) = conj(A_JI(kk, r))
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int e = tid + 256 * u, kk = e & 63, r = e >> 6;
                sA[r * 64 + (kk ^ (2 * (r & 15)))] = make_float2(ra[u].x, -ra[u].y);
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = tid + 256 * u, kk = e >> 5, cc = e & 31;
            sV[kk * BR_B + (cc ^ (16 * (kk & 1)))] = rv[u];
        }
        __syncthreads();
        if (jt + 1 < j1) fetch(jt + 1);
        // the products, one loop per layout (no branch inside), operands
        // read one step ahead
        auto mloop = [&](auto trc) {
            constexpr bool TRT = decltype(trc)::value;
            float2 a[2], v[2][2];
            auto ld = [&](int st, int slot) {
                a[slot] = TRT ? sA[row * 64 + ((4 * st + h) ^ swt)] : sA[aoff + 256 * st];
                v[slot][0] = sV[voff0 + 128 * st];
                v[slot][1] = sV[voff1 + 128 * st];
            };
            ld(0, 0);
#pragma unroll
            for (int st = 0; st < 16; ++st) {
                const int sl = st & 1;
                if (st + 1 < 16) ld(st + 1, sl ^ 1);
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    acc[c][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[sl].x, v[sl][c].x, acc[c][0], 0, 0, 0);
                    acc[c][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[sl].y, v[sl][c].y, acc[c][1], 0, 0, 0);
                    acc[c][2] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[sl].x, v[sl][c].y, acc[c][2], 0, 0, 0);
                    acc[c][3] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[sl].y, v[sl][c].x, acc[c][3], 0, 0, 0);
                }
            }
        };
        if (tr) mloop(std::true_type{});
        else mloop(std::false_type{});
#pragma unroll
        for (int c = 0; c < 2; ++c) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                dR[c][r] += (double)acc[c][0][r] - (double)acc[c][1][r];
                dI[c][r] += (double)acc[c][2][r] + (double)acc[c][3][r];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[c][t] = hemm_f4{0, 0, 0, 0};
        }
    }
    // lane, reg r: X(16 w + 4 h + r, 16 c + l16)
    double2* out = Xp + (int64_t)q * nT * 64 * BR_B + (int64_t)i * 64;
    const int64_t ldx = (int64_t)nT * 64;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            out[(16 * w + 4 * h + r) + (int64_t)(16 * c + l16) * ldx] = cmk(dR[c][r], dI[c][r]);
}

// ---------------------------------------------------------------------------
// Z = X T - V M with M = (1/2) T^H (V^H X T).  Rows of X are tile-aligned in
// Xp (row t of the trailing tile range <-> global row tcs 64 + t); V rows
// start at s.
//   k_band_xt:   XT = (sum of the KSPLIT partials) T for 64 rows per block,
//                into Z (m x b, ld m), and the partial V^H XT of the block
//   k_band_m:    M from the partials (one workgroup, fixed order)
//   k_rows_small (mode 1): Z -= V M
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_band_xt(int n, int s, int k, int b, const double2* __restrict__ Xp,
                                                 const double2* __restrict__ Tp, const double2* __restrict__ V,
                                                 double2* __restrict__ Z, double2* __restrict__ Gp)
{
    __shared__ double2 sT[BR_B][BR_B + 1];
    __shared__ double2 sX[64][BR_B + 1];
    __shared__ double2 sV[64][BR_B + 1];
    const int m = n - s, tcs = s >> 6, nT = tile_count(n) - tcs;
    const int r0 = blockIdx.x * 64;                 // local row of Z (global s + r0)
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, j = e / BR_B;
        sT[i][j] = (i < k && j < k) ? Tp[i + j * b] : cmk(0, 0);
    }
    for (int e = threadIdx.x; e < 64 * BR_B; e += 256) {
        const int rr = e & 63, c = e >> 6;
        const int r = r0 + rr;
        double2 x = cmk(0, 0), v = cmk(0, 0);
        if (r < m && c < k) {
            const int64_t t = (int64_t)(s + r) - (int64_t)tcs * 64;
            for (int qq = 0; qq < HEMM_KSPLIT; ++qq)
                x = cadd(x, Xp[(int64_t)qq * nT * 64 * BR_B + t + (int64_t)c * nT * 64]);
            v = V[r + (int64_t)c * m];
        }
        sX[rr][c] = x;
        sV[rr][c] = v;
    }
    __syncthreads();
    // XT rows: every thread forms its 8 entries from the X rows, then they
    // replace X in LDS once all reads are done
    constexpr int PER = 64 * BR_B / 256;
    double2 xt[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = threadIdx.x + 256 * u, rr = e & 63, c = e >> 6;
        const int r = r0 + rr;
        double2 acc = cmk(0, 0);
        for (int l = 0; l < k; ++l) cfma(acc, sX[rr][l], sT[l][c]);
        if (r < m && c < k) Z[r + (int64_t)c * m] = acc;
        xt[u] = (r < m && c < k) ? acc : cmk(0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int e = threadIdx.x + 256 * u;
        sX[e & 63][e >> 6] = xt[u];
    }
    __syncthreads();
    // partial G = V^H XT over the block's rows (k x k)
    for (int e = threadIdx.x; e < BR_B * BR_B; e += 256) {
        const int i = e % BR_B, j = e / BR_B;
        double2 acc = cmk(0, 0);
        for (int rr = 0; rr < 64; ++rr) cfmac(acc, sV[rr][i], sX[rr][j]);
        Gp[(int64_t)blockIdx.x * BR_B * BR_B + e] = acc;
    }
}

// One workgroup per column j < k of G: the nblk partials summed in eight
// strided groups, the groups in order, then M(:, j) = (1/2) T^H G(:, j).
__global__ __launch_bounds__(256) void k_band_m(int nblk, int k, int b, const double2* __restrict__ Gp,
                                                const double2* __restrict__ Tp, double2* __restrict__ M)
{
    __shared__ double2 part[8][BR_B];
    __shared__ double2 sG[BR_B];
    const int j = blockIdx.x;
    const int l = threadIdx.x & 31, q = threadIdx.x >> 5;
    double2 acc = cmk(0, 0);
    for (int g = q; g < nblk; g += 8) acc = cadd(acc, Gp[(int64_t)g * BR_B * BR_B + l + j * BR_B]);
    part[q][l] = acc;
    __syncthreads();
    if (threadIdx.x < BR_B) {
        double2 t = part[0][l];
#pragma unroll
        for (int u = 1; u < 8; ++u) t = cadd(t, part[u][l]);
        sG[l] = t;
    }
    __syncthreads();
    const int i = threadIdx.x;
    if (i < k) {
        double2 t = cmk(0, 0);
        for (int ll = 0; ll < k; ++ll) cfmac(t, Tp[ll + i * b], sG[ll]);   // (T^H G)(i, j)
        M[i + j * b] = cscale(t, 0.5);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
extern "C" hipError_t bse_band_gather(hipStream_t st, int fp32, int n, int c0, int b, const void* A,
                                      double2* P, double2* AB, int ldab)
{
    const int m = n - c0 - b;
    const unsigned g = (unsigned)std::max<int64_t>(1, ceil_div64(std::max(m, 1), 256));
    if (fp32)
        hipLaunchKernelGGL(k_band_gather<float2>, dim3(g), dim3(256), 0, st, n, c0, b, (const float2*)A, P, AB, ldab);
    else
        hipLaunchKernelGGL(k_band_gather<double2>, dim3(g), dim3(256), 0, st, n, c0, b, (const double2*)A, P, AB, ldab);
    return hipGetLastError();
}

extern "C" hipError_t bse_band_gather_tail(hipStream_t st, int fp32, int n, int c0, const void* A,
                                           double2* AB, int ldab)
{
    if (c0 >= n) return hipSuccess;
    const int w = n - c0;
    const unsigned g = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, ceil_div64((int64_t)w * w, 256)));
    if (fp32)
        hipLaunchKernelGGL(k_band_gather_tail<float2>, dim3(g), dim3(256), 0, st, n, c0, (const float2*)A, AB, ldab);
    else
        hipLaunchKernelGGL(k_band_gather_tail<double2>, dim3(g), dim3(256), 0, st, n, c0, (const double2*)A, AB, ldab);
    return hipGetLastError();
}

// Workspace of one panel QR of an m-row panel (complex elements): leaf R and
// tau, tree levels (V, tau, R per node; X pieces), the thin Q, then the
// CholeskyQR2 partial Gram matrices, R1, R1^{-1}, R, Mp and two flags.
extern "C" int64_t bse_tsqr_workspace(int m, int b) {
    int64_t el = 0;
    int nin = tsqr_nblk(m);
    el += (int64_t)nin * b * b + (int64_t)nin * b;       // leaf R, tau
    el += 2 * (int64_t)nin * b * b;                       // X pieces (two buffers)
    while (nin > 1) {
        const int nn = (int)ceil_div64(nin, TSQR_G);
        el += (int64_t)nn * TSQR_G * b * b + (int64_t)nn * b + (int64_t)nn * b * b;
        nin = nn;
    }
    el += (int64_t)m * b;                                 // Q_thin (Q1 on the fast path)
    el += (int64_t)CQ_G * BR_B * BR_B + 4 * BR_B * BR_B + 1;
    return el + 64;
}

// QR of the m x b panel P (ld m) as I - V T V^H: V (m x k) into Vout (ld m),
// T into Tp (ld b), R' = D^H R into the band store at (c0 + b.., c0..).
// Returns k = min(m, b) in *kout.  Panels of 32 columns and at least 64
// rows take the CholeskyQR2 path.  With tsqr != 0 the TSQR kernels run
// after it and return at once unless it flagged the panel (no host round
// trip decides); with tsqr == 0 they are not launched, and a panel whose
// Gram matrix is not numerically positive definite sets *fail (never
// cleared here) and leaves V, T and R unwritten: the caller then repeats the
// whole reduction with tsqr = 1.  Other panels always take the TSQR.
extern "C" hipError_t bse_band_panel_qr(hipStream_t st, int m, int b, double2* P, double2* Vout,
                                        double2* Tp, double2* ws, double2* AB, int ldab, int c0, int* kout,
                                        int* fail, int tsqr)
{
    if (b > BR_B || m < 1) return hipErrorInvalidValue;
    // the largest in-LDS QR holds TSQR_MAXROWS x 32 complex doubles (128 KB);
    // the limit asked for must leave room for the kernels' static LDS
    static const hipError_t attr = [] {
        const int big = tsqr_ld(TSQR_MAXROWS) * BR_B * (int)sizeof(double2);
        hipError_t e = hipFuncSetAttribute((const void*)k_tsqr_leaf, hipFuncAttributeMaxDynamicSharedMemorySize, big);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_tsqr_node, hipFuncAttributeMaxDynamicSharedMemorySize, big);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_tsqr_node_q, hipFuncAttributeMaxDynamicSharedMemorySize, big);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_tsqr_leaf_q, hipFuncAttributeMaxDynamicSharedMemorySize, big);
        return e;
    }();
    if (attr != hipSuccess) return attr;
    const int k = std::min(m, b);
    *kout = k;
    const int nleaf = tsqr_nblk(m);
    // workspace layout (mirrors bse_tsqr_workspace)
    double2* p = ws;
    double2* Rleaf = p; p += (int64_t)nleaf * b * b;
    double2* tauleaf = p; p += (int64_t)nleaf * b;
    double2* Xa = p; p += (int64_t)nleaf * b * b;
    double2* Xb = p; p += (int64_t)nleaf * b * b;
    struct Lvl { int nin, nn; double2 *V, *tau, *R; };
    Lvl lv[16];
    int nl = 0, nin = nleaf;
    while (nin > 1) {
        Lvl L;
        L.nin = nin;
        L.nn = (int)ceil_div64(nin, TSQR_G);
        L.V = p; p += (int64_t)L.nn * TSQR_G * b * b;
        L.tau = p; p += (int64_t)L.nn * b;
        L.R = p; p += (int64_t)L.nn * b * b;
        lv[nl++] = L;
        nin = L.nn;
    }
    double2* Q = p; p += (int64_t)m * b;
    double2* Gp = p; p += (int64_t)CQ_G * BR_B * BR_B;
    double2* R1 = p; p += BR_B * BR_B;
    double2* R1i = p; p += BR_B * BR_B;
    double2* Rf = p; p += BR_B * BR_B;
    double2* Mp = p; p += BR_B * BR_B;
    int* flags = reinterpret_cast<int*>(p);     // [0]: TSQR needed, [1], [2]: factorisations done
    const bool fast = (b == BR_B && m >= 2 * BR_B);
    const int* gate = nullptr;
    if (fast) {
        const int g = (int)std::min<int64_t>(CQ_G - 1, ceil_div64(m, 64));
        double2* Gs = Gp + (int64_t)(CQ_G - 1) * BR_B * BR_B;    // the summed G (last partial slot is spare)
        hipLaunchKernelGGL(k_cq_gram<0>, dim3((unsigned)g), dim3(256), 0, st, m, (const double2*)P, Gp, flags);
        hipLaunchKernelGGL(k_cq_reduce, dim3(BR_B * BR_B / 256), dim3(256), 0, st, g, (const double2*)Gp, Gs,
                           (const int*)flags, -1);
        hipLaunchKernelGGL(k_cq_chol1, dim3(1), dim3(256), 0, st, 1, (const double2*)Gs, R1, R1i, flags, fail);
        hipLaunchKernelGGL(k_rows_small, dim3((unsigned)ceil_div64(m, 64)), dim3(256), 0, st, m, 0, BR_B, BR_B, BR_B,
                           (const double2*)P, (const double2*)R1i, Q, 0, (const int*)(flags + 1));
        hipLaunchKernelGGL(k_cq_gram<1>, dim3((unsigned)g), dim3(256), 0, st, m, (const double2*)Q, Gp, flags);
        hipLaunchKernelGGL(k_cq_reduce, dim3(BR_B * BR_B / 256), dim3(256), 0, st, g, (const double2*)Gp, Gs,
                           (const int*)flags, 1);
        hipLaunchKernelGGL(k_cq_chol2, dim3(1), dim3(256), 0, st, m, b, 1, (const double2*)Gs,
                           (const double2*)Q, (const double2*)R1, Rf, Vout, Tp, Mp, AB, ldab, c0, flags, fail);
        hipLaunchKernelGGL(k_rows_small, dim3((unsigned)ceil_div64(m - k, 64)), dim3(256), 0, st, m, k, BR_B, k, b,
                           (const double2*)Q, (const double2*)Mp, Vout, 0, (const int*)(flags + 2));
        gate = flags;
        if (!tsqr) return hipGetLastError();
    }
    // TSQR: leaf QRs
    const int maxrows_leaf = (nleaf == 1) ? m : 2 * TSQR_RB - 1;
    if (maxrows_leaf > TSQR_MAXROWS) return hipErrorInvalidValue;
    const size_t lds_leaf = (size_t)tsqr_ld(maxrows_leaf) * b * sizeof(double2);
    hipLaunchKernelGGL(k_tsqr_leaf, dim3((unsigned)nleaf), dim3(256), lds_leaf, st, m, b, P, Rleaf, tauleaf, gate);
    // up the tree
    for (int l = 0; l < nl; ++l) {
        const size_t lds = (size_t)tsqr_ld(TSQR_G * b) * b * sizeof(double2);
        hipLaunchKernelGGL(k_tsqr_node, dim3((unsigned)lv[l].nn), dim3(256), lds, st, lv[l].nin, b,
                           l == 0 ? Rleaf : lv[l - 1].R, lv[l].V, lv[l].tau, lv[l].R, gate);
    }
    const double2* Rfin = nl ? lv[nl - 1].R : Rleaf;
    // down the tree
    double2* Xcur = Xa;
    double2* Xnext = Xb;
    for (int l = nl - 1; l >= 0; --l) {
        const size_t lds = (size_t)tsqr_ld(TSQR_G * b) * k * sizeof(double2);
        hipLaunchKernelGGL(k_tsqr_node_q, dim3((unsigned)lv[l].nn), dim3(256), lds, st, lv[l].nin, b, k,
                           lv[l].V, lv[l].tau, (const double2*)Xcur, l == nl - 1 ? 1 : 0, Xnext, gate);
        std::swap(Xcur, Xnext);
    }
    hipLaunchKernelGGL(k_tsqr_leaf_q, dim3((unsigned)nleaf), dim3(256), (size_t)tsqr_ld(maxrows_leaf) * k * sizeof(double2),
                       st, m, b, k, (const double2*)P, (const double2*)tauleaf, (const double2*)Xcur,
                       nl == 0 ? 1 : 0, Q, gate);
    // Householder form
    double2* Mlow = Xnext;   // b x b scratch
    hipLaunchKernelGGL(k_hh_reconstruct, dim3(1), dim3(256), 0, st, m, b, k, (const double2*)Q, Rfin, Vout, Tp,
                       Mlow, AB, ldab, c0, gate);
    if (m > k)
        hipLaunchKernelGGL(k_rows_small, dim3((unsigned)ceil_div64(m - k, 64)), dim3(256), 0, st, m, k, k, k, b,
                           (const double2*)Q, (const double2*)Mlow, Vout, 0, gate);
    return hipGetLastError();
}

// Workspace of the two-sided update (complex elements) for trailing order m.
extern "C" int64_t bse_band_update_workspace(int n, int s, int b) {
    const int nT = tile_count(n) - (s >> 6);
    const int m = n - s;
    return (int64_t)HEMM_KSPLIT * nT * 64 * BR_B + (int64_t)ceil_div64(m, 64) * BR_B * BR_B + 2 * BR_B * BR_B
           + (int64_t)m * b + ((int64_t)m * BR_B + 1) / 2     // ..., Z, FP32 copy of V
           + (int64_t)m * BR_B + 64;                            // FP32 planes of V and Z
}

// The rank-2k update's operands in FP32 planes: pl[q k m + l m + r] for
// plane q = Re V, Im V, Re Z, Im Z (V, Z: m x k, ld m), rounded as the update
// stages them in LDS.
__global__ __launch_bounds__(256) void k_her2k_pack(int m, int k, const double2* __restrict__ V,
                                                    const double2* __restrict__ Z, float* __restrict__ pl)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x, mk = (int64_t)m * k;
    if (e >= mk) return;
    const double2 v = V[e], z = Z[e];
    pl[e] = (float)v.x;
    pl[mk + e] = (float)v.y;
    pl[2 * mk + e] = (float)z.x;
    pl[3 * mk + e] = (float)z.y;
}

// A22 <- Q^H A22 Q for Q = I - V T V^H (V: m x k, ld m, rows from s).
// her2k: the one-stage trailing-update launcher (kind as tuned); kind 3
// (FP32 storage): the block product and the update both on FP32 matrix cores.
extern "C" hipError_t bse_launch_her2k(hipStream_t, int, int, int, int, void*, const double2*,
                                       const double2*, int64_t, int, int, const double2*);
// The two-sided update's first half: Z = X T - 1/2 V (T^H V^H X T) with
// X = A22 V.  ev (may be null): three events recorded before and after the
// Hermitian block product and after Z.  Returns Z's address in *zout; with
// pout non-null also V and Z packed in FP32 planes (k_her2k_pack) at *pout.
extern "C" hipError_t bse_band_product(hipStream_t st, int fp32, int kind, int n, int s, int k, int b,
                                       const void* A, const double2* V, const double2* Tp, double2* ws,
                                       hipEvent_t* ev, double2** zout, float** pout)
{
    const int m = n - s;
    const int tcs = s >> 6, nT = tile_count(n) - tcs;
    double2* Xp = ws;
    double2* Gp = Xp + (int64_t)HEMM_KSPLIT * nT * 64 * BR_B;
    const int nblk = (int)ceil_div64(m, 64);
    double2* M = Gp + (int64_t)nblk * BR_B * BR_B;
    double2* Z = M + 2 * BR_B * BR_B;
    *zout = Z;
    if (m <= 0 || k <= 0) return hipSuccess;
    const dim3 gh((unsigned)nT, HEMM_KSPLIT);
    if (ev) (void)hipEventRecord(ev[0], st);
    if (fp32 && kind == 3) {
        float2* Vf = reinterpret_cast<float2*>(Z + (int64_t)m * b);     // FP32 copy of V after Z
        hipLaunchKernelGGL(k_v_rows_f32, dim3((unsigned)ceil_div64((int64_t)m * BR_B, 256)), dim3(256), 0, st,
                           m, k, V, Vf);
        hipLaunchKernelGGL(k_band_hemm32, gh, dim3(256), 0, st, n, s, k, (const float2*)A, (const float2*)Vf,
                           (int64_t)m, Xp);
    }
    else if (fp32)
        hipLaunchKernelGGL(k_band_hemm<float2>, gh, dim3(256), 0, st, n, s, k, (const float2*)A, V, (int64_t)m, Xp);
    else
        hipLaunchKernelGGL(k_band_hemm<double2>, gh, dim3(256), 0, st, n, s, k, (const double2*)A, V, (int64_t)m, Xp);
    if (ev) (void)hipEventRecord(ev[1], st);
    hipLaunchKernelGGL(k_band_xt, dim3((unsigned)nblk), dim3(256), 0, st, n, s, k, b, (const double2*)Xp, Tp, V, Z, Gp);
    hipLaunchKernelGGL(k_band_m, dim3((unsigned)k), dim3(256), 0, st, nblk, k, b, (const double2*)Gp, Tp, M);
    hipLaunchKernelGGL(k_rows_small, dim3((unsigned)ceil_div64(m, 64)), dim3(256), 0, st, m, 0, k, k, b, V,
                       (const double2*)M, Z, 1, (const int*)nullptr);
    if (pout) {
        float* pl = reinterpret_cast<float*>(Z + (int64_t)m * b + ((int64_t)m * BR_B + 1) / 2);
        hipLaunchKernelGGL(k_her2k_pack, dim3((unsigned)ceil_div64((int64_t)m * k, 256)), dim3(256), 0, st, m, k,
                           V, (const double2*)Z, pl);
        *pout = pl;
    }
    if (ev) (void)hipEventRecord(ev[2], st);
    return hipGetLastError();
}

// ev (may be null): four events recorded before and after the Hermitian block
// product and before and after the rank-2k update, for per-launch timing.
extern "C" hipError_t bse_band_update(hipStream_t st, int fp32, int kind, int n, int s, int k, int b, void* A,
                                      const double2* V, const double2* Tp, double2* ws, const double2* zeros64,
                                      hipEvent_t* ev)
{
    const int m = n - s;
    if (m <= 0 || k <= 0) return hipSuccess;
    double2* Z = nullptr;
    hipError_t e = bse_band_product(st, fp32, kind, n, s, k, b, A, V, Tp, ws, ev, &Z, nullptr);
    if (e != hipSuccess) return e;
    e = bse_launch_her2k(st, fp32, kind, n, s, A, V, Z, m, s, k, zeros64);
    if (ev) (void)hipEventRecord(ev[3], st);
    return e;
}
