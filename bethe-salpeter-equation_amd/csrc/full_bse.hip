// K12: the full BSE beyond the Tamm-Dancoff approximation (add_coupling_part,
// bse.version13.cpp:2929-3040; eig_gen at 3326), solved by a
// structure-preserving reduction to a Hermitian eigenproblem instead of the
// reference's general (non-Hermitian) eigensolver:
//
//   H = [[A, B], [-conj B, -conj A]] = s3 M,   M = [[A, B], [conj B, conj A]],
//   s3 = diag(1, ..., 1, -1, ..., -1).
//
// For a stable system M is Hermitian positive definite; with M = L L^H
// (Cholesky), W = L^H s3 L is Hermitian and similar to H, so the eigenvalues
// of H (pairs +-E) are those of W, and an eigenvector y of W gives
// x = sqrt(E) L^-H y with x^H s3 x = 1 (Shao, da Jornada, Yang, Deslippe &
// Lin 2016, the method the reference's draft cites at 3371).  The oscillator
// strength of x is rho~^H x = sqrt(E) (L^-1 rho~)^H y, so every exciton's
// strength comes out of the Hermitian pipeline's contraction with
// z = L^-1 rho~, without forming the vectors.
//
// Storage: M (then L), W: tiled lower FP64 storage of order n2 = 2N (the
// layout of the TDA matrix), both halves in the k-major internal order.
// Kernels here: compose M from A and B, the blocked Cholesky (diagonal tile
// factor and inverse in LDS, panel solve; the trailing update is the
// reduction's rank-2k kernel), W = L^H s3 L, and the two triangular solves.
#include "bse_common.hpp"

// ---------------------------------------------------------------------------
// M from the TDA block A and the coupling block B (both tiled lower FP64 of
// order N; B complex symmetric, its lower triangle stored).  One workgroup per
// lower tile of M; padding rows and columns and the strict upper half of
// diagonal tiles are written as zeros.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_compose_full(int64_t N, const double2* __restrict__ A,
                                                      const double2* __restrict__ Bm, double2* __restrict__ M)
{
    const int64_t n2 = 2 * N;
    const int nt2 = tile_count(n2), ntA = tile_count(N);
    int ti, tj;
    tri_tile(blockIdx.x, ti, tj);
    double2* t = M + tile_pos(ti, tj, nt2) * (BSE_TILE * BSE_TILE);
    for (int e = threadIdx.x; e < BSE_TILE * BSE_TILE; e += blockDim.x) {
        const int64_t r = (int64_t)ti * 64 + (e & 63), c = (int64_t)tj * 64 + (e >> 6);
        double2 x = cmk(0, 0);
        if (r < n2 && c < n2 && r >= c) {
            if (r < N) {
                x = A[tix(r, c, ntA)];
            } else if (c < N) {
                const int64_t rr = r - N;
                x = cconj(rr >= c ? Bm[tix(rr, c, ntA)] : Bm[tix(c, rr, ntA)]);
            } else {
                x = cconj(A[tix(r - N, c - N, ntA)]);
            }
            if (r == c) x.y = 0.0;
        }
        t[e] = x;
    }
}

// ---------------------------------------------------------------------------
// Cholesky step J, part 1: factor the diagonal tile (J, J) in LDS (right-
// looking, one column per step), write L_JJ back (zeros above the diagonal)
// and its inverse to Linv[J] (64 x 64, column-major).  Padding rows beyond n2
// get a unit pivot so that the factor of the last tile stays regular; they
// are written back as zeros.  A pivot <= 0 means M is not positive definite:
// err[0] receives the (1-based) row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_potrf_tile(int n2, int J, double2* __restrict__ M,
                                                    double2* __restrict__ Linv, int* __restrict__ err)
{
    __shared__ double2 sL[64][65];    // sL[c][r] = L(r, c)
    __shared__ double2 sI[64][65];    // sI[c][r] = Linv(r, c)
    __shared__ int bad;
    const int nt2 = tile_count(n2);
    double2* t = M + tile_pos(J, J, nt2) * (BSE_TILE * BSE_TILE);
    const int tid = threadIdx.x;
    const int r0 = J * 64;
    if (tid == 0) bad = 0;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        double2 x = (r >= c) ? t[e] : cmk(0, 0);
        if (r == c && r0 + r >= n2) x = cmk(1, 0);
        sL[c][r] = x;
    }
    __syncthreads();
    for (int j = 0; j < 64; ++j) {
        if (tid == 0) {
            const double d = sL[j][j].x;
            if (!(d > 0.0)) {
                if (!bad) bad = r0 + j + 1;
                sL[j][j] = cmk(1, 0);
            } else {
                sL[j][j] = cmk(sqrt(d), 0);
            }
        }
        __syncthreads();
        const double inv = 1.0 / sL[j][j].x;
        for (int i = j + 1 + tid; i < 64; i += 256) sL[j][i] = cscale(sL[j][i], inv);
        __syncthreads();
        // trailing lower update: a(i, k) -= l(i, j) conj(l(k, j)), j < k <= i
        const int m = 63 - j;
        for (int e = tid; e < m * m; e += 256) {
            const int i = j + 1 + e % m, k = j + 1 + e / m;
            if (i >= k) {
                double2 a = sL[k][i];
                const double2 li = sL[j][i], lk = sL[j][k];
                a.x -= li.x * lk.x + li.y * lk.y;
                a.y -= li.y * lk.x - li.x * lk.y;
                sL[k][i] = a;
            }
        }
        __syncthreads();
    }
    // inverse of the lower triangular factor, one column per thread:
    // Linv(:, c) solves L x = e_c by forward substitution
    if (tid < 64) {
        const int c = tid;
        for (int r = 0; r < 64; ++r) {
            double2 s = (r == c) ? cmk(1, 0) : cmk(0, 0);
            if (r > c) {
                for (int k = c; k < r; ++k) {
                    const double2 l = sL[k][r], x = sI[c][k];
                    s.x -= l.x * x.x - l.y * x.y;
                    s.y -= l.x * x.y + l.y * x.x;
                }
            }
            if (r < c) s = cmk(0, 0);
            sI[c][r] = cscale(s, 1.0 / sL[r][r].x);
        }
    }
    __syncthreads();
    double2* li = Linv + (int64_t)J * 64 * 64;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        const bool pad = (r0 + r >= n2) || (r0 + c >= n2);
        t[e] = (r >= c && !pad) ? sL[c][r] : cmk(0, 0);
        li[e] = pad ? cmk(0, 0) : sI[c][r];
    }
    if (tid == 0 && bad && err[0] == 0) err[0] = bad;
}

// ---------------------------------------------------------------------------
// Cholesky step J, part 2: the panel below the diagonal tile,
// L_IJ = A_IJ L_JJ^-H for I > J (one workgroup per tile), also written to the
// dense panel P (m x 64, ld m, rows from s = 64 (J + 1)) and Ph = P / 2 for
// the trailing update A22 -= P Ph^H + Ph P^H = P P^H.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_trsm_panel(int n2, int J, double2* __restrict__ M,
                                                    const double2* __restrict__ Linv,
                                                    double2* __restrict__ P, double2* __restrict__ Ph)
{
    __shared__ double2 sA[64][65];    // sA[k][r] = A_IJ(r, k)
    __shared__ double2 sI[64][65];    // sI[k][c] = Linv(c, k)
    const int nt2 = tile_count(n2);
    const int I = J + 1 + blockIdx.x;
    const int s = (J + 1) * 64, m = n2 - s;
    double2* t = M + tile_pos(I, J, nt2) * (BSE_TILE * BSE_TILE);
    const double2* li = Linv + (int64_t)J * 64 * 64;
    const int tid = threadIdx.x;
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e & 63, c = e >> 6;
        sA[c][r] = t[e];
        sI[c][r] = li[e];   // sI[c][r] = Linv(r, c)
    }
    __syncthreads();
    // X(r, c) = sum_k A(r, k) conj(Linv(c, k)), k <= c
    const int r = tid & 63, cg = tid >> 6;   // 4 groups of 16 columns
    double2 acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = cmk(0, 0);
    for (int k = 0; k < 64; ++k) {
        const double2 a = sA[k][r];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = cg * 16 + q;
            if (k <= c) cfma(acc[q], a, cconj(sI[k][c]));
        }
    }
    const int64_t gr = (int64_t)I * 64 + r;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int c = cg * 16 + q;
        const double2 x = (gr < n2) ? acc[q] : cmk(0, 0);
        t[r + c * 64] = x;
        if (gr < n2) {
            P[(gr - s) + (int64_t)c * m] = x;
            Ph[(gr - s) + (int64_t)c * m] = cscale(x, 0.5);
        }
    }
}

// ---------------------------------------------------------------------------
// W = L^H s3 L, lower tiles: W_IJ = sum_{K >= I} L_KI^H diag(s_K) L_KJ with
// s = +1 on rows < N, -1 on rows N .. n2 - 1, 0 on padding.  One workgroup
// per tile of W; 8 rows of the two L tiles are staged in LDS per step; each
// thread accumulates a 4 x 4 block of the 64 x 64 result in FP64.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lsl(int64_t N, const double2* __restrict__ L, double2* __restrict__ Wt)
{
    const int64_t n2 = 2 * N;
    const int nt2 = tile_count(n2);
    int ti, tj;
    tri_tile(blockIdx.x, ti, tj);
    __shared__ double2 sa[8][64], sb[8][64];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double2 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = cmk(0, 0);
    for (int K = ti; K < nt2; ++K) {
        const double2* tI = L + tile_pos(K, ti, nt2) * (BSE_TILE * BSE_TILE);
        const double2* tJ = L + tile_pos(K, tj, nt2) * (BSE_TILE * BSE_TILE);
        for (int r0 = 0; r0 < 64; r0 += 8) {
            for (int e = threadIdx.x; e < 8 * 64; e += 256) {
                const int rr = e & 7, c = e >> 3;
                const int64_t grow = (int64_t)K * 64 + r0 + rr;
                const double sg = grow < N ? 1.0 : (grow < n2 ? -1.0 : 0.0);
                sa[rr][c] = tI[(r0 + rr) + c * 64];
                sb[rr][c] = cscale(tJ[(r0 + rr) + c * 64], sg);
            }
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < 8; ++rr) {
                double2 va[4], vb[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    va[a] = sa[rr][tx + 16 * a];
                    vb[a] = sb[rr][ty + 16 * a];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) cfmac(acc[a][b], va[a], vb[b]);
            }
            __syncthreads();
        }
    }
    double2* w = Wt + tile_pos(ti, tj, nt2) * (BSE_TILE * BSE_TILE);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int lr = tx + 16 * a, lc = ty + 16 * b;
            const int64_t gr = (int64_t)ti * 64 + lr, gc = (int64_t)tj * 64 + lc;
            double2 x = acc[a][b];
            if (gr >= n2 || gc >= n2 || gr < gc) x = cmk(0, 0);
            if (gr == gc) x.y = 0.0;
            w[lr + lc * 64] = x;
        }
}

// ---------------------------------------------------------------------------
// z = L^-1 b (forward substitution over tile rows; one workgroup).  Row block
// J: t = b_J - sum_{K<J} L_JK z_K (four partial sums per row, then LDS), and
// z_J = Linv_JJ t.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_lower_solve(int n2, const double2* __restrict__ L,
                                                     const double2* __restrict__ Linv,
                                                     const double2* __restrict__ b, double2* __restrict__ z)
{
    __shared__ double2 part[4][64];
    __shared__ double2 tv[64];
    const int nt2 = tile_count(n2);
    const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
    for (int J = 0; J < nt2; ++J) {
        double2 s = cmk(0, 0);
        const int64_t gr = (int64_t)J * 64 + r;
        for (int K = q; K < J; K += 4) {
            const double2* t = L + tile_pos(J, K, nt2) * (BSE_TILE * BSE_TILE);
            const double2* zk = z + (int64_t)K * 64;
            for (int c = 0; c < 64; ++c) cfma(s, t[r + c * 64], zk[c]);
        }
        part[q][r] = s;
        __syncthreads();
        if (q == 0) {
            double2 v = (gr < n2) ? b[gr] : cmk(0, 0);
            for (int p = 0; p < 4; ++p) v = csub(v, part[p][r]);
            tv[r] = v;
        }
        __syncthreads();
        if (q == 0) {
            const double2* li = Linv + (int64_t)J * 64 * 64;
            double2 x = cmk(0, 0);
            for (int c = 0; c <= r; ++c) cfma(x, li[r + c * 64], tv[c]);
            z[gr] = (gr < n2) ? x : cmk(0, 0);
        }
        __syncthreads();
        __threadfence_block();
    }
}

// ---------------------------------------------------------------------------
// X = scale_v * L^-H Y for nv vectors (backward substitution; one workgroup
// per vector).  Row block J from the bottom: t = y_J - sum_{K>J} L_KJ^H x_K,
// x_J = Linv_JJ^H t.  x may alias y.  scale: per vector (sqrt(E)).
// X has n2 rows padded to whole tiles (ldx >= 64 * nt2).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_upper_solve(int n2, const double2* __restrict__ L,
                                                     const double2* __restrict__ Linv,
                                                     const double2* __restrict__ Y, int64_t ldy,
                                                     double2* __restrict__ X, int64_t ldx,
                                                     const double* __restrict__ scale)
{
    __shared__ double2 part[4][64];
    __shared__ double2 tv[64];
    const int nt2 = tile_count(n2);
    const int v = blockIdx.x;
    const double2* y = Y + (int64_t)v * ldy;
    double2* x = X + (int64_t)v * ldx;
    const int r = threadIdx.x & 63, q = threadIdx.x >> 6;
    for (int J = nt2 - 1; J >= 0; --J) {
        double2 s = cmk(0, 0);
        const int64_t gr = (int64_t)J * 64 + r;
        for (int K = J + 1 + q; K < nt2; K += 4) {
            const double2* t = L + tile_pos(K, J, nt2) * (BSE_TILE * BSE_TILE);
            const double2* xk = x + (int64_t)K * 64;
            for (int c = 0; c < 64; ++c) cfmac(s, t[c + r * 64], xk[c]);
        }
        part[q][r] = s;
        __syncthreads();
        if (q == 0) {
            double2 u = (gr < n2) ? y[gr] : cmk(0, 0);
            for (int p = 0; p < 4; ++p) u = csub(u, part[p][r]);
            tv[r] = u;
        }
        __syncthreads();
        if (q == 0) {
            const double2* li = Linv + (int64_t)J * 64 * 64;
            double2 w = cmk(0, 0);
            for (int c = r; c < 64; ++c) cfmac(w, li[c + r * 64], tv[c]);
            x[gr] = (gr < n2) ? w : cmk(0, 0);
        }
        __syncthreads();
        __threadfence_block();
    }
    __syncthreads();
    const double sc = scale ? scale[v] : 1.0;
    for (int64_t i = threadIdx.x; i < (int64_t)nt2 * 64; i += 256) x[i] = cscale(x[i], sc);
}

// rho~ = (rho; -conj(rho)) in the internal order of both halves, zero-padded
// to ntot entries (the anti-resonant pair densities rho_vc = -conj(rho_cv) of
// the point-orbital dipole limit).
__global__ __launch_bounds__(256) void k_full_rho(int64_t N, const double2* __restrict__ rho,
                                                  double2* __restrict__ out, int64_t ntot)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= ntot) return;
    double2 x = cmk(0, 0);
    if (i < N) x = rho[i];
    else if (i < 2 * N) x = cscale(cconj(rho[i - N]), -1.0);
    out[i] = x;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
extern "C" hipError_t bse_launch_full_rho(hipStream_t st, int64_t N, const double2* rho, double2* out,
                                          int64_t ntot)
{
    hipLaunchKernelGGL(k_full_rho, dim3((unsigned)ceil_div64(ntot, 256)), dim3(256), 0, st, N, rho, out, ntot);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_compose_full(hipStream_t st, int64_t N, const double2* A, const double2* Bm,
                                              double2* M)
{
    const int64_t nt2 = tile_count(2 * N);
    hipLaunchKernelGGL(k_compose_full, dim3((unsigned)(nt2 * (nt2 + 1) / 2)), dim3(256), 0, st, N, A, Bm, M);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_potrf_tile(hipStream_t st, int n2, int J, double2* M, double2* Linv, int* err)
{
    hipLaunchKernelGGL(k_potrf_tile, dim3(1), dim3(256), 0, st, n2, J, M, Linv, err);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_trsm_panel(hipStream_t st, int n2, int J, double2* M, const double2* Linv,
                                            double2* P, double2* Ph)
{
    const int nt2 = tile_count(n2);
    if (J + 1 >= nt2) return hipSuccess;
    hipLaunchKernelGGL(k_trsm_panel, dim3((unsigned)(nt2 - J - 1)), dim3(256), 0, st, n2, J, M, Linv, P, Ph);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_lsl(hipStream_t st, int64_t N, const double2* L, double2* Wt)
{
    const int64_t nt2 = tile_count(2 * N);
    hipLaunchKernelGGL(k_lsl, dim3((unsigned)(nt2 * (nt2 + 1) / 2)), dim3(256), 0, st, N, L, Wt);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_lower_solve(hipStream_t st, int n2, const double2* L, const double2* Linv,
                                             const double2* b, double2* z)
{
    hipLaunchKernelGGL(k_lower_solve, dim3(1), dim3(256), 0, st, n2, L, Linv, b, z);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_upper_solve(hipStream_t st, int n2, int nv, const double2* L,
                                             const double2* Linv, const double2* Y, int64_t ldy, double2* X,
                                             int64_t ldx, const double* scale)
{
    if (nv <= 0) return hipSuccess;
    if (ldx < (int64_t)tile_count(n2) * 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_upper_solve, dim3((unsigned)nv), dim3(256), 0, st, n2, L, Linv, Y, ldy, X, ldx, scale);
    return hipGetLastError();
}
