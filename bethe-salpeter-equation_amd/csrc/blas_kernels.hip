// Dense complex FP64 building blocks for K7 (back-transform), K8 (refinement)
// and K9 (oscillator strengths): a tiled GEMM with optional split-K, and the
// compact-WY triangular factor of a panel of Householder reflectors
// (Schreiber & Van Loan; forward, columnwise), which turns the application of
// nb reflectors into three GEMMs: Z <- Z - V (T (V^H Z)).
#include "bse_common.hpp"
#include <algorithm>

// op codes
#define OP_N 0
#define OP_C 1

// C[M x N] (+)= alpha * op(A) * op(B); column-major.  Tile 64 x 64, K step 16,
// 256 threads, 4 x 4 outputs per thread strided by 16 (conflict-free LDS).
// gridDim.z > 1: split-K, partial results to P + z * M * N (ld M), no beta.
template <int OPA, int OPB>
__global__ __launch_bounds__(256) void k_zgemm(int M, int N, int K, double2 alpha,
                                              const double2* __restrict__ A, int64_t lda,
                                              const double2* __restrict__ B, int64_t ldb,
                                              double2 beta, double2* __restrict__ C, int64_t ldc,
                                              double2* __restrict__ P, int kchunk)
{
    __shared__ double2 sA[16][64 + 1];
    __shared__ double2 sB[16][64 + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    double2 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = cmk(0, 0);
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
        // load op(A) tile: 64 (m) x 16 (k)
        for (int e = threadIdx.x; e < 64 * 16; e += 256) {
            int mm, kk;
            double2 val = cmk(0, 0);
            if (OPA == OP_N) {
                mm = e & 63; kk = e >> 6;
                const int gm = m0 + mm, gk = k0 + kk;
                if (gm < M && gk < kend) val = A[gm + (int64_t)gk * lda];
            } else {
                kk = e & 15; mm = e >> 4;
                const int gm = m0 + mm, gk = k0 + kk;
                if (gm < M && gk < kend) val = cconj(A[gk + (int64_t)gm * lda]);
            }
            sA[kk][mm] = val;
        }
        // load op(B) tile: 16 (k) x 64 (n)
        for (int e = threadIdx.x; e < 64 * 16; e += 256) {
            int kk, nn;
            double2 val = cmk(0, 0);
            if (OPB == OP_N) {
                kk = e & 15; nn = e >> 4;
                const int gk = k0 + kk, gn = n0 + nn;
                if (gk < kend && gn < N) val = B[gk + (int64_t)gn * ldb];
            } else {
                nn = e & 63; kk = e >> 6;
                const int gk = k0 + kk, gn = n0 + nn;
                if (gk < kend && gn < N) val = cconj(B[gn + (int64_t)gk * ldb]);
            }
            sB[kk][nn] = val;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double2 av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = sA[kk][tx + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = sB[kk][ty + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) cfma(acc[a][b], av[a], bv[b]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int gm = m0 + tx + 16 * a, gn = n0 + ty + 16 * b;
            if (gm >= M || gn >= N) continue;
            if (gridDim.z > 1) {
                P[(int64_t)blockIdx.z * M * N + gm + (int64_t)gn * M] = acc[a][b];
            } else {
                double2 r = cmul(alpha, acc[a][b]);
                if (beta.x != 0.0 || beta.y != 0.0) r = cadd(r, cmul(beta, C[gm + (int64_t)gn * ldc]));
                C[gm + (int64_t)gn * ldc] = r;
            }
        }
}

__global__ __launch_bounds__(256) void k_splitk_reduce(int M, int N, int nsplit, double2 alpha,
                                                      const double2* __restrict__ P, double2 beta,
                                                      double2* __restrict__ C, int64_t ldc)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)M * N) return;
    const int gm = (int)(idx % M), gn = (int)(idx / M);
    double2 s = cmk(0, 0);
    for (int z = 0; z < nsplit; ++z) s = cadd(s, P[(int64_t)z * M * N + idx]);
    double2 r = cmul(alpha, s);
    if (beta.x != 0.0 || beta.y != 0.0) r = cadd(r, cmul(beta, C[gm + (int64_t)gn * ldc]));
    C[gm + (int64_t)gn * ldc] = r;
}

// Host wrapper.  ws: split-K workspace (>= nsplit*M*N double2); ws_elems its size.
extern "C" hipError_t bse_zgemm(hipStream_t st, int opa, int opb, int M, int N, int K,
                                double ar, double ai, const double2* A, int64_t lda,
                                const double2* B, int64_t ldb, double br, double bi, double2* C,
                                int64_t ldc, double2* ws, int64_t ws_elems)
{
    if (M <= 0 || N <= 0) return hipSuccess;
    const double2 alpha = cmk(ar, ai), beta = cmk(br, bi);
    const int64_t tiles = ceil_div64(M, 64) * ceil_div64(N, 64);
    int nsplit = 1;
    if (K > 0 && tiles < 256 && ws != nullptr) {
        nsplit = (int)std::min<int64_t>(ceil_div64(K, 256), ceil_div64(512, tiles));
        while (nsplit > 1 && (int64_t)nsplit * M * N > ws_elems) --nsplit;
    }
    int kchunk = (int)(ceil_div64(ceil_div64(K > 0 ? K : 1, nsplit), 16) * 16);
    nsplit = (int)ceil_div64(K > 0 ? K : 1, kchunk);
    dim3 grid((unsigned)ceil_div64(M, 64), (unsigned)ceil_div64(N, 64), (unsigned)nsplit);
#define ZG(OA, OB)                                                                         \
    hipLaunchKernelGGL((k_zgemm<OA, OB>), grid, dim3(256), 0, st, M, N, K, alpha, A, lda, B, \
                       ldb, beta, C, ldc, ws, kchunk)
    if (opa == OP_N && opb == OP_N) ZG(OP_N, OP_N);
    else if (opa == OP_C && opb == OP_N) ZG(OP_C, OP_N);
    else if (opa == OP_N && opb == OP_C) ZG(OP_N, OP_C);
    else ZG(OP_C, OP_C);
#undef ZG
    if (nsplit > 1) {
        const int64_t tot = (int64_t)M * N;
        hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)ceil_div64(tot, 256)), dim3(256), 0, st, M,
                           N, nsplit, alpha, (const double2*)ws, beta, C, ldc);
    }
    return hipGetLastError();
}

// Compact-WY factor: T (nb x nb, upper, ld ldt) from tau (nb) and the Gram
// matrix G = V^H V (nb x nb, ld 64; only G(l, i), l < i, is read):
// T(i,i) = tau_i, T(0:i, i) = -tau_i * T(0:i, 0:i) G(0:i, i).
__global__ void k_larft(int nb, const double2* __restrict__ tau, const double2* __restrict__ G,
                        double2* __restrict__ T, int ldt)
{
    __shared__ double2 col[64];
    larft_body(nb, tau, G, T, ldt, col);
}

extern "C" hipError_t bse_launch_larft(hipStream_t st, int nb, const double2* tau, const double2* G,
                                       double2* T, int ldt)
{
    hipLaunchKernelGGL(k_larft, dim3(1), dim3(64), 0, st, nb, tau, G, T, ldt);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K9: TDA dielectric function, eps(w) = 1 - pref sum_m |f_m|^2/(w - E_m + i eta)
// (bse.version13.cpp:3621-3634).  One thread per frequency, oscillator table
// streamed through LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dielectric(int nw, const double* __restrict__ omega,
                                                   double eta, int N, const double* __restrict__ E,
                                                   const double2* __restrict__ f, double pref,
                                                   double* __restrict__ eps_re,
                                                   double* __restrict__ eps_im)
{
    __shared__ double sE[256], sF[256];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const double w = (s < nw) ? omega[s] : 0.0;
    double re = 0.0, im = 0.0;
    for (int m0 = 0; m0 < N; m0 += 256) {
        const int m = m0 + threadIdx.x;
        sE[threadIdx.x] = (m < N) ? E[m] : 0.0;
        sF[threadIdx.x] = (m < N) ? cabs2(f[m]) : 0.0;
        __syncthreads();
        const int cnt = min(256, N - m0);
        for (int j = 0; j < cnt; ++j) {
            const double dw = w - sE[j];
            const double den = dw * dw + eta * eta;
            const double g = sF[j] / den;
            re += g * dw;      // Re 1/(dw + i eta) = dw/den
            im -= g * eta;     // Im 1/(dw + i eta) = -eta/den
        }
        __syncthreads();
    }
    if (s < nw) {
        eps_re[s] = 1.0 - pref * re;
        eps_im[s] = -pref * im;
    }
}

extern "C" hipError_t bse_launch_dielectric(hipStream_t st, int nw, const double* omega, double eta,
                                            int N, const double* E, const double2* f, double pref,
                                            double* eps_re, double* eps_im)
{
    hipLaunchKernelGGL(k_dielectric, dim3((unsigned)ceil_div64(nw, 256)), dim3(256), 0, st, nw,
                       omega, eta, N, E, f, pref, eps_re, eps_im);
    return hipGetLastError();
}

// Permute rows of a complex N x ncol matrix from internal k-major order to the
// reference order: dst[ref(r'), j] = src[r', j], ref(r') = (r' % B) * nk +
// kpos(r' / B), where kpos maps a grid index to the caller's k-list position
// (kmap, or the identity when kmap is null).
__global__ __launch_bounds__(256) void k_rows_to_reference(int64_t N, int B, int nk, int ncol,
                                                          const double2* __restrict__ src,
                                                          int64_t lds, double2* __restrict__ dst,
                                                          int64_t ldd, const int* __restrict__ kmap)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (r >= N || j >= ncol) return;
    const int kg = (int)(r / B);
    const int64_t rr = (r % B) * nk + (kmap ? kmap[kg] : kg);
    dst[rr + (int64_t)j * ldd] = src[r + (int64_t)j * lds];
}

extern "C" hipError_t bse_launch_rows_to_reference(hipStream_t st, int64_t N, int B, int nk,
                                                   int ncol, const double2* src, int64_t lds,
                                                   double2* dst, int64_t ldd, const int* kmap)
{
    if (ncol <= 0) return hipSuccess;
    dim3 grid((unsigned)ceil_div64(N, 256), (unsigned)ncol);
    hipLaunchKernelGGL(k_rows_to_reference, grid, dim3(256), 0, st, N, B, nk, ncol, src, lds, dst,
                       ldd, kmap);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_convert_to_fp32(int64_t n, const double2* __restrict__ src,
                                                        float2* __restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = make_float2((float)src[i].x, (float)src[i].y);
}

extern "C" hipError_t bse_launch_convert_to_fp32(hipStream_t st, int64_t n, const double2* src,
                                                 float2* dst)
{
    hipLaunchKernelGGL(k_convert_to_fp32, dim3((unsigned)ceil_div64(n, 256)), dim3(256), 0, st, n,
                       src, dst);
    return hipGetLastError();
}

// Inverse of k_rows_to_reference: dst[r', j] = src[ref(r'), j].
__global__ __launch_bounds__(256) void k_rows_from_reference(int64_t N, int B, int nk, int ncol,
                                                            const double2* __restrict__ src,
                                                            int64_t lds, double2* __restrict__ dst,
                                                            int64_t ldd, const int* __restrict__ kmap)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (r >= N || j >= ncol) return;
    const int kg = (int)(r / B);
    const int64_t rr = (r % B) * nk + (kmap ? kmap[kg] : kg);
    dst[r + (int64_t)j * ldd] = src[rr + (int64_t)j * lds];
}

extern "C" hipError_t bse_launch_rows_from_reference(hipStream_t st, int64_t N, int B, int nk,
                                                     int ncol, const double2* src, int64_t lds,
                                                     double2* dst, int64_t ldd, const int* kmap)
{
    if (ncol <= 0) return hipSuccess;
    dim3 grid((unsigned)ceil_div64(N, 256), (unsigned)ncol);
    hipLaunchKernelGGL(k_rows_from_reference, grid, dim3(256), 0, st, N, B, nk, ncol, src, lds, dst,
                       ldd, kmap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Application of the reduction's orthogonal factor to a few vectors without
// forming it: Y <- Q^H Y (adjoint = 1; panels in increasing order, T^H) or
// Y <- Q Y (adjoint = 0; decreasing order, T), with Q = prod_p (I - V_p T_p V_p^H)
// from the packed FP64 reflector store (panel p at nb (p n - nb p (p-1)/2),
// ld n - p nb) and the per-panel compact-WY factors (nb x nb, ld nb).
// One workgroup (APQ_THREADS threads; 512 keep the 32 partial dots of a
// thread in registers, 1024 spill them) per vector walks all panels; for
// each it forms w1 = V^H y (wave sums), w2 = T^(H) w1 and y -= V w2.  For the small
// matrices of the batched sweep, where one vector is a few MB of reflector
// reads, this replaces three GEMM launches per panel.  nb <= 32.
// ---------------------------------------------------------------------------
#define APQ_NB 32
#ifndef APQ_THREADS
#define APQ_THREADS 512
#endif
__global__ __launch_bounds__(APQ_THREADS) void k_apply_q(int n, int nb, const double2* __restrict__ Vstore,
                                                 const double2* __restrict__ Tall,
                                                 double2* __restrict__ Y, int64_t ldy, int adjoint)
{
    constexpr int NW = APQ_THREADS / 64;
    __shared__ double2 s_red[NW][APQ_NB];
    __shared__ double2 s_w1[APQ_NB], s_w2[APQ_NB];
    double2* y = Y + (int64_t)blockIdx.x * ldy;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int npan = (int)ceil_div64(n - 1, nb);
    for (int pp = 0; pp < npan; ++pp) {
        const int p = adjoint ? pp : npan - 1 - pp;
        const int j0 = p * nb, m = n - j0, nbp = min(nb, n - 1 - j0);
        const double2* V = Vstore + (int64_t)nb * ((int64_t)p * n - (int64_t)nb * p * (p - 1) / 2);
        const double2* T = Tall + (size_t)p * nb * nb;
        double2 acc[APQ_NB];
#pragma unroll
        for (int l = 0; l < APQ_NB; ++l) acc[l] = cmk(0, 0);
        for (int r = threadIdx.x; r < m; r += APQ_THREADS) {
            const double2 yr = y[j0 + r];
#pragma unroll
            for (int l = 0; l < APQ_NB; ++l)
                if (l < nbp) cfmac(acc[l], V[r + (int64_t)l * m], yr);
        }
#pragma unroll
        for (int l = 0; l < APQ_NB; ++l) {
            const double x = wave_sum(acc[l].x), z = wave_sum(acc[l].y);
            if (lane == 0) s_red[w][l] = cmk(x, z);
        }
        __syncthreads();
        if (threadIdx.x < APQ_NB) {
            const int l = threadIdx.x;
            double2 t = s_red[0][l];
            for (int ww = 1; ww < NW; ++ww) t = cadd(t, s_red[ww][l]);
            s_w1[l] = t;
        }
        __syncthreads();
        if (threadIdx.x < nbp) {
            const int l = threadIdx.x;
            double2 s = cmk(0, 0);
            if (adjoint) {
                for (int k = 0; k <= l; ++k) cfmac(s, T[k + l * nb], s_w1[k]);
            } else {
                for (int k = l; k < nbp; ++k) cfma(s, T[l + k * nb], s_w1[k]);
            }
            s_w2[l] = s;
        }
        __syncthreads();
        for (int r = threadIdx.x; r < m; r += APQ_THREADS) {
            double2 s = cmk(0, 0);
            for (int l = 0; l < nbp; ++l) cfma(s, V[r + (int64_t)l * m], s_w2[l]);
            y[j0 + r] = csub(y[j0 + r], s);
        }
        __syncthreads();
    }
}

extern "C" hipError_t bse_launch_apply_q(hipStream_t st, int n, int nb, const double2* Vstore,
                                         const double2* Tall, int nvec, double2* Y, int64_t ldy,
                                         int adjoint)
{
    if (nb > APQ_NB || nvec <= 0) return hipErrorInvalidValue;
    if (n < 2) return hipSuccess;
    hipLaunchKernelGGL(k_apply_q, dim3((unsigned)nvec), dim3(APQ_THREADS), 0, st, n, nb, Vstore, Tall, Y,
                       ldy, adjoint);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The same application for a large matrix and at most 16 vectors, spread over
// the whole device: per panel (V: m x nbp, ld m; X: the rows j0.. of the
// vectors, ld ldx)
//   (1) k_pq_dots:   partial P_g = V_g^H X_g over row blocks g of PQ_ROWS rows
//   (2) k_pq_small:  W = T^(H) sum_g P_g (one workgroup per vector; fixed summation order)
//   (3) k_pq_update: X -= V W
// Three short launches per panel instead of four GEMM launches; every pass
// reads V once.  The launches are latency-bound rather than bandwidth-bound
// (a panel's V is at most 32 columns): the row blocks are 64 rows, giving up
// to n/64 workgroups; every load of a block is issued before the first LDS
// store; T is staged in LDS; and the update runs one wave per workgroup.
// ---------------------------------------------------------------------------
#define PQ_ROWS 64
#define PQ_MAXK 16
// One block of PQ_ROWS rows per workgroup.  Thread t owns the entries
// (l, j) = (t % 32, t / 32) and (l, j + 8) of the 32 x 16 partial; even and
// odd rows go to separate sums (four independent FMA chains per thread),
// added at the end in a fixed order.
__global__ __launch_bounds__(256) void k_pq_dots(int m, int nbp, int k, const double2* __restrict__ V,
                                                 const double2* __restrict__ X, int64_t ldx,
                                                 double2* __restrict__ P)
{
    __shared__ double2 sV[PQ_ROWS][APQ_NB + 1];
    __shared__ double2 sX[PQ_ROWS][PQ_MAXK];
    const int r0 = blockIdx.x * PQ_ROWS;
    const int t = threadIdx.x;
    // every load is issued before the first LDS store: addresses are clamped
    // into the block and out-of-range entries replaced by zero afterwards, so
    // no load sits behind a branch
    constexpr int NV = PQ_ROWS * APQ_NB / 256, NX = PQ_ROWS * PQ_MAXK / 256;
    double2 rv[NV], rx[NX];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int e = t + 256 * i, rr = e & (PQ_ROWS - 1), l = e / PQ_ROWS, r = r0 + rr;
        const double2 a = V[min(r, m - 1) + (int64_t)min(l, nbp - 1) * m];
        rv[i] = (r < m && l < nbp) ? a : cmk(0, 0);
    }
#pragma unroll
    for (int i = 0; i < NX; ++i) {
        const int e = t + 256 * i, rr = e & (PQ_ROWS - 1), j = e / PQ_ROWS, r = r0 + rr;
        const double2 a = X[min(r, m - 1) + (int64_t)min(j, k - 1) * ldx];
        rx[i] = (r < m && j < k) ? a : cmk(0, 0);
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int e = t + 256 * i;
        sV[e & (PQ_ROWS - 1)][e / PQ_ROWS] = rv[i];
    }
#pragma unroll
    for (int i = 0; i < NX; ++i) {
        const int e = t + 256 * i;
        sX[e & (PQ_ROWS - 1)][e / PQ_ROWS] = rx[i];
    }
    __syncthreads();
    const int l = t & 31, j = t >> 5;
    double2 a0 = cmk(0, 0), a1 = cmk(0, 0), b0 = cmk(0, 0), b1 = cmk(0, 0);
#pragma unroll 8
    for (int rr = 0; rr < PQ_ROWS; rr += 2) {
        const double2 v0 = sV[rr][l], v1 = sV[rr + 1][l];
        cfmac(a0, v0, sX[rr][j]);
        cfmac(a1, v0, sX[rr][j + 8]);
        cfmac(b0, v1, sX[rr + 1][j]);
        cfmac(b1, v1, sX[rr + 1][j + 8]);
    }
    P[(int64_t)blockIdx.x * 512 + t] = cadd(a0, b0);
    P[(int64_t)blockIdx.x * 512 + t + 256] = cadd(a1, b1);
}

// One workgroup per vector column j < k: thread (l, part), part < 16, sums
// entry (l, j) over the partials g = part, part + 16, ... (eight independent
// loads in flight), the sixteen part sums are added in a fixed order, then
// W(:, j) = T^(H) (sum) for the nbp reflector rows.  Fixed summation order,
// identical on every run.
__global__ __launch_bounds__(512) void k_pq_small(int nbp, int k, int G, const double2* __restrict__ P,
                                                  const double2* __restrict__ T, int ldt, int adjoint,
                                                  double2* __restrict__ W)
{
    __shared__ double2 sP[16][32];
    __shared__ double2 sW[32];
    __shared__ double2 sT[32][33];   // sT[b][a] = T(a, b)
    const int j = blockIdx.x, l = threadIdx.x & 31, part = threadIdx.x >> 5;
    if (j >= k) return;
    const int o = l + 32 * j;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int e = threadIdx.x + 512 * i, a = e & 31, b = e >> 5;
        const double2 x = T[min(a, nbp - 1) + (int64_t)min(b, nbp - 1) * ldt];
        sT[b][a] = (a < nbp && b < nbp) ? x : cmk(0, 0);
    }
    double2 acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = cmk(0, 0);
    int g = part;
    if (l < nbp) {
        for (; g + 112 < G; g += 128) {
#pragma unroll
            for (int u = 0; u < 8; ++u) acc[u] = cadd(acc[u], P[(int64_t)(g + 16 * u) * 512 + o]);
        }
        for (; g < G; g += 16) acc[0] = cadd(acc[0], P[(int64_t)g * 512 + o]);
    }
    sP[part][l] = cadd(cadd(cadd(acc[0], acc[1]), cadd(acc[2], acc[3])),
                       cadd(cadd(acc[4], acc[5]), cadd(acc[6], acc[7])));
    __syncthreads();
    if (threadIdx.x < 32)
        sW[l] = cadd(cadd(cadd(cadd(sP[0][l], sP[1][l]), cadd(sP[2][l], sP[3][l])),
                          cadd(cadd(sP[4][l], sP[5][l]), cadd(sP[6][l], sP[7][l]))),
                     cadd(cadd(cadd(sP[8][l], sP[9][l]), cadd(sP[10][l], sP[11][l])),
                          cadd(cadd(sP[12][l], sP[13][l]), cadd(sP[14][l], sP[15][l]))));
    __syncthreads();
    if (threadIdx.x >= 32) return;
    double2 s = cmk(0, 0);
    if (l < nbp) {
        if (adjoint) {
            for (int q = 0; q <= l; ++q) cfmac(s, sT[l][q], sW[q]);
        } else {
            for (int q = l; q < nbp; ++q) cfma(s, sT[q][l], sW[q]);
        }
    }
    W[o] = s;
}

#define PQU_THREADS 64
// X -= V W for one row per thread; all k values of X are loaded before the
// first store, so the loads overlap.
__global__ __launch_bounds__(PQU_THREADS) void k_pq_update(int m, int nbp, int k, const double2* __restrict__ V,
                                                           const double2* __restrict__ W, double2* __restrict__ X,
                                                           int64_t ldx)
{
    __shared__ double2 sW[32 * PQ_MAXK];
    for (int e = threadIdx.x; e < 32 * k; e += PQU_THREADS) sW[e] = W[e];
    __syncthreads();
    const int r = blockIdx.x * PQU_THREADS + threadIdx.x;
    if (r >= m) return;
    double2 v[APQ_NB];
#pragma unroll
    for (int l = 0; l < APQ_NB; ++l) v[l] = (l < nbp) ? V[r + (int64_t)l * m] : cmk(0, 0);
    double2 x[PQ_MAXK];
#pragma unroll
    for (int j = 0; j < PQ_MAXK; ++j) x[j] = (j < k) ? X[r + (int64_t)j * ldx] : cmk(0, 0);
#pragma unroll
    for (int j = 0; j < PQ_MAXK; ++j) {
        if (j < k) {
            double2 s = cmk(0, 0);
#pragma unroll
            for (int l = 0; l < APQ_NB; ++l) cfma(s, v[l], sW[l + 32 * j]);
            X[r + (int64_t)j * ldx] = csub(x[j], s);
        }
    }
}

// Workspace (complex elements) of bse_apply_q_thin for size n.
extern "C" int64_t bse_apply_q_thin_workspace(int n) { return (ceil_div64(n, PQ_ROWS) + 1) * 512 + 512; }

// X <- Q^H X (adjoint = 1) or Q X (adjoint = 0) for k <= 16 columns (ld ldx),
// panels of the packed reflector store as in bse_launch_apply_q.  ws: at
// least bse_apply_q_thin_workspace(n) complex elements.
extern "C" hipError_t bse_apply_q_thin(hipStream_t st, int n, int nb, const double2* Vstore,
                                       const double2* Tall, int k, double2* X, int64_t ldx, int adjoint,
                                       double2* ws)
{
    if (nb > APQ_NB || k <= 0 || k > PQ_MAXK) return hipErrorInvalidValue;
    if (n < 2) return hipSuccess;
    const int npan = (int)ceil_div64(n - 1, nb);
    double2* W = ws;
    double2* P = ws + 512;
    for (int pp = 0; pp < npan; ++pp) {
        const int p = adjoint ? pp : npan - 1 - pp;
        const int j0 = p * nb, m = n - j0, nbp = min(nb, n - 1 - j0);
        const double2* V = Vstore + (int64_t)nb * ((int64_t)p * n - (int64_t)nb * p * (p - 1) / 2);
        const double2* T = Tall + (size_t)p * nb * nb;
        const int G = (int)ceil_div64(m, PQ_ROWS);
        hipLaunchKernelGGL(k_pq_dots, dim3((unsigned)G), dim3(256), 0, st, m, nbp, k, V,
                           (const double2*)(X + j0), ldx, P);
        hipLaunchKernelGGL(k_pq_small, dim3((unsigned)k), dim3(512), 0, st, nbp, k, G, (const double2*)P, T, nb,
                           adjoint, W);
        hipLaunchKernelGGL(k_pq_update, dim3((unsigned)ceil_div64(m, PQU_THREADS)), dim3(PQU_THREADS), 0, st, m,
                           nbp, k, V, (const double2*)W, X + j0, ldx);
    }
    return hipGetLastError();
}

// One panel of any reflector layout: X <- (I - V T^(H) V^H) X for V (m x nbp,
// ld m, nbp <= 32), T (ld ldt), X (m rows, ld ldx, k <= 16 columns).  ws: at
// least bse_apply_q_thin_workspace(m) complex elements.
extern "C" hipError_t bse_apply_q_panel(hipStream_t st, int m, int nbp, const double2* V, const double2* T,
                                        int ldt, int k, double2* X, int64_t ldx, int adjoint, double2* ws)
{
    if (nbp > APQ_NB || k <= 0 || k > PQ_MAXK || m < 1 || nbp < 1) return hipErrorInvalidValue;
    double2* W = ws;
    double2* P = ws + 512;
    const int G = (int)ceil_div64(m, PQ_ROWS);
    hipLaunchKernelGGL(k_pq_dots, dim3((unsigned)G), dim3(256), 0, st, m, nbp, k, V, (const double2*)X, ldx, P);
    hipLaunchKernelGGL(k_pq_small, dim3((unsigned)k), dim3(512), 0, st, nbp, k, G, (const double2*)P, T, ldt,
                       adjoint, W);
    hipLaunchKernelGGL(k_pq_update, dim3((unsigned)ceil_div64(m, PQU_THREADS)), dim3(PQU_THREADS), 0, st, m, nbp,
                       k, V, (const double2*)W, X, ldx);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Copies, fills and time stamps as kernels, so that sequences recorded into a
// graph contain kernel nodes only.
// ---------------------------------------------------------------------------
__global__ void k_copy_words(int64_t n, const uint64_t* __restrict__ src, uint64_t* __restrict__ dst)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_fill_words(int64_t n, uint64_t v, uint64_t* __restrict__ dst)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

// dst[0] = device wall clock (constant-rate counter).
__global__ void k_stamp(unsigned long long* dst)
{
    if (threadIdx.x == 0) dst[0] = wall_clock64();
}

// n8: number of 8-byte words.
extern "C" hipError_t bse_launch_copy_words(hipStream_t st, int64_t n8, const void* src, void* dst)
{
    if (n8 <= 0) return hipSuccess;
    const unsigned g = (unsigned)std::min<int64_t>(ceil_div64(n8, 256), 2048);
    hipLaunchKernelGGL(k_copy_words, dim3(g), dim3(256), 0, st, n8, (const uint64_t*)src, (uint64_t*)dst);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_zero_words(hipStream_t st, int64_t n8, void* dst)
{
    if (n8 <= 0) return hipSuccess;
    const unsigned g = (unsigned)std::min<int64_t>(ceil_div64(n8, 256), 2048);
    hipLaunchKernelGGL(k_fill_words, dim3(g), dim3(256), 0, st, n8, (uint64_t)0, (uint64_t*)dst);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_stamp(hipStream_t st, unsigned long long* dst)
{
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, st, dst);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Spectrum-misfit gradient factors.  With g_m(w) = 1/(w + i eta - E_m), the
// misfit residual r_w = Im eps(w) - target(w) and oscillator strengths f_m:
//   Fbr[m, w] = conj(f_m) g_m(w) r_w,   Fac[m, w] = conj(f_m g_m(w)),
//   sa[m] = f_m sum_w r_w g_m(w),       sb[m] = conj(f_m) sum_w r_w g_m(w).
// One thread per exciton m; column-major N x nw outputs (coalesced per w).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_spec_factors(int N, int nw, const double* __restrict__ omega,
                                                      double eta, const double* __restrict__ E,
                                                      const double2* __restrict__ f,
                                                      const double* __restrict__ r,
                                                      double2* __restrict__ Fbr,
                                                      double2* __restrict__ Fac,
                                                      double2* __restrict__ sa,
                                                      double2* __restrict__ sb)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= N) return;
    const double2 fm = f[m], fc = cconj(fm);
    const double em = E[m];
    double2 acc = cmk(0, 0);
    for (int w = 0; w < nw; ++w) {
        const double dw = omega[w] - em;
        const double den = dw * dw + eta * eta;
        const double2 g = cmk(dw / den, -eta / den);
        const double rw = r[w];
        const int64_t o = m + (int64_t)w * N;
        Fbr[o] = cscale(cmul(fc, g), rw);
        Fac[o] = cconj(cmul(fm, g));
        acc = cadd(acc, cscale(g, rw));
    }
    sa[m] = cmul(fm, acc);
    sb[m] = cmul(fc, acc);
}

// out[0] = sum_i B(i, i) (fixed order; one block of 1024 threads).
__global__ __launch_bounds__(1024) void k_trace(int64_t N, const double2* __restrict__ B, int64_t ldb,
                                                double2* __restrict__ out)
{
    __shared__ double sh[32];
    double2 acc = cmk(0, 0);
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) acc = cadd(acc, B[i + i * ldb]);
    acc = block_sum2(acc, sh);
    if (threadIdx.x == 0) out[0] = cadd(out[0], acc);
}

extern "C" hipError_t bse_launch_spec_factors(hipStream_t st, int N, int nw, const double* omega,
                                              double eta, const double* E, const double2* f,
                                              const double* r, double2* Fbr, double2* Fac,
                                              double2* sa, double2* sb)
{
    hipLaunchKernelGGL(k_spec_factors, dim3((unsigned)ceil_div64(N, 256)), dim3(256), 0, st, N, nw,
                       omega, eta, E, f, r, Fbr, Fac, sa, sb);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_trace(hipStream_t st, int64_t N, const double2* B, int64_t ldb,
                                       double2* out)
{
    hipLaunchKernelGGL(k_trace, dim3(1), dim3(1024), 0, st, N, B, ldb, out);
    return hipGetLastError();
}
