// K5: blocked Householder reduction of a Hermitian matrix to real symmetric
// tridiagonal form, Q^H A Q = T (lower storage).
//
// This is the dense half of the eigensolver that Armadillo's eig_sym hands to
// LAPACK zheevd (bse.version13.cpp:3330).  Algorithm: the panel form of the
// lower-triangular reduction (Golub & Van Loan, 5th ed., sec. 8.3; the
// LAPACK hetrd/latrd structure).  For each panel of nb columns, column by
// column:
//   (a) update the column with the panel's earlier reflectors:
//       a -= V conj(W(c,:))^T + W conj(V(c,:))^T
//   (b) reflector: beta = -sign(Re alpha) ||x||, tau = (beta - alpha)/beta,
//       v = [1; x / (alpha - beta)]
//   (c) y = A22 v          (Hermitian matrix-vector product, lower triangle)
//   (d) y -= V (W^H v) + W (V^H v)
//   (e) y *= tau;  w = y - (tau/2)(y^H v) v
// and after the panel the trailing update A22 -= V W^H + W V^H.
//
// The matrix may be stored in FP64 or FP32 complex; the column, reflectors
// and the V/W panels are always FP64.  Reflectors go to a separate packed
// FP64 store so the back-transform is exact FP64 even for FP32 storage.
//
// All reductions are deterministic and need no atomics: every kernel of the
// column step writes per-block partials, and every block of the kernel that
// consumes the result sums them itself, in a fixed order (identical values
// in every block).  A single "last block" counter would serialise one atomic
// per block across the chip's XCDs, which costs more than the redundant sums.
//
// Per column c = j0 + i the kernels are
//   colupd(i):  alpha2[i-1] from the reduce(i-1) partials; column update;
//               norm partials                                  -> region A
//   hemv(i):    partials of A22 x (x: the updated column below its first
//               subdiagonal entry, before scaling)             -> Yr, Yc;
//               the first workgroups also form partial W^H x,
//               V^H x (region B), and workgroup (0, 0) the
//               reflector scalars from region A                -> d, e, tau
//   reduce(i):  v_i = e_1 + sc x (stored);
//               A22 v = sc (A22 x) + A22 e_1; W^H v, V^H v from region B;
//               W(:, i); partial y^H v                         -> region C
// The reflector's scale sc = 1/(alpha - beta) depends on the column norm (a
// grid-wide sum), so the product is formed with the unscaled column and
// rescaled afterwards: the hemv needs no scalars and starts streaming at
// once.
// and at the panel end panel_close (alpha2 of the last column from region C;
// in the batched path also the compact-WY factor, which the single-matrix
// path forms for all panels after the last column) before the trailing
// update.
#include "bse_common.hpp"
#include <cstdlib>

struct HetrdCol {
    int n;            // matrix size (tiled lower storage, bse_common.hpp)
    int j0;           // panel start column
    int i;            // column inside the panel
    int m;            // n - j0 (panel rows)
    double2* Vp;      // panel reflectors, m x nb, ld m
    double2* Wp;      // panel W without its alpha2 v term, m x nb, ld m
    double2* alpha2;  // per panel column: W(:, l) = Wp(:, l) + alpha2[l] Vp(:, l)
    double2* G;       // nb x nb: G(l, i) = V(:, l)^H v_i  (for the compact-WY factor)
    double2* acol;    // updated column, length m - i (row c .. n-1)
    double* d;
    double* e;
    double2* tau;
    double* scal;     // [0,1] scale (1/(alpha-beta)), [2,3] tau
    double* part;     // reduction partials
    unsigned int* counter;
};

// Partial-sum regions inside HetrdCol::part (doubles), fixed per matrix size.
struct PartLayout {
    int64_t a, b, c;
};
__host__ __device__ inline PartLayout part_layout(int n) {
    PartLayout L;
    const int64_t na = ceil_div64(n, 64) + 1;               // colupd blocks
    const int64_t nb = 4 * 32 * (ceil_div64(n, 64) + 1);    // 2 dots x complex x l < 32 x tile columns
    L.a = 0;
    L.b = na;
    L.c = na + nb;                                          // + 2 (ceil(n/64) + 1): reduce blocks
    return L;
}
__host__ __device__ inline int64_t part_doubles(int n) {
    return part_layout(n).c + 2 * (ceil_div64(n, 64) + 1);
}

// Workgroups of the column update of column c = j0 + i (64 rows each): the
// number of norm partials in region A.
__host__ __device__ inline int colupd_blocks(int m, int i) { return (int)ceil_div64(m - i, 64); }

// W(r, l) of the panel, with the alpha2 v term applied on the fly.
__device__ inline double2 w_at(const HetrdCol& h, int64_t r, int l) {
    return cadd(h.Wp[r + (int64_t)l * h.m], cmul(h.alpha2[l], h.Vp[r + (int64_t)l * h.m]));
}

// Block bodies.  Each kernel's work is written as a device function of the
// block coordinates and of the grid it belongs to, so that the same code runs
// as a single-matrix kernel (coordinates = blockIdx) and inside the batched
// kernels further down, where blockIdx.z selects one matrix of an uneven batch
// and blocks beyond that matrix's own grid return at once.

// Sum of the complex partials p[0..cnt) in a fixed order, in every thread.
__device__ inline double2 block_sum_partials(const double2* __restrict__ p, int cnt, double* sh) {
    double2 x = cmk(0, 0);
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) x = cadd(x, p[k]);
    return block_sum2(x, sh);
}

// (a): column update and its norm partials.  alpha2 of the previous column
// (i - 1) is finished here from the reduce partials (prev_gx of them) by every
// block; block 0 stores it and clears W(0:i, i) and alpha2[i] (rows above the
// reflector are never read, but the whole column is kept defined).  A block
// covers 64 rows; its four waves split the panel columns (l = w + 4k) so that
// each lane issues all of its loads at once, before the partial alpha2 sum.
#define COLUPD_ROWS 64
template <typename T>
__device__ __attribute__((always_inline)) inline void colupd_body(const HetrdCol& h, const T* __restrict__ A, int bx,
                                                                   int prev_gx)
{
    __shared__ double sh[32];
    __shared__ double2 s_wc[32], s_vc[32], s_a2[32];
    __shared__ double2 s_part[4][COLUPD_ROWS];
    const int c = h.j0 + h.i;
    const int L = h.m - h.i;                 // rows c .. n-1
    const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int t = bx * COLUPD_ROWS + lane;
    const bool row_ok = t < L;
    const int64_t lrow = h.i + t;
    const PartLayout pl = part_layout(h.n);
    double2 vr[8], wq[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int l = grp + 4 * k;
        vr[k] = cmk(0, 0);
        wq[k] = cmk(0, 0);
        if (row_ok && l < h.i) {
            vr[k] = h.Vp[lrow + (int64_t)l * h.m];
            wq[k] = h.Wp[lrow + (int64_t)l * h.m];
        }
    }
    double2 a = cmk(0, 0);
    if (grp == 0 && row_ok) a = cplx<T>::load(A + tix(c + t, c, tile_count(h.n)));
    // row i of the panel's V and W (the same for every thread): loaded before
    // the partial sum of alpha2[i-1], which only W(i, i-1) needs
    double2 rv = cmk(0, 0), rw = cmk(0, 0), ra = cmk(0, 0);
    if ((int)threadIdx.x < h.i) {
        const int l = threadIdx.x;
        rv = h.Vp[h.i + (int64_t)l * h.m];
        rw = h.Wp[h.i + (int64_t)l * h.m];
        if (l < h.i - 1) ra = h.alpha2[l];
    }
    double2 a2prev = cmk(0, 0);
    if (h.i > 0) {
        const double2 sp = block_sum_partials((const double2*)(h.part + pl.c), prev_gx, sh);
        a2prev = cscale(cmul(h.tau[c - 1], sp), -0.5);
    }
    if (bx == 0) {
        if ((int)threadIdx.x <= h.i && (int)threadIdx.x < h.m)
            h.Wp[threadIdx.x + (int64_t)h.i * h.m] = cmk(0, 0);
        if (threadIdx.x == 0) {
            h.alpha2[h.i] = cmk(0, 0);
            if (h.i > 0) h.alpha2[h.i - 1] = a2prev;
        }
    }
    if ((int)threadIdx.x < h.i) {
        const int l = threadIdx.x;
        const double2 a2 = (l == h.i - 1) ? a2prev : ra;
        s_vc[l] = rv;
        s_wc[l] = cadd(rw, cmul(a2, rv));
        s_a2[l] = a2;
    }
    __syncthreads();
    double2 acc = cmk(0, 0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int l = grp + 4 * k;
        if (l < h.i) {
            const double2 wr = cadd(wq[k], cmul(s_a2[l], vr[k]));
            acc = cadd(acc, cadd(cmul(vr[k], cconj(s_wc[l])), cmul(wr, cconj(s_vc[l]))));
        }
    }
    s_part[grp][lane] = acc;
    __syncthreads();
    double nrm = 0.0;
    if (grp == 0 && row_ok) {
        a = csub(a, cadd(cadd(s_part[0][lane], s_part[1][lane]), cadd(s_part[2][lane], s_part[3][lane])));
        if (t == 0) a.y = 0.0;
        h.acol[t] = a;
        if (t >= 2) nrm = cabs2(a);
    }
    nrm = block_sum(nrm, sh);
    if (threadIdx.x == 0) h.part[pl.a + bx] = nrm;
}

// (b): reflector scalars from the colupd partials, computed identically by
// every block of the next kernel: beta = -sign(Re alpha) ||x||,
// tau = (beta - alpha)/beta, scale 1/(alpha - beta).
struct Refl {
    double d, e;
    double2 tau, sc;
};
__device__ inline Refl reflector_scalars(const HetrdCol& h, int colupd_gx, double* sh) {
    const PartLayout pl = part_layout(h.n);
    double xp = 0.0;
    for (int b = threadIdx.x; b < colupd_gx; b += blockDim.x) xp += h.part[pl.a + b];
    const double xn2 = block_sum(xp, sh);
    Refl R;
    R.d = h.acol[0].x;
    const double2 alpha = h.acol[1];
    if (xn2 == 0.0 && alpha.y == 0.0) {
        R.e = alpha.x;
        R.tau = cmk(0, 0);
        R.sc = cmk(0, 0);
    } else {
        const double nr = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        const double beta = (alpha.x >= 0.0) ? -nr : nr;
        R.e = beta;
        R.tau = cmk((beta - alpha.x) / beta, -alpha.y / beta);
        R.sc = cdiv(cmk(1, 0), cmk(alpha.x - beta, alpha.y));
    }
    return R;
}


// (c): y = A22 v for the trailing Hermitian block A22 = A[s0:, s0:] (s0 =
// c + 1), over the 64 x 64 storage tiles that intersect it: tile column
// TJ = tc + bj, tile rows TI = tc + bi (tc = s0 / 64).  Workgroup (q, bj)
// handles tile rows bi in [bj + q*CH, bj + (q+1)*CH) of tile column bj; the
// tiles of one tile column are consecutive in memory, so a workgroup streams
// one contiguous run.  Row results of each tile go to Yr[bj][bi] (so that
// the rows of one workgroup are written as one contiguous block: scattered
// 1 KB writes cost the streaming reads about a seventh of their rate); the
// transposed (conjugate) contribution for tile column bj is accumulated across
// the chunk and written to Yc[bj][q].  One wave owns 16 columns of the tile.
// The chunk length CH is chosen per trailing size by the autotuner.
//
// Every tile is read whole and without masks.  This relies on the storage
// invariant that padding entries (rows or columns >= n) and the strict upper
// half of each diagonal tile are zero, so they add nothing; entries of the
// first tile row and column that lie before s0 are finite and meet v = 0 (rows
// and columns of v outside [s0, n) are zero), and the column sums they produce
// belong to columns < s0, which the reduction never reads.  A diagonal tile
// adds a(r, r) v(r) to both the row and the column sum, so Re a(r, r) v(r) is
// taken off the row sum once (which also drops any imaginary part of the
// diagonal).
#define HEMV_CH_MIN 4
// Variant code: ch + 32 u2 + 128 t (ch in {4, 8, 16}; u2 = 1 two tile rows
// per step, u2 = 2 per-slab cross-lane column sums, u2 = 3 the same over
// blocks of four tile columns, which cuts the row partials fourfold; t = 1
// default cache policy
// for the matrix loads, t = 0 non-temporal: a large trailing block is
// streamed once per column and cannot stay in the caches, a small one can).
// AUX is the resulting load cache-policy operand.
#define HEMV_CODE_DEFAULT (8 + 64)
__host__ __device__ inline bool hemv_code_valid(int code) {
    const int ch = code & 31;
    return code >= 0 && code < 256 && (ch == 4 || ch == 8 || ch == 16);
}
// Tile rows per workgroup chunk, and tile columns per workgroup, of a
// variant code.
__host__ __device__ inline int hemv_chunk(int code) { return code & 31; }
__host__ __device__ inline int hemv_rblk(int code) { return ((code >> 5) & 3) == 3 ? 4 : 1; }

// Column sums of sixteen per-lane complex values over the 64 lanes of a wave
// by reduce-scatter: after exchanges at lane distances 32, 16, 8, 4 every lane
// keeps one column, col = 8 b5 + 4 b4 + 2 b3 + b2 (b_k = bit k of the lane),
// and two more exchanges (2, 1) complete its sum in all four lanes holding
// it.  17 complex exchanges instead of 16 x 6 for separate wave sums, and no
// per-column accumulators outlive the slab.
__device__ inline double2 shfl_xor_c(double2 v, int m) {
    return cmk(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}
__device__ __attribute__((always_inline)) inline double2 colsum16(const double2 (&c)[16], int lane) {
    double2 a[8], b[4], e[2];
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8, b2 = lane & 4;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        a[k] = cadd(b5 ? c[k + 8] : c[k], shfl_xor_c(b5 ? c[k] : c[k + 8], 32));
#pragma unroll
    for (int k = 0; k < 4; ++k)
        b[k] = cadd(b4 ? a[k + 4] : a[k], shfl_xor_c(b4 ? a[k] : a[k + 4], 16));
#pragma unroll
    for (int k = 0; k < 2; ++k)
        e[k] = cadd(b3 ? b[k + 2] : b[k], shfl_xor_c(b3 ? b[k] : b[k + 2], 8));
    double2 d = cadd(b2 ? e[1] : e[0], shfl_xor_c(b2 ? e[0] : e[1], 4));
    d = cadd(d, shfl_xor_c(d, 2));
    d = cadd(d, shfl_xor_c(d, 1));
    return d;
}
__device__ inline int colsum16_col(int lane) {
    return 8 * ((lane >> 5) & 1) + 4 * ((lane >> 4) & 1) + 2 * ((lane >> 3) & 1) + ((lane >> 2) & 1);
}

// The vector of the product: an explicit v (indexed from s0), or the updated
// column c = s0 - 1 below its first subdiagonal entry, x(g) = acol(g - c)
// for g > s0 and x(s0) = 0; zero outside [s0, n).
struct HemvVec {
    const double2* v;
    const double2* acol;
    __device__ inline double2 at(int g, int s0, int n) const {
        if (v) return (g >= s0 && g < n) ? v[g - s0] : cmk(0, 0);
        return (g > s0 && g < n) ? acol[g - s0 + 1] : cmk(0, 0);
    }
};

// Reflector scalars of column c = j0 + i from the colupd norm partials,
// formed by one workgroup of the hemv (which does not need them itself):
// d, e, tau, and scal = (sc, tau) for the reduce.
__device__ inline void hemv_scalars(const HetrdCol& h)
{
    __shared__ double sh[32];
    const Refl R = reflector_scalars(h, colupd_blocks(h.m, h.i), sh);
    if (threadIdx.x == 0) {
        const int c = h.j0 + h.i;
        h.d[c] = R.d;
        h.e[c] = R.e;
        h.tau[c] = R.tau;
        h.scal[0] = R.sc.x; h.scal[1] = R.sc.y; h.scal[2] = R.tau.x; h.scal[3] = R.tau.y;
    }
}

// Partial dots W(:, l)^H x and V(:, l)^H x (l < i) over the 64 rows of
// storage tile row TJ = tc + t: region B entry ((l nbk + t) 2 + {0: W, 1: V}).
// Done by the first nbk workgroups in dispatch order, before their own tiles,
// so that this work never lands in the tail of the grid.  Wave w takes
// l = w + 4k (k < 8, so i <= 32), and one cross-lane reduce-scatter sums the
// 16 products of a lane.
__device__ __attribute__((always_inline)) inline void hemv_dots(const HetrdCol& h, const HemvVec& vs,
                                                                 int TJ, int bj, int nbk)
{
    const int c = h.j0 + h.i, s0 = c + 1, n = h.n;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = TJ * 64 + lane;
    if (h.i == 0) return;
    const bool ok = g > s0 && g < n;
    const double2 vg = vs.at(g, s0, n);
    const int64_t lr = g - h.j0;
    double2 cc[16];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int l = w + 4 * k;
        cc[k] = cmk(0, 0);
        cc[8 + k] = cmk(0, 0);
        if (ok && l < h.i) {
            const double2 vv = h.Vp[lr + (int64_t)l * h.m];
            const double2 ww = cadd(h.Wp[lr + (int64_t)l * h.m], cmul(h.alpha2[l], vv));
            cfmac(cc[k], ww, vg);
            cfmac(cc[8 + k], vv, vg);
        }
    }
    const double2 sum = colsum16(cc, lane);
    const int col = colsum16_col(lane), l = w + 4 * (col & 7);
    if ((lane & 3) == 0 && l < h.i) {
        double2* pp = (double2*)(h.part + part_layout(n).b);
        pp[((int64_t)l * nbk + bj) * 2 + (col >> 3)] = sum;
    }
}

// Re a(row, row) of a diagonal tile from this wave's 16 loaded columns (zero
// in lanes whose diagonal column belongs to another wave).
template <typename T>
__device__ __attribute__((always_inline)) inline double hemv_diag_re(const T (&a)[16], int lane, int w) {
    double d = 0.0;
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) d = (w * 16 + jj == lane) ? (double)a[jj].x : d;
    return d;
}

template <typename T, int CH, int U2, int AUX>
__device__ __attribute__((always_inline)) inline void hemv_body(int n, int s0, const T* __restrict__ A,
                                                                 const double2* __restrict__ v,
                                                                 double2* __restrict__ Yr,
                                                                 double2* __restrict__ Yc, int nbk,
                                                                 int nq, int q, int bj,
                                                                 const HetrdCol& h)
{
    HemvVec vs;
    vs.v = v;
    vs.acol = h.acol;
    if (!v && q == 0 && bj == 0) hemv_scalars(h);
    if (!v && bj * nq + q < nbk) hemv_dots(h, vs, (s0 >> 6) + bj * nq + q, bj * nq + q, nbk);
    const int bi0 = bj + q * CH;
    if (bi0 >= nbk) return;
    const int bi1 = min(nbk, bi0 + CH);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntile = tile_count(n), tc = s0 >> 6, TJ = tc + bj;
    __shared__ double2 s_v[64];
    __shared__ double2 s_part[4][CH][64];
    if (threadIdx.x < 64) s_v[threadIdx.x] = vs.at(TJ * 64 + threadIdx.x, s0, n);
    __syncthreads();
    const double2* sv = s_v + w * 16;         // this wave's 16 entries of v (LDS)
    double2 cacc[16];
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) cacc[jj] = cmk(0, 0);
    // the chunk's tiles are one contiguous run starting at tile (tc+bi0, TJ);
    // descriptor base = that run + this wave's first column, lane offset
    // (tile, row) and column offset jj * 64 elements in a scalar register
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const T* run = A + tile_pos(tc + bi0, TJ, ntile) * (BSE_TILE * BSE_TILE);
    const __amdgpu_buffer_rsrc_t rs = bse_rsrc(run + wu * 16 * 64, 0x7fffffff);
    const int cstride = 64 * (int)sizeof(T);
    for (int bi = bi0; bi < bi1;) {
        const int TI = tc + bi;
        const bool diag = TI == TJ;
        if (U2 && sizeof(T) == 8 && !diag && bi + 1 < bi1) {
            // FP32 storage, two off-diagonal tiles: lane l reads the
            // consecutive rows 2l, 2l+1 of the 128-row slab (tile TI for
            // l < 32, TI+1 above) in one 16-byte load
            const int hb = lane >> 5, rl = (2 * lane) & 63;
            const int vo2 = ((bi - bi0 + hb) * 4096 + rl) * (int)sizeof(T);
            bse_i4v qv[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) qv[jj] = __builtin_amdgcn_raw_buffer_load_b128(rs, vo2, jj * cstride, AUX);
            const int r2 = (TI + hb) * 64 + rl;
            const double2 v0 = vs.at(r2, s0, n), v1 = vs.at(r2 + 1, s0, n);
            double2 r0 = cmk(0, 0), r1 = cmk(0, 0);
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const int q0 = qv[jj][0], q1 = qv[jj][1], q2 = qv[jj][2], q3 = qv[jj][3];
                const double2 x = cmk((double)bse_i2f(q0), (double)bse_i2f(q1));
                const double2 y = cmk((double)bse_i2f(q2), (double)bse_i2f(q3));
                cfma(r0, x, sv[jj]);
                cfmac(cacc[jj], x, v0);
                cfma(r1, y, sv[jj]);
                cfmac(cacc[jj], y, v1);
            }
            s_part[w][bi + hb - bi0][rl] = r0;
            s_part[w][bi + hb - bi0][rl + 1] = r1;
            bi += 2;
            continue;
        }
        const int vo = ((bi - bi0) * 4096 + lane) * (int)sizeof(T);
        const int row = TI * 64 + lane;
        if (U2 && sizeof(T) == 16 && !diag && bi + 1 < bi1) {
            // two off-diagonal tiles: 32 loads in flight per lane
            T a[16], b[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                a[jj] = bufld<T>::template ld<AUX>(rs, vo, jj * cstride);
                b[jj] = bufld<T>::template ld<AUX>(rs, vo + 4096 * (int)sizeof(T), jj * cstride);
            }
            const double2 vr0 = vs.at(row, s0, n), vr1 = vs.at(row + 64, s0, n);
            double2 r0 = cmk(0, 0), r1 = cmk(0, 0);
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const double2 x = cmk((double)a[jj].x, (double)a[jj].y);
                const double2 y = cmk((double)b[jj].x, (double)b[jj].y);
                cfma(r0, x, sv[jj]);
                cfmac(cacc[jj], x, vr0);
                cfma(r1, y, sv[jj]);
                cfmac(cacc[jj], y, vr1);
            }
            s_part[w][bi - bi0][lane] = r0;
            s_part[w][bi + 1 - bi0][lane] = r1;
            bi += 2;
            continue;
        }
        const double2 vr = vs.at(row, s0, n);
        double2 racc = cmk(0, 0);
        T a[16];
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) a[jj] = bufld<T>::template ld<AUX>(rs, vo, jj * cstride);
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const double2 x = cmk((double)a[jj].x, (double)a[jj].y);
            cfma(racc, x, sv[jj]);
            cfmac(cacc[jj], x, vr);
        }
        if (diag) racc = csub(racc, cscale(vr, hemv_diag_re(a, lane, w)));
        s_part[w][bi - bi0][lane] = racc;
        ++bi;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (bi1 - bi0) * 64; e += 256) {
        const int bb = e >> 6, ln = e & 63;
        const double2 s = cadd(cadd(s_part[0][bb][ln], s_part[1][bb][ln]),
                               cadd(s_part[2][bb][ln], s_part[3][bb][ln]));
        Yr[((int64_t)bj * nbk + bi0 + bb) * 64 + ln] = s;
    }
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
        const double x = wave_sum(cacc[jj].x), y = wave_sum(cacc[jj].y);
        if (lane == 0) Yc[((int64_t)bj * nq + q) * 64 + w * 16 + jj] = cmk(x, y);
    }
}



// Hemv variant with few live registers: the column partials of each slab (128
// rows for FP32 storage read as row pairs, 64 rows otherwise) are reduced
// across the wave at once and added to one FP64 accumulator per lane, so the
// kernel keeps four or more waves per SIMD while a full slab of loads is in
// flight.  Same tiles, masks and partial outputs as hemv_body.
template <typename T, int CH, int R, int AUX>
__device__ __attribute__((always_inline)) inline void hemv_bfly_body(int n, int s0, const T* __restrict__ A,
                                                                      const double2* __restrict__ v,
                                                                      double2* __restrict__ Yr,
                                                                      double2* __restrict__ Yc, int nbk,
                                                                      int nq, int q, int pb, int npb,
                                                                      const HetrdCol& h)
{
    HemvVec vs;
    vs.v = v;
    vs.acol = h.acol;
    if (!v && q == 0 && pb == 0) hemv_scalars(h);
    if (!v)
        for (int t = pb * nq + q; t < nbk; t += nq * npb) hemv_dots(h, vs, (s0 >> 6) + t, t, nbk);
    // workgroup (q, pb): tile columns bj = pb R + r (r < R), tile rows
    // [row0, row1) with row0 = pb R + q CH; lower tiles only (bi >= bj)
    const int row0 = pb * R + q * CH;
    if (row0 >= nbk) return;
    const int row1 = min(nbk, row0 + CH);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ntile = tile_count(n), tc = s0 >> 6;
    __shared__ double2 s_v[R][64];
    __shared__ double2 s_part[4][CH][64];
    for (int e = threadIdx.x; e < R * 64; e += 256)
        s_v[e >> 6][e & 63] = vs.at((tc + pb * R + (e >> 6)) * 64 + (e & 63), s0, n);
    if (R > 1) {
        // row sums accumulate over the R tile columns: this wave's slots start at zero
#pragma unroll
        for (int bb = 0; bb < CH; ++bb) s_part[w][bb][lane] = cmk(0, 0);
    }
    __syncthreads();
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const int cstride = 64 * (int)sizeof(T);
#pragma unroll 1
    for (int r = 0; r < R; ++r) {
        double2 colacc = cmk(0, 0);
        const int bj = pb * R + r, TJ = tc + bj;
        const int bstart = max(row0, bj);
        if (bj >= nbk || bstart >= row1) continue;
        const double2* sv = s_v[r] + w * 16;
        const T* run = A + tile_pos(tc + bstart, TJ, ntile) * (BSE_TILE * BSE_TILE);
        const __amdgpu_buffer_rsrc_t rs = bse_rsrc(run + wu * 16 * 64, 0x7fffffff);
        for (int bi = bstart; bi < row1;) {
            const int TI = tc + bi;
            const bool diag = TI == TJ;
            double2 c[16];
            if (sizeof(T) == 8 && !diag && bi + 1 < row1) {
                const int hb = lane >> 5, rl = (2 * lane) & 63;
                const int vo2 = ((bi - bstart + hb) * 4096 + rl) * (int)sizeof(T);
                bse_i4v qv[16];
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) qv[jj] = __builtin_amdgcn_raw_buffer_load_b128(rs, vo2, jj * cstride, AUX);
                const int r2 = (TI + hb) * 64 + rl;
                const double2 v0 = vs.at(r2, s0, n), v1 = vs.at(r2 + 1, s0, n);
                double2 r0 = cmk(0, 0), r1 = cmk(0, 0);
#pragma unroll
                for (int jj = 0; jj < 16; ++jj) {
                    const int q0 = qv[jj][0], q1 = qv[jj][1], q2 = qv[jj][2], q3 = qv[jj][3];
                    const double2 x = cmk((double)bse_i2f(q0), (double)bse_i2f(q1));
                    const double2 y = cmk((double)bse_i2f(q2), (double)bse_i2f(q3));
                    const double2 sj = sv[jj];
                    cfma(r0, x, sj);
                    cfma(r1, y, sj);
                    double2 cc = cmk(0, 0);
                    cfmac(cc, x, v0);
                    cfmac(cc, y, v1);
                    c[jj] = cc;
                }
                colacc = cadd(colacc, colsum16(c, lane));
                double2* sp = &s_part[w][bi + hb - row0][rl];
                if (R > 1) {
                    sp[0] = cadd(sp[0], r0);
                    sp[1] = cadd(sp[1], r1);
                } else {
                    sp[0] = r0;
                    sp[1] = r1;
                }
                bi += 2;
                continue;
            }
            const int row = TI * 64 + lane;
            const double2 vr = vs.at(row, s0, n);
            double2 racc = cmk(0, 0);
            const int vo = ((bi - bstart) * 4096 + lane) * (int)sizeof(T);
            T a[16];
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) a[jj] = bufld<T>::template ld<AUX>(rs, vo, jj * cstride);
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const double2 x = cmk((double)a[jj].x, (double)a[jj].y);
                cfma(racc, x, sv[jj]);
                double2 cc = cmk(0, 0);
                cfmac(cc, x, vr);
                c[jj] = cc;
            }
            if (diag) racc = csub(racc, cscale(vr, hemv_diag_re(a, lane, w)));
            colacc = cadd(colacc, colsum16(c, lane));
            double2* sp = &s_part[w][bi - row0][lane];
            *sp = (R > 1) ? cadd(*sp, racc) : racc;
            ++bi;
        }
        if ((lane & 3) == 0) Yc[((int64_t)bj * nq + q) * 64 + w * 16 + colsum16_col(lane)] = colacc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (row1 - row0) * 64; e += 256) {
        const int bb = e >> 6, ln = e & 63;
        const double2 s = cadd(cadd(s_part[0][bb][ln], s_part[1][bb][ln]),
                               cadd(s_part[2][bb][ln], s_part[3][bb][ln]));
        Yr[((int64_t)pb * nbk + row0 + bb) * 64 + ln] = s;
    }
}

// Occupancy bound of the hemv kernels: two waves per SIMD (up to 256 VGPRs)
// with 16 to 32 loads in flight per lane; tighter bounds make the compiler
// spill the column accumulators.
#ifndef HEMV_WPE_FOR
#define HEMV_WPE_FOR(U2) 2
#endif
template <typename T, int CH, int U2, int AUX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HEMV_WPE_FOR(U2), 8)))
void k_hemv_lower(int n, int s0, const T* __restrict__ A, const double2* __restrict__ v,
                  double2* __restrict__ Yr, double2* __restrict__ Yc, int nbk, int nq, HetrdCol h)
{
    if (U2 == 3) hemv_bfly_body<T, CH, 4, AUX>(n, s0, A, v, Yr, Yc, nbk, nq, blockIdx.x, blockIdx.y, gridDim.y, h);
    else if (U2 == 2) hemv_bfly_body<T, CH, 1, AUX>(n, s0, A, v, Yr, Yc, nbk, nq, blockIdx.x, blockIdx.y, gridDim.y, h);
    else hemv_body<T, CH, U2 & 1, AUX>(n, s0, A, v, Yr, Yc, nbk, nq, blockIdx.x, blockIdx.y, h);
}

// Instantiates M(CH, U2, AUX) for the variant of a (valid) code.
#define HEMV_DISPATCH_CH(ch_, UU, AX, M) \
    { if (ch_ == 4) M(4, UU, AX); else if (ch_ == 16) M(16, UU, AX); else M(8, UU, AX); }
#define HEMV_DISPATCH_U2(ch_, u2_, AX, M) \
    { if (u2_ == 3) HEMV_DISPATCH_CH(ch_, 3, AX, M) else if (u2_ == 2) HEMV_DISPATCH_CH(ch_, 2, AX, M) \
      else if (u2_ == 1) HEMV_DISPATCH_CH(ch_, 1, AX, M) else HEMV_DISPATCH_CH(ch_, 0, AX, M) }
#define HEMV_DISPATCH(code, M) \
    { const int ch_ = (code) & 31, u2_ = ((code) >> 5) & 3; \
      if ((code) & 128) HEMV_DISPATCH_U2(ch_, u2_, 0, M) else HEMV_DISPATCH_U2(ch_, u2_, BSE_LOAD_STREAM, M) }

// Reduce the hemv partials of one block row (64 rows) per workgroup and
// finish the reflector step of column c = j0 + i, with sc = 1/(alpha - beta)
// and tau from the hemv (scal, tau[c]).  t1 = W^H v and t2 = V^H v come from
// the hemv's region-B partials of x (one per tile column; sixteen lanes per
// sum and a fixed butterfly):
//   t1 = conj(W(s0, :)) + sc (W^H x),   t2 = conj(V(s0, :)) + sc (V^H x)
// (block 0 also stores t2 as column i of the Gram matrix).  Per row r >= s0:
//   v(r) = [r == s0] + sc x(r)                  -> V(:, i)
//   y(r) = tau (sc (A22 x)(r) + a(r, s0) - V(r, :) t1 - W(r, :) t2)
// (the panel corrections (d) and the tau scaling (e)), stored in Wp(:, i),
// with the partial y^H v in region C; alpha2[i] = -tau/2 (y^H v) is finished
// by the next kernel.  RED_G lane groups split the partial sums so that
// many loads are in flight; every load that does not depend on t1, t2 is
// issued before the one barrier that publishes them.
// Lane groups of the reduce (workgroup of 64 RED_G threads).  Eight keep the
// workgroup at 512 threads, so that two fit a CU and a trailing block of up
// to 512 tile rows reduces in one round (one 1024-thread workgroup per CU
// left half of the blocks for a second round above 256 tile rows).
#define RED_G 8
#define RED_NL (32 / RED_G)   // panel columns per lane group (i <= 32)
template <typename T>
__device__ __attribute__((always_inline)) inline void reduce_body(const HetrdCol& h, const T* __restrict__ A,
                                                                   const double2* __restrict__ Yr,
                                                                   const double2* __restrict__ Yc,
                                                                   int nbk, int nq, int ch, int rbk, int b)
{
    __shared__ double2 sacc[RED_G][64];
    __shared__ double2 s_t[2][64];
    __shared__ double sh[32];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int c = h.j0 + h.i, s0 = c + 1;
    const PartLayout pl = part_layout(h.n);
    const double2 sc = cmk(h.scal[0], h.scal[1]);
    const double2 tau = h.tau[c];
    // block b covers storage tile row (s0 / 64) + b of the trailing block
    const int grow = ((s0 >> 6) + b) * 64 + lane;       // matrix row
    const bool valid = grow >= s0 && grow < h.n;
    const int lr = grow - h.j0;                         // panel-local row
    double2 acc0 = cmk(0, 0), acc1 = cmk(0, 0);
    double2 vv[RED_NL], ww[RED_NL];
#pragma unroll
    for (int u = 0; u < RED_NL; ++u) vv[u] = ww[u] = cmk(0, 0);
    double2 a_rs0 = cmk(0, 0), xr = cmk(0, 0);
    if (valid) {
        // partials of this block row: Yr[pb][b] (pb <= b / rbk, one per
        // block of rbk tile columns) and Yc[b][0..nqb), several independent
        // loads issued before any of them is summed
        const double2* yr = Yr + (int64_t)b * 64 + lane;
        const int64_t ys = (int64_t)nbk * 64;
        const int pmax = b / rbk;
        int bj = g;
        for (; bj + 7 * RED_G <= pmax; bj += 8 * RED_G) {
            double2 x[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = yr[(bj + u * RED_G) * ys];
            acc0 = cadd(acc0, cadd(cadd(x[0], x[1]), cadd(x[2], x[3])));
            acc1 = cadd(acc1, cadd(cadd(x[4], x[5]), cadd(x[6], x[7])));
        }
        for (; bj + 3 * RED_G <= pmax; bj += 4 * RED_G) {
            const double2 x0 = yr[bj * ys], x1 = yr[(bj + RED_G) * ys];
            const double2 x2 = yr[(bj + 2 * RED_G) * ys], x3 = yr[(bj + 3 * RED_G) * ys];
            acc0 = cadd(acc0, cadd(x0, x1));
            acc1 = cadd(acc1, cadd(x2, x3));
        }
        for (; bj <= pmax; bj += RED_G) acc0 = cadd(acc0, yr[bj * ys]);
        const int nqb = (int)ceil_div64(nbk - pmax * rbk, ch);
        const double2* yc = Yc + (int64_t)b * nq * 64 + lane;
        int q = g;
        for (; q + RED_G < nqb; q += 2 * RED_G) {
            const double2 x0 = yc[(int64_t)q * 64], x1 = yc[(int64_t)(q + RED_G) * 64];
            acc1 = cadd(acc1, cadd(x0, x1));
        }
        if (q < nqb) acc1 = cadd(acc1, yc[(int64_t)q * 64]);
        // panel rows of this row for the corrections (l = g + RED_G u < i <= 32)
#pragma unroll
        for (int u = 0; u < RED_NL; ++u) {
            const int l = g + u * RED_G;
            if (l < h.i) {
                vv[u] = h.Vp[lr + (int64_t)l * h.m];
                ww[u] = w_at(h, lr, l);
            }
        }
        if (g == 0) {
            // A22 e_1: column s0 of the stored lower triangle (real diagonal)
            a_rs0 = cplx<T>::load(A + tix(grow, s0, tile_count(h.n)));
            if (grow == s0) a_rs0.y = 0.0;
            if (grow > s0) xr = h.acol[grow - c];
        }
    }
    // t1, t2 from the region-B partials
    for (int pr = threadIdx.x >> 4; pr < 64; pr += (int)(blockDim.x >> 4)) {
        const int k = threadIdx.x & 15;           // (l, which) = (pr / 2, pr % 2)
        const int ll = pr >> 1, which = pr & 1;
        double2 psum = cmk(0, 0), e1 = cmk(0, 0);
        if (ll < h.i) {
            // eight independent loads per step (the partial count is nbk)
            const double2* pp = (const double2*)(h.part + pl.b) + (int64_t)ll * nbk * 2 + which;
            int rb = k;
            for (; rb + 7 * 16 < nbk; rb += 8 * 16) {
                double2 x[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) x[u] = pp[(int64_t)(rb + 16 * u) * 2];
                psum = cadd(psum, cadd(cadd(cadd(x[0], x[1]), cadd(x[2], x[3])),
                                       cadd(cadd(x[4], x[5]), cadd(x[6], x[7]))));
            }
            for (; rb < nbk; rb += 16) psum = cadd(psum, pp[(int64_t)rb * 2]);
            if (k == 0) {
                const int64_t r0 = s0 - h.j0;
                e1 = which ? h.Vp[r0 + (int64_t)ll * h.m] : w_at(h, r0, ll);
            }
        }
#pragma unroll
        for (int m = 8; m >= 1; m >>= 1) psum = cadd(psum, shfl_xor_c(psum, m));
        if (k == 0 && ll < h.i) {
            const double2 t = cadd(cconj(e1), cmul(sc, psum));
            s_t[which][ll] = t;
            if (b == 0 && which) h.G[ll + (int64_t)h.i * 64] = t;
        }
    }
    __syncthreads();
    double2 corr = cmk(0, 0);
#pragma unroll
    for (int u = 0; u < RED_NL; ++u) {
        const int l = g + u * RED_G;
        if (valid && l < h.i) corr = cadd(corr, cadd(cmul(vv[u], s_t[0][l]), cmul(ww[u], s_t[1][l])));
    }
    sacc[g][lane] = csub(cadd(cmul(sc, cadd(acc0, acc1)), a_rs0), corr);
    __syncthreads();
    double2 dotp = cmk(0, 0);
    if (g == 0 && grow >= h.j0 && grow < h.n) {
        double2 vr = cmk(0, 0);
        if (valid) {
            double2 y = cmk(0, 0);
#pragma unroll
            for (int kk = 0; kk < RED_G; ++kk) y = cadd(y, sacc[kk][lane]);
            y = cmul(y, tau);
            h.Wp[lr + (int64_t)h.i * h.m] = y;
            vr = grow == s0 ? cmk(1, 0) : cmul(xr, sc);
            cfmac(dotp, y, vr);
        }
        h.Vp[lr + (int64_t)h.i * h.m] = vr;
    }
    // panel rows above the first tile row of the trailing block: V(:, i) = 0
    if (b == 0)
        for (int r = threadIdx.x; r < ((s0 >> 6) << 6) - h.j0; r += blockDim.x)
            h.Vp[r + (int64_t)h.i * h.m] = cmk(0, 0);
    dotp = block_sum2(dotp, sh);
    if (threadIdx.x == 0) ((double2*)(h.part + pl.c))[b] = dotp;
}

// Panel close: alpha2 of the panel's last column (from its reduce partials,
// red_gx of them) and, if form_t, the compact-WY factor of the panel.  One
// block of 64.
__device__ __attribute__((always_inline)) inline void panel_close_body(const HetrdCol& h, int nbp, int red_gx,
                                                                        double2* __restrict__ T, int ldt,
                                                                        bool form_t)
{
    __shared__ double sh[32];
    __shared__ double2 col[64];
    const int cl = h.j0 + nbp - 1;
    const double2 s = block_sum_partials((const double2*)(h.part + part_layout(h.n).c), red_gx, sh);
    if (threadIdx.x == 0) h.alpha2[nbp - 1] = cscale(cmul(h.tau[cl], s), -0.5);
    if (form_t) larft_body(nbp, h.tau + h.j0, h.G, T, ldt, col);
}

// Compact-WY factors of every panel of an n x n reduction after it (panel p
// = blockIdx.x, Gram block at G + p g_stride).  One block of 64 per panel.
__global__ __launch_bounds__(64) void k_larft_all(int n, int nb, const double2* __restrict__ tau,
                                                  const double2* __restrict__ G, int64_t g_stride,
                                                  double2* __restrict__ Tall)
{
    __shared__ double2 col[64];
    const int p = blockIdx.x, j0 = p * nb;
    if (j0 >= n - 1) return;
    const int nbp = min(nb, n - 1 - j0);
    larft_body(nbp, tau + j0, G + (int64_t)p * g_stride, Tall + (size_t)p * nb * nb, nb, col);
}


// Trailing update A22 -= V W^H + W V^H, K = nb, for the trailing block
// A[s:, s:] (s = j0 + nb) over the 64 x 64 storage tiles that intersect it:
// block idx is tile (tcs + ti, tcs + tj) with (ti, tj) the idx-th lower tile
// (tcs = s / 64); rows and columns below s are masked.  V and W are the panel
// (global row j0 at index 0, ld ldv).
template <typename T>
__device__ __attribute__((always_inline)) inline void her2k_valu_body(int n, int s, T* __restrict__ A,
                                       const double2* __restrict__ V,
                                       const double2* __restrict__ W, int64_t ldv, int j0, int nb,
                                       const double2* __restrict__ alpha2, int64_t idx)
{
    int ti, tj;
    tri_tile(idx, ti, tj);
    const int tcs = s >> 6;
    T* tile = A + tile_pos(tcs + ti, tcs + tj, tile_count(n)) * (BSE_TILE * BSE_TILE);
    __shared__ double2 sVr[8][64], sWr[8][64], sVc[8][64], sWc[8][64];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;   // 16 x 16 threads
    const int r0 = (tcs + ti) * 64, c0 = (tcs + tj) * 64;
    double2 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = cmk(0, 0);
    for (int l0 = 0; l0 < nb; l0 += 8) {
        for (int e = threadIdx.x; e < 8 * 64; e += 256) {
            const int rr = e & 63, ll = e >> 6;
            const int l = l0 + ll;
            const bool lok = l < nb;
            const int gr = r0 + rr, gc = c0 + rr;
            const bool rok = lok && gr >= s && gr < n, cok = lok && gc >= s && gc < n;
            const double2 a2 = lok ? alpha2[l] : cmk(0, 0);
            const double2 vr = rok ? V[(gr - j0) + (int64_t)l * ldv] : cmk(0, 0);
            const double2 vc = cok ? V[(gc - j0) + (int64_t)l * ldv] : cmk(0, 0);
            sVr[ll][rr] = vr;
            sWr[ll][rr] = rok ? cadd(W[(gr - j0) + (int64_t)l * ldv], cmul(a2, vr)) : cmk(0, 0);
            sVc[ll][rr] = vc;
            sWc[ll][rr] = cok ? cadd(W[(gc - j0) + (int64_t)l * ldv], cmul(a2, vc)) : cmk(0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int ll = 0; ll < 8; ++ll) {
            double2 vr[4], wr[4], vc[4], wc[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                vr[a] = sVr[ll][tx + 16 * a];
                wr[a] = sWr[ll][tx + 16 * a];
                vc[a] = sVc[ll][ty + 16 * a];
                wc[a] = sWc[ll][ty + 16 * a];
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    cfma(acc[a][b], vr[a], cconj(wc[b]));
                    cfma(acc[a][b], wr[a], cconj(vc[b]));
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int gr = r0 + tx + 16 * a, gc = c0 + ty + 16 * b;
            if (gc >= s && gr < n && gr >= gc) {
                T* p = tile + (tx + 16 * a) + (ty + 16 * b) * 64;
                double2 x = cplx<T>::load(p);
                x = csub(x, acc[a][b]);
                if (gr == gc) x.y = 0.0;
                cplx<T>::store(p, x);
            }
        }
}

// MFMA form of the trailing update, A22 -= V W^H + W V^H, lower 64x64 tiles.
// With P = [V W] and Q = [W V] (complex, K = 2 nb columns) and the real
// concatenations Pcat = [Re P | Im P], Pimc = [Im P | -Re P], Qcat = [Re Q | Im Q]:
//   Re(P Q^H) = Pcat Qcat^T,   Im(P Q^H) = Pimc Qcat^T.
// The kernel computes the transposed products Qcat Pcat^T so that the 16 lanes
// holding one MFMA output column map to 16 consecutive rows of the
// column-major matrix (coalesced read-modify-write).  Both storage types use
// v_mfma_f64_16x16x4 with the FP64 panels and FP64 accumulation, so an FP32
// matrix is rounded once per update, exactly as the FP64-VALU form does; an
// FP32 MFMA update (FP32 operands and accumulation) loses enough accuracy over
// N/nb updates to slow the FP64 refinement of the lowest excitons.
typedef double her2k_d4 __attribute__((ext_vector_type(4)));

struct her2k_f64 {
    typedef double real;
    typedef her2k_d4 acc;
    __device__ static inline acc mfma(real a, real b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int r) { return mfma16_row_f64(lane, r); }
};

// FP32 operands and accumulation (v_mfma_f32_16x16x4_f32): the band
// reduction's update of an FP32-stored matrix, whose products then carry the
// same relative rounding as the storage itself.
typedef float her2k_f4 __attribute__((ext_vector_type(4)));
struct her2k_f32 {
    typedef float real;
    typedef her2k_f4 acc;
    __device__ static inline acc mfma(real a, real b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int r) { return mfma16_row_f32(lane, r); }
};

// PK: the operands come packed in FP32 planes pl (real V, imaginary V, real
// W, imaginary W; plane q at pl + q nb ldv, element (row r, column l) at
// l ldv + r - j0) and alpha2 is taken as zero -- the band reduction's
// update, whose panels are then read at 4 bytes per real part instead of 8,
// and V once instead of twice; the values staged in LDS are the same.
template <typename T, typename MM = her2k_f64, bool PK = false>
__device__ __attribute__((always_inline)) inline void her2k_mfma_body(int n, int s, T* __restrict__ A,
                                       const double2* __restrict__ V,
                                       const double2* __restrict__ W, int64_t ldv, int j0, int nb,
                                       const double2* __restrict__ alpha2, int64_t idx,
                                       const float* __restrict__ pl = nullptr)
{
    typedef typename MM::real R;
    typedef typename MM::acc ACC;
    int ti, tj;
    tri_tile(idx, ti, tj);
    const int tcs = s >> 6;
    T* tile = A + tile_pos(tcs + ti, tcs + tj, tile_count(n)) * (BSE_TILE * BSE_TILE);
    // chunk of 16 complex K columns: real and imaginary parts, [kcol][row];
    // rows padded so that the four 16-lane groups of an operand read (four
    // consecutive kcol) start 16 banks apart
    constexpr int LPAD = 64 + 64 / (int)sizeof(R);
    __shared__ R sPr[16][LPAD], sPi[16][LPAD], sQr[16][LPAD], sQi[16][LPAD];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r0 = (tcs + ti) * 64, c0 = (tcs + tj) * 64;
    const int rowbase = (wv >> 1) * 32, colbase = (wv & 1) * 32;
    ACC accR[2][2], accI[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accR[a][b] = ACC{0, 0, 0, 0};
            accI[a][b] = ACC{0, 0, 0, 0};
        }
    const int K2 = 2 * nb;                    // complex K columns of P = [V W''], Q = [W V]
    // Only Re(alpha2) enters the Hermitian update: with W' = W + V a,
    //   V W'^H + W' V^H = V W^H + W V^H + 2 Re(a) V V^H = P Q^H
    // for P = [V, W + 2 Re(a) V] and Q = [W, V]; the column side is raw.
    // Each thread stages elements e = tid + 256 u (u < 4) of a chunk; the raw
    // operands of the next chunk are loaded into registers while the current
    // chunk's products run.
    const int rr = threadIdx.x & 63;
    const int gr = r0 + rr, gc = c0 + rr;
    const bool rok = gr >= s && gr < n, cok = gc >= s && gc < n;
    const int64_t ro = (int64_t)(gr - j0), co = (int64_t)(gc - j0);
    double2 pvr[4], pwr[4], pc[4];           // row V, row W, column operand
    float fpr[4], fpi[4], fqr[4], fqi[4];    // PK: row P, column Q (real, imaginary)
    const int64_t pstride = (int64_t)nb * ldv;
    auto fetch = [&](int l0) {
        if constexpr (PK) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int kk = l0 + (threadIdx.x >> 6) + 4 * u;
                fpr[u] = fpi[u] = fqr[u] = fqi[u] = 0.0f;
                if (kk < K2) {
                    const bool first = kk < nb;
                    const int64_t lo = (int64_t)(first ? kk : kk - nb) * ldv;
                    const float* P = pl + (first ? 0 : 2 * pstride);     // P = [V, W]
                    const float* Q = pl + (first ? 2 * pstride : 0);     // Q = [W, V]
                    if (rok) {
                        fpr[u] = P[ro + lo];
                        fpi[u] = P[pstride + ro + lo];
                    }
                    if (cok) {
                        fqr[u] = Q[co + lo];
                        fqi[u] = Q[pstride + co + lo];
                    }
                }
            }
            return;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kk = l0 + (threadIdx.x >> 6) + 4 * u;
            pvr[u] = cmk(0, 0);
            pwr[u] = cmk(0, 0);
            pc[u] = cmk(0, 0);
            if (kk < K2) {
                const bool first = kk < nb;
                const int64_t lo = (int64_t)(first ? kk : kk - nb) * ldv;
                if (rok) {
                    pvr[u] = V[ro + lo];
                    if (!first) pwr[u] = W[ro + lo];
                }
                if (cok) pc[u] = first ? W[co + lo] : V[co + lo];
            }
        }
    };
    fetch(0);
    // the tile entries this lane updates, loaded before the products so that
    // their latency hides under them
    T old[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
                const int lr = rowbase + a * 16 + (lane & 15);
                const int lc = colbase + b * 16 + MM::row(lane, reg);
                old[a][b][reg] = tile[lr + lc * 64];
            }
    for (int l0 = 0; l0 < K2; l0 += 16) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kc = (threadIdx.x >> 6) + 4 * u, kk = l0 + kc;
            if constexpr (PK) {
                sPr[kc][rr] = (R)fpr[u];
                sPi[kc][rr] = (R)fpi[u];
                sQr[kc][rr] = (R)fqr[u];
                sQi[kc][rr] = (R)fqi[u];
                continue;
            }
            double2 pv = pvr[u];
            if (kk < K2 && kk >= nb) {
                const double ar = 2.0 * alpha2[kk - nb].x;
                pv = cmk(fma(ar, pvr[u].x, pwr[u].x), fma(ar, pvr[u].y, pwr[u].y));
            }
            sPr[kc][rr] = (R)pv.x;
            sPi[kc][rr] = (R)pv.y;
            sQr[kc][rr] = (R)pc[u].x;
            sQi[kc][rr] = (R)pc[u].y;
        }
        __syncthreads();
        if (l0 + 16 < K2) fetch(l0 + 16);
        // real K index kappa in [0, 32): < 16 real parts, >= 16 imaginary parts
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const int kap = ks * 4 + (lane >> 4);
            const int kc = kap & 15;
            const bool re = kap < 16;
            R bP[2], bPi[2], aQ[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int row = rowbase + a * 16 + (lane & 15);
                bP[a] = re ? sPr[kc][row] : sPi[kc][row];
                bPi[a] = re ? sPi[kc][row] : -sPr[kc][row];
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int col = colbase + b * 16 + (lane & 15);
                aQ[b] = re ? sQr[kc][col] : sQi[kc][col];
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    accR[a][b] = MM::mfma(aQ[b], bP[a], accR[a][b]);
                    accI[a][b] = MM::mfma(aQ[b], bPi[a], accI[a][b]);
                }
        }
        __syncthreads();
    }
    // lane holds D'[i][j = lane % 16] with i = MM::row(lane, reg):
    // column i, row j of the subtile
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
                const int lr = rowbase + a * 16 + (lane & 15);
                const int lc = colbase + b * 16 + MM::row(lane, reg);
                const int gr = r0 + lr, gc = c0 + lc;
                if (gc >= s && gr < n && gr >= gc) {
                    T* p = tile + lr + lc * 64;
                    double2 x = cplx<T>::load(&old[a][b][reg]);
                    x.x -= (double)accR[a][b][reg];
                    x.y -= (double)accI[a][b][reg];
                    if (gr == gc) x.y = 0.0;
                    cplx<T>::store(p, x);
                }
            }
}

// The same MFMA update over 128 x 128 macro tiles (2 x 2 storage tiles):
// every V/W row staged in LDS feeds twice as many products, which halves the
// panel traffic from L2 per flop (the 64 x 64 form is limited by it).  Wave
// w = 2 rh + ch computes rows 64 rh.. and columns 64 ch.. of the macro tile
// (4 x 4 MFMA subtiles); on a diagonal macro tile the upper storage tile
// (rh = 0, ch = 1) is skipped.  idx runs over the lower macro tiles of the
// trailing block A[s:, s:] (tile rows and columns tcs + 2 mi, tcs + 2 mj).
template <typename T>
__device__ __attribute__((always_inline)) inline void her2k_mfma2_body(int n, int s, T* __restrict__ A,
                                       const double2* __restrict__ V,
                                       const double2* __restrict__ W, int64_t ldv, int j0, int nb,
                                       const double2* __restrict__ alpha2, int64_t idx)
{
    typedef her2k_f64::real R;
    typedef her2k_f64::acc ACC;
    int mi, mj;
    tri_tile(idx, mi, mj);
    const int tcs = s >> 6, ntile = tile_count(n);
    const int R0 = (tcs + 2 * mi) * 64, C0 = (tcs + 2 * mj) * 64;
    __shared__ R sPr[16][128], sPi[16][128], sQr[16][128], sQi[16][128];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rh = wv >> 1, chf = wv & 1;
    const bool idle = (mi == mj && rh == 0 && chf == 1);
    ACC accR[4][4], accI[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            accR[a][b] = ACC{0, 0, 0, 0};
            accI[a][b] = ACC{0, 0, 0, 0};
        }
    const int K2 = 2 * nb;
    for (int l0 = 0; l0 < K2; l0 += 16) {
        for (int e = threadIdx.x; e < 16 * 128; e += 256) {
            const int rr = e & 127, kc = e >> 7;
            const int kk = l0 + kc;
            double2 pv = cmk(0, 0), qv = cmk(0, 0);
            if (kk < K2) {
                const bool first = kk < nb;
                const int l = first ? kk : kk - nb;
                const double2 a2 = alpha2[l];
                const int gr = R0 + rr, gc = C0 + rr;
                if (gr >= s && gr < n) {
                    const double2 v = V[(gr - j0) + (int64_t)l * ldv];
                    const double2 w = cadd(W[(gr - j0) + (int64_t)l * ldv], cmul(a2, v));
                    pv = first ? v : w;
                }
                if (gc >= s && gc < n) {
                    const double2 v = V[(gc - j0) + (int64_t)l * ldv];
                    const double2 w = cadd(W[(gc - j0) + (int64_t)l * ldv], cmul(a2, v));
                    qv = first ? w : v;
                }
            }
            sPr[kc][rr] = (R)pv.x;
            sPi[kc][rr] = (R)pv.y;
            sQr[kc][rr] = (R)qv.x;
            sQi[kc][rr] = (R)qv.y;
        }
        __syncthreads();
        if (!idle) {
#pragma unroll 2
            for (int ks = 0; ks < 8; ++ks) {
                const int kap = ks * 4 + (lane >> 4);
                const int kc = kap & 15;
                const bool re = kap < 16;
                R bP[4], bPi[4], aQ[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int row = 64 * rh + a * 16 + (lane & 15);
                    const R pr = sPr[kc][row], pi = sPi[kc][row];
                    bP[a] = re ? pr : pi;
                    bPi[a] = re ? pi : -pr;
                }
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int col = 64 * chf + b * 16 + (lane & 15);
                    aQ[b] = re ? sQr[kc][col] : sQi[kc][col];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        accR[a][b] = her2k_f64::mfma(aQ[b], bP[a], accR[a][b]);
                        accI[a][b] = her2k_f64::mfma(aQ[b], bPi[a], accI[a][b]);
                    }
            }
        }
        __syncthreads();
    }
    if (idle) return;
    // lane holds D'[i][j = lane % 16] with i = her2k_f64::row(lane, reg):
    // column i, row j of the subtile
    const int TI = tcs + 2 * mi + rh, TJ = tcs + 2 * mj + chf;
    if (TI >= ntile || TJ >= ntile) return;
    T* tile = A + tile_pos(TI, TJ, ntile) * (BSE_TILE * BSE_TILE);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
                const int lr = a * 16 + (lane & 15);
                const int lc = b * 16 + her2k_f64::row(lane, reg);
                const int gr = TI * 64 + lr, gc = TJ * 64 + lc;
                if (gc >= s && gr < n && gr >= gc) {
                    T* p = tile + lr + lc * 64;
                    double2 x = cplx<T>::load(p);
                    x.x -= (double)accR[a][b][reg];
                    x.y -= (double)accI[a][b][reg];
                    if (gr == gc) x.y = 0.0;
                    cplx<T>::store(p, x);
                }
            }
}

// Number of work items (blocks) of a trailing update over nT storage tile
// rows and columns: lower 64 x 64 tiles, or lower 128 x 128 macro tiles
// (kind 2).
__host__ __device__ inline int64_t her2k_blocks(int64_t nT, int kind) {
    const int64_t m = (kind == 2) ? (nT + 1) / 2 : nT;
    return m * (m + 1) / 2;
}

template <typename T>
__global__ __launch_bounds__(256) void k_her2k_mfma2(int n, int s, T* __restrict__ A,
                                                    const double2* __restrict__ V,
                                                    const double2* __restrict__ W, int64_t ldv,
                                                    int j0, int nb, const double2* __restrict__ alpha2)
{
    her2k_mfma2_body<T>(n, s, A, V, W, ldv, j0, nb, alpha2, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(256) void k_her2k_lower(int n, int s, T* __restrict__ A,
                                                    const double2* __restrict__ V,
                                                    const double2* __restrict__ W, int64_t ldv,
                                                    int j0, int nb, const double2* __restrict__ alpha2)
{
    her2k_valu_body<T>(n, s, A, V, W, ldv, j0, nb, alpha2, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(256) void k_her2k_mfma(int n, int s, T* __restrict__ A,
                                                   const double2* __restrict__ V,
                                                   const double2* __restrict__ W, int64_t ldv,
                                                   int j0, int nb, const double2* __restrict__ alpha2)
{
    her2k_mfma_body<T>(n, s, A, V, W, ldv, j0, nb, alpha2, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_her2k_mfma32(int n, int s, float2* __restrict__ A,
                                                     const double2* __restrict__ V,
                                                     const double2* __restrict__ W, int64_t ldv,
                                                     int j0, int nb, const double2* __restrict__ alpha2)
{
    her2k_mfma_body<float2, her2k_f32>(n, s, A, V, W, ldv, j0, nb, alpha2, blockIdx.x);
}

__global__ void k_last_diag(int n, const void* A, int fp32, double* d)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t o = tix(n - 1, n - 1, tile_count(n));
        if (fp32) d[n - 1] = (double)((const float2*)A)[o].x;
        else d[n - 1] = ((const double2*)A)[o].x;
    }
}

// ---------------------------------------------------------------------------
// Batched lockstep form: one launch per step serves every matrix of an uneven
// batch.  Matrix b (blockIdx.z) has its own storage and workspace (HetrdSys);
// column c of the reduction is processed for every matrix with c < n_b - 1,
// and each block first checks that it lies inside its own matrix's grid.
// The panel geometry is derived from (n, nb, c) exactly as in the
// single-matrix driver: panel p = c / nb starts at j0 = p nb, its reflectors
// sit at offset nb (p n - nb p (p-1)/2) of the packed store.
// ---------------------------------------------------------------------------

__host__ __device__ inline int64_t panel_voff(int n, int nb, int p) {
    return (int64_t)nb * ((int64_t)p * n - (int64_t)nb * p * (p - 1) / 2);
}

__device__ inline bool batch_col(const HetrdSys& s, int c, int nb, HetrdCol& h) {
    if (c >= s.n - 1) return false;
    const int p = c / nb;
    h.n = s.n; h.j0 = p * nb; h.i = c - h.j0; h.m = s.n - h.j0;
    h.Vp = s.Vstore + panel_voff(s.n, nb, p);
    h.Wp = s.Wp; h.alpha2 = s.alpha2; h.G = s.G + (int64_t)p * s.g_stride; h.acol = s.acol; h.d = s.d; h.e = s.e;
    h.tau = s.tau; h.scal = s.scal; h.part = s.part; h.counter = s.counter;
    return true;
}

template <typename T>
__global__ __launch_bounds__(256) void kb_colupd(const HetrdSys* __restrict__ sys, int c, int nb)
{
    HetrdCol h;
    const HetrdSys s = sys[blockIdx.z];
    if (!batch_col(s, c, nb, h)) return;
    if ((int)blockIdx.x >= colupd_blocks(h.m, h.i)) return;
    const int prev_gx = (h.i > 0) ? tile_count(s.n) - (c >> 6) : 0;
    colupd_body<T>(h, (const T*)s.A, blockIdx.x, prev_gx);
}


template <typename T, int CH, int U2, int AUX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HEMV_WPE_FOR(U2), 8)))
void kb_hemv(const HetrdSys* __restrict__ sys, int c, int nb)
{
    HetrdCol h;
    const HetrdSys s = sys[blockIdx.z];
    if (!batch_col(s, c, nb, h)) return;
    constexpr int RB = U2 == 3 ? 4 : 1;
    const int nbk = tile_count(s.n) - ((c + 1) >> 6), nq = (int)ceil_div64(nbk, CH);
    const int npb = (int)ceil_div64(nbk, RB);
    if ((int)blockIdx.x >= nq || (int)blockIdx.y >= npb) return;
    const T* A = (const T*)s.A;
    if (U2 >= 2) hemv_bfly_body<T, CH, RB, AUX>(s.n, c + 1, A, nullptr, s.Yr, s.Yc, nbk, nq, blockIdx.x, blockIdx.y, npb, h);
    else hemv_body<T, CH, U2, AUX>(s.n, c + 1, A, nullptr, s.Yr, s.Yc, nbk, nq, blockIdx.x, blockIdx.y, h);
}

template <typename T>
__global__ __launch_bounds__(64 * RED_G) void kb_reduce(const HetrdSys* __restrict__ sys, int c, int nb,
                                                  int ch, int rbk)
{
    HetrdCol h;
    const HetrdSys s = sys[blockIdx.z];
    if (!batch_col(s, c, nb, h)) return;
    const int nbk = tile_count(s.n) - ((c + 1) >> 6), nq = (int)ceil_div64(nbk, ch);
    if ((int)blockIdx.x >= nbk) return;
    reduce_body<T>(h, (const T*)s.A, s.Yr, s.Yc, nbk, nq, ch, rbk, blockIdx.x);
}

// Column c closes the matrix's current panel when c = j0 + nbp - 1.
__device__ inline bool batch_panel_end(const HetrdSys& s, int c, int nb, int& j0, int& nbp, int& p) {
    if (c >= s.n - 1) return false;
    p = c / nb;
    j0 = p * nb;
    nbp = min(nb, s.n - 1 - j0);
    return c == j0 + nbp - 1;
}

__global__ __launch_bounds__(64) void kb_larft(const HetrdSys* __restrict__ sys, int c, int nb)
{
    const HetrdSys s = sys[blockIdx.z];
    int j0, nbp, p;
    if (!batch_panel_end(s, c, nb, j0, nbp, p)) return;
    HetrdCol h;
    batch_col(s, c, nb, h);
    panel_close_body(h, nbp, tile_count(s.n) - ((c + 1) >> 6), s.Tall + (size_t)p * nb * nb, nb, s.g_stride == 0);
}

// KIND: 0 FP64 MFMA on 64 x 64 tiles, 1 FP64 VALU, 2 FP64 MFMA on 128 x 128
// macro tiles.
template <typename T, int KIND>
__global__ __launch_bounds__(256) void kb_her2k(const HetrdSys* __restrict__ sys, int c, int nb)
{
    const HetrdSys s = sys[blockIdx.z];
    int j0, nbp, p;
    if (!batch_panel_end(s, c, nb, j0, nbp, p)) return;
    const int m = s.n - j0, st = j0 + nbp;
    if (st >= s.n) return;
    const int64_t nT = tile_count(s.n) - (st >> 6);
    if ((int64_t)blockIdx.x >= her2k_blocks(nT, KIND)) return;
    const double2* Vp = s.Vstore + panel_voff(s.n, nb, p);
    if (KIND == 0)
        her2k_mfma_body<T>(s.n, st, (T*)s.A, Vp, s.Wp, (int64_t)m, j0, nbp, s.alpha2, blockIdx.x);
    else if (KIND == 2)
        her2k_mfma2_body<T>(s.n, st, (T*)s.A, Vp, s.Wp, (int64_t)m, j0, nbp, s.alpha2, blockIdx.x);
    else
        her2k_valu_body<T>(s.n, st, (T*)s.A, Vp, s.Wp, (int64_t)m, j0, nbp, s.alpha2, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Host drivers (run on the caller's stream; no host synchronisation).
// ---------------------------------------------------------------------------
extern "C" hipError_t bse_launch_larft(hipStream_t st, int nb, const double2* tau, const double2* G,
                                       double2* T, int ldt);

// Hemv variant code for trailing size nt: ch + 32 u2 (chunk length ch in
// {4, 8, 16}, u2 = two row blocks per step), from a table indexed by
// floor(log2(nt)) (autotuned per device); 0 entries mean the default 8.
static inline int hemv_ch_for(const unsigned char* table, int nt) {
    return hemv_ch_table_lookup(table, nt);
}

// Number of storage tile rows of the trailing block A[s0:, s0:].
static inline int hemv_nbk(int n, int s0) { return tile_count(n) - (s0 >> 6); }

// v given: plain product (autotuner timing); v null: the reflector step of
// column h.j0 + h.i (s0 = that column + 1) is fused in.
template <typename T>
static void hemv_launch(hipStream_t st, int code, int n, int s0, const T* A, const double2* v,
                        double2* Yr, double2* Yc, const HetrdCol& h)
{
    const int ch = hemv_chunk(code);
    const int nbk = hemv_nbk(n, s0);
    const int nq = (int)ceil_div64(nbk, ch), npb = (int)ceil_div64(nbk, hemv_rblk(code));
#define HL(CC, UU, AX) hipLaunchKernelGGL((k_hemv_lower<T, CC, UU, AX>), dim3(nq, npb), dim3(256), 0, st, n, s0, A, v, Yr, Yc, nbk, nq, h)
    HEMV_DISPATCH(code, HL)
#undef HL
}

// Workspace sizes for a given n, nb (Yc sized for the shortest hemv chunk).
extern "C" void bse_hetrd_workspace(int n, int nb, int64_t* vstore_elems, int64_t* yr_elems,
                                    int64_t* yc_elems, int64_t* part_doubles)
{
    int64_t vs = 0;
    for (int j0 = 0; j0 < n - 1; j0 += nb) vs += (int64_t)(n - j0) * nb;
    *vstore_elems = vs > 0 ? vs : 1;
    const int64_t nbk = ceil_div64(n, 64);
    const int64_t nq = ceil_div64(nbk, HEMV_CH_MIN);
    *yr_elems = nbk * nbk * 64;
    *yc_elems = nbk * nq * 64;
    *part_doubles = ::part_doubles(n);
}

extern "C" hipError_t bse_hetrd_batch_panel(hipStream_t st, int fp32, int c, int nb, int nact,
                                            const void* d_sys, int64_t max_nt, int her2k_kind);
static hipError_t column_launch(hipStream_t st, int fp32, int c, int nb, int nact, int n0,
                                const void* d_sys, int code, hipEvent_t e0, hipEvent_t e1);

static_assert(sizeof(HetrdSys) <= 256, "HetrdSys must fit the 256-byte device slot");

// Single-matrix reduction, nb <= 32, A in tiled lower storage: the batched
// kernels with one matrix, whose description (HetrdSys) is copied to d_sys
// (at least 256 bytes of device memory) first, so that every launch carries
// only a pointer and the column index (large kernel arguments cost about a
// microsecond per launch).  Tall receives one nb x nb compact-WY factor per panel,
// formed in one launch after the last column from the per-panel Gram blocks
// (G: ceil((n-1)/nb) blocks of 64 x 64); prof_ev (optional) gets 2 events per
// column around the hemv launch.
extern "C" hipError_t bse_hetrd(hipStream_t st, int fp32, int n, void* A, int nb,
                                double* d, double* e, double2* tau, double2* Vstore,
                                double2* Tall, double2* Wp, double2* alpha2,
                                double2* G, double2* acol, double2* Yr, double2* Yc, double2* t12,
                                double* part, double* scal, unsigned int* counter,
                                hipEvent_t* prof_ev, const unsigned char* ch_table, int her2k_kind,
                                void* d_sys)
{
    if (nb < 1 || nb > 32 || n < 1 || !d_sys) return hipErrorInvalidValue;
    hipError_t err;
    if (n > 1) {
        HetrdSys x;
        x.n = n; x.g_stride = 64 * 64; x.A = A; x.Vstore = Vstore; x.Tall = Tall; x.Wp = Wp; x.alpha2 = alpha2; x.G = G;
        x.acol = acol; x.d = d; x.e = e; x.tau = tau; x.scal = scal; x.part = part;
        x.counter = counter; x.Yr = Yr; x.Yc = Yc; x.t12 = t12;
        if ((err = hipMemcpyAsync(d_sys, &x, sizeof(x), hipMemcpyHostToDevice, st)) != hipSuccess)
            return err;
        const int ntile = tile_count(n);
        for (int c = 0; c < n - 1; ++c) {
            const int code = hemv_ch_for(ch_table, n - c - 1);
            err = column_launch(st, fp32, c, nb, 1, n, d_sys, code, prof_ev ? prof_ev[2 * c] : nullptr,
                                prof_ev ? prof_ev[2 * c + 1] : nullptr);
            if (err != hipSuccess) return err;
            const int j0 = (c / nb) * nb, nbp = min(nb, n - 1 - j0);
            if (c != j0 + nbp - 1) continue;
            const int64_t nT = (j0 + nbp < n) ? ntile - ((j0 + nbp) >> 6) : 0;
            if ((err = bse_hetrd_batch_panel(st, fp32, c, nb, 1, d_sys, nT, her2k_kind)) != hipSuccess)
                return err;
        }
        const int npan = (int)ceil_div64(n - 1, nb);
        hipLaunchKernelGGL(k_larft_all, dim3((unsigned)npan), dim3(64), 0, st, n, nb, (const double2*)tau,
                           (const double2*)G, (int64_t)64 * 64, Tall);
    }
    hipLaunchKernelGGL(k_last_diag, dim3(1), dim3(64), 0, st, n, (const void*)A, fp32, d);
    return hipGetLastError();
}

// One hemv launch (autotuner timing): y partials of A[s0:, s0:] v for A in
// tiled lower storage of size n (v indexed from row s0).
extern "C" hipError_t bse_launch_hemv(hipStream_t st, int fp32, int code, int n, int s0,
                                      const void* A, const double2* v, double2* Yr, double2* Yc)
{
    if (s0 < 0 || s0 >= n || !hemv_code_valid(code)) return hipErrorInvalidValue;
    if (!v) return hipErrorInvalidValue;
    HetrdCol h0 = {};
    if (fp32) hemv_launch<float2>(st, code, n, s0, (const float2*)A, v, Yr, Yc, h0);
    else hemv_launch<double2>(st, code, n, s0, (const double2*)A, v, Yr, Yc, h0);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_last_diag(hipStream_t st, int fp32, int n, const void* A, double* d)
{
    hipLaunchKernelGGL(k_last_diag, dim3(1), dim3(64), 0, st, n, A, fp32, d);
    return hipGetLastError();
}

// Column c of the batched reduction for the first nact matrices of d_sys
// (sorted by decreasing n; n0 = n of matrix 0 fixes the grid extents).  The
// optional events bracket the hemv launch.
static hipError_t column_launch(hipStream_t st, int fp32, int c, int nb, int nact, int n0,
                                const void* d_sys, int code, hipEvent_t e0, hipEvent_t e1)
{
    if (!hemv_code_valid(code) || nb < 1 || nb > 32) return hipErrorInvalidValue;
    const int ch = hemv_chunk(code);
    if (nact <= 0 || c >= n0 - 1) return hipSuccess;
    const HetrdSys* sys = (const HetrdSys*)d_sys;
    const int j0 = (c / nb) * nb, i = c - j0, m = n0 - j0;
    const int nbk = hemv_nbk(n0, c + 1);
    const unsigned z = (unsigned)nact;
    if (fp32)
        hipLaunchKernelGGL(kb_colupd<float2>, dim3((unsigned)colupd_blocks(m, i), 1, z), dim3(256), 0, st, sys, c, nb);
    else
        hipLaunchKernelGGL(kb_colupd<double2>, dim3((unsigned)colupd_blocks(m, i), 1, z), dim3(256), 0, st, sys, c, nb);
    const dim3 hg((unsigned)ceil_div64(nbk, ch), (unsigned)ceil_div64(nbk, hemv_rblk(code)), z);
    if (e0) (void)hipEventRecord(e0, st);
#define KBF(CC, UU, AX) hipLaunchKernelGGL((kb_hemv<float2, CC, UU, AX>), hg, dim3(256), 0, st, sys, c, nb)
#define KBD(CC, UU, AX) hipLaunchKernelGGL((kb_hemv<double2, CC, UU, AX>), hg, dim3(256), 0, st, sys, c, nb)
    if (fp32) HEMV_DISPATCH(code, KBF) else HEMV_DISPATCH(code, KBD)
#undef KBF
#undef KBD
    if (e1) (void)hipEventRecord(e1, st);
    const int rbk = hemv_rblk(code);
    if (fp32) hipLaunchKernelGGL(kb_reduce<float2>, dim3((unsigned)nbk, 1, z), dim3(64 * RED_G), 0, st, sys, c, nb, ch, rbk);
    else hipLaunchKernelGGL(kb_reduce<double2>, dim3((unsigned)nbk, 1, z), dim3(64 * RED_G), 0, st, sys, c, nb, ch, rbk);
    return hipGetLastError();
}

extern "C" hipError_t bse_hetrd_batch_column(hipStream_t st, int fp32, int c, int nb, int nact,
                                             int n0, const void* d_sys, int code)
{
    return column_launch(st, fp32, c, nb, nact, n0, d_sys, code, nullptr, nullptr);
}

// Panel close at column c: compact-WY factor and trailing update for every
// matrix whose current panel ends at c.  max_nt: largest trailing tile count
// among them (grid extent); her2k_kind: 0 FP64 MFMA 64 x 64, 1 FP64 VALU,
// 2 FP64 MFMA 128 x 128.
extern "C" hipError_t bse_hetrd_batch_panel(hipStream_t st, int fp32, int c, int nb, int nact,
                                            const void* d_sys, int64_t max_nt, int her2k_kind)
{
    if (nact <= 0) return hipSuccess;
    if (her2k_kind < 0 || her2k_kind > 2) return hipErrorInvalidValue;
    const HetrdSys* sys = (const HetrdSys*)d_sys;
    const unsigned z = (unsigned)nact;
    hipLaunchKernelGGL(kb_larft, dim3(1, 1, z), dim3(64), 0, st, sys, c, nb);
    if (max_nt > 0) {
        const dim3 g((unsigned)her2k_blocks(max_nt, her2k_kind), 1, z);
#define KBK(TT, KK) hipLaunchKernelGGL((kb_her2k<TT, KK>), g, dim3(256), 0, st, sys, c, nb)
        if (fp32) {
            if (her2k_kind == 0) KBK(float2, 0); else if (her2k_kind == 2) KBK(float2, 2); else KBK(float2, 1);
        } else {
            if (her2k_kind == 0) KBK(double2, 0); else if (her2k_kind == 2) KBK(double2, 2); else KBK(double2, 1);
        }
#undef KBK
    }
    return hipGetLastError();
}

// One trailing-update launch on A[s:, s:] of a tiled lower matrix of size n
// (autotuner timing); V, W start at global row j0 with leading dimension ldv.
// kind 3 (FP32 storage only): FP32 operands and accumulation on matrix cores.
extern "C" hipError_t bse_launch_her2k(hipStream_t st, int fp32, int kind, int n, int s, void* A,
                                       const double2* V, const double2* W, int64_t ldv, int j0,
                                       int nb, const double2* alpha2)
{
    if (s >= n) return hipSuccess;
    if (s < j0 || ldv < n - j0 || kind < 0 || kind > 3 || (kind == 3 && !fp32)) return hipErrorInvalidValue;
    const int64_t nT = tile_count(n) - (s >> 6);
    const dim3 g((unsigned)her2k_blocks(nT, kind));
    if (kind == 3) {
        hipLaunchKernelGGL(k_her2k_mfma32, g, dim3(256), 0, st, n, s, (float2*)A, V, W, ldv, j0, nb, alpha2);
        return hipGetLastError();
    }
#define KH(K, TT) hipLaunchKernelGGL(K<TT>, g, dim3(256), 0, st, n, s, (TT*)A, V, W, ldv, j0, nb, alpha2)
    if (fp32) {
        if (kind == 1) KH(k_her2k_lower, float2); else if (kind == 2) KH(k_her2k_mfma2, float2); else KH(k_her2k_mfma, float2);
    } else {
        if (kind == 1) KH(k_her2k_lower, double2); else if (kind == 2) KH(k_her2k_mfma2, double2); else KH(k_her2k_mfma, double2);
    }
#undef KH
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The trailing update in two parts, for the band reduction's look-ahead:
// PART 1 updates the first tile column of the trailing block A[s:, s:] (it
// holds the next panel's columns, so the next panel can be factored while
// PART 2 updates every other lower tile).  Lower 64 x 64 tiles only (kinds
// 0, 1 and 3, and 4: kind 3 with the operands packed in FP32 planes, passed
// in place of V, W unused, alpha2 zero); block x of PART 1 is tile (x, 0), of
// PART 2 tile (a+1, c+1) for the x-th lower tile (a, c) of the remaining order.
// ---------------------------------------------------------------------------
__device__ inline int64_t her2k_part_idx(int part, int64_t x) {
    if (part == 1) return x * (x + 1) / 2;
    int a, c;
    tri_tile(x, a, c);
    return (int64_t)(a + 1) * (a + 2) / 2 + c + 1;
}

template <typename T, int KIND, int PART>
__global__ __launch_bounds__(256) void k_her2k_part(int n, int s, T* __restrict__ A, const double2* __restrict__ V,
                                                   const double2* __restrict__ W, int64_t ldv, int j0, int nb,
                                                   const double2* __restrict__ alpha2)
{
    const int64_t idx = her2k_part_idx(PART, blockIdx.x);
    if constexpr (KIND == 4)
        her2k_mfma_body<float2, her2k_f32, true>(n, s, (float2*)A, V, W, ldv, j0, nb, alpha2, idx,
                                                 reinterpret_cast<const float*>(V));
    else if constexpr (KIND == 3) her2k_mfma_body<float2, her2k_f32>(n, s, (float2*)A, V, W, ldv, j0, nb, alpha2, idx);
    else if constexpr (KIND == 1) her2k_valu_body<T>(n, s, A, V, W, ldv, j0, nb, alpha2, idx);
    else her2k_mfma_body<T>(n, s, A, V, W, ldv, j0, nb, alpha2, idx);
}

extern "C" hipError_t bse_launch_her2k_part(hipStream_t st, int fp32, int kind, int part, int n, int s, void* A,
                                            const double2* V, const double2* W, int64_t ldv, int j0, int nb,
                                            const double2* alpha2)
{
    if (s >= n) return hipSuccess;
    if (s < j0 || ldv < n - j0 || (kind != 0 && kind != 1 && kind != 3 && kind != 4) || (kind >= 3 && !fp32) ||
        (part != 1 && part != 2))
        return hipErrorInvalidValue;
    const int64_t nT = tile_count(n) - (s >> 6);
    const int64_t nblk = part == 1 ? nT : (nT - 1) * nT / 2;
    if (nblk <= 0) return hipSuccess;
    const dim3 g((unsigned)nblk);
#define KP(TT, K, P) hipLaunchKernelGGL((k_her2k_part<TT, K, P>), g, dim3(256), 0, st, n, s, (TT*)A, V, W, ldv, j0, nb, alpha2)
    if (kind == 4) {
        if (part == 1) KP(float2, 4, 1); else KP(float2, 4, 2);
    } else if (kind == 3) {
        if (part == 1) KP(float2, 3, 1); else KP(float2, 3, 2);
    } else if (fp32) {
        if (kind == 1) { if (part == 1) KP(float2, 1, 1); else KP(float2, 1, 2); }
        else { if (part == 1) KP(float2, 0, 1); else KP(float2, 0, 2); }
    } else {
        if (kind == 1) { if (part == 1) KP(double2, 1, 1); else KP(double2, 1, 2); }
        else { if (part == 1) KP(double2, 0, 1); else KP(double2, 0, 2); }
    }
#undef KP
    return hipGetLastError();
}
