// K4: assembly of the q = 0 Tamm-Dancoff BSE matrix.
//
// Replaces the direct and exchange loops of pull_resonant_part_and_rcv
// (bse.version13.cpp:2796-2897):
//   H_rs = delta_rs (eps_c(k) - eps_v(k-q))
//        - (1/(A Nk)) sum_ij X_{cc',i} W_ij(k - k') Y_{vv',j}        (direct)
//        + (f_s/(A Nk)) sum_i (P V)_{r,i} conj(P_{s,i})              (exchange)
// with X_{cc',i} = conj(c_{ck,i}) c_{c'k',i}, Y_{vv',j} = c_{v,k-q,j}
// conj(c_{v',k'-q,j}), r = (c,v,k), s = (c',v',k').  The direct row is kept in
// its textbook orientation and the exchange in its Hermitian form (the
// reference stores the conjugate of the direct row, 2878, and uses
// conj(rho) conj(rho) in the exchange, 2806-2816).
//
// Device layout: k-major internal order r' = k*B + c*nv + v (B = nc*nv), so
// every (k, k') pair owns a contiguous B x B block of the column-major
// matrix.  A workgroup owns a tile of TK x TK k-pairs (TK*B rows and columns)
// of the lower triangle; it first forms the pair intermediates
//   Apair[(k,k'), c, c', j] = sum_i conj(c_{ck,i}) c_{c'k',i} W_ij(k - k')
// in LDS (one read of W per pair), then contracts them with the valence
// coefficients for every (v, v') of the tile and adds exchange and diagonal.
// Four coefficient arrays (Cl, Cr, Vl, Vr) are separate inputs so the same
// kernel evaluates product-rule derivative terms.
#include "bse_common.hpp"
#include <algorithm>

struct AsmArgs {
    int nk, n1, n2, n3, nv, nc, norb, TK;
    int64_t N;
    int nt;              // tile rows of the tiled lower storage
    const double2* Cl;   // coef table [k][slot][orb] for the row (left) conduction factor
    const double2* Cr;   // ... column (right) conduction factor
    const double2* Vl;   // ... row valence factor
    const double2* Vr;   // ... column valence factor
    const double2* wtab; // (2n1-1) x (2n2-1) x (2n3-1) x norb x norb
    const double2* PVr;  // N x norb exchange row factor (may be null)
    const double2* Pc;   // N x norb exchange column factor
    const double* dE;    // N diagonal (may be null)
    double alpha;        // direct prefactor  -direct_scale/(A Nk)
    double beta;         // exchange prefactor f_s exchange_scale/(A Nk)
};

// Triangular tile index -> (ti, tj) with ti >= tj.
__device__ inline void tri_index(int64_t idx, int& ti, int& tj) {
    int64_t r = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > idx) --r;
    while ((r + 1) * (r + 2) / 2 <= idx) ++r;
    ti = (int)r;
    tj = (int)(idx - r * (r + 1) / 2);
}

// CPL = true: the coupling block B of the full (non-Tamm-Dancoff) BSE instead
// (add_coupling_part, 2929-3040): the primed conduction and valence states
// trade places in the direct term,
//   Bd_rs = -(1/(A Nk)) sum_ij conj(c_{ck,i}) c_{v',k'-q,i} W_ij(k - k')
//                              c_{v,k-q,j} conj(c_{c'k',j}),
// and the exchange is sum_i (P V)_{r,i} P_{s,i} (P V P^T, no conjugate);
// no diagonal.  B is complex symmetric, so its lower triangle is stored.
template <typename T, bool CPL>
__global__ __launch_bounds__(256) void k_assemble_tile(AsmArgs a, T* __restrict__ H, int lower_only)
{
    extern __shared__ double2 smem[];
    const int B = a.nc * a.nv, ns = a.nv + a.nc, no = a.norb, TK = a.TK;
    const int np = CPL ? a.nv : a.nc;   // primed partner of c in the pair intermediates
    const int TR = TK * B;
    int ti, tj;
    if (lower_only) {
        tri_index(blockIdx.x, ti, tj);
    } else {
        const int nt = (int)ceil_div64(a.nk, TK);
        ti = blockIdx.x % nt;
        tj = blockIdx.x / nt;
    }
    const int k0 = ti * TK, kp0 = tj * TK;
    // LDS carve-up
    double2* sCl = smem;                               // TK x nc x no  (row k)
    double2* sCr = sCl + TK * a.nc * no;               // TK x nc x no  (col k')
    double2* sVl = sCr + TK * a.nc * no;               // TK x nv x no
    double2* sVr = sVl + TK * a.nv * no;               // TK x nv x no
    double2* sA = sVr + TK * a.nv * no;                // TK*TK x nc x np x no
    double2* sPV = sA + TK * TK * a.nc * np * no;      // TR x no
    double2* sP = sPV + TR * no;                       // TR x no
    const int tid = threadIdx.x;

    for (int e = tid; e < TK * a.nc * no; e += blockDim.x) {
        const int i = e % no, c = (e / no) % a.nc, kk = e / (no * a.nc);
        const int k = k0 + kk, kp = kp0 + kk;
        sCl[e] = (k < a.nk) ? a.Cl[((size_t)k * ns + a.nv + c) * no + i] : cmk(0, 0);
        sCr[e] = (kp < a.nk) ? a.Cr[((size_t)kp * ns + a.nv + c) * no + i] : cmk(0, 0);
    }
    for (int e = tid; e < TK * a.nv * no; e += blockDim.x) {
        const int i = e % no, v = (e / no) % a.nv, kk = e / (no * a.nv);
        const int k = k0 + kk, kp = kp0 + kk;
        sVl[e] = (k < a.nk) ? a.Vl[((size_t)k * ns + v) * no + i] : cmk(0, 0);
        sVr[e] = (kp < a.nk) ? a.Vr[((size_t)kp * ns + v) * no + i] : cmk(0, 0);
    }
    if (a.PVr != nullptr) {
        for (int e = tid; e < TR * no; e += blockDim.x) {
            const int64_t row = (int64_t)k0 * B + e / no, col = (int64_t)kp0 * B + e / no;
            sPV[e] = (row < a.N) ? a.PVr[row * no + e % no] : cmk(0, 0);
            sP[e] = (col < a.N) ? a.Pc[col * no + e % no] : cmk(0, 0);
        }
    }
    __syncthreads();

    // Pair intermediates Apair[p][c][c'][j] (coupling: [p][c][v'][j]), p = kk*TK + kq.
    const int nA = TK * TK * a.nc * np * no;
    for (int e = tid; e < nA; e += blockDim.x) {
        const int j = e % no;
        const int cp = (e / no) % np;
        const int c = (e / (no * np)) % a.nc;
        const int p = e / (no * np * a.nc);
        const int kk = p / TK, kq = p % TK;
        const int k = k0 + kk, kp = kp0 + kq;
        double2 acc = cmk(0, 0);
        if (k < a.nk && kp < a.nk) {
            const double2* wq = a.wtab + (size_t)ktable_index(k, kp, a.n1, a.n2, a.n3) * no * no;
            const double2* cl = sCl + (kk * a.nc + c) * no;
            const double2* cr = CPL ? sVr + (kq * a.nv + cp) * no : sCr + (kq * a.nc + cp) * no;
            for (int i = 0; i < no; ++i) {
                const double2 x = cmulc(cl[i], cr[i]);
                cfma(acc, x, wq[i * no + j]);
            }
        }
        sA[e] = acc;
    }
    __syncthreads();

    const int64_t row0 = (int64_t)k0 * B, col0 = (int64_t)kp0 * B;
    for (int e = tid; e < TR * TR; e += blockDim.x) {
        const int rr = e % TR, cc = e / TR;
        const int64_t grow = row0 + rr, gcol = col0 + cc;
        if (grow >= a.N || gcol >= a.N) continue;
        const int kk = rr / B, b = rr % B, kq = cc / B, bp = cc % B;
        const int c = b / a.nv, v = b % a.nv, cp = bp / a.nv, vp = bp % a.nv;
        const double2* Ap = sA + (((kk * TK + kq) * a.nc + c) * np + (CPL ? vp : cp)) * no;
        const double2* vl = sVl + (kk * a.nv + v) * no;
        const double2* vr = CPL ? sCr + (kq * a.nc + cp) * no : sVr + (kq * a.nv + vp) * no;
        double2 d = cmk(0, 0);
        for (int j = 0; j < no; ++j) {
            const double2 y = cmul(vl[j], cconj(vr[j]));
            cfma(d, Ap[j], y);
        }
        double2 h = cscale(d, a.alpha);
        if (a.PVr != nullptr) {
            double2 x = cmk(0, 0);
            const double2* pv = sPV + rr * no;
            const double2* pc = sP + cc * no;
            for (int j = 0; j < no; ++j) cfma(x, pv[j], CPL ? pc[j] : cconj(pc[j]));
            h = cadd(h, cscale(x, a.beta));
        }
        if (!CPL && grow == gcol && a.dE != nullptr) h.x += a.dE[grow];
        if (grow >= gcol) cplx<T>::store(H + tix(grow, gcol, a.nt), h);
    }
}

// LDS bytes for a given tile (the coupling form's pair intermediates are
// nc x nv instead of nc x nc; the larger of the two is reserved).
extern "C" size_t bse_assemble_lds_bytes(int TK, int nv, int nc, int norb) {
    const int B = nv * nc;
    size_t n = (size_t)TK * nc * norb * 2 + (size_t)TK * nv * norb * 2 +
               (size_t)TK * TK * nc * std::max(nc, nv) * norb + (size_t)2 * TK * B * norb;
    return n * sizeof(double2);
}

// Pick the largest power-of-two pair tile with TK*B <= 64 that fits 60 KB of LDS.
extern "C" int bse_assemble_pick_tk(int nv, int nc, int norb) {
    const int B = nv * nc;
    int tk = 1;
    while (tk * 2 * B <= 64 && bse_assemble_lds_bytes(tk * 2, nv, nc, norb) <= 60 * 1024) tk *= 2;
    return tk;
}

template <typename T, bool CPL = false>
static hipError_t launch_assemble(hipStream_t st, const AsmArgs& a, T* H, int lower_only) {
    const int64_t nt = ceil_div64(a.nk, a.TK);
    const int64_t nblocks = lower_only ? nt * (nt + 1) / 2 : nt * nt;
    const size_t lds = bse_assemble_lds_bytes(a.TK, a.nv, a.nc, a.norb);
    hipLaunchKernelGGL((k_assemble_tile<T, CPL>), dim3((unsigned)nblocks), dim3(256), lds, st, a, H,
                       lower_only);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_assemble(hipStream_t st, int store_fp32, int nk, int n1, int n2,
                                          int n3, int nv, int nc, int norb, int TK, int64_t ld,
                                          const double2* Cl, const double2* Cr,
                                          const double2* Vl, const double2* Vr,
                                          const double2* wtab, const double2* PVr,
                                          const double2* Pc, const double* dE, double alpha,
                                          double beta, void* H, int lower_only)
{
    AsmArgs a;
    a.nk = nk; a.n1 = n1; a.n2 = n2; a.n3 = n3; a.nv = nv; a.nc = nc; a.norb = norb; a.TK = TK;
    a.N = (int64_t)nk * nv * nc;
    a.nt = tile_count(a.N);
    (void)ld;            // the matrix is stored in tiles (bse_common.hpp)
    a.Cl = Cl; a.Cr = Cr; a.Vl = Vl; a.Vr = Vr; a.wtab = wtab; a.PVr = PVr; a.Pc = Pc;
    a.dE = dE; a.alpha = alpha; a.beta = beta;
    if (store_fp32) return launch_assemble<float2>(st, a, (float2*)H, lower_only);
    return launch_assemble<double2>(st, a, (double2*)H, lower_only);
}

// Coupling block B (FP64, lower triangle in tiled storage of order N): the
// direct term with alpha, the exchange P V P^T with beta (PVr = P V, Pc = P).
extern "C" hipError_t bse_launch_assemble_coupling(hipStream_t st, int nk, int n1, int n2, int n3, int nv,
                                                   int nc, int norb, int TK, const double2* coef,
                                                   const double2* wtab, const double2* PVr,
                                                   const double2* Pc, double alpha, double beta, double2* Bm)
{
    AsmArgs a;
    a.nk = nk; a.n1 = n1; a.n2 = n2; a.n3 = n3; a.nv = nv; a.nc = nc; a.norb = norb; a.TK = TK;
    a.N = (int64_t)nk * nv * nc;
    a.nt = tile_count(a.N);
    a.Cl = coef; a.Cr = coef; a.Vl = coef; a.Vr = coef; a.wtab = wtab; a.PVr = PVr; a.Pc = Pc;
    a.dE = nullptr; a.alpha = alpha; a.beta = beta;
    return launch_assemble<double2, true>(st, a, Bm, 1);
}

// ---------------------------------------------------------------------------
// Copy-out: internal k-major tiled lower storage -> full Hermitian matrix
// (column-major, ldd) in reference or internal order.  dst element type
// matches src.
// ---------------------------------------------------------------------------
// kinv (may be null): caller's k-list position -> grid index.
// SYM = true: the stored triangle belongs to a complex symmetric matrix (the
// coupling block), mirrored without conjugation.
template <typename T, bool SYM = false>
__global__ __launch_bounds__(256) void k_expand(int64_t N, int B, int nk, const T* __restrict__ src,
                                                T* __restrict__ dst, int64_t ldd, int ref_order,
                                                const int* __restrict__ kinv)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // destination row
    const int64_t s = blockIdx.y;                                     // destination column
    if (r >= N) return;
    const int nt = tile_count(N);
    // map destination (reference) indices to internal ones
    int64_t ri = r, si = s;
    if (ref_order) {
        const int kr = (int)(r % nk), ks = (int)(s % nk);
        ri = (int64_t)(kinv ? kinv[kr] : kr) * B + r / nk;
        si = (int64_t)(kinv ? kinv[ks] : ks) * B + s / nk;
    }
    T v;
    if (ri >= si) {
        v = src[tix(ri, si, nt)];
    } else {
        T u = src[tix(si, ri, nt)];
        v = u;
        if (!SYM) v.y = -u.y;
    }
    if (!SYM && ri == si) v.y = 0;
    dst[r + s * ldd] = v;
}

// The coupling block (FP64, complex symmetric) to a dense column-major matrix.
extern "C" hipError_t bse_launch_expand_sym(hipStream_t st, int64_t N, int B, int nk, const double2* src,
                                            double2* dst, int64_t ldd, int ref_order, const int* kinv)
{
    dim3 grid((unsigned)ceil_div64(N, 256), (unsigned)N);
    hipLaunchKernelGGL((k_expand<double2, true>), grid, dim3(256), 0, st, N, B, nk, src, dst, ldd, ref_order,
                       kinv);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_expand(hipStream_t st, int fp32, int64_t N, int B, int nk,
                                        const void* src, void* dst, int64_t ldd, int ref_order,
                                        const int* kinv)
{
    dim3 grid((unsigned)ceil_div64(N, 256), (unsigned)N);
    if (fp32)
        hipLaunchKernelGGL(k_expand<float2>, grid, dim3(256), 0, st, N, B, nk,
                           (const float2*)src, (float2*)dst, ldd, ref_order, kinv);
    else
        hipLaunchKernelGGL(k_expand<double2>, grid, dim3(256), 0, st, N, B, nk,
                           (const double2*)src, (double2*)dst, ldd, ref_order, kinv);
    return hipGetLastError();
}

// Lower triangle of a column-major FP64 matrix (n x n, lds) into tiled lower
// storage of FP64 or FP32 elements (dense eigensolver input).  Each
// workgroup fills one tile; padding rows and columns beyond n are zero.
template <typename T>
__global__ __launch_bounds__(256) void k_to_tiles(int64_t n, const double2* __restrict__ src,
                                                  int64_t lds, T* __restrict__ dst)
{
    const int nt = tile_count(n);
    int ti, tj;
    tri_tile(blockIdx.x, ti, tj);
    T* t = dst + tile_pos(ti, tj, nt) * (BSE_TILE * BSE_TILE);
    for (int e = threadIdx.x; e < BSE_TILE * BSE_TILE; e += blockDim.x) {
        const int64_t r = (int64_t)ti * 64 + (e & 63), c = (int64_t)tj * 64 + (e >> 6);
        double2 x = cmk(0, 0);
        if (r < n && c < n && r >= c) x = src[r + c * lds];
        cplx<T>::store(t + e, x);
    }
}

extern "C" hipError_t bse_launch_to_tiles(hipStream_t st, int fp32, int64_t n, const double2* src,
                                          int64_t lds, void* dst)
{
    const int64_t nt = tile_count(n);
    const unsigned g = (unsigned)(nt * (nt + 1) / 2);
    if (fp32) hipLaunchKernelGGL(k_to_tiles<float2>, dim3(g), dim3(256), 0, st, n, src, lds, (float2*)dst);
    else hipLaunchKernelGGL(k_to_tiles<double2>, dim3(g), dim3(256), 0, st, n, src, lds, (double2*)dst);
    return hipGetLastError();
}

// Storage invariant of the tiled lower layout: padding entries (rows or
// columns >= n) and the strict upper half of every diagonal tile are zero
// (the tridiagonalisation reads whole tiles without masks).  Nothing else
// writes these entries, so clearing them once after allocation is enough.
// Blocks 0 .. nt-1 clear the upper halves of the diagonal tiles, blocks
// nt .. 2nt-1 the rows >= n of the last tile row.
template <typename T>
__global__ __launch_bounds__(256) void k_tile_pad_zero(int64_t n, T* __restrict__ dst)
{
    const int nt = tile_count(n);
    const int b = blockIdx.x;
    const bool diag = b < nt;
    const int ti = diag ? b : nt - 1, tj = diag ? b : b - nt;
    T* t = dst + tile_pos(ti, tj, nt) * (BSE_TILE * BSE_TILE);
    T z;
    z.x = 0;
    z.y = 0;
    for (int e = threadIdx.x; e < BSE_TILE * BSE_TILE; e += blockDim.x) {
        const int r = e & 63, c = e >> 6;
        const bool clear = diag ? (r < c || (int64_t)ti * 64 + r >= n) : ((int64_t)ti * 64 + r >= n);
        if (clear) t[e] = z;
    }
}

extern "C" hipError_t bse_launch_tile_pad_zero(hipStream_t st, int fp32, int64_t n, void* dst)
{
    if (n <= 0) return hipSuccess;
    const int64_t nt = tile_count(n);
    const unsigned g = (unsigned)((n & 63) ? 2 * nt : nt);
    if (fp32) hipLaunchKernelGGL(k_tile_pad_zero<float2>, dim3(g), dim3(256), 0, st, n, (float2*)dst);
    else hipLaunchKernelGGL(k_tile_pad_zero<double2>, dim3(g), dim3(256), 0, st, n, (double2*)dst);
    return hipGetLastError();
}
