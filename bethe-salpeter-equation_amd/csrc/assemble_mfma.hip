// K4 on matrix cores: the direct term of the TDA matrix as two chained complex
// contractions per (k, k') pair, realified onto 16x16 MFMA tiles.
//
//   step 1:  A_p[(c,c'), j] = sum_i X_p[(c,c'), i] W_ij(k - k')
//            X_p[(c,c'), i] = conj(C[k,c,i]) C[k',c',i]
//   step 2:  D_p[(c,c'), (v,v')] = sum_j A_p[(c,c'), j] Y_p[(v,v'), j]
//            Y_p[(v,v'), j] = V[k,v,j] conj(V[k',v',j])
//   H[(c,v,k), (c',v',k')] = alpha D_p + exchange + diagonal     (2796-2897)
//
// Realification (K = 16 real for norb <= 8):
//   [Ar | Ai] = [Xr | Xi] [[Wr, Wi], [-Wi, Wr]]
//   Dr = [Ar | Ai] [Yr | -Yi]^T,   Di = [Ar | Ai] [Yi | Yr]^T
// so one pair costs 3 real 16x16x16 GEMMs: 12 v_mfma_*_16x16x4 (FP64 or FP32)
// or, in the compensated BF16 mode, 6 v_mfma_f32_16x16x32_bf16 where every
// FP32 operand is split a = a_hi + a_lo (bf16) and the three significant
// products a_hi b_hi + a_hi b_lo + a_lo b_hi are laid end to end along K
// (3 x 16 = 48, padded to 64), accumulated in FP32.
//
// Shapes: nc*nc <= 16, nv*nv <= 16, norb <= 8 (padded with zeros).  Each
// wave processes whole pairs; the step-1 result is transposed through a
// per-wave LDS tile into the A-operand layout of step 2.  The exchange
// (rank-norb, sum_j PV[r,j] conj(P[s,j])) and the diagonal are added in the
// epilogue with VALU FP64 arithmetic.
#include "bse_common.hpp"
#include <algorithm>

#define PREC_F64 0
#define PREC_F32 1
#define PREC_BF16 2

typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 mf_b8 __attribute__((ext_vector_type(8)));

template <int P> struct mf_t;
template <> struct mf_t<PREC_F64> {
    typedef double real; typedef mf_d4 acc; typedef double2 out;
    __device__ static inline int row(int lane, int r) { return mfma16_row_f64(lane, r); }
};
template <> struct mf_t<PREC_F32> {
    typedef float real; typedef mf_f4 acc; typedef float2 out;
    __device__ static inline int row(int lane, int r) { return mfma16_row_f32(lane, r); }
};
template <> struct mf_t<PREC_BF16> {
    typedef float real; typedef mf_f4 acc; typedef float2 out;
    __device__ static inline int row(int lane, int r) { return mfma16_row_f32(lane, r); }
};

// One real 16x16x16 product  acc += A(16x16) B(16x16).  a[ks], b[ks] hold the
// lane's operands for K step ks (A[l%16][4ks + l/16], B[4ks + l/16][l%16]).
__device__ inline mf_d4 mm16(const double* a, const double* b, mf_d4 acc) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ks], b[ks], acc, 0, 0, 0);
    return acc;
}
__device__ inline mf_f4 mm16(const float* a, const float* b, mf_f4 acc) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ks], b[ks], acc, 0, 0, 0);
    return acc;
}

// Compensated BF16 product.  The operands are the lane's 16 K values in the
// f32 layout (a[ks] = A[l%16][4ks + l/16]); they are regathered through LDS
// into the bf16 layout (8 consecutive K per lane).
__device__ inline mf_f4 mm16_bf16x2(const float* a, const float* b, mf_f4 acc, float* sbuf)
{
    const int lane = threadIdx.x & 63;
    const int i = lane & 15, q = lane >> 4;
    // sbuf: A[16][17] then B[16][17] (row = i for A, col = j for B)
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        sbuf[i * 17 + 4 * ks + q] = a[ks];
        sbuf[272 + i * 17 + 4 * ks + q] = b[ks];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // K' = [hi*hi (0..15) | hi*lo (16..31) | lo*hi (32..47) | 0 (48..63)]
    mf_b8 A0, B0, A1, B1;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int k0 = 8 * q + t;          // first instruction: K' in [0, 32)
        const int kk = k0 & 15;
        const float av = sbuf[i * 17 + kk];
        const float bv = sbuf[272 + i * 17 + kk];
        const __bf16 ah = (__bf16)av, bh = (__bf16)bv;
        const __bf16 bl = (__bf16)(bv - (float)bh);
        A0[t] = ah;
        B0[t] = (k0 < 16) ? bh : bl;
        const int k1 = 32 + 8 * q + t;     // second instruction: K' in [32, 64)
        const __bf16 al = (__bf16)(av - (float)ah);
        A1[t] = (k1 < 48) ? al : (__bf16)0.0f;
        B1[t] = (k1 < 48) ? bh : (__bf16)0.0f;
    }
    __builtin_amdgcn_wave_barrier();
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B1, acc, 0, 0, 0);
    return acc;
}

struct MfArgs {
    int nk, n1, n2, n3, nv, nc, norb, TK;
    int64_t N;
    int nt;              // tile rows of the tiled lower storage
    const double2* coef;
    const double2* wtab;
    const double2* PV;
    const double2* P;
    const double* dE;
    double alpha, beta;
};

template <int PREC>
__global__ __launch_bounds__(256) void k_assemble_mfma(MfArgs a, typename mf_t<PREC>::out* __restrict__ H)
{
    typedef typename mf_t<PREC>::real R;
    typedef typename mf_t<PREC>::acc ACC;
    typedef typename mf_t<PREC>::out OUT;
    extern __shared__ double2 smem[];
    const int B = a.nc * a.nv, ns = a.nv + a.nc, no = a.norb, TK = a.TK;
    const int TR = TK * B;
    int ti, tj;
    {
        const int64_t idx = blockIdx.x;
        int64_t r = (int64_t)((sqrt(8.0 * (double)idx + 1.0) - 1.0) * 0.5);
        while (r * (r + 1) / 2 > idx) --r;
        while ((r + 1) * (r + 2) / 2 <= idx) ++r;
        ti = (int)r;
        tj = (int)(idx - r * (r + 1) / 2);
    }
    const int k0 = ti * TK, kp0 = tj * TK;
    double2* sCr = smem;                       // [TK][ns][8]  row k-points, all slots
    double2* sCc = sCr + TK * ns * 8;          // [TK][ns][8]  column k-points
    double2* sPV = sCc + TK * ns * 8;          // [TR][8]
    double2* sP = sPV + TR * 8;                // [TR][8]
    R* sT = (R*)(sP + TR * 8);                 // per-wave 16x17 (x2 for bf16) transposition
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int e = tid; e < TK * ns * 8; e += 256) {
        const int i = e & 7, sl = (e >> 3) % ns, kk = e / (8 * ns);
        const int k = k0 + kk, kp = kp0 + kk;
        sCr[e] = (k < a.nk && i < no) ? a.coef[((size_t)k * ns + sl) * no + i] : cmk(0, 0);
        sCc[e] = (kp < a.nk && i < no) ? a.coef[((size_t)kp * ns + sl) * no + i] : cmk(0, 0);
    }
    for (int e = tid; e < TR * 8; e += 256) {
        const int i = e & 7, rr = e >> 3;
        const int64_t row = (int64_t)k0 * B + rr, col = (int64_t)kp0 * B + rr;
        sPV[e] = (row < a.N && i < no) ? a.PV[row * no + i] : cmk(0, 0);
        sP[e] = (col < a.N && i < no) ? a.P[col * no + i] : cmk(0, 0);
    }
    __syncthreads();
    R* tw = sT + wv * (PREC == PREC_BF16 ? 544 : 272);
    const int li = lane & 15, lq = lane >> 4;
    for (int p = wv; p < TK * TK; p += 4) {
        const int kk = p / TK, kq = p % TK;
        const int k = k0 + kk, kp = kp0 + kq;
        if (k >= a.nk || kp >= a.nk) continue;
        const double2* wq = a.wtab + (size_t)ktable_index(k, kp, a.n1, a.n2, a.n3) * no * no;
        // ---- step 1 operands: A1[rho][kappa] (rho = c*nc + c'), B1[kappa][j']
        R a1[4], b1[4];
        const int rho = li, c = rho / a.nc, cp = rho % a.nc;
        const bool rok = rho < a.nc * a.nc;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int kap = 4 * ks + lq;
            const int i = kap & 7;
            double2 x = cmk(0, 0);
            if (rok && i < no)
                x = cmulc(sCr[(kk * ns + a.nv + c) * 8 + i], sCc[(kq * ns + a.nv + cp) * 8 + i]);
            a1[ks] = (R)(kap < 8 ? x.x : x.y);
            const int jp = li;                        // output column of [Ar | Ai]
            const int jj = jp & 7, ii = kap & 7;
            double wv2 = 0.0;
            if (ii < no && jj < no) {
                const double2 w = wq[ii * no + jj];
                if (kap < 8) wv2 = (jp < 8) ? w.x : w.y;
                else wv2 = (jp < 8) ? -w.y : w.x;
            }
            b1[ks] = (R)wv2;
        }
        ACC d1 = ACC{0, 0, 0, 0};
        if constexpr (PREC == PREC_BF16) d1 = mm16_bf16x2((const float*)a1, (const float*)b1, d1, (float*)tw);
        else d1 = mm16(a1, b1, d1);
        // ---- transpose D1 (lane: row mf_t::row(lane, reg), col li) into the A2
        // layout (row li, col 4ks + lq)
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int reg = 0; reg < 4; ++reg) tw[mf_t<PREC>::row(lane, reg) * 17 + li] = d1[reg];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        R a2[4], b2r[4], b2i[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) a2[ks] = tw[li * 17 + 4 * ks + lq];
        __builtin_amdgcn_wave_barrier();
        // ---- step 2 operands: Y[sigma][j], sigma = v*nv + v'
        const int sg = li, v = sg / a.nv, vp = sg % a.nv;
        const bool sok = sg < a.nv * a.nv;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int kap = 4 * ks + lq;
            const int j = kap & 7;
            double2 y = cmk(0, 0);
            if (sok && j < no)
                y = cmul(sCr[(kk * ns + v) * 8 + j], cconj(sCc[(kq * ns + vp) * 8 + j]));
            b2r[ks] = (R)(kap < 8 ? y.x : -y.y);
            b2i[ks] = (R)(kap < 8 ? y.y : y.x);
        }
        ACC dr = ACC{0, 0, 0, 0}, di = ACC{0, 0, 0, 0};
        if constexpr (PREC == PREC_BF16) {
            dr = mm16_bf16x2((const float*)a2, (const float*)b2r, dr, (float*)tw);
            di = mm16_bf16x2((const float*)a2, (const float*)b2i, di, (float*)tw);
        } else {
            dr = mm16(a2, b2r, dr);
            di = mm16(a2, b2i, di);
        }
        // ---- epilogue: lane holds D[rho = mf_t::row(lane, reg)][sigma = li]
#pragma unroll
        for (int reg = 0; reg < 4; ++reg) {
            const int rh = mf_t<PREC>::row(lane, reg);
            if (rh >= a.nc * a.nc || !sok) continue;
            const int cc = rh / a.nc, ccp = rh % a.nc;
            const int rl = kk * B + cc * a.nv + v, cl = kq * B + ccp * a.nv + vp;
            const int64_t grow = (int64_t)k0 * B + rl, gcol = (int64_t)kp0 * B + cl;
            double2 h = cmk(a.alpha * (double)dr[reg], a.alpha * (double)di[reg]);
            double2 x = cmk(0, 0);
            for (int j = 0; j < no; ++j) cfma(x, sPV[rl * 8 + j], cconj(sP[cl * 8 + j]));
            h = cadd(h, cscale(x, a.beta));
            if (grow == gcol) { h.x += a.dE[grow]; h.y = 0.0; }
            if (grow >= gcol) cplx<OUT>::store(H + tix(grow, gcol, a.nt), h);
        }
    }
}

extern "C" int bse_assemble_mfma_tk(int nv, int nc);
extern "C" int bse_assemble_mfma_supported(int nv, int nc, int norb);
// ---------------------------------------------------------------------------
// Matrix-free FP64 product Y = H X for k <= 16 vectors on matrix cores (the
// FP64 refinement of the mixed path).  Workgroup (ti, g) owns the TR rows of
// k-point block ti and the column blocks tj = g, g + ncg, ...: for each it
// forms the TR x TR tile of the screened direct term of H exactly as the
// FP64 assembly does (same two chained MFMA contractions per (k, k') pair)
// but into LDS, and multiplies it by the tile's rows of X with
// v_mfma_f64_16x16x4 (wave w: rows 16w .. 16w+15, all 16 vectors):
//   Yr += Hr Xr - Hi Xi,   Yi += Hr Xi + Hi Xr.
// Partial results per column group go to Yp[g] (N x 16).  The exchange term
// has rank norb, H^x = beta PV P^H, so it is applied as PV (P^H X) outside
// the tile loop: k_apply_px / k_apply_pxsum form Z = P^H X (norb x k, fixed
// summation order), and k_apply_sum adds, in a fixed order, the partials,
// beta PV Z and the diagonal dE X.  Requires TR <= 64 (bse_assemble_mfma_tk).
// ---------------------------------------------------------------------------
#define AM_K 32      // vectors per launch: KG groups of 16, one MFMA column block each
template <int KG>
__global__ __launch_bounds__(256) void k_apply_mfma(MfArgs a, const double2* __restrict__ X, int64_t ldx, int k,
                                                    int ncg, double2* __restrict__ Yp)
{
    constexpr int KX = 16 * KG;               // vectors staged per tile
    extern __shared__ double2 smem[];
    const int B = a.nc * a.nv, ns = a.nv + a.nc, no = a.norb, TK = a.TK;
    const int TR = TK * B;
    const int ti = blockIdx.x, g = blockIdx.y;
    const int nt = (int)ceil_div64(a.nk, TK);
    const int k0 = ti * TK;
    double2* sCr = smem;                       // [TK][ns][8]
    double2* sCc = sCr + TK * ns * 8;          // [TK][ns][8]
    double* sT = (double*)(sCc + TK * ns * 8); // per-wave 16x17 transposition
    double* sHr = sT + 4 * 272;                // [64][65]
    double* sHi = sHr + 64 * 65;
    double* sXr = sHi + 64 * 65;               // [64][KX]
    double* sXi = sXr + 64 * KX;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int e = tid; e < TK * ns * 8; e += 256) {
        const int i = e & 7, sl = (e >> 3) % ns, kk = e / (8 * ns);
        const int kq = k0 + kk;
        sCr[e] = (kq < a.nk && i < no) ? a.coef[((size_t)kq * ns + sl) * no + i] : cmk(0, 0);
    }
    // rows and columns of the 64 x 64 LDS tile beyond TR stay zero
    for (int e = tid; e < 64 * 65; e += 256) { sHr[e] = 0.0; sHi[e] = 0.0; }
    double* tw = sT + wv * 272;
    const int li = lane & 15, lq = lane >> 4;
    mf_d4 yr[KG], yi[KG];
#pragma unroll
    for (int q = 0; q < KG; ++q) { yr[q] = mf_d4{0, 0, 0, 0}; yi[q] = mf_d4{0, 0, 0, 0}; }
    for (int tj = g; tj < nt; tj += ncg) {
        const int kp0 = tj * TK;
        __syncthreads();   // previous tile fully consumed
        for (int e = tid; e < TK * ns * 8; e += 256) {
            const int i = e & 7, sl = (e >> 3) % ns, kk = e / (8 * ns);
            const int kp = kp0 + kk;
            sCc[e] = (kp < a.nk && i < no) ? a.coef[((size_t)kp * ns + sl) * no + i] : cmk(0, 0);
        }
        for (int e = tid; e < 64 * KX; e += 256) {
            const int rr = e / KX, j = e % KX;
            const int64_t col = (int64_t)kp0 * B + rr;
            double2 x = cmk(0, 0);
            if (rr < TR && col < a.N && j < k) x = X[col + (int64_t)j * ldx];
            sXr[e] = x.x;
            sXi[e] = x.y;
        }
        __syncthreads();
        for (int p = wv; p < TK * TK; p += 4) {
            const int kk = p / TK, kq = p % TK;
            const int kx = k0 + kk, kp = kp0 + kq;
            const bool pok = kx < a.nk && kp < a.nk;
            const double2* wq = a.wtab + (pok ? (size_t)ktable_index(kx, kp, a.n1, a.n2, a.n3) * no * no : 0);
            double a1[4], b1[4];
            const int rho = li, c = rho / a.nc, cp = rho % a.nc;
            const bool rok = rho < a.nc * a.nc;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const int kap = 4 * ks + lq;
                const int i = kap & 7;
                double2 x = cmk(0, 0);
                if (pok && rok && i < no)
                    x = cmulc(sCr[(kk * ns + a.nv + c) * 8 + i], sCc[(kq * ns + a.nv + cp) * 8 + i]);
                a1[ks] = kap < 8 ? x.x : x.y;
                const int jp = li;
                const int jj = jp & 7, ii = kap & 7;
                double wv2 = 0.0;
                if (pok && ii < no && jj < no) {
                    const double2 w = wq[ii * no + jj];
                    if (kap < 8) wv2 = (jp < 8) ? w.x : w.y;
                    else wv2 = (jp < 8) ? -w.y : w.x;
                }
                b1[ks] = wv2;
            }
            mf_d4 d1 = mm16(a1, b1, mf_d4{0, 0, 0, 0});
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) tw[mfma16_row_f64(lane, reg) * 17 + li] = d1[reg];
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            double a2[4], b2r[4], b2i[4];
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) a2[ks] = tw[li * 17 + 4 * ks + lq];
            __builtin_amdgcn_wave_barrier();
            const int sg = li, v = sg / a.nv, vp = sg % a.nv;
            const bool sok = sg < a.nv * a.nv;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                const int kap = 4 * ks + lq;
                const int j = kap & 7;
                double2 y = cmk(0, 0);
                if (pok && sok && j < no)
                    y = cmul(sCr[(kk * ns + v) * 8 + j], cconj(sCc[(kq * ns + vp) * 8 + j]));
                b2r[ks] = kap < 8 ? y.x : -y.y;
                b2i[ks] = kap < 8 ? y.y : y.x;
            }
            const mf_d4 dr = mm16(a2, b2r, mf_d4{0, 0, 0, 0});
            const mf_d4 di = mm16(a2, b2i, mf_d4{0, 0, 0, 0});
#pragma unroll
            for (int reg = 0; reg < 4; ++reg) {
                const int rh = mfma16_row_f64(lane, reg);
                if (rh >= a.nc * a.nc || !sok) continue;
                const int cc = rh / a.nc, ccp = rh % a.nc;
                const int rl = kk * B + cc * a.nv + v, cl = kq * B + ccp * a.nv + vp;
                const int64_t grow = (int64_t)k0 * B + rl, gcol = (int64_t)kp0 * B + cl;
                double2 hv = cmk(0, 0);
                if (pok) {
                    hv = cmk(a.alpha * dr[reg], a.alpha * di[reg]);
                    if (grow == gcol) hv.y = 0.0;   // the direct term's diagonal is real
                }
                sHr[rl * 65 + cl] = hv.x;
                sHi[rl * 65 + cl] = hv.y;
            }
        }
        __syncthreads();
        // Y rows 16 wv .. 16 wv + 15 += H(rows, 0:64) X(0:64, 0:KX), one H
        // tile formed for all KX vectors
        const int hr0 = (16 * wv + li) * 65;
#pragma unroll 4
        for (int kc = 0; kc < 64; kc += 4) {
            const double hr = sHr[hr0 + kc + lq], hi = sHi[hr0 + kc + lq];
#pragma unroll
            for (int q = 0; q < KG; ++q) {
                const double xr = sXr[(kc + lq) * KX + 16 * q + li], xi = sXi[(kc + lq) * KX + 16 * q + li];
                yr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(hr, xr, yr[q], 0, 0, 0);
                yr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-hi, xi, yr[q], 0, 0, 0);
                yi[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(hr, xi, yi[q], 0, 0, 0);
                yi[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(hi, xr, yi[q], 0, 0, 0);
            }
        }
    }
    // lane holds Y[row 16 wv + mfma16_row_f64(lane, reg)][vector 16 q + li]
#pragma unroll
    for (int q = 0; q < KG; ++q)
#pragma unroll
        for (int reg = 0; reg < 4; ++reg) {
            const int rl = 16 * wv + mfma16_row_f64(lane, reg);
            const int64_t grow = (int64_t)k0 * B + rl;
            if (rl < TR && grow < a.N) Yp[((int64_t)g * a.N + grow) * AM_K + 16 * q + li] = cmk(yr[q][reg], yi[q][reg]);
        }
}

// Z partials: block g sums conj(P(r, i)) X(r, j) over its rows; 256
// entries (i < 8, j < 32), one thread each.
#define APX_BLOCKS 256
__global__ __launch_bounds__(256) void k_apply_px(int64_t N, int no, int k, const double2* __restrict__ P,
                                                  const double2* __restrict__ X, int64_t ldx,
                                                  double2* __restrict__ Zp)
{
    const int t = threadIdx.x, i = t & 7, j = t >> 3;
    const int64_t per = ceil_div64(N, (int64_t)gridDim.x);
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(N, r0 + per);
    double2 acc = cmk(0, 0);
    if (i < no && j < k)
        for (int64_t r = r0; r < r1; ++r) cfmac(acc, P[r * no + i], X[r + (int64_t)j * ldx]);
    Zp[(int64_t)blockIdx.x * 256 + t] = acc;
}

// Z(i, j) = sum over the G partials in order (one thread per entry).
__global__ __launch_bounds__(256) void k_apply_pxsum(int G, const double2* __restrict__ Zp,
                                                     double2* __restrict__ Z)
{
    double2 z = cmk(0, 0);
    for (int g = 0; g < G; ++g) z = cadd(z, Zp[(int64_t)g * 256 + threadIdx.x]);
    Z[threadIdx.x] = z;
}

// Y(r, j) = sum_g Yp[g](r, j) + beta sum_i PV(r, i) Z(i, j) + dE(r) X(r, j).
__global__ __launch_bounds__(256) void k_apply_sum(int64_t N, int k, int ncg, int no,
                                                   const double2* __restrict__ Yp,
                                                   const double2* __restrict__ PV,
                                                   const double2* __restrict__ Z, const double* __restrict__ dE,
                                                   double beta, const double2* __restrict__ X, int64_t ldx,
                                                   double2* __restrict__ Y, int64_t ldy)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= N * AM_K) return;
    const int64_t r = e / AM_K;
    const int j = (int)(e % AM_K);
    if (j >= k) return;
    double2 s = cmk(0, 0);
    for (int gg = 0; gg < ncg; ++gg) s = cadd(s, Yp[(int64_t)gg * N * AM_K + e]);
    double2 x = cmk(0, 0);
    for (int i = 0; i < no; ++i) cfma(x, PV[r * no + i], Z[i + 8 * j]);
    const double2 xr = X[r + (int64_t)j * ldx];
    s = cadd(s, cscale(x, beta));
    s = cadd(s, cscale(xr, dE[r]));
    Y[r + (int64_t)j * ldy] = s;
}

static size_t apply_mfma_lds_bytes(int TK, int nv, int nc, int kg) {
    const int ns = nv + nc, B = nv * nc;
    (void)B;
    return ((size_t)2 * TK * ns * 8) * sizeof(double2) +
           ((size_t)4 * 272 + 2 * 64 * 65 + 2 * 64 * 16 * kg) * sizeof(double);
}

// Column groups per row block of the matrix-core product (about four
// workgroups per compute unit), and the size of its partial buffer.
extern "C" int bse_apply_mfma_groups(int nk, int nv, int nc) {
    const int nt = (int)ceil_div64(nk, bse_assemble_mfma_tk(nv, nc));
    return (int)std::max<int64_t>(1, std::min<int64_t>(nt, ceil_div64(1024, nt)));
}
extern "C" int64_t bse_apply_mfma_workspace(int nk, int nv, int nc) {
    return (int64_t)bse_apply_mfma_groups(nk, nv, nc) * nk * nv * nc * AM_K + (APX_BLOCKS + 1) * 256;
}

extern "C" hipError_t bse_launch_apply_mfma(hipStream_t st, int nk, int n1, int n2, int n3, int nv, int nc,
                                            int norb, const double2* coef, const double2* wtab,
                                            const double2* PV, const double2* P, const double* dE,
                                            double alpha, double beta, int k, const double2* X,
                                            int64_t ldx, double2* Y, int64_t ldy, double2* ws)
{
    if (!bse_assemble_mfma_supported(nv, nc, norb) || k <= 0 || k > AM_K) return hipErrorInvalidValue;
    MfArgs a;
    a.nk = nk; a.n1 = n1; a.n2 = n2; a.n3 = n3; a.nv = nv; a.nc = nc; a.norb = norb;
    a.TK = bse_assemble_mfma_tk(nv, nc);
    if (a.TK * nv * nc > 64) return hipErrorInvalidValue;
    a.N = (int64_t)nk * nv * nc;
    a.nt = tile_count(a.N);
    a.coef = coef; a.wtab = wtab; a.PV = PV; a.P = P; a.dE = dE; a.alpha = alpha; a.beta = beta;
    const int nt = (int)ceil_div64(nk, a.TK);
    const int ncg = bse_apply_mfma_groups(nk, nv, nc);
    const int kg = k > 16 ? 2 : 1;
    const size_t lds = apply_mfma_lds_bytes(a.TK, nv, nc, kg);
    double2* Zp = ws + (int64_t)ncg * a.N * AM_K;   // exchange: Z = P^H X
    double2* Z = Zp + (int64_t)APX_BLOCKS * 256;
    hipLaunchKernelGGL(k_apply_px, dim3(APX_BLOCKS), dim3(256), 0, st, a.N, norb, k, P, X, ldx, Zp);
    hipLaunchKernelGGL(k_apply_pxsum, dim3(1), dim3(256), 0, st, APX_BLOCKS, (const double2*)Zp, Z);
    // the LDS beyond 64 KB is asked for explicitly (once per process)
    static const hipError_t attr = [] {
        const int big = 160 * 1024 - 1024;
        hipError_t e = hipFuncSetAttribute((const void*)k_apply_mfma<1>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
        if (e == hipSuccess)
            e = hipFuncSetAttribute((const void*)k_apply_mfma<2>, hipFuncAttributeMaxDynamicSharedMemorySize, big);
        return e;
    }();
    if (attr != hipSuccess) return attr;
    if (kg == 2)
        hipLaunchKernelGGL(k_apply_mfma<2>, dim3((unsigned)nt, (unsigned)ncg), dim3(256), lds, st, a, X, ldx, k,
                           ncg, ws);
    else
        hipLaunchKernelGGL(k_apply_mfma<1>, dim3((unsigned)nt, (unsigned)ncg), dim3(256), lds, st, a, X, ldx, k,
                           ncg, ws);
    hipLaunchKernelGGL(k_apply_sum, dim3((unsigned)ceil_div64(a.N * AM_K, 256)), dim3(256), 0, st, a.N, k,
                       ncg, norb, (const double2*)ws, PV, (const double2*)Z, dE, beta, X, ldx, Y, ldy);
    return hipGetLastError();
}

extern "C" int bse_assemble_mfma_supported(int nv, int nc, int norb) {
    return nc * nc <= 16 && nv * nv <= 16 && norb <= 8 && nv * nc >= 4;
}

extern "C" int bse_assemble_mfma_tk(int nv, int nc) {
    const int B = nv * nc;
    int tk = 1;
    while (tk * 2 * B <= 64) tk *= 2;
    return tk;
}

static size_t mfma_lds_bytes(int TK, int nv, int nc, int prec) {
    const int ns = nv + nc, B = nv * nc;
    const size_t c2 = (size_t)2 * TK * ns * 8 + (size_t)2 * TK * B * 8;
    const size_t rb = (prec == PREC_F64 ? 8 : 4) * (size_t)4 * (prec == PREC_BF16 ? 544 : 272);
    return c2 * sizeof(double2) + rb;
}

extern "C" hipError_t bse_launch_assemble_mfma(hipStream_t st, int prec, int nk, int n1, int n2,
                                               int n3, int nv, int nc, int norb, int64_t ld,
                                               const double2* coef, const double2* wtab,
                                               const double2* PV, const double2* P,
                                               const double* dE, double alpha, double beta,
                                               void* H)
{
    if (!bse_assemble_mfma_supported(nv, nc, norb)) return hipErrorInvalidValue;
    MfArgs a;
    a.nk = nk; a.n1 = n1; a.n2 = n2; a.n3 = n3; a.nv = nv; a.nc = nc; a.norb = norb;
    a.TK = bse_assemble_mfma_tk(nv, nc);
    a.N = (int64_t)nk * nv * nc;
    a.nt = tile_count(a.N);
    (void)ld;            // the matrix is stored in tiles (bse_common.hpp)
    a.coef = coef; a.wtab = wtab; a.PV = PV; a.P = P; a.dE = dE; a.alpha = alpha; a.beta = beta;
    const int64_t nt = ceil_div64(nk, a.TK);
    const dim3 grid((unsigned)(nt * (nt + 1) / 2));
    const size_t lds = mfma_lds_bytes(a.TK, nv, nc, prec);
    if (prec == PREC_F64)
        hipLaunchKernelGGL(k_assemble_mfma<PREC_F64>, grid, dim3(256), lds, st, a, (double2*)H);
    else if (prec == PREC_F32)
        hipLaunchKernelGGL(k_assemble_mfma<PREC_F32>, grid, dim3(256), lds, st, a, (float2*)H);
    else
        hipLaunchKernelGGL(k_assemble_mfma<PREC_BF16>, grid, dim3(256), lds, st, a, (float2*)H);
    return hipGetLastError();
}
