// Matrix-free FP64 products with the BSE matrix (or with one of its parameter
// derivatives): Y += H X for a block of k vectors, with every element of H
// recomputed on the fly from the interaction tables -- no N x N FP64 matrix is
// stored.  Used by the FP64 Rayleigh-Ritz refinement of the lowest excitons
// (mixed-precision path) and by the Hellmann-Feynman gradient contraction
// A^H (dH/dtheta) A.
//
// Element formula (same as the assembly, bse.version13.cpp:2796-2897 in its
// textbook/Hermitian form):
//   H_rs = delta_rs dE_r + alpha sum_j Apair_{(k,k'),c,c',j} Vl_{k,v,j} conj(Vr_{k',v',j})
//        + beta sum_j PVr_{r,j} conj(Pc_{s,j})
// in the k-major order r = k*B + c*nv + v.  A workgroup owns TK k-points of
// rows (TR = TK*B rows) and sweeps over all column tiles; each thread owns one
// row and a residue class of columns and accumulates its share of Y in
// registers, so the only LDS traffic is the tile's pair intermediates and the
// X rows of the current column tile.
//
// Trace mode (Bm != null, no vectors): instead of H X the kernel accumulates
// tr(H Bm) = sum_rs H_rs Bm_sr over the same recomputed tiles and adds each
// workgroup's share to tr_part[blockIdx.x] (fixed order, deterministic).
#include "bse_common.hpp"

#define APPLY_KMAX 32   // vectors per pass

struct ApplyArgs {
    int nk, n1, n2, n3, nv, nc, norb, TK, k;
    int64_t N, ldx, ldy;
    const double2* Cl;
    const double2* Cr;
    const double2* Vl;
    const double2* Vr;
    const double2* wtab;
    const double2* PVr;
    const double2* Pc;
    const double* dE;
    double alpha, beta;
    const double2* X;
    double2* Y;
    const double2* Bm;   // trace mode: N x N, ld ldb
    int64_t ldb;
    double2* tr_part;
};

__global__ __launch_bounds__(256) void k_apply(ApplyArgs a)
{
    extern __shared__ double2 smem[];
    const int B = a.nc * a.nv, ns = a.nv + a.nc, no = a.norb, TK = a.TK;
    const int TR = TK * B;
    const int k0 = blockIdx.x * TK;
    double2* sCl = smem;                               // TK x nc x no
    double2* sCr = sCl + TK * a.nc * no;
    double2* sVl = sCr + TK * a.nc * no;               // TK x nv x no
    double2* sVr = sVl + TK * a.nv * no;
    double2* sA = sVr + TK * a.nv * no;                // TK*TK x nc x nc x no
    double2* sPV = sA + TK * TK * a.nc * a.nc * no;    // TR x no
    double2* sP = sPV + TR * no;                       // TR x no
    double2* sX = sP + TR * no;                        // TR x k  (column-tile rows of X)
    double2* sT = sX + TR * a.k;                       // TR x TR tile of H
    const int tid = threadIdx.x;
    const int ngrp = blockDim.x / TR;
    const int rr = tid % TR, grp = tid / TR;
    const int64_t grow = (int64_t)k0 * B + rr;

    for (int e = tid; e < TK * a.nc * no; e += blockDim.x) {
        const int i = e % no, c = (e / no) % a.nc, kk = e / (no * a.nc);
        const int k = k0 + kk;
        sCl[e] = (k < a.nk) ? a.Cl[((size_t)k * ns + a.nv + c) * no + i] : cmk(0, 0);
    }
    for (int e = tid; e < TK * a.nv * no; e += blockDim.x) {
        const int i = e % no, v = (e / no) % a.nv, kk = e / (no * a.nv);
        const int k = k0 + kk;
        sVl[e] = (k < a.nk) ? a.Vl[((size_t)k * ns + v) * no + i] : cmk(0, 0);
    }
    if (a.PVr != nullptr) {
        for (int e = tid; e < TR * no; e += blockDim.x) {
            const int64_t row = (int64_t)k0 * B + e / no;
            sPV[e] = (row < a.N) ? a.PVr[row * no + e % no] : cmk(0, 0);
        }
    }
    double2 tacc = cmk(0, 0);
    double2 acc[APPLY_KMAX / 2];
#pragma unroll
    for (int v = 0; v < APPLY_KMAX / 2; ++v) acc[v] = cmk(0, 0);

    const int nt = (int)ceil_div64(a.nk, TK);
    for (int tj = 0; tj < nt; ++tj) {
        const int kp0 = tj * TK;
        __syncthreads();
        for (int e = tid; e < TK * a.nc * no; e += blockDim.x) {
            const int i = e % no, cc = (e / no) % a.nc, q = e / (no * a.nc);
            const int kp = kp0 + q;
            sCr[e] = (kp < a.nk) ? a.Cr[((size_t)kp * ns + a.nv + cc) * no + i] : cmk(0, 0);
        }
        for (int e = tid; e < TK * a.nv * no; e += blockDim.x) {
            const int i = e % no, v = (e / no) % a.nv, q = e / (no * a.nv);
            const int kp = kp0 + q;
            sVr[e] = (kp < a.nk) ? a.Vr[((size_t)kp * ns + v) * no + i] : cmk(0, 0);
        }
        if (a.PVr != nullptr) {
            for (int e = tid; e < TR * no; e += blockDim.x) {
                const int64_t col = (int64_t)kp0 * B + e / no;
                sP[e] = (col < a.N) ? a.Pc[col * no + e % no] : cmk(0, 0);
            }
        }
        for (int e = tid; a.Bm == nullptr && e < TR * a.k; e += blockDim.x) {
            const int s2 = e % TR, v = e / TR;
            const int64_t col = (int64_t)kp0 * B + s2;
            sX[s2 * a.k + v] = (col < a.N) ? a.X[col + (int64_t)v * a.ldx] : cmk(0, 0);
        }
        __syncthreads();
        const int nA = TK * TK * a.nc * a.nc * no;
        for (int e = tid; e < nA; e += blockDim.x) {
            const int j = e % no;
            const int cp = (e / no) % a.nc;
            const int cq = (e / (no * a.nc)) % a.nc;
            const int p = e / (no * a.nc * a.nc);
            const int pk = p / TK, pq = p % TK;
            const int k = k0 + pk, kp = kp0 + pq;
            double2 s2 = cmk(0, 0);
            if (k < a.nk && kp < a.nk) {
                const double2* wq = a.wtab + (size_t)ktable_index(k, kp, a.n1, a.n2, a.n3) * no * no;
                const double2* cl = sCl + (pk * a.nc + cq) * no;
                const double2* cr = sCr + (pq * a.nc + cp) * no;
                for (int i = 0; i < no; ++i) cfma(s2, cmulc(cl[i], cr[i]), wq[i * no + j]);
            }
            sA[e] = s2;
        }
        __syncthreads();
        for (int e = tid; e < TR * TR; e += blockDim.x) {
            const int r = e / TR, s2 = e % TR;
            const int64_t gr = (int64_t)k0 * B + r, gc = (int64_t)kp0 * B + s2;
            double2 h = cmk(0, 0);
            if (gr < a.N && gc < a.N) {
                const int kk = r / B, b = r % B, c = b / a.nv, vv = b % a.nv;
                const int kq = s2 / B, bp = s2 % B, cp = bp / a.nv, vp = bp % a.nv;
                const double2* Ap = sA + (((kk * TK + kq) * a.nc + c) * a.nc + cp) * no;
                const double2* vl = sVl + (kk * a.nv + vv) * no;
                const double2* vr = sVr + (kq * a.nv + vp) * no;
                double2 d = cmk(0, 0);
                for (int j = 0; j < no; ++j) cfma(d, Ap[j], cmul(vl[j], cconj(vr[j])));
                h = cscale(d, a.alpha);
                if (a.PVr != nullptr) {
                    double2 x = cmk(0, 0);
                    const double2* pv = sPV + r * no;
                    const double2* pc = sP + s2 * no;
                    for (int j = 0; j < no; ++j) cfma(x, pv[j], cconj(pc[j]));
                    h = cadd(h, cscale(x, a.beta));
                }
                if (gc == gr && a.dE != nullptr) h.x += a.dE[gr];
            }
            sT[e] = h;
        }
        __syncthreads();
        if (a.Bm != nullptr) {
            for (int e = tid; e < TR * TR; e += blockDim.x) {
                const int r = e / TR, s2 = e % TR;
                const int64_t gr = (int64_t)k0 * B + r, gc = (int64_t)kp0 * B + s2;
                if (gr < a.N && gc < a.N) cfma(tacc, sT[e], a.Bm[gc + gr * a.ldb]);
            }
        } else if (grp < ngrp) {
            const double2* trow = sT + rr * TR;
#pragma unroll
            for (int vi = 0; vi < APPLY_KMAX / 2; ++vi) {
                const int v = grp + vi * ngrp;
                if (v >= a.k) break;
                double2 s2 = acc[vi];
                for (int s3 = 0; s3 < TR; ++s3) cfma(s2, trow[s3], sX[s3 * a.k + v]);
                acc[vi] = s2;
            }
        }
    }
    if (a.Bm != nullptr) {
        __shared__ double sh[32];
        tacc = block_sum2(tacc, sh);
        if (tid == 0) a.tr_part[blockIdx.x] = cadd(a.tr_part[blockIdx.x], tacc);
        return;
    }
    if (grp < ngrp && grow < a.N) {
#pragma unroll
        for (int vi = 0; vi < APPLY_KMAX / 2; ++vi) {
            const int v = grp + vi * ngrp;
            if (v >= a.k) break;
            double2* y = a.Y + grow + (int64_t)v * a.ldy;
            *y = cadd(*y, acc[vi]);
        }
    }
}

static size_t apply_lds_elems(int TK, int nv, int nc, int norb, int k) {
    const int B = nv * nc, TR = TK * B;
    return (size_t)TK * nc * norb * 2 + (size_t)TK * nv * norb * 2 +
           (size_t)TK * TK * nc * nc * norb + (size_t)2 * TR * norb + (size_t)TR * k +
           (size_t)TR * TR;
}

// Largest power-of-two k tile with 16 <= TR = TK*B <= 64 (so every thread's
// vector slice fits APPLY_KMAX/2 accumulators) and <= 60 KB of LDS.
extern "C" int bse_apply_pick_tk(int nv, int nc, int norb) {
    const int B = nv * nc;
    int tk = 1;
    while (2 * tk * B <= 64 &&
           apply_lds_elems(2 * tk, nv, nc, norb, APPLY_KMAX) * sizeof(double2) <= 60 * 1024)
        tk *= 2;
    return tk;
}

extern "C" hipError_t bse_launch_apply(hipStream_t st, int nk, int n1, int n2, int n3, int nv, int nc,
                                       int norb, int TK, const double2* Cl, const double2* Cr,
                                       const double2* Vl, const double2* Vr, const double2* wtab,
                                       const double2* PVr, const double2* Pc, const double* dE,
                                       double alpha, double beta, int k, const double2* X,
                                       int64_t ldx, double2* Y, int64_t ldy, const double2* Bm,
                                       int64_t ldb, double2* tr_part)
{
    ApplyArgs a;
    a.nk = nk; a.n1 = n1; a.n2 = n2; a.n3 = n3; a.nv = nv; a.nc = nc; a.norb = norb; a.TK = TK;
    a.N = (int64_t)nk * nv * nc;
    a.Cl = Cl; a.Cr = Cr; a.Vl = Vl; a.Vr = Vr; a.wtab = wtab; a.PVr = PVr; a.Pc = Pc; a.dE = dE;
    a.alpha = alpha; a.beta = beta;
    const int TR = TK * nv * nc;
    // each of the 256/TR vector groups holds at most APPLY_KMAX/2 accumulators
    if (TR > 256 || (int64_t)APPLY_KMAX > (int64_t)(256 / TR) * (APPLY_KMAX / 2))
        return hipErrorInvalidValue;
    a.Bm = Bm; a.ldb = ldb; a.tr_part = tr_part;
    if (Bm != nullptr) {
        if (tr_part == nullptr) return hipErrorInvalidValue;
        a.k = 0; a.X = nullptr; a.ldx = 0; a.Y = nullptr; a.ldy = 0;
        const size_t lds = apply_lds_elems(TK, nv, nc, norb, 0) * sizeof(double2);
        hipLaunchKernelGGL(k_apply, dim3((unsigned)ceil_div64(nk, TK)), dim3(256), lds, st, a);
        return hipGetLastError();
    }
    for (int v0 = 0; v0 < k; v0 += APPLY_KMAX) {
        a.k = (k - v0 < APPLY_KMAX) ? (k - v0) : APPLY_KMAX;
        a.X = X + (int64_t)v0 * ldx; a.ldx = ldx;
        a.Y = Y + (int64_t)v0 * ldy; a.ldy = ldy;
        const size_t lds = apply_lds_elems(TK, nv, nc, norb, a.k) * sizeof(double2);
        hipLaunchKernelGGL(k_apply, dim3((unsigned)ceil_div64(nk, TK)), dim3(256), lds, st, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
