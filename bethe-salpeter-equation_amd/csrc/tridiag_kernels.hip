// K6: eigenpairs of the real symmetric tridiagonal T (diagonal d, off-diagonal
// e) produced by K5.
//
// Eigenvalues: Sturm-count bisection (Golub & Van Loan, sec. 8.4.1) run as a
// (W+1)-way multisection -- W lanes of a wavefront (W = 2, 4 or 8, fewer for
// larger matrices) evaluate the count at W interior points of the current
// interval, so each round narrows it (W+1)-fold.
// The recurrence divides by a reciprocal refined with one Newton step
// instead of the correctly rounded division sequence: the hardware
// reciprocal is good to 4.6e-8 relative, one step brings it to 2.2e-15
// (about 20 ulp; both maxima over 2^20 normal inputs on gfx950), a
// perturbation of each step that moves the eigenvalues by O(20 eps |T|).
// The matrix is first split at negligible off-diagonals
// (e_i^2 <= eps^2 |d_i d_{i+1}| + safmin) and each unreduced block is solved
// on its own, so exactly degenerate eigenvalues of the BSE matrix (which show
// up as splits) get orthogonal eigenvectors by construction.
//
// Eigenvectors: inverse iteration (Wilkinson; G&VL sec. 8.2.2) with the
// tridiagonal LU factorisation with partial pivoting, three solves from a
// deterministic pseudo-random start, one thread per eigenvalue; the
// factorisation and iterate are interleaved across threads so that every
// step is a coalesced load.  Eigenvalues closer than a cluster tolerance are
// re-orthogonalised by modified Gram-Schmidt (twice) in a second kernel.
#include "bse_common.hpp"
#include <cfloat>
#include <climits>

// 1/b from the hardware reciprocal and one Newton step (|b| >= pivmin is a
// normal number here).
__device__ inline double fast_rcp(double b) {
    const double r = __builtin_amdgcn_rcp(b);
    return fma(fma(-b, r, 1.0), r, r);
}

// Count of eigenvalues of block [bs, bs+nbk) strictly below x.  The
// recurrence is a chain of dependent operations; its inputs are read sixteen
// steps at a time, the next sixteen while the current ones are consumed.
__device__ inline int sturm_count(const double* __restrict__ d, const double* __restrict__ e2,
                                  int bs, int nbk, double x, double pivmin)
{
    constexpr int SB = 16;
    double q = d[bs] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    int cnt = q < 0.0;
    int i = 1;
    if (i + SB <= nbk) {
        double dv[SB], ev[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            dv[u] = d[bs + i + u];
            ev[u] = e2[bs + i + u - 1];
        }
        for (; i + 2 * SB <= nbk; i += SB) {
            double dn[SB], en[SB];
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                dn[u] = d[bs + i + SB + u];
                en[u] = e2[bs + i + SB + u - 1];
            }
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                q = (dv[u] - x) - ev[u] * fast_rcp(q);
                if (fabs(q) <= pivmin) q = -pivmin;
                cnt += q < 0.0;
            }
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                dv[u] = dn[u];
                ev[u] = en[u];
            }
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            q = (dv[u] - x) - ev[u] * fast_rcp(q);
            if (fabs(q) <= pivmin) q = -pivmin;
            cnt += q < 0.0;
        }
        i += SB;
    }
    for (; i < nbk; ++i) {
        q = (d[bs + i] - x) - e2[bs + i - 1] * fast_rcp(q);
        if (fabs(q) <= pivmin) q = -pivmin;
        cnt += q < 0.0;
    }
    return cnt;
}

// items: int4 (block start, block size, local index, block id); bounds per
// block id: (lo, hi).  One W-lane group per item: (W + 1)-way multisection.
// prm (may be null): device copy of (pivmin, tnorm) that replaces the
// by-value arguments, for launches recorded before those are known.
template <int W>
__global__ __launch_bounds__(256) void k_bisect(int nitems, const int4* __restrict__ items,
                                               const double2* __restrict__ bounds,
                                               const double* __restrict__ d,
                                               const double* __restrict__ e2, double pivmin,
                                               double tnorm, const double* __restrict__ prm,
                                               double* __restrict__ lam)
{
    if (prm) {
        pivmin = prm[0];
        tnorm = prm[1];
    }
    const int gid = (blockIdx.x * blockDim.x + threadIdx.x);
    static_assert(W == 2 || W == 4 || W == 8, "lanes per item");
    const int item = gid / W, g = gid % W;
    const int lane = threadIdx.x & 63;
    const int base = lane & ~(W - 1);
    const bool active = item < nitems;
    int4 it = make_int4(0, 1, 0, 0);
    double lo = 0.0, hi = 0.0;
    if (active) {
        it = items[item];
        const double2 b = bounds[it.w];
        lo = b.x;
        hi = b.y;
    }
    if (active && it.y == 1) {
        if (g == 0) lam[item] = d[it.x];
    }
    const bool work = active && it.y > 1;
    const double eps = DBL_EPSILON;
    // the interval shrinks (W + 1)-fold per round from at most 2 |T| to
    // eps |T|: 17 rounds for W = 8, 23 for W = 4, 34 for W = 2
    for (int round = 0; round < 80; ++round) {
        const double tol = 2.0 * eps * fmax(fabs(lo), fabs(hi)) + eps * tnorm;
        const bool conv = !work || (hi - lo) <= tol;
        // group-uniform; the whole wave keeps looping until every group converged
        const unsigned long long notdone = __ballot(!conv);
        if (notdone == 0ull) break;
        const double x = lo + (double)(g + 1) * (hi - lo) / (double)(W + 1);
        int cnt = 0;
        if (!conv) cnt = sturm_count(d, e2, it.x, it.y, x, pivmin);
        const unsigned long long below = __ballot(!conv && cnt <= it.z);
        // the split point is the leading run of lanes whose counts mark a
        // lower bound; the reciprocal-based recurrence does not guarantee a
        // count monotone in x, so a set bit after a clear one is not trusted
        const unsigned long long bits = (below >> base) & ((1ull << W) - 1);
        const int nlo = __builtin_ctzll(~bits);
        const double xlo = __shfl(x, base + max(nlo - 1, 0), 64);
        const double xhi = __shfl(x, base + min(nlo, W - 1), 64);
        if (!conv) {
            if (nlo > 0) lo = xlo;
            if (nlo < W) hi = xhi;
        }
    }
    if (work && g == 0) lam[item] = 0.5 * (lo + hi);
}

__device__ inline double hash_uniform(uint64_t a, uint64_t b) {
    uint64_t z = a * 0x9E3779B97F4A7C15ull + b * 0xD1B54A32D192ED03ull + 0x94D049BB133111EBull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Rows per memory round trip of the inverse iteration: with one thread per
// eigenvalue the kernel runs about one wave per CU, so its rate is the number
// of loads each wave keeps in flight.
#ifndef IIB
#define IIB 16
#endif

// Inverse iteration.  Thread t of the batch owns sorted eigenvalue m0 + t in
// block (bs, nb).  Scratch arrays are indexed [global row * bsz + t]; the
// pivot flags take one byte each at the start of the pf region.  The kernel
// is bound by its scratch traffic, so the flags are bytes and the random
// start vector is generated inside the first forward pass, never stored.
__global__ __launch_bounds__(64) void k_inverse_iter(
    int bsz, int cnt, int m0, const double* __restrict__ lam, const int* __restrict__ bsv,
    const int* __restrict__ nbv, const double* __restrict__ d, const double* __restrict__ e,
    double tnorm, const double* __restrict__ prm, double* __restrict__ u0,
    double* __restrict__ u1, double* __restrict__ u2, double* __restrict__ mul,
    double* __restrict__ pf, double* __restrict__ x)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    if (prm) tnorm = prm[1];
    const double lm = lam[m0 + t];
    const int bs = bsv[m0 + t], nb = nbv[m0 + t];
    const int64_t S = bsz;
    const double tiny = DBL_EPSILON * tnorm + DBL_MIN;
    unsigned char* __restrict__ pb = reinterpret_cast<unsigned char*>(pf);
    if (nb == 1) {
        x[(int64_t)bs * S + t] = 1.0;
        return;
    }
    // LU with partial pivoting of T_block - lam I.  The step takes e(k),
    // d(k+1), e(k+1); they are read IIB steps at a time (every thread of the
    // block reads the same ones) so that the chain does not wait on them.
    double cur_d = d[bs] - lm;
    double cur_s = e[bs];
    auto lu_step = [&](int k, double sub, double nd, double ns) {
        const int64_t gi = bs + k;
        double a0, a1, a2, m;
        unsigned char fl;
        if (fabs(cur_d) >= fabs(sub)) {
            fl = 0;
            m = (cur_d != 0.0) ? sub / cur_d : 0.0;
            a0 = cur_d; a1 = cur_s; a2 = 0.0;
            cur_d = nd - m * cur_s;
            cur_s = ns;
        } else {
            fl = 1;
            m = cur_d / sub;
            a0 = sub; a1 = nd; a2 = ns;
            cur_d = cur_s - m * nd;
            cur_s = -m * ns;
        }
        if (fabs(a0) < tiny) a0 = (a0 < 0.0) ? -tiny : tiny;
        u0[gi * S + t] = a0;
        u1[gi * S + t] = a1;
        u2[gi * S + t] = a2;
        mul[gi * S + t] = m;
        pb[gi * S + t] = fl;
    };
    {
        int k = 0;
        for (; k + IIB <= nb - 1; k += IIB) {
            double ev[IIB + 1], dv[IIB];
#pragma unroll
            for (int u = 0; u < IIB; ++u) {
                ev[u] = e[bs + k + u];
                dv[u] = d[bs + k + u + 1] - lm;
            }
            ev[IIB] = (k + IIB < nb - 1) ? e[bs + k + IIB] : 0.0;
#pragma unroll
            for (int u = 0; u < IIB; ++u) lu_step(k + u, ev[u], dv[u], ev[u + 1]);
        }
        for (; k < nb - 1; ++k)
            lu_step(k, e[bs + k], d[bs + k + 1] - lm, (k + 1 < nb - 1) ? e[bs + k + 1] : 0.0);
    }
    {
        double a0 = cur_d;
        if (fabs(a0) < tiny) a0 = (a0 < 0.0) ? -tiny : tiny;
        u0[(int64_t)(bs + nb - 1) * S + t] = a0;
    }
    // Random start; three solves.  The normalisation of each iterate is folded
    // into the next forward pass (scale s), and the last back substitution
    // records the largest component so the final pass also fixes the sign.
    // random start x(k) = hash_uniform(m0 + t, k), read in pass 0 instead of x
    double s = 1.0, big = 0.0;
    // The solves walk the column in steps of IIB rows: all loads of a step are
    // issued before its recurrence runs (the compiler cannot move them above
    // the stores into x), so each step costs one memory round trip.
    for (int it = 0; it < 3; ++it) {
        const bool first = (it == 0);
        double cur = (first ? hash_uniform(m0 + t, 0) : x[(int64_t)bs * S + t]) * s;
        int k = 0;
        for (; k + IIB <= nb - 1; k += IIB) {
            double xn[IIB], mu[IIB];
            unsigned char pv[IIB];
            if (first) {
#pragma unroll
                for (int u = 0; u < IIB; ++u) xn[u] = hash_uniform(m0 + t, k + u + 1);
            } else {
#pragma unroll
                for (int u = 0; u < IIB; ++u) xn[u] = x[(bs + k + u + 1) * S + t];
            }
#pragma unroll
            for (int u = 0; u < IIB; ++u) {
                const int64_t gi = bs + k + u;
                mu[u] = mul[gi * S + t];
                pv[u] = pb[gi * S + t];
            }
#pragma unroll
            for (int u = 0; u < IIB; ++u) {
                double nxt = xn[u] * s;
                if (pv[u] != 0) { const double tmp = cur; cur = nxt; nxt = tmp; }
                nxt -= mu[u] * cur;
                x[(bs + k + u) * S + t] = cur;
                cur = nxt;
            }
        }
        for (; k < nb - 1; ++k) {
            const int64_t gi = bs + k;
            double nxt = (first ? hash_uniform(m0 + t, k + 1) : x[(gi + 1) * S + t]) * s;
            if (pb[gi * S + t] != 0) { const double tmp = cur; cur = nxt; nxt = tmp; }
            nxt -= mul[gi * S + t] * cur;
            x[gi * S + t] = cur;
            cur = nxt;
        }
        const int64_t gl = bs + nb - 1;
        double x1 = cur / u0[gl * S + t];
        x[gl * S + t] = x1;
        double x2 = 0.0, nrm = x1 * x1;
        big = x1;
        int kb = nb - 2;
        for (; kb - IIB + 1 >= 0; kb -= IIB) {
            double xv[IIB], a0[IIB], a1[IIB], a2[IIB];
#pragma unroll
            for (int u = 0; u < IIB; ++u) {
                const int64_t gi = bs + kb - u;
                xv[u] = x[gi * S + t];
                a0[u] = u0[gi * S + t];
                a1[u] = u1[gi * S + t];
                a2[u] = u2[gi * S + t];
            }
#pragma unroll
            for (int u = 0; u < IIB; ++u) {
                const double xk = (xv[u] - a1[u] * x1 - a2[u] * x2) / a0[u];
                x[(bs + kb - u) * S + t] = xk;
                nrm += xk * xk;
                if (fabs(xk) > fabs(big)) big = xk;
                x2 = x1;
                x1 = xk;
            }
        }
        for (; kb >= 0; --kb) {
            const int64_t gi = bs + kb;
            const double xk = (x[gi * S + t] - u1[gi * S + t] * x1 - u2[gi * S + t] * x2) / u0[gi * S + t];
            x[gi * S + t] = xk;
            nrm += xk * xk;
            if (fabs(xk) > fabs(big)) big = xk;
            x2 = x1;
            x1 = xk;
        }
        s = 1.0 / sqrt(nrm);
    }
    const double f = (big < 0.0) ? -s : s;
    for (int k = 0; k < nb; ++k) x[(int64_t)(bs + k) * S + t] *= f;
}

// Modified Gram-Schmidt (two passes) inside clusters of close eigenvalues.
// clusters: int2 (offset, size); one workgroup per cluster.  With members
// (non-null) the cluster's eigenvalues are the sorted positions
// members[offset .. offset + size) (ascending, all in this batch); without,
// they are the consecutive batch items offset .. offset + size - 1.  All
// members of a cluster share the block (bs, nb): clusters are formed inside
// one unreduced block, where eigenvalues of other blocks may lie between them
// in the global order.
// ncl (may be null): device count of valid clusters; blocks beyond it return.
__global__ __launch_bounds__(256) void k_cluster_mgs(int bsz, const int2* __restrict__ clusters,
                                                     const int* __restrict__ ncl, int m0,
                                                     const int* __restrict__ bsv,
                                                     const int* __restrict__ nbv,
                                                     const int* __restrict__ members,
                                                     double* __restrict__ x)
{
    __shared__ double sh[32];
    if (ncl && (int)blockIdx.x >= *ncl) return;
    const int2 cl = clusters[blockIdx.x];
    const int csz = cl.y;
    auto item = [&](int i) { return members ? members[cl.x + i] - m0 : cl.x + i; };
    const int t0 = item(0);
    const int bs = bsv[m0 + t0], nb = nbv[m0 + t0];
    const int64_t S = bsz;
    for (int j = 0; j < csz; ++j) {
        const int tj = item(j);
        for (int pass = 0; pass < 2; ++pass) {
            for (int i = 0; i < j; ++i) {
                const int ti = item(i);
                double dot = 0.0;
                for (int k = threadIdx.x; k < nb; k += blockDim.x)
                    dot += x[(int64_t)(bs + k) * S + ti] * x[(int64_t)(bs + k) * S + tj];
                dot = block_sum(dot, sh);
                for (int k = threadIdx.x; k < nb; k += blockDim.x)
                    x[(int64_t)(bs + k) * S + tj] -= dot * x[(int64_t)(bs + k) * S + ti];
                __syncthreads();
            }
        }
        double nrm = 0.0;
        for (int k = threadIdx.x; k < nb; k += blockDim.x) {
            const double v = x[(int64_t)(bs + k) * S + tj];
            nrm += v * v;
        }
        nrm = block_sum(nrm, sh);
        const double inv = 1.0 / sqrt(nrm);
        for (int k = threadIdx.x; k < nb; k += blockDim.x) x[(int64_t)(bs + k) * S + tj] *= inv;
        __syncthreads();
    }
}

// Commit the batch: Z[:, col0 + t] = x_t (complex, imaginary part zero) for
// t < nstore, via a 64x64 LDS transpose; rows outside the vector's block are
// zero.
__global__ __launch_bounds__(256) void k_commit(int n, int bsz, int cnt, int nstore, int m0,
                                                const int* __restrict__ bsv,
                                                const int* __restrict__ nbv,
                                                const double* __restrict__ x,
                                                double2* __restrict__ Z, int64_t ldz, int col0)
{
    __shared__ double tile[64][65];
    const int r0 = blockIdx.x * 64, t0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int rr = ty; rr < 64; rr += 4) {
        const int row = r0 + rr, t = t0 + tx;
        double v = 0.0;
        if (row < n && t < cnt) {
            const int bs = bsv[m0 + t], nb = nbv[m0 + t];
            if (row >= bs && row < bs + nb) v = x[(int64_t)row * bsz + t];
        }
        tile[rr][tx] = v;
    }
    __syncthreads();
    for (int tt = ty; tt < 64; tt += 4) {
        const int t = t0 + tt, row = r0 + tx;
        if (t < nstore && row < n) Z[row + (int64_t)(col0 + t) * ldz] = cmk(tile[tx][tt], 0.0);
    }
}

// Contract the batch with a complex vector: f[m0 + t] = sum_rows x_t[row] conj(y[row]).
// With y = Q^H rho this is f_m = sum_r conj(rho_r) (Q x_m)_r, the oscillator
// strength of the back-transformed eigenvector, without forming it.
__global__ __launch_bounds__(256) void k_contract(int bsz, int cnt, int m0,
                                                  const int* __restrict__ bsv,
                                                  const int* __restrict__ nbv,
                                                  const double* __restrict__ x,
                                                  const double2* __restrict__ yc,
                                                  double2* __restrict__ f)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int bs = bsv[m0 + t], nb = nbv[m0 + t];
    double2 acc = cmk(0, 0);
    for (int k = 0; k < nb; ++k) {
        const double v = x[(int64_t)(bs + k) * bsz + t];
        const double2 y = yc[bs + k];
        acc.x += v * y.x;
        acc.y -= v * y.y;
    }
    f[m0 + t] = acc;
}

// The same contraction for large blocks: the walk over each eigenvector is
// cut into CT_SEG segments, one per grid row, each item's segment partial
// written to part (CT_SEG x cnt), then k_contract_sum adds the partials in
// segment order (fixed, identical on every run).  One wave per workgroup, so
// a block of order 20736 runs 324 x 8 workgroups instead of 81.
#define CT_SEG 8
__global__ __launch_bounds__(64) void k_contract_part(int bsz, int cnt, int m0, const int* __restrict__ bsv,
                                                      const int* __restrict__ nbv, const double* __restrict__ x,
                                                      const double2* __restrict__ yc, double2* __restrict__ part)
{
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= cnt) return;
    const int bs = bsv[m0 + t], nb = nbv[m0 + t];
    const int k0 = (int)((int64_t)nb * blockIdx.y / CT_SEG), k1 = (int)((int64_t)nb * (blockIdx.y + 1) / CT_SEG);
    double2 a0 = cmk(0, 0), a1 = cmk(0, 0);
    int k = k0;
#pragma unroll 4
    for (; k + 1 < k1; k += 2) {
        const double v0 = x[(int64_t)(bs + k) * bsz + t], v1 = x[(int64_t)(bs + k + 1) * bsz + t];
        const double2 y0 = yc[bs + k], y1 = yc[bs + k + 1];
        a0.x += v0 * y0.x;
        a0.y -= v0 * y0.y;
        a1.x += v1 * y1.x;
        a1.y -= v1 * y1.y;
    }
    if (k < k1) {
        const double v0 = x[(int64_t)(bs + k) * bsz + t];
        const double2 y0 = yc[bs + k];
        a0.x += v0 * y0.x;
        a0.y -= v0 * y0.y;
    }
    part[(int64_t)blockIdx.y * cnt + t] = cadd(a0, a1);
}

__global__ __launch_bounds__(256) void k_contract_sum(int cnt, int m0, const double2* __restrict__ part,
                                                      double2* __restrict__ f)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= cnt) return;
    double2 acc = part[t];
#pragma unroll
    for (int s = 1; s < CT_SEG; ++s) acc = cadd(acc, part[(int64_t)s * cnt + t]);
    f[m0 + t] = acc;
}

extern "C" hipError_t bse_launch_bisect(hipStream_t st, int nitems, const int4* items,
                                        const double2* bounds, const double* d, const double* e2,
                                        double pivmin, double tnorm, double* lam)
{
    // Fewer lanes per eigenvalue do less redundant Sturm work (W lanes run
    // W * log_{W+1}(2/eps) counts) at the price of more rounds; once the
    // items alone give a wave per CU, W = 2 is fastest (eigenvalue stage at
    // order 20736: 32.7 ms against 37.2 ms for W = 4 and 38.2 ms for W = 8).
    const int W = nitems >= 8192 ? 2 : nitems >= 4096 ? 4 : 8;
    const int64_t threads = (int64_t)nitems * W;
    if (W == 2)
        hipLaunchKernelGGL(k_bisect<2>, dim3((unsigned)ceil_div64(threads, 256)), dim3(256), 0, st,
                           nitems, items, bounds, d, e2, pivmin, tnorm, (const double*)nullptr, lam);
    else if (W == 4)
        hipLaunchKernelGGL(k_bisect<4>, dim3((unsigned)ceil_div64(threads, 256)), dim3(256), 0, st,
                           nitems, items, bounds, d, e2, pivmin, tnorm, (const double*)nullptr, lam);
    else
        hipLaunchKernelGGL(k_bisect<8>, dim3((unsigned)ceil_div64(threads, 256)), dim3(256), 0, st,
                           nitems, items, bounds, d, e2, pivmin, tnorm, (const double*)nullptr, lam);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_inverse_iter(hipStream_t st, int bsz, int cnt, int m0,
                                              const double* lam, const int* bsv, const int* nbv,
                                              const double* d, const double* e, double tnorm,
                                              double* scratch /* 5 * n * bsz */, int n,
                                              double* x)
{
    const int64_t A = (int64_t)n * bsz;
    hipLaunchKernelGGL(k_inverse_iter, dim3((unsigned)ceil_div64(cnt, 64)), dim3(64), 0, st, bsz,
                       cnt, m0, lam, bsv, nbv, d, e, tnorm, (const double*)nullptr, scratch,
                       scratch + A, scratch + 2 * A, scratch + 3 * A, scratch + 4 * A, x);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_cluster_mgs(hipStream_t st, int bsz, int ncl, const int2* clusters,
                                             int m0, const int* bsv, const int* nbv,
                                             const int* members, double* x)
{
    if (ncl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cluster_mgs, dim3(ncl), dim3(256), 0, st, bsz, clusters, (const int*)nullptr,
                       m0, bsv, nbv, members, x);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_commit(hipStream_t st, int n, int bsz, int cnt, int nstore, int m0,
                                        const int* bsv, const int* nbv, const double* x,
                                        double2* Z, int64_t ldz, int col0)
{
    if (nstore <= 0) return hipSuccess;
    dim3 grid((unsigned)ceil_div64(n, 64), (unsigned)ceil_div64(nstore, 64));
    hipLaunchKernelGGL(k_commit, grid, dim3(256), 0, st, n, bsz, cnt, nstore, m0, bsv, nbv, x, Z,
                       ldz, col0);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_contract(hipStream_t st, int bsz, int cnt, int m0, const int* bsv,
                                          const int* nbv, const double* x, const double2* yc,
                                          double2* f)
{
    hipLaunchKernelGGL(k_contract, dim3((unsigned)ceil_div64(cnt, 256)), dim3(256), 0, st, bsz, cnt,
                       m0, bsv, nbv, x, yc, f);
    return hipGetLastError();
}

// ws: at least 2 * CT_SEG * cnt doubles (the segment partials).
extern "C" hipError_t bse_launch_contract_split(hipStream_t st, int bsz, int cnt, int m0, const int* bsv,
                                                const int* nbv, const double* x, const double2* yc,
                                                double2* f, double* ws)
{
    if (cnt <= 0) return hipSuccess;
    double2* part = reinterpret_cast<double2*>(ws);
    hipLaunchKernelGGL(k_contract_part, dim3((unsigned)ceil_div64(cnt, 64), CT_SEG), dim3(64), 0, st, bsz, cnt,
                       m0, bsv, nbv, x, yc, part);
    hipLaunchKernelGGL(k_contract_sum, dim3((unsigned)ceil_div64(cnt, 256)), dim3(256), 0, st, cnt, m0,
                       (const double2*)part, f);
    return hipGetLastError();
}
extern "C" int64_t bse_contract_split_workspace(int cnt) { return (int64_t)2 * CT_SEG * cnt; }

// ---------------------------------------------------------------------------
// Device-only bookkeeping of the tridiagonal stage, for pipelines that run
// without host round trips (the batched sweep inside one graph).  Same rules
// as the host bookkeeping of the single-matrix path: split where
// e_i^2 <= eps^2 |d_i d_{i+1}| + safmin, eigenvalues of every unreduced block
// by bisection, ascending sort, clusters of eigenvalues closer than
// 1e-6 ||T|| inside one block.  One workgroup per matrix, n <= TRI_DEV_MAXN.
// ---------------------------------------------------------------------------
#define TRI_DEV_MAXN 4096

__device__ inline double block_max(double v, double* sh /* >= 16 */) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = sh[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) t = fmax(t, sh[i]);
    return t;
}

// Norm, pivot threshold, splits (e and e2 updated in place), one bisection
// item per row (block start, block size, local index, 0) and a common
// Gershgorin interval bounds[0].  prm <- (pivmin, tnorm).
__global__ __launch_bounds__(1024) void k_tri_prepare(int n, const double* __restrict__ d,
                                                      double* __restrict__ e,
                                                      double* __restrict__ e2,
                                                      int4* __restrict__ items,
                                                      double2* __restrict__ bounds,
                                                      double* __restrict__ prm)
{
    __shared__ double sh[32];
    __shared__ unsigned char s_split[TRI_DEV_MAXN];
    __shared__ int s_bs[TRI_DEV_MAXN];
    const int tid = threadIdx.x, bd = blockDim.x;
    double tn = 0.0, em = 0.0;
    for (int i = tid; i < n; i += bd) {
        double r = fabs(d[i]);
        if (i > 0) r += fabs(e[i - 1]);
        if (i < n - 1) {
            r += fabs(e[i]);
            em = fmax(em, e[i] * e[i]);
        }
        tn = fmax(tn, r);
    }
    tn = block_max(tn, sh);
    em = block_max(em, sh);
    const double tnorm = (tn == 0.0) ? 1.0 : tn;
    const double pivmin = DBL_MIN * fmax(1.0, em);
    __syncthreads();
    for (int i = tid; i < n - 1; i += bd) {
        const double ee = e[i] * e[i];
        const bool sp = fabs(d[i] * d[i + 1]) * DBL_EPSILON * DBL_EPSILON + DBL_MIN >= ee;
        s_split[i] = sp ? 1 : 0;
        e2[i] = sp ? 0.0 : ee;
    }
    __syncthreads();
    double lo = DBL_MAX, hi = -DBL_MAX;
    for (int i = tid; i < n; i += bd) {
        double r = 0.0;
        if (i > 0 && !s_split[i - 1]) r += fabs(e[i - 1]);
        if (i < n - 1 && !s_split[i]) r += fabs(e[i]);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
    }
    hi = block_max(hi, sh);
    lo = -block_max(-lo, sh);
    __syncthreads();
    for (int i = tid; i < n - 1; i += bd)
        if (s_split[i]) e[i] = 0.0;
    if (tid == 0) {
        int bs = 0;
        for (int i = 0; i < n; ++i) {
            if (i > 0 && s_split[i - 1]) bs = i;
            s_bs[i] = bs;
        }
        const double fud = 2.1 * DBL_EPSILON * tnorm * n + 4.2 * pivmin;
        bounds[0] = make_double2(lo - fud, hi + fud);
        prm[0] = pivmin;
        prm[1] = tnorm;
    }
    __syncthreads();
    for (int i = tid; i < n; i += bd) {
        const int bs = s_bs[i];
        int be = i + 1;
        while (be < n && !s_split[be - 1]) ++be;
        items[i] = make_int4(bs, be - bs, i - bs, 0);
    }
}

// Ascending sort of the bisection results (ties by row), block data of each
// sorted eigenvalue, and the cluster list (offset into members, size) with
// its count.  members holds the sorted positions ordered by (block start,
// position), so each block's eigenvalues are contiguous and ascending there;
// a cluster is a run of them closer than the tolerance.
__global__ __launch_bounds__(1024) void k_tri_sort(int n, const double* __restrict__ lami,
                                                   const int4* __restrict__ items,
                                                   const double* __restrict__ prm,
                                                   double* __restrict__ lam, int* __restrict__ bsv,
                                                   int* __restrict__ nbv, int2* __restrict__ clusters,
                                                   int* __restrict__ members, int* __restrict__ ncl)
{
    __shared__ double key[TRI_DEV_MAXN];
    __shared__ int idx[TRI_DEV_MAXN];
    __shared__ int s_ncl;
    const int tid = threadIdx.x, bd = blockDim.x;
    int P = 1;
    while (P < n) P <<= 1;
    for (int i = tid; i < P; i += bd) {
        key[i] = (i < n) ? lami[i] : DBL_MAX;
        idx[i] = i;
    }
    if (tid == 0) s_ncl = 0;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < P; i += bd) {
                const int l = i ^ j;
                if (l > i) {
                    const bool asc = (i & k) == 0;
                    const bool gt = key[i] > key[l] || (key[i] == key[l] && idx[i] > idx[l]);
                    if (gt == asc) {
                        const double tk = key[i]; key[i] = key[l]; key[l] = tk;
                        const int ti = idx[i]; idx[i] = idx[l]; idx[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    int myb[TRI_DEV_MAXN / 1024];
    for (int i = tid, q = 0; i < n; i += bd, ++q) {
        const int4 it = items[idx[i]];
        lam[i] = key[i];
        bsv[i] = it.x;
        nbv[i] = it.y;
        myb[q] = it.x;
    }
    __syncthreads();
    // second sort: positions by (block start, position); the block start is
    // below n <= 4096, so the composite key fits an int
    __syncthreads();
    for (int i = tid, q = 0; i < P; i += bd, ++q)
        idx[i] = (i < n) ? myb[q] * TRI_DEV_MAXN + i : INT_MAX;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < P; i += bd) {
                const int l = i ^ j;
                if (l > i) {
                    const bool asc = (i & k) == 0;
                    if ((idx[i] > idx[l]) == asc) {
                        const int ti = idx[i]; idx[i] = idx[l]; idx[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int q = tid; q < n; q += bd) members[q] = idx[q] % TRI_DEV_MAXN;
    __syncthreads();
    const double tolc = 1e-6 * prm[1];
    for (int q = tid; q < n; q += bd) {
        const int pq = idx[q] % TRI_DEV_MAXN, bq = idx[q] / TRI_DEV_MAXN;
        const bool start = (q == 0) || idx[q - 1] / TRI_DEV_MAXN != bq ||
                           key[pq] - key[idx[q - 1] % TRI_DEV_MAXN] > tolc;
        if (!start) continue;
        int j = q + 1;
        while (j < n && idx[j] / TRI_DEV_MAXN == bq &&
               key[idx[j] % TRI_DEV_MAXN] - key[idx[j - 1] % TRI_DEV_MAXN] <= tolc)
            ++j;
        if (j - q > 1) {
            const int slot = atomicAdd(&s_ncl, 1);
            clusters[slot] = make_int2(q, j - q);
        }
    }
    __syncthreads();
    if (tid == 0) *ncl = s_ncl;
}

extern "C" hipError_t bse_launch_tri_prepare(hipStream_t st, int n, const double* d, double* e,
                                             double* e2, int4* items, double2* bounds, double* prm)
{
    if (n < 1 || n > TRI_DEV_MAXN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_tri_prepare, dim3(1), dim3(1024), 0, st, n, d, e, e2, items, bounds, prm);
    return hipGetLastError();
}

extern "C" hipError_t bse_launch_tri_sort(hipStream_t st, int n, const double* lami,
                                          const int4* items, const double* prm, double* lam,
                                          int* bsv, int* nbv, int2* clusters, int* members, int* ncl)
{
    if (n < 1 || n > TRI_DEV_MAXN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_tri_sort, dim3(1), dim3(1024), 0, st, n, lami, items, prm, lam, bsv, nbv,
                       clusters, members, ncl);
    return hipGetLastError();
}

// Device-parameter forms of the bisection, inverse iteration and cluster
// orthogonalisation (all n eigenpairs of one matrix in one batch of n).
extern "C" hipError_t bse_launch_tri_vectors_dev(hipStream_t st, int n, const int4* items,
                                                 const double2* bounds, const double* d,
                                                 const double* e, const double* e2,
                                                 const double* prm, double* lami,
                                                 const double* lam, const int* bsv, const int* nbv,
                                                 const int2* clusters, const int* members,
                                                 const int* ncl, double* scratch /* 5 n^2 */,
                                                 double* x /* n^2 */, int stage)
{
    if (stage == 0) {
        const int64_t threads = (int64_t)n * 8;
        hipLaunchKernelGGL(k_bisect<8>, dim3((unsigned)ceil_div64(threads, 256)), dim3(256), 0, st, n,
                           items, bounds, d, e2, 0.0, 1.0, prm, lami);
    } else {
        const int64_t A = (int64_t)n * n;
        hipLaunchKernelGGL(k_inverse_iter, dim3((unsigned)ceil_div64(n, 64)), dim3(64), 0, st, n, n,
                           0, lam, bsv, nbv, d, e, 1.0, prm, scratch, scratch + A, scratch + 2 * A,
                           scratch + 3 * A, scratch + 4 * A, x);
        const int maxcl = (n + 1) / 2;
        hipLaunchKernelGGL(k_cluster_mgs, dim3((unsigned)maxcl), dim3(256), 0, st, n, clusters, ncl,
                           0, bsv, nbv, members, x);
    }
    return hipGetLastError();
}
