// K5, two-stage form, second stage: Hermitian band (half-bandwidth b) ->
// real symmetric tridiagonal by bulge chasing, and eigenvectors of the band
// matrix by inverse iteration.
//
// Bulge chasing (the successive band reduction of Bischof, Lang & Sun, in
// the task form of the LAPACK hb2st kernels).  Sweep s (s = 0 .. n-2)
// annihilates column s below its first subdiagonal and chases the bulge this
// creates down the band.  With blocks D_k = [s+1+k b, s+(k+1) b] (cut at
// n-1), its tasks are, in order,
//   D_0        reflector H_0 from A(D_0, s) (H_0^H x = beta e_1, beta real:
//              the subdiagonal e_s), then A(D_0, D_0) <- H_0^H A H_0;
//   X_k (k>=1) A(D_k, D_{k-1}) <- A(D_k, D_{k-1}) H_{k-1}; reflector H_k from
//              its first column; A(D_k, D_{k-1}) <- H_k^H A(D_k, D_{k-1});
//   D_k (k>=1) A(D_k, D_k) <- H_k^H A(D_k, D_k) H_k.
// The rows of D_k have entries only in the column blocks D_{k-1}, D_k,
// D_{k+1}, so these three tasks cover each two-sided product exactly.
// Task j of sweep s+1 touches only entries that tasks <= j+2 of sweep s
// have finished with, and none that sweep s's tasks >= j+3 touch: sweep
// s+1 starts task j once sweep s has completed j+3 tasks.  (A numpy
// restatement of this schedule, run in an adversarial order, reproduces
// the sequential reduction; with a lag of 2 it does not.)  The same
// reflectors are applied to a vector y as they are generated, which turns
// y = Q1^H rho into Q^H rho for the oscillator strengths.
//
// Mapping: one wavefront runs one sweep at a time, its current block staged
// in LDS (row i <-> lane i, b <= 32).  A workgroup holds CH_W waves that run
// consecutive sweeps; workgroup g takes sweep chunks g, g + grid, ....  Each
// sweep publishes its completed task count; the next sweep polls it with
// bounded spins (workgroup-scope ordering inside a chunk, agent scope from
// one chunk to the next): a wave that waits too long (or sees the error
// flag) sets the flag and stops, so every wave exits.
// A waiter only ever waits on a lower sweep, and workgroups are dispatched
// in index order, so the lowest unfinished sweep always makes progress.
//
// Band storage: AB[c * ldab + d] = A(c + d, c), 0 <= d <= 2b (FP64).
#include "bse_common.hpp"
#include <cfloat>

#define CH_B 32                  // largest band half-width
#define CH_W 4                   // sweeps (waves) per workgroup
#define CHASE_SPIN_LIMIT (1u << 26)

__device__ inline double2 wave_sum_c(double2 v) { return cmk(wave_sum_dpp(v.x), wave_sum_dpp(v.y)); }

// Hand-over of LDS data between the lanes of one wave: an LDS-only fence
// (s_waitcnt lgkmcnt(0)), so global loads and stores already in flight stay
// in flight (a full workgroup fence would also wait for every outstanding
// global access).  The wave's global stores are ordered by chase_post's
// release fence before any other wave reads them.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
}

// Reflector (LAPACK zlarfg): from alpha (lane 0's x) and ss = sum |x_i|^2
// over i >= 1, H = I - tau v v^H with v_0 = 1, v_i = x_i scal, H^H x = beta e_1.
struct Refl { double2 tau, scal; double beta; };
__device__ inline Refl make_refl(double2 alpha, double ss) {
    Refl r;
    r.tau = cmk(0, 0);
    r.scal = cmk(0, 0);
    r.beta = alpha.x;
    if (!(ss == 0.0 && alpha.y == 0.0)) {
        const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + ss);
        r.beta = (alpha.x >= 0.0) ? -nrm : nrm;
        r.tau = cmk((r.beta - alpha.x) / r.beta, -alpha.y / r.beta);
        r.scal = cdiv(cmk(1.0, 0.0), cmk(alpha.x - r.beta, alpha.y));
    }
    return r;
}

struct ChaseArgs {
    int n, b, ldab, nchunks;
    double2* AB;
    double2* y;          // may be null
    double* d;           // n
    double* e;           // n - 1
    int* prog;           // n - 1 completed-task counters (zeroed)
    int* err;            // error flag (zeroed)
};

__device__ inline int chase_ntask(int n, int b, int s) { return 2 * ((n - 2 - s) / b + 1) - 1; }

// Sweep s waits for sweep s - 1.  Consecutive sweeps of one chunk run as the
// waves of one workgroup (one CU, one L1 and one L2), so their hand-over
// needs only workgroup-scope ordering (local = true); the first wave of a
// chunk waits on the last wave of the previous chunk, in another workgroup
// and possibly on another XCD, with agent-scope acquire / release (L2
// write-back and invalidation).  Agent scope is cumulative: what earlier
// waves of a workgroup wrote (ordered before its last wave's release) is
// written back by that release too.
// COH = 0: the band and y go through the caches; a hand-over inside a chunk
// is ordered at workgroup scope, one across chunks (another workgroup,
// possibly another XCD) at agent scope, whose release writes back the XCD's
// L2 and whose acquire invalidates it.
// COH = 1: every band and y access is device-coherent (sc1 loads and stores,
// which do not leave stale or dirty lines in a non-coherent L2), so a
// hand-over needs only the completion of the wave's stores (s_waitcnt
// vmcnt(0)) before the progress counter is written, and no cache
// maintenance at all.
// COH = 1 inside a chunk: the waves of one workgroup hand over through
// counters in LDS (lprog[w], wave w's completed tasks plus base = the
// workgroup's chunk iteration << 20, so the values only grow and a slot is
// never reset); the band data still goes through device-coherent memory,
// and the producer's stores have completed (vmcnt(0)) before its counter
// is written.  Global counters are polled with the error flag read only
// every 32 spins, so a poll costs one memory round trip.
template <int COH>
__device__ inline bool chase_wait(const ChaseArgs& a, int s, int need, int lane, bool local, int* lprog, int w,
                                  int base) {
    int ok = 1;
    if (lane == 0) {
        unsigned spins = 0;
        for (;;) {
            int got;
            if (COH && local) got = __hip_atomic_load(lprog + (w - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) - base;
            else if (local) got = __hip_atomic_load(a.prog + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            else got = __hip_atomic_load(a.prog + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (got >= need) break;
            ++spins;
            if ((spins & 31) == 0 && __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                ok = 0;
                break;
            }
            if (spins > CHASE_SPIN_LIMIT) {
                __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    ok = __shfl(ok, 0, 64);
    if (COH || local) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok != 0;
}

template <int COH>
__device__ inline void chase_post(const ChaseArgs& a, int s, int val, int lane, bool local, int* lprog, int w,
                                  int base) {
    if (COH) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0x0F70);      // vmcnt(0): the wave's sc1 stores are done
        if (lane == 0) {
            if (local) __hip_atomic_store(lprog + w, base + val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            else __hip_atomic_store(a.prog + s, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (local) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_store(a.prog + s, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane == 0) __hip_atomic_store(a.prog + s, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Band and y accesses through buffer descriptors, COH = 1 with the
// device-scope cache policy.  A masked-off element gets an offset past the
// descriptor's range, which the hardware turns into a zero load or a
// dropped store, so no access needs a branch.
#define CH_SC1 16
#define CH_OOB ((int)0xFFFFFFF0u)
template <int COH> struct ChaseMem {
    __amdgpu_buffer_rsrc_t rab, ry;
    __device__ ChaseMem(const ChaseArgs& a) {
        rab = bse_rsrc(a.AB, (int)((int64_t)a.n * a.ldab * 16));
        ry = bse_rsrc(a.y ? (void*)a.y : (void*)a.AB, a.y ? a.n * 16 : 0);
    }
    static constexpr int AUX = COH ? CH_SC1 : 0;
    // element index i < 2^27 (the launcher checks the band store's size)
    __device__ inline double2 ld(int i, bool ok = true) const {
        return bufld<double2>::ld<AUX>(rab, ok ? i * 16 : CH_OOB, 0);
    }
    __device__ inline void st(int i, double2 v, bool ok = true) const {
        const bse_d2v w = {v.x, v.y};
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(bse_i4v, w), rab, ok ? i * 16 : CH_OOB, 0, AUX);
    }
    __device__ inline double2 ldy(int i, bool ok = true) const {
        return bufld<double2>::ld<AUX>(ry, ok ? i * 16 : CH_OOB, 0);
    }
    __device__ inline void sty(int i, double2 v, bool ok = true) const {
        const bse_d2v w = {v.x, v.y};
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(bse_i4v, w), ry, ok ? i * 16 : CH_OOB, 0, AUX);
    }
};

template <int COH>
__global__ __launch_bounds__(64 * CH_W) void k_chase(ChaseArgs a)
{
    const ChaseMem<COH> mem(a);
    __shared__ double2 sBlk[CH_W][CH_B][CH_B + 1];
    __shared__ double2 sV[CH_W][CH_B];
    __shared__ double2 sU[CH_W][CH_B];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & (CH_B - 1);          // in-range LDS row / column index
    const int n = a.n, b = a.b, ldab = a.ldab;
    double2 (*X)[CH_B + 1] = sBlk[w];
    double2* sv = sV[w];
    double2* su = sU[w];
    __shared__ int lprog[CH_W];
    if (threadIdx.x < CH_W) lprog[threadIdx.x] = 0;
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) a.d[0] = mem.ld(0).x;
    int iter = 0;
    for (int chunk = blockIdx.x; chunk < a.nchunks; chunk += gridDim.x, ++iter) {
        const int s = chunk * CH_W + w;
        const int base = (iter + 1) << 20;     // above any count of an earlier iteration
        if (s > n - 2) break;
        const int ntask = chase_ntask(n, b, s);
        const int nprev = (s > 0) ? chase_ntask(n, b, s - 1) : 0;
        double2 tau = cmk(0, 0);
        bool alive = true;
        for (int j = 0; j < ntask && alive; ++j) {
            if (s > 0 && !chase_wait<COH>(a, s - 1, min(j + 3, nprev), lane, w > 0, lprog, w, base)) { alive = false; break; }
            if ((j & 1) == 0) {
                // ---- D_k: (k = 0: reflector from column s) two-sided update
                const int k = j >> 1;
                const int lo = s + 1 + k * b, hi = min(lo + b - 1, n - 1), L = hi - lo + 1;
                const bool row = lane < L;
                // the lower triangle of A(D_k, D_k), row lane in registers;
                // every load (and for k = 0 the column s and the y segment)
                // is issued before the first is waited for
                // element (lo + lane, lo + c) of the band store: dbase + c (ldab - 1) + lane
                const int dbase = lo * ldab;
                double2 tmp[CH_B];
#pragma unroll
                for (int c = 0; c < CH_B; ++c)
                    tmp[c] = mem.ld(dbase + c * (ldab - 1) + lane, c < L && lane >= c && row);
                double2 yy = cmk(0, 0);
                if (k == 0) {
                    const double2 x = mem.ld(s * ldab + 1 + lane, row);
                    if (a.y) yy = mem.ldy(lo + lane, row);
                    const double ss = wave_sum_dpp((lane >= 1 && row) ? cabs2(x) : 0.0);
                    const double2 alpha = cmk(__shfl(x.x, 0, 64), __shfl(x.y, 0, 64));
                    const Refl r = make_refl(alpha, ss);
                    tau = r.tau;
                    const double2 v0 = (lane == 0) ? cmk(1.0, 0.0) : cmul(x, r.scal);
                    if (row) sv[lane] = v0;
                    mem.st(s * ldab + 1 + lane, (lane == 0) ? cmk(r.beta, 0.0) : cmk(0, 0), row);
                    if (lane == 0) a.e[s] = r.beta;
                    if (a.y) {
                        const double2 dot = wave_sum_c(row ? cmul(cconj(v0), yy) : cmk(0, 0));
                        mem.sty(lo + lane, csub(yy, cmul(cmul(cconj(tau), v0), dot)), row);
                    }
                }
                if (lane < CH_B) {
#pragma unroll
                    for (int c = 0; c < CH_B; ++c) X[lane][c] = tmp[c];
                }
                wave_lds_sync();
                // w = A v over the Hermitian block: A(i, c) = tmp[c] for c <= i,
                // conj(X[c][i]) above; LDS reads in batches of 8
                double2 wv = cmk(0, 0);
#pragma unroll
                for (int c0 = 0; c0 < CH_B; c0 += 8) {
                    double2 xc[8], vc[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) { xc[q] = X[c0 + q][li]; vc[q] = sv[c0 + q]; }
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int c = c0 + q;
                        const double2 h = (lane >= c) ? tmp[c] : cconj(xc[q]);
                        const double2 vq = (c < L) ? vc[q] : cmk(0, 0);
                        cfma(wv, h, vq);
                    }
                }
                if (!row) wv = cmk(0, 0);
                const double2 v = row ? sv[li] : cmk(0, 0);
                const double gam = wave_sum_dpp(v.x * wv.x + v.y * wv.y);
                const double t2 = cabs2(tau);
                const double2 u = csub(cmul(tau, wv), cscale(v, 0.5 * t2 * gam));
                if (row) su[lane] = u;
                wave_lds_sync();
                // A <- A - u v^H - v u^H on the lower triangle
#pragma unroll
                for (int c0 = 0; c0 < CH_B; c0 += 8) {
                    double2 vc[8], uc[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) { vc[q] = sv[c0 + q]; uc[q] = su[c0 + q]; }
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int c = c0 + q;
                        double2 h = csub(tmp[c], cadd(cmul(u, cconj(vc[q])), cmul(v, cconj(uc[q]))));
                        if (lane == c) h.y = 0.0;
                        if (c == 0 && k == 0 && lane == 0) a.d[lo] = h.x;
                        mem.st(dbase + c * (ldab - 1) + lane, h, c < L && lane >= c && row);
                    }
                }
                wave_lds_sync();
            } else {
                // ---- X_k: right product with H_{k-1}, new reflector, left
                // product; row lane of A(D_k, D_{k-1}) lives in registers
                const int k = (j + 1) >> 1;
                const int lo = s + 1 + k * b, hi = min(lo + b - 1, n - 1), Lr = hi - lo + 1;
                const int plo = lo - b;
                const bool row = lane < Lr;
                // element (lo + lane, plo + c): xbase + c (ldab - 1) + lane
                const int xbase = plo * ldab + b;
                double2 xr[CH_B];
#pragma unroll
                for (int c = 0; c < CH_B; ++c) xr[c] = mem.ld(xbase + c * (ldab - 1) + lane, c < b && row);
                const double2 yy = a.y ? mem.ldy(lo + lane, row) : cmk(0, 0);
                // (1) X <- X H_{k-1}:  X -= tau (X v) v^H
                double2 t = cmk(0, 0);
#pragma unroll
                for (int c0 = 0; c0 < CH_B; c0 += 8) {
                    double2 vc[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) vc[q] = sv[c0 + q];
#pragma unroll
                    for (int q = 0; q < 8; ++q) cfma(t, xr[c0 + q], (c0 + q < b) ? vc[q] : cmk(0, 0));
                }
                const double2 tt = cmul(tau, t);
#pragma unroll
                for (int c0 = 0; c0 < CH_B; c0 += 8) {
                    double2 vc[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) vc[q] = sv[c0 + q];
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (c0 + q < b) xr[c0 + q] = csub(xr[c0 + q], cmul(tt, cconj(vc[q])));
                }
                wave_lds_sync();     // every lane is done with H_{k-1}'s vector
                // (2) reflector from column 0
                const double2 x = row ? xr[0] : cmk(0, 0);
                const double ss = wave_sum_dpp((lane >= 1 && row) ? cabs2(x) : 0.0);
                const double2 alpha = cmk(__shfl(x.x, 0, 64), __shfl(x.y, 0, 64));
                const Refl r = make_refl(alpha, ss);
                const double2 v = (lane == 0) ? cmk(1.0, 0.0) : (row ? cmul(x, r.scal) : cmk(0, 0));
                if (lane < CH_B) {
                    sv[lane] = v;
#pragma unroll
                    for (int c = 0; c < CH_B; ++c) X[lane][c] = xr[c];
                }
                wave_lds_sync();
                // (3) X(:, 1:) <- H^H X(:, 1:):  s_c = v^H X(:, c)  (lane = column c)
                double2 sc = cmk(0, 0);
#pragma unroll
                for (int i0 = 0; i0 < CH_B; i0 += 8) {
                    double2 xi[8], vi[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) { xi[q] = X[i0 + q][li]; vi[q] = sv[i0 + q]; }
#pragma unroll
                    for (int q = 0; q < 8; ++q) cfmac(sc, vi[q], xi[q]);
                }
                if (lane < CH_B) su[lane] = sc;
                wave_lds_sync();
                const double2 ctau = cconj(r.tau);
                const double2 f = cmul(ctau, v);
#pragma unroll
                for (int c0 = 0; c0 < CH_B; c0 += 8) {
                    double2 uc[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) uc[q] = su[c0 + q];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int c = c0 + q;
                        double2 h = (c == 0) ? ((lane == 0) ? cmk(r.beta, 0.0) : cmk(0, 0)) : csub(xr[c], cmul(f, uc[q]));
                        mem.st(xbase + c * (ldab - 1) + lane, h, c < b && row);
                    }
                }
                if (a.y) {
                    const double2 dot = wave_sum_c(row ? cmul(cconj(v), yy) : cmk(0, 0));
                    mem.sty(lo + lane, csub(yy, cmul(cmul(ctau, v), dot)), row);
                }
                tau = r.tau;
                wave_lds_sync();
            }
            chase_post<COH>(a, s, j + 1, lane, w < CH_W - 1, lprog, w, base);
        }
        if (!alive) break;
    }
}

extern "C" int64_t bse_chase_counters(int n) { return n > 1 ? n - 1 : 1; }

// prog (n-1 ints) and err (1 int) must be zero.  grid: workgroups.
// Device-coherent form (COH = 1) unless BSE_CHASE_COH=0.  The band store
// must stay below 2 GiB (32-bit buffer offsets; n < 2^21 for b = 32).
extern "C" hipError_t bse_launch_chase(hipStream_t st, int n, int b, double2* AB, int ldab, double2* y,
                                       double* d, double* e, int* prog, int* err, int grid)
{
    if (b < 1 || b > CH_B || ldab < 2 * b + 1) return hipErrorInvalidValue;
    const char* ce = getenv("BSE_CHASE_COH");
    // the band store is addressed with 32-bit byte offsets
    if ((int64_t)n * ldab * 16 >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    const bool coh = !(ce && ce[0] == '0');
    const int nchunks = (n < 2) ? 0 : (int)ceil_div64(n - 1, CH_W);
    const int g = (n < 2) ? 1 : std::max(1, std::min(grid, nchunks));
    const ChaseArgs args{n, b, ldab, nchunks, AB, y, d, e, prog, err};
    if (coh)
        hipLaunchKernelGGL(k_chase<1>, dim3((unsigned)g), dim3(64 * CH_W), 0, st, args);
    else
        hipLaunchKernelGGL(k_chase<0>, dim3((unsigned)g), dim3(64 * CH_W), 0, st, args);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Inverse iteration on the Hermitian band matrix B (band store AB0, the copy
// taken before the chase) for eigenvalue lam[m] (one workgroup each):
// LU with partial pivoting of B - lam I in a sliding LDS window of rows
// j .. j+b and columns j .. j+2b (circular slots), U rows (2b + 1 entries)
// and the b multipliers of each step to global memory, then two solves
// from a deterministic pseudo-random start (the first a backward solve
// only), each normalised.  A pivot below
// eps ||T|| is replaced by eps ||T|| (the shift is an eigenvalue).
// ---------------------------------------------------------------------------
__device__ inline double bii_hash(uint64_t a, uint64_t b) {
    uint64_t z = a * 0x9E3779B97F4A7C15ull + b * 0xD1B54A32D192ED03ull + 0x94D049BB133111EBull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

#define BII_W (2 * CH_B + 1)
#define BII_BLK 64          // steps whose inputs are staged in LDS at once (one per lane of the wave)
static_assert(BII_BLK == 64, "the staging loads one entering row per lane");

// B(R, C) - lam delta_RC without branches (0 outside the band and the
// matrix): a masked-off element reads past the descriptor's range.
__device__ inline double2 band_elem_nb(__amdgpu_buffer_rsrc_t rab, int ldab, int n, int b, int R, int C,
                                       double lam) {
    const bool lower = C <= R;
    const bool ok = R >= 0 && R < n && C >= 0 && C < n && (lower ? R - C : C - R) <= b;
    const int idx = lower ? C * ldab + (R - C) : R * ldab + (C - R);
    double2 x = bufld<double2>::ld<0>(rab, ok ? idx * 16 : CH_OOB, 0);
    if (!lower) x.y = -x.y;
    if (R == C) x = cmk(x.x - lam, 0.0);
    return x;
}

// One wave per eigenvalue.  LU: lane c updates window column j+1+c; the
// band rows entering over the next BII_BLK steps are staged in LDS in one
// batch of loads, so a step touches only LDS and registers (its U row,
// multipliers and pivot index are stored, never waited for).  Solves: the
// multipliers (forward) and U rows (backward) of BII_BLK steps are staged
// the same way.  The band store is addressed with 32-bit byte offsets
// (n (2b + 1) 16 < 2^31).
// BB > 0: the half-bandwidth as a compile-time constant (the band
// reduction's b = 32), so the per-element bounds of the elimination vanish;
// BB = 0: b at run time.
template <int BB>
__global__ __launch_bounds__(64) void k_band_ii(int n, int b_rt, int ldab, const double2* __restrict__ AB0,
                                                const double* __restrict__ lam, double tnorm,
                                                double2* __restrict__ Ub, double2* __restrict__ Lm,
                                                int* __restrict__ piv, double2* __restrict__ X, int64_t ldx)
{
    const int b = BB > 0 ? BB : b_rt;
    __shared__ double2 W[CH_B + 1][BII_W];
    __shared__ double2 buf[BII_BLK * BII_W];
    __shared__ double2 lb[CH_B];
    __shared__ double2 sx[BII_BLK];
    __shared__ int spv[BII_BLK];
    const int m = blockIdx.x, lane = threadIdx.x;
    const double lm = lam[m];
    const int wb = 2 * b + 1;                    // U row length
    const int nr = b + 1;                        // window rows
    double2* U = Ub + (int64_t)m * n * wb;
    double2* Lmul = Lm + (int64_t)m * n * b;
    int* pv = piv + (int64_t)m * n;
    const double tiny = DBL_EPSILON * tnorm + DBL_MIN;
    const __amdgpu_buffer_rsrc_t rab = bse_rsrc(AB0, (int)((int64_t)n * ldab * 16));
    // initial window: rows 0 .. b, columns 0 .. 2b
    for (int e = lane; e < nr * wb; e += 64) {
        const int r = e / wb, c = e % wb;
        W[r % nr][c % wb] = band_elem_nb(rab, ldab, n, b, r, c, lm);
    }
    // circular slots of row j (mod b + 1) and column j (mod 2b + 1), carried
    // from step to step; a slot offset by d < modulus wraps at most once
    int rj = 0, cj = 0;
    auto wrap = [](int v, int mod) { return v >= mod ? v - mod : v; };
    const bool act = lane < 2 * b;           // lane c: column j+1+c (lanes beyond the window idle)
    for (int j0 = 0; j0 < n; j0 += BII_BLK) {
        // band rows R = j0 + b + 1 + r entering at steps j0 + r: entry c
        // (0 .. 2b) is B(R, R - b + c); lane c loads entry c, lane 0 also 2b
        wave_lds_sync();
        for (int r0 = 0; r0 < BII_BLK; r0 += 16) {
            double2 t[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int R = j0 + b + 1 + r0 + r;
                t[r] = band_elem_nb(rab, ldab, n, b, R, R - b + lane, lm);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) buf[(r0 + r) * BII_W + lane] = t[r];
        }
        {
            // entry 2b of every entering row: one row per lane, all loads at once
            const int R = j0 + b + 1 + lane;
            buf[lane * BII_W + 2 * b] = band_elem_nb(rab, ldab, n, b, R, R + b, lm);
        }
        wave_lds_sync();
        const int jend = min(n, j0 + BII_BLK);
        for (int j = j0; j < jend; ++j) {
            const int rmax = min(j + b, n - 1);
            const int ccol = wrap(cj + 1 + lane, wb);   // slot of column j+1+lane (lanes < 2b)
            // 1. pivot among rows j .. rmax: the largest |.|^2 (wave maximum),
            // ties to the lower row (first set bit of the ballot)
            const bool cand = lane <= b && j + lane <= rmax;
            const double2 wj = W[wrap(rj + (lane <= b ? lane : 0), nr)][cj];
            const double a2 = cand ? cabs2(wj) : 0.0;
            const double amax = wave_max_dpp(a2);
            const int p = j + (int)__builtin_ctzll(__ballot(cand && a2 == amax));
            const int rp = wrap(rj + (p - j), nr);
            // 2. pivot and multipliers (column j is not touched by the swap)
            double2 uu = W[rp][cj];
            if (cabs2(uu) < tiny * tiny) uu = cmk(uu.x >= 0.0 ? tiny : -tiny, 0.0);
            const double2 uinv = crecip(uu);
            if (lane < b) {
                const int r = j + 1 + lane;
                double2 l = cmk(0, 0);
                if (r <= rmax) l = cmul((r == p) ? W[rj][cj] : W[wrap(rj + 1 + lane, nr)][cj], uinv);
                lb[lane] = l;
                Lmul[(int64_t)j * b + lane] = l;
            }
            // 3. swap rows j and p over columns j+1 .. j+2b (lane c: column
            // j+1+c), 4. U row j, from the swapped pivot row
            const int col = j + 1 + lane;
            double2 prv = act ? W[rj][ccol] : cmk(0, 0);
            if (p != j && act) {
                const double2 t = W[rp][ccol];
                W[rp][ccol] = prv;
                prv = t;
            }
            if (act) U[(int64_t)j * wb + 1 + lane] = (col < n) ? prv : cmk(0, 0);
            if (lane == 0) {
                U[(int64_t)j * wb] = uinv;           // the backward solve multiplies by it
                pv[j] = p;
            }
            wave_lds_sync();
            // 5. elimination of rows j+1 .. rmax in column j+1+lane: all
            // reads of the column issued, then the updates written
            // (rows past the matrix end have zero multipliers, so only the
            // uniform t < b decides; the reads of a batch precede its writes)
            if (act) {
#pragma unroll
                for (int t0 = 0; t0 < CH_B; t0 += 8) {
                    double2 wv8[8], lv8[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        wv8[q] = W[wrap(rj + 1 + t0 + q, nr)][ccol];
                        lv8[q] = lb[t0 + q];
                    }
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (t0 + q < b) W[wrap(rj + 1 + t0 + q, nr)][ccol] = csub(wv8[q], cmul(lv8[q], prv));
                }
            }
            // 6. the next window: row j+b+1 takes row j's slot (columns j+1 ..
            // j+2b+1); column j+2b+1 takes column j's slot, zero in rows j+1 .. j+b
            const int r = j - j0;
            if (act) W[rj][ccol] = buf[r * BII_W + lane];
            if (lane == 0) W[rj][cj] = buf[r * BII_W + 2 * b];
            if (lane < b) W[wrap(rj + 1 + lane, nr)][cj] = cmk(0, 0);
            rj = wrap(rj + 1, nr);
            cj = wrap(cj + 1, wb);
            wave_lds_sync();
        }
    }
    // ---- solves: x <- U^-1 L^-1 P x, twice.  The first starts from a
    // pseudo-random vector taken as L^-1 P x0 itself (x0 is as random as
    // it is), so it is a backward solve alone; the second is a full solve.
    double2* x = X + (int64_t)m * ldx;
    for (int i = lane; i < n; i += 64) x[i] = cmk(bii_hash(m, 2 * i), bii_hash(m, 2 * i + 1));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");   // the LU's stores and x are complete
    const int bw = 2 * b;
    for (int it = 0; it < 2; ++it) {
        // forward: window x[j .. j+b] in lanes 0 .. b; a block's pivots,
        // multipliers and entering x[j+b+1] staged in LDS
        double2 xw = (lane <= b && lane < n) ? x[lane] : cmk(0, 0);
        for (int j0 = 0; j0 < (it > 0 ? n : 0); j0 += BII_BLK) {
            const int jend = min(n, j0 + BII_BLK);
            wave_lds_sync();
            for (int q0 = 0; q0 < BII_BLK * CH_B / 64; q0 += 16) {
                double2 t[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int e = lane + 64 * (q0 + q), u = e / CH_B, tt = e % CH_B;
                    const int j = j0 + u;
                    t[q] = (j < n && tt < b) ? Lmul[(int64_t)j * b + tt] : cmk(0, 0);
                }
#pragma unroll
                for (int q = 0; q < 16; ++q) buf[lane + 64 * (q0 + q)] = t[q];
            }
            {
                const int jp = j0 + lane;
                spv[lane] = (jp < n) ? pv[jp] : jp;
                sx[lane] = (j0 + b + 1 + lane < n) ? x[j0 + b + 1 + lane] : cmk(0, 0);
            }
            wave_lds_sync();
            for (int j = j0; j < jend; ++j) {
                const int u = j - j0;
                const int lp = spv[u] - j;             // 0 .. b
                const double2 xj = cmk(__shfl(xw.x, 0, 64), __shfl(xw.y, 0, 64));
                const double2 xp = cmk(__shfl(xw.x, lp, 64), __shfl(xw.y, lp, 64));
                if (lane == lp) xw = xj;
                if (lane == 0) xw = xp;
                // x[j+t] -= l_t x[j]  (t = 1 .. b)
                const double2 l = buf[u * CH_B + ((lane - 1) & (CH_B - 1))];
                if (lane >= 1 && lane <= b && j + lane < n) xw = csub(xw, cmul(l, xp));
                if (lane == 0) x[j] = xw;
                // shift the window down by one lane; x[j+b+1] enters at lane b
                const double2 nx = cmk(__shfl_down(xw.x, 1, 64), __shfl_down(xw.y, 1, 64));
                xw = nx;
                if (lane == b) xw = sx[u];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        // backward: lane t holds x[j+1+t] (t < 2b); a block's U rows and x[j]
        // (final from the forward pass) staged in LDS, rows in descending order
        double2 xr = cmk(0, 0);
        double nrm = 0.0;
        for (int j0 = n - 1; j0 >= 0; j0 -= BII_BLK) {
            const int jlo = max(0, j0 - BII_BLK + 1);
            wave_lds_sync();
            for (int u0 = 0; u0 < BII_BLK; u0 += 16) {
                double2 t[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int j = j0 - u0 - u;
                    t[u] = (j >= jlo && lane < bw && j + 1 + lane < n) ? U[(int64_t)j * wb + 1 + lane] : cmk(0, 0);
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) buf[(u0 + u) * 64 + lane] = t[u];
            }
            {
                const int jd = j0 - lane;
                buf[BII_BLK * 64 + lane] = (jd >= jlo) ? U[(int64_t)jd * wb] : cmk(1, 0);   // 1 / U(j, j)
                sx[lane] = (jd >= jlo) ? x[jd] : cmk(0, 0);
            }
            wave_lds_sync();
            for (int j = j0; j >= jlo; --j) {
                const int u = j0 - j;
                const double2 acc = wave_sum_c(cmul(buf[u * 64 + lane], xr));
                const double2 xj = cmul(csub(sx[u], acc), buf[BII_BLK * 64 + u]);
                nrm += cabs2(xj);
                if (lane == 0) x[j] = xj;
                // shift up: new lane t holds old lane t-1; lane 0 gets x[j]
                const double2 up = cmk(__shfl_up(xr.x, 1, 64), __shfl_up(xr.y, 1, 64));
                xr = (lane == 0) ? xj : up;
                if (lane >= bw) xr = cmk(0, 0);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        const double inv = 1.0 / sqrt(nrm);
        for (int i = lane; i < n; i += 64) x[i] = cscale(x[i], inv);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
}

// Modified Gram-Schmidt (two passes) of the k columns of X (n x k, ld ldx) in
// order; the eigenvectors of clusters of close eigenvalues become
// orthonormal, well separated ones change at rounding level.  One workgroup.
__global__ __launch_bounds__(1024) void k_mgs_columns(int n, int k, double2* __restrict__ X, int64_t ldx)
{
    __shared__ double sh[64];
    for (int j = 0; j < k; ++j) {
        double2* xj = X + (int64_t)j * ldx;
        for (int pass = 0; pass < 2; ++pass) {
            for (int i = 0; i < j; ++i) {
                const double2* xi = X + (int64_t)i * ldx;
                double2 dot = cmk(0, 0);
                for (int r = threadIdx.x; r < n; r += blockDim.x) cfmac(dot, xi[r], xj[r]);
                dot = block_sum2(dot, sh);
                for (int r = threadIdx.x; r < n; r += blockDim.x) xj[r] = csub(xj[r], cmul(xi[r], dot));
                __syncthreads();
            }
        }
        double nr = 0.0;
        for (int r = threadIdx.x; r < n; r += blockDim.x) nr += cabs2(xj[r]);
        nr = block_sum(nr, sh);
        const double inv = 1.0 / sqrt(nr);
        for (int r = threadIdx.x; r < n; r += blockDim.x) xj[r] = cscale(xj[r], inv);
        __syncthreads();
    }
}

extern "C" int64_t bse_band_ii_workspace(int n, int b, int nvec) {
    return (int64_t)nvec * n * (2 * b + 1) + (int64_t)nvec * n * b + ((int64_t)nvec * n + 1) / 4 + 64;
}

// X (n x nvec, ld ldx) <- eigenvectors of the band matrix for lam[0 .. nvec).
extern "C" hipError_t bse_launch_band_ii(hipStream_t st, int n, int b, int ldab, const double2* AB0,
                                         const double* lam, int nvec, double tnorm, double2* ws,
                                         double2* X, int64_t ldx)
{
    if (b < 1 || b > CH_B || nvec < 1) return hipErrorInvalidValue;
    double2* Ub = ws;
    double2* Lm = Ub + (int64_t)nvec * n * (2 * b + 1);
    int* piv = reinterpret_cast<int*>(Lm + (int64_t)nvec * n * b);
    if (b == CH_B)
        hipLaunchKernelGGL(k_band_ii<CH_B>, dim3((unsigned)nvec), dim3(64), 0, st, n, b, ldab, AB0, lam, tnorm, Ub,
                           Lm, piv, X, ldx);
    else
        hipLaunchKernelGGL(k_band_ii<0>, dim3((unsigned)nvec), dim3(64), 0, st, n, b, ldab, AB0, lam, tnorm, Ub,
                           Lm, piv, X, ldx);
    hipLaunchKernelGGL(k_mgs_columns, dim3(1), dim3(1024), 0, st, n, nvec, X, ldx);
    return hipGetLastError();
}
