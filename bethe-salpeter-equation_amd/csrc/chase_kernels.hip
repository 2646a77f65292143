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
#define CH_W 8                   // sweeps (waves) per workgroup
#define CHASE_SPIN_LIMIT (1u << 26)

__device__ inline double2 wave_sum_c(double2 v) { return cmk(wave_sum(v.x), wave_sum(v.y)); }

__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
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
__device__ inline bool chase_wait(const ChaseArgs& a, int s, int need, int lane, bool local) {
    int ok = 1;
    if (lane == 0) {
        unsigned spins = 0;
        for (;;) {
            const int got = local ? __hip_atomic_load(a.prog + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                                  : __hip_atomic_load(a.prog + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (got >= need) break;
            if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = 0; break; }
            if (++spins > CHASE_SPIN_LIMIT) {
                __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    ok = __shfl(ok, 0, 64);
    if (local) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok != 0;
}

__device__ inline void chase_post(const ChaseArgs& a, int s, int val, int lane, bool local) {
    if (local) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_store(a.prog + s, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (lane == 0) __hip_atomic_store(a.prog + s, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ __launch_bounds__(64 * CH_W) void k_chase(ChaseArgs a)
{
    __shared__ double2 sBlk[CH_W][CH_B][CH_B + 1];
    __shared__ double2 sV[CH_W][CH_B];
    __shared__ double2 sU[CH_W][CH_B];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = a.n, b = a.b, ldab = a.ldab;
    double2 (*X)[CH_B + 1] = sBlk[w];
    double2* sv = sV[w];
    double2* su = sU[w];
    if (blockIdx.x == 0 && threadIdx.x == 0) a.d[0] = a.AB[0].x;
    for (int chunk = blockIdx.x; chunk < a.nchunks; chunk += gridDim.x) {
        const int s = chunk * CH_W + w;
        if (s > n - 2) break;
        const int ntask = chase_ntask(n, b, s);
        const int nprev = (s > 0) ? chase_ntask(n, b, s - 1) : 0;
        double2 tau = cmk(0, 0);
        bool alive = true;
        for (int j = 0; j < ntask && alive; ++j) {
            if (s > 0 && !chase_wait(a, s - 1, min(j + 3, nprev), lane, w > 0)) { alive = false; break; }
            if ((j & 1) == 0) {
                // ---- D_k: (k = 0: reflector from column s) two-sided update
                const int k = j >> 1;
                const int lo = s + 1 + k * b, hi = min(lo + b - 1, n - 1), L = hi - lo + 1;
                if (k == 0) {
                    const double2 x = (lane < L) ? a.AB[(int64_t)s * ldab + 1 + lane] : cmk(0, 0);
                    const double ss = wave_sum((lane >= 1 && lane < L) ? cabs2(x) : 0.0);
                    const double2 alpha = cmk(__shfl(x.x, 0, 64), __shfl(x.y, 0, 64));
                    const Refl r = make_refl(alpha, ss);
                    tau = r.tau;
                    const double2 v = (lane == 0) ? cmk(1.0, 0.0) : cmul(x, r.scal);
                    if (lane < L) {
                        sv[lane] = v;
                        a.AB[(int64_t)s * ldab + 1 + lane] = (lane == 0) ? cmk(r.beta, 0.0) : cmk(0, 0);
                    }
                    if (lane == 0) a.e[s] = r.beta;
                    if (a.y) {
                        const double2 yy = (lane < L) ? a.y[lo + lane] : cmk(0, 0);
                        const double2 dot = wave_sum_c((lane < L) ? cmul(cconj(v), yy) : cmk(0, 0));
                        if (lane < L) a.y[lo + lane] = csub(yy, cmul(cmul(cconj(tau), v), dot));
                    }
                    wave_lds_sync();
                }
                // stage the lower triangle of A(D_k, D_k): all loads are
                // issued before the first one is waited for
                {
                    double2 tmp[CH_B];
#pragma unroll
                    for (int c = 0; c < CH_B; ++c)
                        tmp[c] = (c < L && lane >= c && lane < L) ? a.AB[(int64_t)(lo + c) * ldab + (lane - c)]
                                                                  : cmk(0, 0);
#pragma unroll
                    for (int c = 0; c < CH_B; ++c)
                        if (c < L && lane >= c && lane < L) X[lane][c] = tmp[c];
                }
                wave_lds_sync();
                double2 wv = cmk(0, 0);
                if (lane < L) {
#pragma unroll
                    for (int c = 0; c < CH_B; ++c) {
                        if (c < L) {
                            const double2 h = (lane >= c) ? X[lane][c] : cconj(X[c][lane]);
                            cfma(wv, h, sv[c]);
                        }
                    }
                }
                const double2 v = (lane < L) ? sv[lane] : cmk(0, 0);
                const double gam = wave_sum((lane < L) ? (v.x * wv.x + v.y * wv.y) : 0.0);
                const double t2 = cabs2(tau);
                const double2 u = csub(cmul(tau, wv), cscale(v, 0.5 * t2 * gam));
                if (lane < L) su[lane] = u;
                wave_lds_sync();
                for (int c = 0; c < L; ++c) {
                    if (lane >= c && lane < L) {
                        double2 h = X[lane][c];
                        h = csub(h, cadd(cmul(u, cconj(sv[c])), cmul(v, cconj(su[c]))));
                        if (lane == c) h.y = 0.0;
                        a.AB[(int64_t)(lo + c) * ldab + (lane - c)] = h;
                        if (k == 0 && lane == 0) a.d[lo] = h.x;
                    }
                }
                wave_lds_sync();
            } else {
                // ---- X_k: right product with H_{k-1}, new reflector, left product
                const int k = (j + 1) >> 1;
                const int lo = s + 1 + k * b, hi = min(lo + b - 1, n - 1), Lr = hi - lo + 1;
                const int plo = lo - b;
                {
                    double2 tmp[CH_B];
#pragma unroll
                    for (int c = 0; c < CH_B; ++c)
                        tmp[c] = (c < b && lane < Lr) ? a.AB[(int64_t)(plo + c) * ldab + (b + lane - c)]
                                                      : cmk(0, 0);
#pragma unroll
                    for (int c = 0; c < CH_B; ++c)
                        if (c < b && lane < Lr) X[lane][c] = tmp[c];
                }
                wave_lds_sync();
                // (1) X <- X H_{k-1}:  X -= tau (X v) v^H
                if (lane < Lr) {
                    double2 t = cmk(0, 0);
                    for (int c = 0; c < b; ++c) cfma(t, X[lane][c], sv[c]);
                    const double2 tt = cmul(tau, t);
                    for (int c = 0; c < b; ++c) X[lane][c] = csub(X[lane][c], cmul(tt, cconj(sv[c])));
                }
                wave_lds_sync();
                // (2) reflector from column 0
                const double2 x = (lane < Lr) ? X[lane][0] : cmk(0, 0);
                const double ss = wave_sum((lane >= 1 && lane < Lr) ? cabs2(x) : 0.0);
                const double2 alpha = cmk(__shfl(x.x, 0, 64), __shfl(x.y, 0, 64));
                const Refl r = make_refl(alpha, ss);
                const double2 v = (lane == 0) ? cmk(1.0, 0.0) : cmul(x, r.scal);
                if (lane < Lr) sv[lane] = v;
                wave_lds_sync();
                // (3) X(:, 1:) <- H^H X(:, 1:):  s_c = v^H X(:, c)
                if (lane >= 1 && lane < b) {
                    double2 sc = cmk(0, 0);
                    for (int i = 0; i < Lr; ++i) cfmac(sc, sv[i], X[i][lane]);
                    su[lane] = sc;
                }
                wave_lds_sync();
                const double2 ctau = cconj(r.tau);
                if (lane < Lr) {
                    const double2 f = cmul(ctau, v);
                    for (int c = 1; c < b; ++c) X[lane][c] = csub(X[lane][c], cmul(f, su[c]));
                    X[lane][0] = (lane == 0) ? cmk(r.beta, 0.0) : cmk(0, 0);
                }
                wave_lds_sync();
                for (int c = 0; c < b; ++c)
                    if (lane < Lr) a.AB[(int64_t)(plo + c) * ldab + (b + lane - c)] = X[lane][c];
                if (a.y) {
                    const double2 yy = (lane < Lr) ? a.y[lo + lane] : cmk(0, 0);
                    const double2 dot = wave_sum_c((lane < Lr) ? cmul(cconj(v), yy) : cmk(0, 0));
                    if (lane < Lr) a.y[lo + lane] = csub(yy, cmul(cmul(ctau, v), dot));
                }
                tau = r.tau;
                wave_lds_sync();
            }
            chase_post(a, s, j + 1, lane, w < CH_W - 1);
        }
        if (!alive) break;
    }
}

extern "C" int64_t bse_chase_counters(int n) { return n > 1 ? n - 1 : 1; }

// prog (n-1 ints) and err (1 int) must be zero.  grid: workgroups.
extern "C" hipError_t bse_launch_chase(hipStream_t st, int n, int b, double2* AB, int ldab, double2* y,
                                       double* d, double* e, int* prog, int* err, int grid)
{
    if (b < 1 || b > CH_B || ldab < 2 * b + 1) return hipErrorInvalidValue;
    if (n < 2) {
        hipLaunchKernelGGL(k_chase, dim3(1), dim3(64 * CH_W), 0, st,
                           ChaseArgs{n, b, ldab, 0, AB, y, d, e, prog, err});
        return hipGetLastError();
    }
    const int nchunks = (int)ceil_div64(n - 1, CH_W);
    const int g = std::max(1, std::min(grid, nchunks));
    hipLaunchKernelGGL(k_chase, dim3((unsigned)g), dim3(64 * CH_W), 0, st,
                       ChaseArgs{n, b, ldab, nchunks, AB, y, d, e, prog, err});
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Inverse iteration on the Hermitian band matrix B (band store AB0, the copy
// taken before the chase) for eigenvalue lam[m] (one workgroup each):
// LU with partial pivoting of B - lam I in a sliding LDS window of rows
// j .. j+b and columns j .. j+2b (circular slots), U rows (2b + 1 entries)
// and the b multipliers of each step to global memory, then three solves
// from a deterministic pseudo-random start, each normalised.  A pivot below
// eps ||T|| is replaced by eps ||T|| (the shift is an eigenvalue).
// ---------------------------------------------------------------------------
__device__ inline double bii_hash(uint64_t a, uint64_t b) {
    uint64_t z = a * 0x9E3779B97F4A7C15ull + b * 0xD1B54A32D192ED03ull + 0x94D049BB133111EBull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// B(r, c) - lam delta_rc for 0 <= r, c < n (zero outside the band)
__device__ inline double2 band_elem(const double2* AB0, int ldab, int n, int b, int r, int c, double lam) {
    if (r >= n || c >= n || r < 0 || c < 0) return cmk(0, 0);
    const int dd = r - c;
    if (dd > b || dd < -b) return cmk(0, 0);
    double2 x = (dd >= 0) ? AB0[(int64_t)c * ldab + dd] : cconj(AB0[(int64_t)r * ldab - dd]);
    if (dd == 0) x = cmk(x.x - lam, 0.0);
    return x;
}

#define BII_W (2 * CH_B + 1)
__global__ __launch_bounds__(256) void k_band_ii(int n, int b, int ldab, const double2* __restrict__ AB0,
                                                 const double* __restrict__ lam, double tnorm,
                                                 double2* __restrict__ Ub, double2* __restrict__ Lm,
                                                 int* __restrict__ piv, double2* __restrict__ X, int64_t ldx)
{
    __shared__ double2 W[CH_B + 1][BII_W];
    __shared__ double2 lb[CH_B];
    __shared__ int s_p;
    const int m = blockIdx.x, tid = threadIdx.x;
    const double lm = lam[m];
    const int wb = 2 * b + 1;                    // U row length
    const int nr = b + 1;                        // window rows
    double2* U = Ub + (int64_t)m * n * wb;
    double2* Lmul = Lm + (int64_t)m * n * b;
    int* pv = piv + (int64_t)m * n;
    const double tiny = DBL_EPSILON * tnorm + DBL_MIN;
    auto rs = [&](int r) { return r % nr; };
    auto cs = [&](int c) { return c % wb; };
    // initial window: rows 0 .. b, columns 0 .. 2b
    for (int e = tid; e < nr * wb; e += 256) {
        const int r = e / wb, c = e % wb;
        W[rs(r)][cs(c)] = band_elem(AB0, ldab, n, b, r, c, lm);
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        const int rmax = min(j + b, n - 1);
        // prefetch of the row entering after this step: row j+b+1, columns j+1 .. j+2b+1
        const int rin = j + b + 1;
        double2 pre = cmk(0, 0);
        if (tid < wb && rin < n) pre = band_elem(AB0, ldab, n, b, rin, j + 1 + tid, lm);
        // 1. pivot among rows j .. rmax
        if (tid < 64) {
            double best = -1.0;
            int bi = j;
            if (j + tid <= rmax) { best = cabs2(W[rs(j + tid)][cs(j)]); bi = j + tid; }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (tid == 0) s_p = bi;
        }
        __syncthreads();
        const int p = s_p;
        // 2. swap rows j and p over columns j .. j+2b
        if (p != j && tid < wb) {
            const int c = cs(j + tid);
            const double2 t = W[rs(j)][c];
            W[rs(j)][c] = W[rs(p)][c];
            W[rs(p)][c] = t;
        }
        __syncthreads();
        // 3. pivot and multipliers
        double2 u = W[rs(j)][cs(j)];
        if (cabs2(u) < tiny * tiny) u = cmk(u.x >= 0.0 ? tiny : -tiny, 0.0);
        if (tid < b) {
            const int r = j + 1 + tid;
            const double2 l = (r <= rmax) ? cdiv(W[rs(r)][cs(j)], u) : cmk(0, 0);
            lb[tid] = l;
            Lmul[(int64_t)j * b + tid] = l;
        }
        __syncthreads();
        // 4. elimination of rows j+1 .. rmax over columns j+1 .. j+2b
        for (int e = tid; e < b * 2 * b; e += 256) {
            const int rr = e / (2 * b), cc = e % (2 * b);
            const int r = j + 1 + rr, c = j + 1 + cc;
            if (r <= rmax) W[rs(r)][cs(c)] = csub(W[rs(r)][cs(c)], cmul(lb[rr], W[rs(j)][cs(c)]));
        }
        // 5. U row j
        if (tid < wb) {
            const int c = j + tid;
            U[(int64_t)j * wb + tid] = (tid == 0) ? u : (c < n ? W[rs(j)][cs(c)] : cmk(0, 0));
        }
        if (tid == 0) pv[j] = p;
        __syncthreads();
        // 6. the next window: row j+b+1 replaces row j; column j+2b+1 replaces column j
        if (tid < wb) W[rs(rin)][cs(j + 1 + tid)] = pre;
        if (tid < b) W[rs(j + 1 + tid)][cs(j + 2 * b + 1)] = (j + 1 + tid < rin) ? cmk(0, 0) : W[rs(j + 1 + tid)][cs(j + 2 * b + 1)];
        __syncthreads();
    }
    // ---- solves (one wave): x <- U^-1 L^-1 P x, three times from a random start
    if (tid >= 64) return;
    const int lane = tid;
    double2* x = X + (int64_t)m * ldx;
    for (int i = lane; i < n; i += 64) x[i] = cmk(bii_hash(m, 2 * i), bii_hash(m, 2 * i + 1));
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    for (int it = 0; it < 3; ++it) {
        // forward: window x[j .. j+b] in lanes 0 .. b
        double2 xw = (lane <= b && lane < n) ? x[lane] : cmk(0, 0);
        for (int j = 0; j < n; ++j) {
            const int p = pv[j];
            const int lp = p - j;                 // 0 .. b
            // swap x[j] and x[p]
            const double2 xj = cmk(__shfl(xw.x, 0, 64), __shfl(xw.y, 0, 64));
            const double2 xp = cmk(__shfl(xw.x, lp, 64), __shfl(xw.y, lp, 64));
            if (lane == lp) xw = xj;
            if (lane == 0) xw = xp;
            // x[j+t] -= l_t x[j]  (t = 1 .. b)
            if (lane >= 1 && lane <= b && j + lane < n) xw = csub(xw, cmul(Lmul[(int64_t)j * b + lane - 1], xp));
            if (lane == 0) x[j] = xw;
            // shift the window down by one lane and load x[j+b+1]
            const double2 nx = cmk(__shfl_down(xw.x, 1, 64), __shfl_down(xw.y, 1, 64));
            xw = nx;
            if (lane == b) xw = (j + b + 1 < n) ? x[j + b + 1] : cmk(0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        // backward: lane t holds x[j+1+t] (t < 2b)
        const int bw = 2 * b;
        double2 xr = cmk(0, 0);                  // x[j + 1 + lane] for the current j
        double nrm = 0.0;
        for (int j = n - 1; j >= 0; --j) {
            const double2 uu = (lane < bw && j + 1 + lane < n) ? U[(int64_t)j * wb + 1 + lane] : cmk(0, 0);
            const double2 acc = wave_sum_c(cmul(uu, xr));
            const double2 xj = cdiv(csub(x[j], acc), U[(int64_t)j * wb]);
            nrm += cabs2(xj);
            if (lane == 0) x[j] = xj;
            // shift up: new lane t holds old lane t-1; lane 0 gets x[j]
            const double2 up = cmk(__shfl_up(xr.x, 1, 64), __shfl_up(xr.y, 1, 64));
            xr = (lane == 0) ? xj : up;
            if (lane >= bw) xr = cmk(0, 0);
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
    hipLaunchKernelGGL(k_band_ii, dim3((unsigned)nvec), dim3(256), 0, st, n, b, ldab, AB0, lam, tnorm, Ub, Lm,
                       piv, X, ldx);
    hipLaunchKernelGGL(k_mgs_columns, dim3(1), dim3(1024), 0, st, n, nvec, X, ldx);
    return hipGetLastError();
}
