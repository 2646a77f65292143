// K11: batched sweep of many small, uneven excitonic problems in one graph.
//
// A parameter sweep (for example 64 honeycomb systems on k-grids from 12x12
// to 40x40, N = 144 .. 1600) is far too small per system to fill the chip,
// and the Householder reduction of one system is a chain of ~4 dependent
// launches per column.  Here every system gets its own slice of one device
// arena, the systems are ordered by decreasing N, and the reduction runs in
// lockstep: column c of all systems with c < N - 1 is one launch per step
// (blockIdx.z = system), so the launch count follows the largest system, not
// the sum.  Everything per system that is not a column step (TB states,
// tables, assembly before the reduction; tridiagonal eigenpairs, oscillator
// strengths, lowest vectors, FP64 refinement and spectrum after it) is
// recorded on side streams that fork from and join the lockstep chain, and
// the whole sequence is captured once into a hipGraph.  One bse_batch_run is
// one graph launch; the per-system results follow the same algorithms as
// bse_build + bse_eigh + bse_spectrum of the single-system path.
#include "bse_internal.hpp"

#include <numeric>

namespace {

// side streams for per-system work: the tails (tridiagonal eigenpairs, Q
// applications, refinement, spectrum) are latency-bound one-workgroup chains,
// so a graph replay with more branches overlaps more of them (64-system
// sweep: 3 streams 260 ms, 6 or 10 streams 213 ms, 16 to 64 streams 156 ms)
const int kSide = 16;
const int kRefineIters = 2;   // Rayleigh-Ritz rounds, as the single-system path

struct SysDev {
    double *rcart, *tau, *kpts, *gvec;
    double2* hr;
    double* eps;
    double2 *coef, *wtab, *vx, *P, *PV, *rho;
    double* dE;
    void* H;
    double *d, *e, *e2, *lami, *lam, *prm, *part, *scal;
    double2 *tauh, *Vst, *Tall, *Wp, *alpha2, *G, *acol, *Yr, *Yc, *t12;
    unsigned int* counter;
    int4* items;
    double2* bounds;
    int *bsv, *nbv, *ncl, *members;
    int2* clusters;
    double *iiscr, *iix;
    double2 *y, *f, *Z;
    double2 *RS, *RHS, *RX, *RG, *RHG, *RC, *skw;
    double* small;
    int64_t sk_el;
    double* eps_out;     // re (n_omega) | im (n_omega)
    double* evals_out;   // n_lowest
};

struct SysHost {
    Geom g;
    std::vector<double> rcart, hr, tau;
    int orig = 0;         // index in the caller's order
    int TK = 1;
};

}  // namespace

struct BatchState {
    int nsys = 0, precision = 0, fp32 = 0, nlow = 0, refine = 0, n_omega = 0;
    double eta = 0.0;
    std::vector<SysHost> hs;   // sorted by decreasing N
    std::vector<SysDev> ds;
    void* arena = nullptr;
    size_t arena_bytes = 0, arena_cap = 0;
    HetrdSys* d_hsys = nullptr;
    double* d_omega = nullptr;
    unsigned int* counters = nullptr;       // 16 per system, contiguous
    unsigned long long* stamps = nullptr;   // [0] start, [1 + s] end of system s (sorted order)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t side[kSide] = {};
    std::vector<hipEvent_t> evs;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int her2k_valu = 0;
    int ch_top = 8;
    int64_t nodes = 0;
    int wall_khz = 0;
    bool ran = false;
};

namespace {

void batch_release_graph(BatchState* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Bump allocator over one arena; pass 1 (base == nullptr) only counts.
struct Bump {
    char* base;
    size_t off = 0;
    template <typename T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? (T*)(base + off) : nullptr;
        off += std::max<size_t>(count, 1) * sizeof(T);
        return p;
    }
};

void carve(Bump& bp, const SysHost& h, int fp32, int nlow, int n_omega, SysDev& d,
           unsigned int* counter) {
    const Geom& g = h.g;
    const int no = g.norb, n = (int)g.N, nb = kPanel;
    const int64_t N = g.N;
    d.rcart = bp.take<double>(3 * (size_t)g.nr);
    d.hr = bp.take<double2>((size_t)g.nr * no * no);
    d.tau = bp.take<double>(3 * (size_t)no);
    d.kpts = bp.take<double>(3 * (size_t)g.nk);
    d.gvec = bp.take<double>(3 * (size_t)g.ng);
    d.eps = bp.take<double>((size_t)g.nk * g.ns);
    d.coef = bp.take<double2>((size_t)g.nk * g.ns * no);
    d.wtab = bp.take<double2>((size_t)ktable_size(g.sys.n1, g.sys.n2, g.n3) * no * no);
    d.vx = bp.take<double2>((size_t)no * no);
    d.P = bp.take<double2>((size_t)N * no);
    d.PV = bp.take<double2>((size_t)N * no);
    d.rho = bp.take<double2>(N);
    d.dE = bp.take<double>(N);
    d.H = bp.take<char>((size_t)tiled_elems(N) * (fp32 ? sizeof(float2) : sizeof(double2)));
    int64_t vs_el, yr_el, yc_el, part_d;
    bse_hetrd_workspace(n, nb, &vs_el, &yr_el, &yc_el, &part_d);
    const int npan = (int)ceil_div64(std::max(n - 1, 1), nb);
    d.d = bp.take<double>(N);
    d.e = bp.take<double>(N);
    d.tauh = bp.take<double2>(N);
    d.Vst = bp.take<double2>(vs_el);
    d.Tall = bp.take<double2>((size_t)npan * nb * nb);
    d.Wp = bp.take<double2>((size_t)N * nb);
    d.alpha2 = bp.take<double2>(64);
    d.G = bp.take<double2>(64 * 64);
    d.acol = bp.take<double2>(N);
    d.Yr = bp.take<double2>(yr_el);
    d.Yc = bp.take<double2>(yc_el);
    d.t12 = bp.take<double2>(128);
    d.part = bp.take<double>(part_d);
    d.scal = bp.take<double>(8);
    d.counter = counter;
    d.e2 = bp.take<double>(N);
    d.items = bp.take<int4>(N);
    d.bounds = bp.take<double2>(1);
    d.lami = bp.take<double>(N);
    d.lam = bp.take<double>(N);
    d.prm = bp.take<double>(8);
    d.bsv = bp.take<int>(N);
    d.nbv = bp.take<int>(N);
    d.clusters = bp.take<int2>(N);
    d.members = bp.take<int>(N);
    d.ncl = bp.take<int>(4);
    d.iiscr = bp.take<double>((size_t)5 * N * N);
    d.iix = bp.take<double>((size_t)N * N);
    d.y = bp.take<double2>(N);
    d.f = bp.take<double2>(N);
    d.Z = bp.take<double2>((size_t)N * nlow);
    const int m = 2 * nlow;
    d.RS = bp.take<double2>((size_t)N * m);
    d.RHS = bp.take<double2>((size_t)N * m);
    d.RX = bp.take<double2>((size_t)N * m);
    d.RG = bp.take<double2>((size_t)m * m);
    d.RHG = bp.take<double2>((size_t)m * m);
    d.RC = bp.take<double2>((size_t)m * nlow);
    d.small = bp.take<double>(4 * (size_t)nlow + 8);
    d.sk_el = (int64_t)64 * m * m;
    d.skw = bp.take<double2>(d.sk_el);
    d.eps_out = bp.take<double>(2 * (size_t)std::max(n_omega, 1));
    d.evals_out = bp.take<double>(nlow);
}

}  // namespace

#define BCK(expr)                                                                       \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            err = fail(ctx, BSE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
            goto capture_failed;                                                        \
        }                                                                               \
    } while (0)

// Per-system work before the reduction: TB states, W/V tables, transitions and
// the assembled matrix (same kernels and arguments as bse_build).
static hipError_t record_front(BatchState* b, int s, hipStream_t st) {
    const SysHost& h = b->hs[s];
    const SysDev& d = b->ds[s];
    const Geom& g = h.g;
    const bse_system& sy = g.sys;
    const int no = g.norb;
    hipError_t e;
    if ((e = bse_launch_tb_states(st, g.nk, no, g.nr, d.rcart, d.hr, d.kpts, sy.q, sy.n_occ, sy.nv,
                                  sy.nc, sy.scissor, nullptr, d.eps, d.coef, nullptr, nullptr)))
        return e;
    if ((e = bse_launch_w_table(st, sy.n1, sy.n2, g.n3, g.dim, no, g.ng, d.gvec, g.recip, d.tau, sy.r0,
                                sy.eps_env, sy.radius, kC, g.wbar, 0, d.wtab)))
        return e;
    if ((e = bse_launch_v_exchange(st, no, g.ng, g.dim, d.gvec, d.tau, kC, sy.radius, sy.q, d.vx))) return e;
    if ((e = bse_launch_transitions(st, g.nk, sy.nv, sy.nc, no, d.eps, d.coef, nullptr, nullptr,
                                    d.tau, g.qhat, d.vx, d.P, d.PV, d.rho, d.dE)))
        return e;
    const double alpha = -sy.direct_scale / (g.cell * g.nk);
    const double beta = sy.f_s * sy.exchange_scale / (g.cell * g.nk);
    if ((e = bse_launch_tile_pad_zero(st, b->fp32, g.N, d.H))) return e;
    if (bse_assemble_mfma_supported(sy.nv, sy.nc, no))
        return bse_launch_assemble_mfma(st, b->precision, g.nk, sy.n1, sy.n2, g.n3, sy.nv, sy.nc, no, g.N,
                                        d.coef, d.wtab, d.PV, d.P, d.dE, alpha, beta, d.H);
    return bse_launch_assemble(st, b->fp32, g.nk, sy.n1, sy.n2, g.n3, sy.nv, sy.nc, no, h.TK, g.N, d.coef,
                               d.coef, d.coef, d.coef, d.wtab, d.PV, d.P, d.dE, alpha, beta, d.H, 1);
}

// Matrix-free FP64 product with system s (as the single-system apply_H).
static hipError_t record_apply(BatchState* b, int s, hipStream_t st, int k, const double2* X,
                               double2* Y) {
    const SysHost& h = b->hs[s];
    const SysDev& d = b->ds[s];
    const Geom& g = h.g;
    const bse_system& sy = g.sys;
    const double alpha = -sy.direct_scale / (g.cell * g.nk);
    const double beta = sy.f_s * sy.exchange_scale / (g.cell * g.nk);
    hipError_t e;
    if ((e = bse_launch_zero_words(st, 2 * g.N * k, Y))) return e;
    return bse_launch_apply(st, g.nk, sy.n1, sy.n2, g.n3, sy.nv, sy.nc, g.norb,
                            bse_apply_pick_tk(sy.nv, sy.nc, g.norb), d.coef, d.coef, d.coef, d.coef,
                            d.wtab, d.PV, d.P, d.dE, alpha, beta, k, X, g.N, Y, g.N, nullptr, 0, nullptr);
}

// Per-system work after the reduction: tridiagonal eigenpairs, oscillator
// strengths of all excitons, the lowest n_lowest vectors (back-transformed
// and optionally refined in FP64), the spectrum and the completion stamp.
static hipError_t record_tail(BatchState* b, int s, hipStream_t st) {
    const SysHost& h = b->hs[s];
    const SysDev& d = b->ds[s];
    const Geom& g = h.g;
    const int n = (int)g.N, k = b->nlow, nb = kPanel;
    const int64_t N = g.N;
    hipError_t e;
    if ((e = bse_launch_last_diag(st, b->fp32, n, d.H, d.d))) return e;
    if ((e = bse_launch_tri_prepare(st, n, d.d, d.e, d.e2, d.items, d.bounds, d.prm))) return e;
    if ((e = bse_launch_tri_vectors_dev(st, n, d.items, d.bounds, d.d, d.e, d.e2, d.prm, d.lami,
                                        d.lam, d.bsv, d.nbv, d.clusters, d.members, d.ncl, d.iiscr,
                                        d.iix, 0)))
        return e;
    if ((e = bse_launch_tri_sort(st, n, d.lami, d.items, d.prm, d.lam, d.bsv, d.nbv, d.clusters,
                                 d.members, d.ncl)))
        return e;
    if ((e = bse_launch_tri_vectors_dev(st, n, d.items, d.bounds, d.d, d.e, d.e2, d.prm, d.lami,
                                        d.lam, d.bsv, d.nbv, d.clusters, d.members, d.ncl, d.iiscr,
                                        d.iix, 1)))
        return e;
    // oscillator strengths of all excitons: f_m = sum_i x_m[i] conj((Q^H rho)_i)
    if ((e = bse_launch_copy_words(st, 2 * N, d.rho, d.y))) return e;
    if ((e = bse_launch_apply_q(st, n, nb, d.Vst, d.Tall, 1, d.y, N, 1))) return e;
    if ((e = bse_launch_contract(st, n, n, 0, d.bsv, d.nbv, d.iix, d.y, d.f))) return e;
    // lowest k vectors: Z = Q x
    if ((e = bse_launch_commit(st, n, n, n, k, 0, d.bsv, d.nbv, d.iix, d.Z, N, 0))) return e;
    if ((e = bse_launch_apply_q(st, n, nb, d.Vst, d.Tall, k, d.Z, N, 0))) return e;
    if (b->refine) {
        const int m = 2 * k;
        double2 *S = d.RS, *HS = d.RHS, *T = d.RX;
        if ((e = bse_launch_copy_words(st, 2 * N * k, d.Z, S))) return e;
        if ((e = record_apply(b, s, st, k, S, HS))) return e;
        for (int it = 0; it < kRefineIters; ++it) {
            if ((e = bse_launch_residual(st, N, k, S, HS, N, S + N * k, d.small + k, nullptr))) return e;
            if ((e = record_apply(b, s, st, k, S + N * k, HS + N * k))) return e;
            if ((e = bse_zgemm(st, 1, 0, m, m, n, 1, 0, S, N, S, N, 0, 0, d.RG, m, d.skw, d.sk_el))) return e;
            if ((e = bse_zgemm(st, 1, 0, m, m, n, 1, 0, S, N, HS, N, 0, 0, d.RHG, m, d.skw, d.sk_el))) return e;
            if ((e = bse_launch_small_rr(st, m, k, d.RG, d.RHG, 1e-10, d.small, d.RC, (int*)(d.small + 2 * k))))
                return e;
            if ((e = bse_zgemm(st, 0, 0, n, k, m, 1, 0, S, N, d.RC, m, 0, 0, T, N, nullptr, 0))) return e;
            if ((e = bse_zgemm(st, 0, 0, n, k, m, 1, 0, HS, N, d.RC, m, 0, 0, T + N * k, N, nullptr, 0)))
                return e;
            if ((e = bse_launch_copy_words(st, 2 * N * k, T, S))) return e;
            if ((e = bse_launch_copy_words(st, 2 * N * k, T + N * k, HS))) return e;
        }
        if ((e = bse_launch_copy_words(st, 2 * N * k, S, d.Z))) return e;
        if ((e = bse_launch_copy_words(st, k, d.small, d.lam))) return e;
        if ((e = bse_zgemm(st, 1, 0, 1, k, n, 1, 0, d.rho, N, S, N, 0, 0, d.f, 1, d.skw, d.sk_el))) return e;
    }
    if (b->n_omega > 0) {
        const double pref = kC / (g.volume * g.nk);
        if ((e = bse_launch_dielectric(st, b->n_omega, b->d_omega, b->eta, n, d.lam, d.f, pref,
                                       d.eps_out, d.eps_out + b->n_omega)))
            return e;
    }
    if ((e = bse_launch_copy_words(st, k, d.lam, d.evals_out))) return e;
    return bse_launch_stamp(st, b->stamps + 1 + s);
}

// Time the two trailing-update kernels on the largest system's (not yet
// assembled) matrix and keep the faster one; tune the hemv table there too.
static int batch_autotune(bse_ctx* ctx, BatchState* b) {
    const SysDev& d = b->ds[0];
    const int n = (int)b->hs[0].g.N, nb = kPanel;
    const int64_t N = n;
    hipStream_t st = ctx->st;
    const size_t es = b->fp32 ? sizeof(float2) : sizeof(double2);
    HCK(hipMemsetAsync(d.H, 0, (size_t)tiled_elems(N) * es, st));
    HCK(hipMemsetAsync(d.Vst, 0, sizeof(double2) * (size_t)N * nb, st));
    HCK(hipMemsetAsync(d.Wp, 0, sizeof(double2) * (size_t)N * nb, st));
    HCK(hipMemsetAsync(d.alpha2, 0, sizeof(double2) * 64, st));
    int rc = autotune_hemv(ctx, b->fp32, n, d.H, d.acol, d.Yr, d.Yc);
    if (rc) return rc;
    b->her2k_valu = 0;
    if (ctx->autotune && n > nb) {
        float best = 0.f;
        for (int kind : {0, 1, 2}) {
            HCK(bse_launch_her2k(st, b->fp32, kind, n, nb, d.H, d.Vst, d.Wp, N, 0, nb, d.alpha2));
            HCK(hipEventRecord(ctx->tev[0], st));
            for (int r = 0; r < 5; ++r)
                HCK(bse_launch_her2k(st, b->fp32, kind, n, nb, d.H, d.Vst, d.Wp, N, 0, nb, d.alpha2));
            HCK(hipEventRecord(ctx->tev[1], st));
            HCK(hipEventSynchronize(ctx->tev[1]));
            float ms = 0.f;
            HCK(hipEventElapsedTime(&ms, ctx->tev[0], ctx->tev[1]));
            if (kind == 0 || ms < 0.97f * best) {
                b->her2k_valu = kind;
                best = ms;
            }
        }
    }
    return BSE_OK;
}

static int batch_capture(bse_ctx* ctx, BatchState* b) {
    int err = BSE_OK;
    hipStream_t st = ctx->st;
    const int nsys = b->nsys, nb = kPanel;
    const int n0 = (int)b->hs[0].g.N;
    std::vector<int> nsz(nsys);
    for (int s = 0; s < nsys; ++s) nsz[s] = (int)b->hs[s].g.N;
    // one event per fork/join point
    const size_t nev = (size_t)nsys + 2 * kSide + 2;
    while (b->evs.size() < nev) {
        hipEvent_t ev;
        HCK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b->evs.push_back(ev);
    }
    size_t ei = 0;
    batch_release_graph(b);
    HCK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    BCK(bse_launch_stamp(st, b->stamps));
    BCK(bse_launch_zero_words(st, (int64_t)nsys * 16 / 2, b->counters));
    // ---- per-system front half on the side streams
    {
        hipEvent_t fork = b->evs[ei++];
        BCK(hipEventRecord(fork, st));
        for (int k = 0; k < kSide; ++k) BCK(hipStreamWaitEvent(b->side[k], fork, 0));
        for (int s = 0; s < nsys; ++s) BCK(record_front(b, s, b->side[s % kSide]));
        for (int k = 0; k < kSide; ++k) {
            hipEvent_t j = b->evs[ei++];
            BCK(hipEventRecord(j, b->side[k]));
            BCK(hipStreamWaitEvent(st, j, 0));
        }
    }
    // ---- lockstep reduction; each system's tail forks after its last column
    {
        int tail_stream = 0;
        for (int c = 0; c < n0 - 1; ++c) {
            int nact = 0;
            while (nact < nsys && nsz[nact] - 1 > c) ++nact;
            const int ch = hemv_ch_table_lookup(ctx->hemv_ch[b->fp32], n0 - c - 1);
            if (c == 0) b->ch_top = ch;
            BCK(bse_hetrd_batch_column(st, b->fp32, c, nb, nact, n0, b->d_hsys, ch));
            int64_t max_nt = 0;
            bool any_end = false;
            for (int s = 0; s < nact; ++s) {
                const int p = c / nb, j0 = p * nb, nbp = std::min(nb, nsz[s] - 1 - j0);
                if (c != j0 + nbp - 1) continue;
                any_end = true;
                if (j0 + nbp >= nsz[s]) continue;
                const int64_t nT = tile_count(nsz[s]) - ((j0 + nbp) >> 6);
                max_nt = std::max(max_nt, nT);
            }
            if (any_end) BCK(bse_hetrd_batch_panel(st, b->fp32, c, nb, nact, b->d_hsys, max_nt, b->her2k_valu));
            for (int s = 0; s < nact; ++s) {
                if (nsz[s] - 2 != c) continue;
                hipEvent_t fork = b->evs[ei++];
                BCK(hipEventRecord(fork, st));
                hipStream_t ss = b->side[tail_stream++ % kSide];
                BCK(hipStreamWaitEvent(ss, fork, 0));
                BCK(record_tail(b, s, ss));
            }
        }
        for (int k = 0; k < kSide; ++k) {
            hipEvent_t j = b->evs[ei++];
            BCK(hipEventRecord(j, b->side[k]));
            BCK(hipStreamWaitEvent(st, j, 0));
        }
    }
    HCK(hipStreamEndCapture(st, &b->graph));
    {
        size_t nn = 0;
        HCK(hipGraphGetNodes(b->graph, nullptr, &nn));
        b->nodes = (int64_t)nn;
    }
    HCK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return BSE_OK;
capture_failed: {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(st, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return err;
}
}

extern "C" int bse_batch_prepare(bse_ctx* ctx, int n_systems, const bse_tb_model* models,
                                 const bse_system* systems, const bse_batch_options* opt) {
    if (!ctx) return BSE_ERR_ARG;
    if (n_systems < 1 || !models || !systems || !opt) return fail(ctx, BSE_ERR_ARG, "null argument or no systems");
    if (opt->precision != BSE_PREC_FP64 && opt->precision != BSE_PREC_FP32 &&
        opt->precision != BSE_PREC_BF16X2)
        return fail(ctx, BSE_ERR_ARG, "unknown precision");
    if (opt->n_lowest < 1 || opt->n_lowest > 16) return fail(ctx, BSE_ERR_ARG, "n_lowest must be in [1, 16]");
    if (opt->n_omega < 0 || (opt->n_omega > 0 && !opt->omega))
        return fail(ctx, BSE_ERR_ARG, "n_omega > 0 needs omega");
    if (opt->n_omega > 0 && !(opt->eta > 0.0)) return fail(ctx, BSE_ERR_ARG, "eta must be > 0");
    (void)hipSetDevice(ctx->device);
    if (!ctx->batch) ctx->batch = new BatchState();
    BatchState* b = ctx->batch;
    b->ran = false;
    batch_release_graph(b);
    b->hs.assign(n_systems, SysHost());
    for (int i = 0; i < n_systems; ++i) {
        std::string msg;
        const int vrc = make_geom(&models[i], &systems[i], b->hs[i].g, msg);
        if (vrc) return fail(ctx, vrc, "system " + std::to_string(i) + ": " + msg);
        const Geom& g = b->hs[i].g;
        if (systems[i].screening != 0)
            return fail(ctx, BSE_ERR_UNSUPPORTED, "system " + std::to_string(i) + ": the batch uses model screening");
        if (systems[i].dipole_model != 0)
            return fail(ctx, BSE_ERR_UNSUPPORTED, "system " + std::to_string(i) + ": the batch uses point-orbital dipoles");
        if (g.N < 2 || g.N > 4096)
            return fail(ctx, BSE_ERR_UNSUPPORTED, "system " + std::to_string(i) + ": batched N must be in [2, 4096]");
        if (opt->n_lowest > g.N) return fail(ctx, BSE_ERR_ARG, "n_lowest exceeds a system's N");
        const bse_tb_model& m = models[i];
        b->hs[i].rcart.assign(m.r_cart, m.r_cart + 3 * (size_t)m.nr);
        b->hs[i].hr.assign(m.h_r, m.h_r + 2 * (size_t)m.nr * m.norb * m.norb);
        b->hs[i].tau.assign(m.tau, m.tau + 3 * (size_t)m.norb);
        b->hs[i].orig = i;
        b->hs[i].TK = bse_assemble_pick_tk(systems[i].nv, systems[i].nc, m.norb);
    }
    std::stable_sort(b->hs.begin(), b->hs.end(),
                     [](const SysHost& x, const SysHost& y) { return x.g.N > y.g.N; });
    b->nsys = n_systems;
    b->precision = opt->precision;
    b->fp32 = opt->precision != BSE_PREC_FP64;
    b->nlow = opt->n_lowest;
    b->refine = opt->refine != 0;
    b->n_omega = opt->n_omega;
    b->eta = opt->eta;
    // ---- arena: sizes (pass 1), allocation, carving (pass 2)
    auto layout = [&](char* base) {
        Bump bp{base};
        unsigned int* counters = bp.take<unsigned int>((size_t)n_systems * 16);
        HetrdSys* hsys = bp.take<HetrdSys>(n_systems);
        double* om = bp.take<double>(std::max(opt->n_omega, 1));
        unsigned long long* stamps = bp.take<unsigned long long>(n_systems + 1);
        b->ds.assign(n_systems, SysDev());
        for (int s = 0; s < n_systems; ++s)
            carve(bp, b->hs[s], b->fp32, b->nlow, b->n_omega, b->ds[s], counters ? counters + 16 * s : nullptr);
        b->counters = counters;
        b->d_hsys = hsys;
        b->d_omega = om;
        b->stamps = stamps;
        return bp.off;
    };
    const size_t need = layout(nullptr);
    if (need > b->arena_cap) {
        if (b->arena) (void)hipFree(b->arena);
        b->arena = nullptr;
        b->arena_cap = 0;
        if (hipMalloc(&b->arena, need) != hipSuccess) {
            b->arena = nullptr;
            return fail(ctx, BSE_ERR_ALLOC, "batch arena of " + std::to_string(need) + " bytes");
        }
        b->arena_cap = need;
    }
    b->arena_bytes = need;
    layout((char*)b->arena);
    // ---- inputs and descriptors
    hipStream_t st = ctx->st;
    std::vector<HetrdSys> hsys(n_systems);
    for (int s = 0; s < n_systems; ++s) {
        const SysHost& h = b->hs[s];
        const SysDev& d = b->ds[s];
        HCK(hipMemcpyAsync(d.rcart, h.rcart.data(), sizeof(double) * h.rcart.size(), hipMemcpyHostToDevice, st));
        HCK(hipMemcpyAsync(d.hr, h.hr.data(), sizeof(double) * h.hr.size(), hipMemcpyHostToDevice, st));
        HCK(hipMemcpyAsync(d.tau, h.tau.data(), sizeof(double) * h.tau.size(), hipMemcpyHostToDevice, st));
        HCK(hipMemcpyAsync(d.kpts, h.g.kpts.data(), sizeof(double) * h.g.kpts.size(), hipMemcpyHostToDevice, st));
        HCK(hipMemcpyAsync(d.gvec, h.g.gvec.data(), sizeof(double) * h.g.gvec.size(), hipMemcpyHostToDevice, st));
        HetrdSys& x = hsys[s];
        x.n = (int)h.g.N; x.g_stride = 0; x.A = d.H; x.Vstore = d.Vst; x.Tall = d.Tall; x.Wp = d.Wp;
        x.alpha2 = d.alpha2; x.G = d.G; x.acol = d.acol; x.d = d.d; x.e = d.e; x.tau = d.tauh;
        x.scal = d.scal; x.part = d.part; x.counter = d.counter; x.Yr = d.Yr; x.Yc = d.Yc; x.t12 = d.t12;
    }
    HCK(hipMemcpyAsync(b->d_hsys, hsys.data(), sizeof(HetrdSys) * n_systems, hipMemcpyHostToDevice, st));
    if (opt->n_omega > 0)
        HCK(hipMemcpyAsync(b->d_omega, opt->omega, sizeof(double) * opt->n_omega, hipMemcpyHostToDevice, st));
    HCK(hipMemsetAsync(b->counters, 0, sizeof(unsigned int) * 16 * n_systems, st));
    HCK(hipStreamSynchronize(st));
    for (int k = 0; k < kSide; ++k)
        if (!b->side[k]) HCK(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking));
    if (!b->t0) HCK(hipEventCreate(&b->t0));
    if (!b->t1) HCK(hipEventCreate(&b->t1));
    HCK(hipDeviceGetAttribute(&b->wall_khz, hipDeviceAttributeWallClockRate, ctx->device));
    int rc = batch_autotune(ctx, b);
    if (rc) return rc;
    HCK(hipStreamSynchronize(st));
    rc = batch_capture(ctx, b);
    if (rc) return rc;
    // the single-system matrix slots were not touched; the context's problem state is unchanged
    return BSE_OK;
}

extern "C" int bse_batch_run(bse_ctx* ctx, double* total_ms, double* system_ms) {
    if (!ctx) return BSE_ERR_ARG;
    BatchState* b = ctx->batch;
    if (!b || !b->exec) return fail(ctx, BSE_ERR_STATE, "no prepared batch");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->st;
    HCK(hipEventRecord(b->t0, st));
    HCK(hipGraphLaunch(b->exec, st));
    HCK(hipEventRecord(b->t1, st));
    HCK(hipEventSynchronize(b->t1));
    float ms = 0.f;
    HCK(hipEventElapsedTime(&ms, b->t0, b->t1));
    if (total_ms) *total_ms = ms;
    if (system_ms) {
        std::vector<unsigned long long> sv(b->nsys + 1);
        HCK(hipMemcpy(sv.data(), b->stamps, sizeof(unsigned long long) * sv.size(), hipMemcpyDeviceToHost));
        const double khz = b->wall_khz > 0 ? (double)b->wall_khz : 1.0e5;
        for (int s = 0; s < b->nsys; ++s)
            system_ms[b->hs[s].orig] = (double)(sv[1 + s] - sv[0]) / khz;
    }
    b->ran = true;
    return BSE_OK;
}

extern "C" int bse_batch_results(bse_ctx* ctx, double* evals, double* eps_re, double* eps_im) {
    if (!ctx) return BSE_ERR_ARG;
    BatchState* b = ctx->batch;
    if (!b || !b->ran) return fail(ctx, BSE_ERR_STATE, "batch has not run");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->st;
    for (int s = 0; s < b->nsys; ++s) {
        const SysDev& d = b->ds[s];
        const int o = b->hs[s].orig;
        if (evals)
            HCK(hipMemcpyAsync(evals + (size_t)o * b->nlow, d.evals_out, sizeof(double) * b->nlow,
                               hipMemcpyDeviceToHost, st));
        if (b->n_omega > 0 && eps_re)
            HCK(hipMemcpyAsync(eps_re + (size_t)o * b->n_omega, d.eps_out, sizeof(double) * b->n_omega,
                               hipMemcpyDeviceToHost, st));
        if (b->n_omega > 0 && eps_im)
            HCK(hipMemcpyAsync(eps_im + (size_t)o * b->n_omega, d.eps_out + b->n_omega,
                               sizeof(double) * b->n_omega, hipMemcpyDeviceToHost, st));
    }
    HCK(hipStreamSynchronize(st));
    return BSE_OK;
}

extern "C" int bse_batch_info(const bse_ctx* ctx, int64_t info[6]) {
    if (!ctx || !info) return BSE_ERR_ARG;
    const BatchState* b = ctx->batch;
    if (!b || !b->exec) return BSE_ERR_STATE;
    info[0] = b->nodes;
    info[1] = (int64_t)b->arena_bytes;
    info[2] = b->hs.empty() ? 0 : b->hs[0].g.N;
    info[3] = b->nsys;
    info[4] = b->ch_top;
    info[5] = b->her2k_valu;
    return BSE_OK;
}

// Called by bse_destroy.
void batch_destroy(bse_ctx* ctx) {
    BatchState* b = ctx->batch;
    if (!b) return;
    batch_release_graph(b);
    if (b->arena) (void)hipFree(b->arena);
    for (hipStream_t s : b->side)
        if (s) (void)hipStreamDestroy(s);
    for (hipEvent_t e : b->evs) (void)hipEventDestroy(e);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    delete b;
    ctx->batch = nullptr;
}
