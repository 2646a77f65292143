// C ABI of the MI355X excitonic hot path (include/bse_mi355x.h).
//
// Host-side orchestration of the kernels; every heavy step runs on the
// context's HIP stream.  The only host work is O(N) bookkeeping of the
// tridiagonal stage (split points, sort of eigenvalues, cluster and batch
// boundaries), which needs the tridiagonal (2N doubles) and the eigenvalues
// (N doubles) on the host.
#include "bse_internal.hpp"

// ===========================================================================
extern "C" int bse_abi_version(void) { return BSE_ABI_VERSION; }

extern "C" const char* bse_last_error(const bse_ctx* ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int bse_create(bse_ctx** out, int device) {
    if (!out) return BSE_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BSE_ERR_HIP;
    if (device < 0 || device >= ndev) return BSE_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return BSE_ERR_HIP;
    bse_ctx* ctx = new bse_ctx();
    ctx->device = device;
    ctx->bufs.resize(S_NSLOTS);
    if (hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return BSE_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) {
        (void)hipStreamDestroy(ctx->st);
        delete ctx;
        return BSE_ERR_HIP;
    }
    for (int i = 0; i <= ST_N; ++i) (void)hipEventCreate(&ctx->ev[i]);
    for (int i = 0; i < 3; ++i) (void)hipEventCreate(&ctx->tev[i]);
    (void)hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming);
    for (int i = 0; i < 2; ++i) (void)hipEventCreateWithFlags(&ctx->la_ev[i], hipEventDisableTiming);
    const char* at = getenv("BSE_AUTOTUNE");
    ctx->autotune = !(at && at[0] == '0');
    const char* ts = getenv("BSE_TWO_STAGE");
    if (ts) ctx->two_stage = (ts[0] == '0') ? -1 : 1;
    const char* am = getenv("BSE_ASSEMBLY");
    ctx->force_valu_assembly = am && strcmp(am, "valu") == 0;
    *out = ctx;
    return BSE_OK;
}

extern "C" int bse_destroy(bse_ctx* ctx) {
    if (!ctx) return BSE_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->st);
    (void)hipStreamSynchronize(ctx->side);
    batch_destroy(ctx);
    for (auto& b : ctx->bufs)
        if (b.p) (void)hipFree(b.p);
    for (int i = 0; i <= ST_N; ++i)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    for (hipEvent_t e : ctx->prof_ev) (void)hipEventDestroy(e);
    for (int i = 0; i < 3; ++i)
        if (ctx->tev[i]) (void)hipEventDestroy(ctx->tev[i]);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
    for (int i = 0; i < 2; ++i)
        if (ctx->la_ev[i]) (void)hipEventDestroy(ctx->la_ev[i]);
    (void)hipStreamDestroy(ctx->side);
    (void)hipStreamDestroy(ctx->st);
    delete ctx;
    return BSE_OK;
}

extern "C" int bse_stage_times(const bse_ctx* ctx, double* ms, int n) {
    if (!ctx || !ms || n < 0 || n > 16) return BSE_ERR_ARG;
    for (int i = 0; i < n; ++i) ms[i] = ctx->stage_ms[i];
    return BSE_OK;
}

// Kernel-level profile of the last tridiagonalisation.  One-stage: out[0] =
// total hemv time (ms, HIP events around every launch), out[1] = algorithmic
// bytes read (lower triangle of each trailing block), out[2] = launches.
// Two-stage (out[4] = 1): out[0], out[3], out[2] = rank-2k update time,
// algorithmic flops, launches; out[5], out[6], out[7] the same for the
// Hermitian block product; out[1] = 1 when both ran on FP32 matrix cores.
extern "C" int bse_profile(bse_ctx* ctx, int enable, double* out) {
    if (!ctx) return BSE_ERR_ARG;
    ctx->prof = enable != 0;
    if (out) {
        for (int i = 0; i < 9; ++i) out[i] = 0.0;
        if (ctx->prof_two_stage) {
            out[0] = ctx->b2k_ms;
            out[1] = ctx->band_fp32prod ? 1.0 : 0.0;
            out[2] = (double)ctx->band_launches;
            out[3] = ctx->b2k_flops;
            out[4] = 1.0;
            out[5] = ctx->hemm_ms;
            out[6] = ctx->hemm_flops;
            out[7] = (double)ctx->band_launches;
            out[8] = ctx->band_lookahead ? (ctx->band_pack ? 4.0 : 3.0) : 2.0;
        } else {
            out[0] = ctx->hemv_ms;
            out[1] = ctx->hemv_bytes;
            out[2] = (double)ctx->hemv_launches;
        }
    }
    return BSE_OK;
}

extern "C" int bse_problem_sizes(const bse_ctx* ctx, int64_t out[4]) {
    if (!ctx || !out) return BSE_ERR_ARG;
    out[0] = ctx->nk; out[1] = ctx->N; out[2] = ctx->norb; out[3] = ctx->ng;
    return BSE_OK;
}

// ---------------------------------------------------------------------------
// Problem upload: geometry (Crystal_Lattice 85-124, K_points 183-242,
// G_points 319-385) is derived on the host and copied to the device.
// ---------------------------------------------------------------------------
// Upload of a problem whose geometry g is derived; kpos (may be null): the
// caller's k-list position of each grid index.
static int load_problem(bse_ctx* ctx, const bse_tb_model* m, const Geom& g, const int* kpos) {
    (void)hipSetDevice(ctx->device);
    const bse_system* s = &g.sys;
    ctx->sys = g.sys;
    ctx->norb = g.norb;
    ctx->nr = g.nr;
    ctx->nk = g.nk;
    ctx->B = g.B;
    ctx->ns = g.ns;
    ctx->N = g.N;
    for (int i = 0; i < 9; ++i) ctx->recip[i] = g.recip[i];
    ctx->volume = g.volume;
    ctx->area = g.area;
    ctx->cell = g.cell;
    ctx->n3 = g.n3;
    ctx->dim = g.dim;
    ctx->kpts_h = g.kpts;
    ctx->gvec_h = g.gvec;
    ctx->ng = g.ng;
    ctx->wbar = g.wbar;
    ctx->vbar = g.vbar;
    ctx->TK = bse_assemble_pick_tk(s->nv, s->nc, m->norb);
    const size_t nh = (size_t)m->nr * m->norb * m->norb * 2;

    double* d_rcart = BUF(double, S_RCART, 3 * (size_t)m->nr);
    double2* d_hr = BUF(double2, S_HR, (size_t)m->nr * m->norb * m->norb);
    double* d_tau = BUF(double, S_TAU, 3 * (size_t)m->norb);
    double* d_k = BUF(double, S_KPTS, 3 * (size_t)ctx->nk);
    double* d_g = BUF(double, S_GVEC, 3 * (size_t)ctx->ng);
    double* d_q = BUF(double, S_Q, 3);
    ctx->hr_h.assign(m->h_r, m->h_r + nh);
    ctx->rcart_h.assign(m->r_cart, m->r_cart + 3 * (size_t)m->nr);
    HCK(hipMemcpyAsync(d_rcart, m->r_cart, 3 * sizeof(double) * m->nr, hipMemcpyHostToDevice, ctx->st));
    HCK(hipMemcpyAsync(d_hr, m->h_r, nh * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    HCK(hipMemcpyAsync(d_tau, m->tau, 3 * sizeof(double) * m->norb, hipMemcpyHostToDevice, ctx->st));
    HCK(hipMemcpyAsync(d_k, ctx->kpts_h.data(), 3 * sizeof(double) * ctx->nk, hipMemcpyHostToDevice, ctx->st));
    HCK(hipMemcpyAsync(d_g, ctx->gvec_h.data(), 3 * sizeof(double) * ctx->ng, hipMemcpyHostToDevice, ctx->st));
    HCK(hipMemcpyAsync(d_q, s->q, 3 * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    ctx->dip_nr = 0;
    if (s->dipole_model == 1) {
        // the Wannier position table is copied in (the caller's arrays need not outlive the call)
        const int nr = s->dipole_nr;
        double* d_dr = BUF(double, S_DIPR, 3 * (size_t)nr);
        double* d_dd = BUF(double, S_DIPD, (size_t)nr * m->norb * m->norb * 3);
        HCK(hipMemcpyAsync(d_dr, s->dipole_r, 3 * sizeof(double) * nr, hipMemcpyHostToDevice, ctx->st));
        HCK(hipMemcpyAsync(d_dd, s->dipole_d, sizeof(double) * nr * m->norb * m->norb * 3,
                           hipMemcpyHostToDevice, ctx->st));
        ctx->dip_nr = nr;
        ctx->sys.dipole_r = nullptr;       // the host pointers are not kept
        ctx->sys.dipole_d = nullptr;
    }
    ctx->kmap_h.clear();
    ctx->kinv_h.clear();
    if (kpos) {
        ctx->kmap_h.assign(kpos, kpos + ctx->nk);
        ctx->kinv_h.assign(ctx->nk, 0);
        for (int k = 0; k < ctx->nk; ++k) ctx->kinv_h[kpos[k]] = k;
        int* d_kmap = BUF(int, S_KMAP, ctx->nk);
        int* d_kinv = BUF(int, S_KINV, ctx->nk);
        HCK(hipMemcpyAsync(d_kmap, ctx->kmap_h.data(), sizeof(int) * ctx->nk, hipMemcpyHostToDevice, ctx->st));
        HCK(hipMemcpyAsync(d_kinv, ctx->kinv_h.data(), sizeof(int) * ctx->nk, hipMemcpyHostToDevice, ctx->st));
    }
    for (int x = 0; x < 3; ++x) ctx->qhat[x] = g.qhat[x];
    HCK(hipStreamSynchronize(ctx->st));
    ctx->have_problem = true;
    ctx->have_tb = ctx->have_tables = ctx->built = ctx->have_eig = ctx->have_f = false;
    ctx->full_built = ctx->full_eig = ctx->have_coupling = false;
    return BSE_OK;
}

// Device maps of the caller's k order (null for the plain grid order).
static const int* kmap_dev(const bse_ctx* ctx) {
    return ctx->kmap_h.empty() ? nullptr : (const int*)ctx->bufs[S_KMAP].p;
}
static const int* kinv_dev(const bse_ctx* ctx) {
    return ctx->kinv_h.empty() ? nullptr : (const int*)ctx->bufs[S_KINV].p;
}

extern "C" int bse_set_problem(bse_ctx* ctx, const bse_tb_model* m, const bse_system* s) {
    if (!ctx) return BSE_ERR_ARG;
    Geom g;
    std::string msg;
    const int vrc = make_geom(m, s, g, msg);
    if (vrc) return fail(ctx, vrc, msg);
    return load_problem(ctx, m, g, nullptr);
}

// A k-list in crystal coordinates as a regular grid: along every axis the
// fractional parts (k - floor k) must take n equally spaced values v0 +
// j/n, the offset v0 n is the grid shift in steps, and the n1 n2 n3 index
// triples must each appear once.  2D systems need a zero third coordinate
// (mod 1).  grid[f] receives the grid index of list entry f.
static bool klist_to_grid(const double* kc, int nk, int dim, int n[3], double shift[3],
                          std::vector<int>& grid, std::string& msg) {
    const double tol = 1e-7;
    std::vector<int> idx[3];
    for (int a = 0; a < 3; ++a) {
        std::vector<double> fr(nk);
        for (int f = 0; f < nk; ++f) {
            double x = kc[3 * (size_t)f + a] - floor(kc[3 * (size_t)f + a]);
            if (x > 1.0 - tol) x = 0.0;
            fr[f] = x;
        }
        if (a == 2 && dim == 2) {
            for (int f = 0; f < nk; ++f)
                if (fr[f] > tol) { msg = "2D k-list with a component along b3"; return false; }
            n[2] = 1;
            shift[2] = 0.0;
            idx[2].assign(nk, 0);
            continue;
        }
        std::vector<double> vals(fr);
        std::sort(vals.begin(), vals.end());
        std::vector<double> u;
        for (double v : vals)
            if (u.empty() || v - u.back() > tol) u.push_back(v);
        const int na = (int)u.size();
        for (int j = 0; j < na; ++j)
            if (fabs(u[j] - u[0] - (double)j / na) > tol) {
                msg = "k-list coordinates are not equally spaced along an axis";
                return false;
            }
        n[a] = na;
        shift[a] = u[0] * na;
        idx[a].resize(nk);
        for (int f = 0; f < nk; ++f) {
            const double t = (fr[f] - u[0]) * na;
            const int j = (int)lround(t);
            if (fabs(t - j) > 1e-5 || j < 0 || j >= na) { msg = "k-list point off the grid"; return false; }
            idx[a][f] = j;
        }
    }
    if ((int64_t)n[0] * n[1] * n[2] != nk) { msg = "k-list is not a complete grid"; return false; }
    grid.assign(nk, 0);
    std::vector<char> seen(nk, 0);
    for (int f = 0; f < nk; ++f) {
        const int gi = (idx[0][f] * n[1] + idx[1][f]) * n[2] + idx[2][f];
        if (seen[gi]) { msg = "k-list repeats a grid point"; return false; }
        seen[gi] = 1;
        grid[f] = gi;
    }
    return true;
}

extern "C" int bse_set_problem_klist(bse_ctx* ctx, const bse_tb_model* m, const bse_system* s,
                                     const double* k_cryst, int nk) {
    if (!ctx) return BSE_ERR_ARG;
    if (!s || !k_cryst || nk < 1) return fail(ctx, BSE_ERR_ARG, "null k-list");
    if (!finite_all(k_cryst, 3 * (size_t)nk)) return fail(ctx, BSE_ERR_NUMERIC, "non-finite k-list");
    const int dim = s->dimension == 3 ? 3 : 2;
    int n[3];
    double sh[3];
    std::vector<int> grid;
    std::string msg;
    if (!klist_to_grid(k_cryst, nk, dim, n, sh, grid, msg)) return fail(ctx, BSE_ERR_UNSUPPORTED, msg);
    bse_system s2 = *s;
    s2.n1 = n[0];
    s2.n2 = n[1];
    s2.n3 = n[2];
    s2.shift[0] = sh[0];
    s2.shift[1] = sh[1];
    s2.shift3 = sh[2];
    Geom g;
    const int vrc = make_geom(m, &s2, g, msg);
    if (vrc) return fail(ctx, vrc, msg);
    std::vector<int> kpos(nk);
    for (int f = 0; f < nk; ++f) kpos[grid[f]] = f;
    bool ident = true;
    for (int k = 0; k < nk; ++k) ident = ident && kpos[k] == k;
    return load_problem(ctx, m, g, ident ? nullptr : kpos.data());
}

// ---------------------------------------------------------------------------
static int run_tb(bse_ctx* ctx) {
    const bse_system& s = ctx->sys;
    double* eps = BUF(double, S_EPS, (size_t)ctx->nk * ctx->ns);
    double2* coef = BUF(double2, S_COEF, (size_t)ctx->nk * ctx->ns * ctx->norb);
    stage_begin(ctx, ST_TB);
    HCK(bse_launch_tb_states(ctx->st, ctx->nk, ctx->norb, ctx->nr,
                             (const double*)ctx->bufs[S_RCART].p, (const double2*)ctx->bufs[S_HR].p,
                             (const double*)ctx->bufs[S_KPTS].p, s.q, s.n_occ, s.nv, s.nc,
                             s.scissor, nullptr, eps, coef, nullptr, nullptr));
    stage_end(ctx, ST_TB);
    ctx->have_tb = true;
    return BSE_OK;
}

extern "C" int bse_tb_states(bse_ctx* ctx, double* eps, double* coef) {
    if (!ctx) return BSE_ERR_ARG;
    if (!ctx->have_problem) return fail(ctx, BSE_ERR_STATE, "no problem loaded");
    (void)hipSetDevice(ctx->device);
    int rc = run_tb(ctx);
    if (rc) return rc;
    const size_t ne = (size_t)ctx->nk * ctx->ns, nc = ne * ctx->norb;
    std::vector<double> he(eps ? ne : 0);
    std::vector<double2> hc(coef ? nc : 0);
    if (eps) HCK(hipMemcpyAsync(he.data(), ctx->bufs[S_EPS].p, sizeof(double) * ne, hipMemcpyDeviceToHost, ctx->st));
    if (coef) HCK(hipMemcpyAsync(hc.data(), ctx->bufs[S_COEF].p, sizeof(double2) * nc, hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    // rows in the caller's k order
    const size_t rs = (size_t)ctx->ns, rc2 = rs * ctx->norb;
    for (int k = 0; k < ctx->nk; ++k) {
        const size_t kd = ctx->kmap_h.empty() ? (size_t)k : (size_t)ctx->kmap_h[k];
        if (eps) std::memcpy(eps + kd * rs, he.data() + (size_t)k * rs, sizeof(double) * rs);
        if (coef) std::memcpy(coef + 2 * kd * rc2, hc.data() + (size_t)k * rc2, sizeof(double2) * rc2);
    }
    return BSE_OK;
}

// States at caller-given k (pull_ks_states / pull_ks_states_subset for a
// list of k).  The valence half of K1+K2 is asked for at k itself (q = 0).
extern "C" int bse_tb_states_at(bse_ctx* ctx, const bse_tb_model* m, const double* kpts, int nk,
                                int n_occ, int nv, int nc, double scissor, double* eps, double* coef) {
    if (!ctx || !kpts || nk < 1) return BSE_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    int norb, nr;
    const double* d_rc;
    const double2* d_hr;
    if (m) {
        if (!m->r_cart || !m->h_r || !m->tau) return fail(ctx, BSE_ERR_ARG, "null model array");
        if (m->norb < 1 || m->norb > BSE_MAXORB) return fail(ctx, BSE_ERR_UNSUPPORTED, "norb must be in [1, 16]");
        if (m->nr < 1) return fail(ctx, BSE_ERR_ARG, "nr must be >= 1");
        norb = m->norb;
        nr = m->nr;
        const size_t nh = (size_t)nr * norb * norb * 2;
        if (!finite_all(m->h_r, nh) || !finite_all(m->r_cart, 3 * (size_t)nr))
            return fail(ctx, BSE_ERR_NUMERIC, "non-finite model input");
        double* r = BUF(double, S_TBR, 3 * (size_t)nr);
        double2* h = BUF(double2, S_TBH, (size_t)nr * norb * norb);
        HCK(hipMemcpyAsync(r, m->r_cart, 3 * sizeof(double) * nr, hipMemcpyHostToDevice, ctx->st));
        HCK(hipMemcpyAsync(h, m->h_r, nh * sizeof(double), hipMemcpyHostToDevice, ctx->st));
        d_rc = r;
        d_hr = h;
    } else {
        if (!ctx->have_problem) return fail(ctx, BSE_ERR_STATE, "no model given and no problem loaded");
        norb = ctx->norb;
        nr = ctx->nr;
        d_rc = (const double*)ctx->bufs[S_RCART].p;
        d_hr = (const double2*)ctx->bufs[S_HR].p;
    }
    const bool all = nv == 0 && nc == 0;
    if (all) {
        n_occ = 0;
        nc = norb;
        scissor = 0.0;
    } else if (nv < 0 || nc < 0 || n_occ < nv || n_occ + nc > norb) {
        return fail(ctx, BSE_ERR_ARG, "band window outside the Hamiltonian (nv <= n_occ, n_occ + nc <= norb)");
    }
    if (!finite_all(kpts, 3 * (size_t)nk)) return fail(ctx, BSE_ERR_NUMERIC, "non-finite k");
    const int ns = nv + nc;
    double* d_k = BUF(double, S_TBK, 3 * (size_t)nk);
    double* d_e = BUF(double, S_TBE, (size_t)nk * ns);
    double2* d_c = BUF(double2, S_TBC, (size_t)nk * ns * norb);
    HCK(hipMemcpyAsync(d_k, kpts, 3 * sizeof(double) * nk, hipMemcpyHostToDevice, ctx->st));
    const double q0[3] = {0.0, 0.0, 0.0};
    HCK(bse_launch_tb_states(ctx->st, nk, norb, nr, d_rc, d_hr, d_k, q0, n_occ, nv, nc, scissor, nullptr,
                             d_e, d_c, nullptr, nullptr));
    if (eps) HCK(hipMemcpyAsync(eps, d_e, sizeof(double) * nk * ns, hipMemcpyDeviceToHost, ctx->st));
    if (coef) HCK(hipMemcpyAsync(coef, d_c, sizeof(double2) * nk * ns * norb, hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    return BSE_OK;
}

static int run_tables(bse_ctx* ctx) {
    const bse_system& s = ctx->sys;
    const int no = ctx->norb;
    const int64_t nw = ktable_size(s.n1, s.n2, ctx->n3) * no * no;
    double2* wtab = BUF(double2, S_WTAB, nw);
    double2* vx = BUF(double2, S_VX, (size_t)no * no);
    double2* P = BUF(double2, S_P, (size_t)ctx->N * no);
    double2* PV = BUF(double2, S_PV, (size_t)ctx->N * no);
    double2* rho = BUF(double2, S_RHO, (size_t)ctx->N);
    double* dE = BUF(double, S_DE, (size_t)ctx->N);
    const double* gvec = (const double*)ctx->bufs[S_GVEC].p;
    const double* tau = (const double*)ctx->bufs[S_TAU].p;
    stage_begin(ctx, ST_TABLES);
    if (s.screening == 1) {
        // static RPA: the screening window's states at the grid points (no
        // optical shift), Pi(Q) and the screened table (rpa_kernels.hip)
        const int nvs = s.rpa_nv > 0 ? std::min(s.rpa_nv, s.n_occ) : s.n_occ;
        const int ncs = s.rpa_nc > 0 ? std::min(s.rpa_nc, no - s.n_occ) : no - s.n_occ;
        if (nvs < 1 || ncs < 1) return fail(ctx, BSE_ERR_ARG, "RPA screening needs occupied and empty bands");
        const int nss = nvs + ncs;
        double* reps = BUF(double, S_RPAEPS, (size_t)ctx->nk * nss);
        double2* rcoef = BUF(double2, S_RPACOEF, (size_t)ctx->nk * nss * no);
        double2* pi = BUF(double2, S_RPAPI, bse_rpa_pi_elems(s.n1, s.n2, ctx->n3, no));
        static const double q0[3] = {0.0, 0.0, 0.0};
        HCK(bse_launch_tb_states(ctx->st, ctx->nk, no, ctx->nr, (const double*)ctx->bufs[S_RCART].p,
                                 (const double2*)ctx->bufs[S_HR].p, (const double*)ctx->bufs[S_KPTS].p, q0,
                                 s.n_occ, nvs, ncs, s.scissor, nullptr, reps, rcoef, nullptr, nullptr));
        HCK(bse_launch_rpa_table(ctx->st, s.n1, s.n2, ctx->n3, no, ctx->ng, ctx->dim, ctx->recip, nvs, ncs,
                                 s.f_s / (ctx->cell * ctx->nk), ctx->vbar, s.radius, kC, reps, rcoef, gvec, tau,
                                 pi, wtab));
    } else {
        HCK(bse_launch_w_table(ctx->st, s.n1, s.n2, ctx->n3, ctx->dim, no, ctx->ng, gvec, ctx->recip, tau,
                               s.r0, s.eps_env, s.radius, kC, ctx->wbar, 0, wtab));
    }
    HCK(bse_launch_v_exchange(ctx->st, no, ctx->ng, ctx->dim, gvec, tau, kC, s.radius, s.q, vx));
    const double2* dq = nullptr;
    if (ctx->dip_nr > 0) {
        // Wannier position matrix on the grid: Dq_ij(k) = sum_R e^{-ik.R} qhat.D_ij(R)
        double2* dqb = BUF(double2, S_DIPQ, (size_t)ctx->nk * no * no);
        HCK(bse_launch_dipole_table(ctx->st, ctx->nk, no, ctx->dip_nr, (const double*)ctx->bufs[S_KPTS].p,
                                    (const double*)ctx->bufs[S_DIPR].p, (const double*)ctx->bufs[S_DIPD].p,
                                    ctx->qhat, dqb));
        dq = dqb;
    }
    HCK(bse_launch_transitions(ctx->st, ctx->nk, s.nv, s.nc, no, (const double*)ctx->bufs[S_EPS].p,
                               (const double2*)ctx->bufs[S_COEF].p, nullptr, nullptr, tau, ctx->qhat,
                               vx, P, PV, rho, dE, dq));
    stage_end(ctx, ST_TABLES);
    ctx->have_tables = true;
    return BSE_OK;
}

extern "C" int bse_build(bse_ctx* ctx, int precision) {
    if (!ctx) return BSE_ERR_ARG;
    if (!ctx->have_problem) return fail(ctx, BSE_ERR_STATE, "no problem loaded");
    if (precision != BSE_PREC_FP64 && precision != BSE_PREC_FP32 && precision != BSE_PREC_BF16X2)
        return fail(ctx, BSE_ERR_ARG, "unknown precision");
    (void)hipSetDevice(ctx->device);
    int rc;
    if ((rc = run_tb(ctx))) return rc;
    if ((rc = run_tables(ctx))) return rc;
    const bse_system& s = ctx->sys;
    const int fp32 = precision != BSE_PREC_FP64;
    const int64_t N = ctx->N;
    // tiled lower storage: 64 x 64 tiles of the lower triangle, each column-major,
    // ordered by tile column and then tile row (one tile column is contiguous)
    void* H = buf(ctx, S_H, (size_t)tiled_elems(N) * (fp32 ? sizeof(float2) : sizeof(double2)), &rc);
    if (!H) return rc;
    const double alpha = -s.direct_scale / (ctx->cell * ctx->nk);
    const double beta = s.f_s * s.exchange_scale / (ctx->cell * ctx->nk);
    const double2* coef = (const double2*)ctx->bufs[S_COEF].p;
    stage_begin(ctx, ST_ASM);
    HCK(bse_launch_tile_pad_zero(ctx->st, fp32, N, H));
    if (!ctx->force_valu_assembly && bse_assemble_mfma_supported(s.nv, s.nc, ctx->norb)) {
        // matrix cores: FP64 MFMA (FP64 storage), FP32 MFMA or compensated BF16 MFMA (FP32 storage)
        HCK(bse_launch_assemble_mfma(ctx->st, precision, ctx->nk, s.n1, s.n2, ctx->n3, s.nv, s.nc, ctx->norb,
                                     N, coef, (const double2*)ctx->bufs[S_WTAB].p,
                                     (const double2*)ctx->bufs[S_PV].p,
                                     (const double2*)ctx->bufs[S_P].p,
                                     (const double*)ctx->bufs[S_DE].p, alpha, beta, H));
        ctx->asm_kernel = 1 + precision;
    } else {
        // small band windows (e.g. 1v+1c): FP64 VALU tiles, FP64 or FP32 storage
        HCK(bse_launch_assemble(ctx->st, fp32, ctx->nk, s.n1, s.n2, ctx->n3, s.nv, s.nc, ctx->norb, ctx->TK,
                                N, coef, coef, coef, coef, (const double2*)ctx->bufs[S_WTAB].p,
                                (const double2*)ctx->bufs[S_PV].p, (const double2*)ctx->bufs[S_P].p,
                                (const double*)ctx->bufs[S_DE].p, alpha, beta, H, 1));
        ctx->asm_kernel = 0;
    }
    stage_end(ctx, ST_ASM);
    ctx->precision = precision;
    ctx->built = true;
    ctx->have_eig = ctx->have_f = false;
    ctx->full_built = ctx->full_eig = ctx->have_coupling = false;
    return BSE_OK;
}

extern "C" int bse_get_matrix(bse_ctx* ctx, void* dst, int64_t ld, int dst_on_device, int order) {
    if (!ctx || !dst) return BSE_ERR_ARG;
    if (!ctx->built) return fail(ctx, BSE_ERR_STATE, "matrix not built");
    if (ctx->have_eig) return fail(ctx, BSE_ERR_STATE, "matrix storage was overwritten by bse_eigh; build again");
    const int64_t N = ctx->N;
    if (ld < N) return fail(ctx, BSE_ERR_ARG, "ld < N");
    (void)hipSetDevice(ctx->device);
    const int fp32 = ctx->precision != BSE_PREC_FP64;
    const size_t es = fp32 ? sizeof(float2) : sizeof(double2);
    void* out = dst;
    int rc = BSE_OK;
    if (!dst_on_device) {
        out = buf(ctx, S_Z, (size_t)ld * N * es, &rc);
        if (!out) return rc;
    }
    HCK(bse_launch_expand(ctx->st, fp32, N, ctx->B, ctx->nk, ctx->bufs[S_H].p, out, ld,
                          order == BSE_ORDER_REFERENCE, kinv_dev(ctx)));
    if (!dst_on_device) HCK(hipMemcpyAsync(dst, out, (size_t)ld * N * es, hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    ctx->have_eig = false;
    return BSE_OK;
}

extern "C" int bse_get_dipoles(bse_ctx* ctx, double* rho) {
    if (!ctx || !rho) return BSE_ERR_ARG;
    if (!ctx->have_tables) return fail(ctx, BSE_ERR_STATE, "dipoles not built");
    (void)hipSetDevice(ctx->device);
    double2* tmp = BUF(double2, S_W1, ctx->N);
    HCK(bse_launch_rows_to_reference(ctx->st, ctx->N, ctx->B, ctx->nk, 1,
                                     (const double2*)ctx->bufs[S_RHO].p, ctx->N, tmp, ctx->N, kmap_dev(ctx)));
    HCK(hipMemcpyAsync(rho, tmp, sizeof(double2) * ctx->N, hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    return BSE_OK;
}

// ---------------------------------------------------------------------------
// Eigen pipeline on a device matrix H (n x n, lower triangle, ld) of element
// type FP64 (fp32 = 0) or FP32 complex.  rho (may be null): contraction vector
// for the oscillator strengths.
// ---------------------------------------------------------------------------
static bool hemv_code_ok(int code) {
    const int ch = code & 31;
    return code > 0 && code < 256 && (ch == 4 || ch == 8 || ch == 16);
}

// Partial-sum vectors one hemv launch of variant `code` writes for the
// trailing block A[s0:, s0:]: a row partial per (tile row, block of rbk tile
// columns) and a column partial per (tile row, chunk of ch tile rows below it).
static int64_t hemv_partials(int code, int n, int s0) {
    const int ch = code & 31, rbk = ((code >> 5) & 3) == 3 ? 4 : 1;
    const int nbk = (n + 63) / 64 - (s0 >> 6);
    int64_t cnt = 0;
    for (int b = 0; b < nbk; ++b) {
        const int pmax = b / rbk;
        cnt += pmax + 1 + (nbk - pmax * rbk + ch - 1) / ch;
    }
    return cnt;
}

// ---------------------------------------------------------------------------
// Autotuner for the tridiagonalisation's Hermitian matrix-vector kernel: for
// every floor(log2(nt)) bucket up to the current size that is not tuned yet,
// time chunk lengths 4, 8 and 16 on two trailing blocks of the matrix about
// to be reduced (the kernel only reads it) and keep the fastest, each as the
// plain variant (code ch), the two-row-blocks-per-step variant (ch + 32) and
// the variant with per-slab cross-lane column sums (ch + 64), the variant
// with waves owning rows (ch + 96, chunks of 2 ch tiles), with
// non-temporal or default-policy loads (+ 128).  Repeated launches over the
// same block see the cache reuse that consecutive columns see.  The table is
// per context and storage precision; BSE_AUTOTUNE=0 keeps the default
// (ch 8, column sums, non-temporal).  Each candidate's time includes the
// reduce's read of the partial sums it writes (a variant that writes more of
// them makes the column step's reduce slower).
// ---------------------------------------------------------------------------
int autotune_hemv(bse_ctx* ctx, int fp32, int n, const void* H, double2* v, double2* Yr,
                  double2* Yc) {
    unsigned char* tab = ctx->hemv_ch[fp32];
    // BSE_HEMV_CODE=<code> pins one variant for every trailing size (for
    // measurements of a single variant)
    const char* hc = getenv("BSE_HEMV_CODE");
    if (hc && hemv_code_ok(atoi(hc))) {
        for (int b = 0; b < BSE_HEMV_BUCKETS; ++b) tab[b] = (unsigned char)atoi(hc);
        return BSE_OK;
    }
    if (!ctx->autotune || n < 128) return BSE_OK;
    hipStream_t st = ctx->st;
    bool zeroed = false;
    for (int b = 6; b < BSE_HEMV_BUCKETS && (1 << b) <= n - 1; ++b) {
        if (tab[b] != 0) continue;
        if (!zeroed) {
            HCK(hipMemsetAsync(v, 0, sizeof(double2) * n, st));
            zeroed = true;
        }
        // the bucket holds trailing orders [2^b, top]; every candidate is
        // timed at a quarter and at three quarters of that range (a variant
        // with long chunks can win at the top of a bucket and underfill the
        // chip lower in it) and the per-launch times are summed
        const int top = std::min((2 << b) - 1, n - 1);
        const int span = top - (1 << b);
        const int sizes[2] = {(1 << b) + span / 4, (1 << b) + (3 * span) / 4};
        static const int base[12] = {8 + 64, 4 + 64, 16 + 64, 8 + 32, 4 + 32, 16 + 32, 8, 4, 16,
                                     4 + 96, 8 + 96, 16 + 96};
        float tk[24];
        for (int k = 0; k < 24; ++k) tk[k] = 0.f;
        for (int si = 0; si < 2; ++si) {
            const int nt = sizes[si];
            const int s0 = n - nt;             // trailing block A[s0:, s0:]
            const int reps = nt > 4096 ? 5 : 16;
            // two timing passes over the candidates, each keeping its faster one
            float tp[24];
            for (int pass = 0; pass < 2; ++pass)
                for (int k = 0; k < 24; ++k) {
                    const int ch = base[k % 12] + (k >= 12 ? 128 : 0);
                    HCK(bse_launch_hemv(st, fp32, ch, n, s0, H, v, Yr, Yc));
                    HCK(hipEventRecord(ctx->tev[0], st));
                    for (int r = 0; r < reps; ++r) HCK(bse_launch_hemv(st, fp32, ch, n, s0, H, v, Yr, Yc));
                    HCK(hipEventRecord(ctx->tev[1], st));
                    HCK(hipEventSynchronize(ctx->tev[1]));
                    float ms = 0.f;
                    HCK(hipEventElapsedTime(&ms, ctx->tev[0], ctx->tev[1]));
                    ms /= (float)reps;
                    tp[k] = (pass == 0) ? ms : std::min(tp[k], ms);
                }
            // each candidate is charged for the partial sums that the column
            // step's reduce reads back (64 double2 per partial), at the
            // 2.5 TB/s that the reduce sustains on them
            for (int k = 0; k < 24; ++k)
                tk[k] += tp[k] + (float)(1024.0 * (double)hemv_partials(base[k % 12], n, s0) / 2.5e9);
        }
        float best = 0.f;
        int best_ch = 8 + 64;
        for (int k = 0; k < 24; ++k)
            if (best == 0.f || tk[k] < 0.97f * best) { best = tk[k]; best_ch = base[k % 12] + (k >= 12 ? 128 : 0); }
        tab[b] = (unsigned char)best_ch;
        if (getenv("BSE_VERBOSE")) {
            float second = 0.f;
            int second_ch = 0;
            for (int k = 0; k < 24; ++k) {
                const int code = base[k % 12] + (k >= 12 ? 128 : 0);
                if (code != best_ch && (second == 0.f || tk[k] < second)) { second = tk[k]; second_ch = code; }
            }
            fprintf(stderr, "hemv tuner fp32=%d m=%d,%d: code %d %.4f ms (next %d %.4f)\n", fp32, sizes[0],
                    sizes[1], best_ch, best, second_ch, second);
        }
    }
    return BSE_OK;
}

// Trailing-update kernel choice (FP64 VALU, FP64 MFMA on 64 x 64 tiles or on
// 128 x 128 macro tiles; VALU and MFMA have the same FP64 peak on this part):
// all timed once per context and storage precision on a scratch matrix carved
// from the hemv partial buffer (contents irrelevant).  BSE_HER2K=valu / mfma /
// mfma2 forces the choice.
static int autotune_her2k(bse_ctx* ctx, int fp32, void* scratch, size_t scratch_bytes,
                          double2* V, double2* W, int64_t vw_elems, double2* alpha2) {
    if (ctx->her2k_valu[fp32] >= 0) return BSE_OK;
    const char* hv = getenv("BSE_HER2K");
    if (hv && (!strcmp(hv, "valu") || !strcmp(hv, "mfma") || !strcmp(hv, "mfma2"))) {
        ctx->her2k_valu[fp32] = !strcmp(hv, "valu") ? 1 : (!strcmp(hv, "mfma2") ? 2 : 0);
        return BSE_OK;
    }
    const int nb = kPanel;
    const size_t es = fp32 ? sizeof(float2) : sizeof(double2);
    // a tiled n2 x n2 matrix updated below row and column nb (one panel)
    int n2 = 4096;
    while (n2 > 0 && ((size_t)tiled_elems(n2) * es > scratch_bytes || (int64_t)n2 * nb > vw_elems))
        n2 -= 64;
    if (!ctx->autotune || n2 < 512) { ctx->her2k_valu[fp32] = 1; return BSE_OK; }
    hipStream_t st = ctx->st;
    HCK(hipMemsetAsync(scratch, 0, (size_t)tiled_elems(n2) * es, st));
    HCK(hipMemsetAsync(V, 0, sizeof(double2) * n2 * nb, st));
    HCK(hipMemsetAsync(W, 0, sizeof(double2) * n2 * nb, st));
    HCK(hipMemsetAsync(alpha2, 0, sizeof(double2) * 64, st));
    float best = 0.f;
    int choice = 1;
    for (int kind : {1, 2, 0}) {
        HCK(bse_launch_her2k(st, fp32, kind, n2, nb, scratch, V, W, n2, 0, nb, alpha2));
        HCK(hipEventRecord(ctx->tev[0], st));
        for (int r = 0; r < 4; ++r) HCK(bse_launch_her2k(st, fp32, kind, n2, nb, scratch, V, W, n2, 0, nb, alpha2));
        HCK(hipEventRecord(ctx->tev[1], st));
        HCK(hipEventSynchronize(ctx->tev[1]));
        float ms = 0.f;
        HCK(hipEventElapsedTime(&ms, ctx->tev[0], ctx->tev[1]));
        if (kind == 1 || ms < 0.97f * best) { best = ms; choice = kind; }
    }
    ctx->her2k_valu[fp32] = choice;
    return BSE_OK;
}

// ---------------------------------------------------------------------------
// Two-stage reduction to the real tridiagonal (d, e) in S_D, S_E:
// dense -> band of half-width kBand (Householder panels; band_kernels.hip),
// band -> tridiagonal (bulge chasing; chase_kernels.hip).  With rho,
// y = Q^H rho lands in S_Y (Q1^H by panels, Q2^H inside the chase).  The
// band matrix is kept in S_BAB0 for inverse iteration on it, the panel
// reflectors in S_VSTORE / S_TALL (ctx->band_panels).
// ---------------------------------------------------------------------------
static const int kBand = 32;

// Automatic choice (two_stage == 0) when at most 20 vectors are wanted: the
// two-stage reduction for FP64 storage from order 256 up and for FP32
// storage from order 4096 up (measured at N = 20736: FP64 lowest 16 in
// 3.0 s against 5.9 s one-stage; FP32 storage 1.8 s against 3.1-3.6 s, the
// one-stage hemv being bound by streaming the trailing block once per
// column).  BSE_TWO_STAGE=1 / =0 forces it on / off for any size.
static bool use_two_stage(const bse_ctx* ctx, int n, int n_vectors, int max_vectors, int fp32) {
    const int want = (n_vectors < 0) ? n : std::max(n_vectors, max_vectors);
    if (want > 20 || ctx->two_stage < 0 || n < 2) return false;
    return ctx->two_stage > 0 || n >= (fp32 ? 4096 : 256);
}

static int twostage_tridiag(bse_ctx* ctx, int n, void* H, int fp32, const double2* rho) {
    hipStream_t st = ctx->st;
    const int b = kBand, ldab = 2 * b + 1;
    double* d = BUF(double, S_D, n);
    double* e = BUF(double, S_E, std::max(n - 1, 1));
    double2* AB = BUF(double2, S_BAB, (size_t)n * ldab);
    double2* AB0 = BUF(double2, S_BAB0, (size_t)n * ldab);
    HCK(bse_launch_zero_words(st, 2 * (int64_t)n * ldab, AB));
    ctx->band_panels.clear();
    int64_t voff = 0;
    for (int c0 = 0;; c0 += b) {
        const int s = c0 + b, m = n - s;
        if (m < 2) break;
        ctx->band_panels.push_back(bse_ctx::BandPanel{s, m, std::min(m, b), voff});
        voff += (int64_t)m * b;
    }
    const int npan = (int)ctx->band_panels.size();
    double2* Vst = BUF(double2, S_VSTORE, std::max<int64_t>(voff, 1));
    double2* Tall = BUF(double2, S_TALL, (size_t)std::max(npan, 1) * b * b);
    double2* zeros = BUF(double2, S_BZERO, 64);
    HCK(bse_launch_zero_words(st, 128, zeros));
    hipEvent_t* tv = ctx->tev;
    double2* y = nullptr;
    if (npan > 0) {
        const int m0 = n - b;
        double2* P = BUF(double2, S_BPAN, (size_t)m0 * b);
        double2* tw = BUF(double2, S_BTSQR, bse_tsqr_workspace(m0, b));
        const int64_t uel = bse_band_update_workspace(n, b, b);
        double2* uw = BUF(double2, S_BUPD, uel);
        // FP32 storage: the block product and the rank-2k update on FP32
        // matrix cores (kind 3), whose rounding matches the storage's; FP64
        // storage (or BSE_BAND_FP64PROD=1): FP64 products, kind as tuned
        const char* fp64prod = getenv("BSE_BAND_FP64PROD");
        const bool mfma32 = fp32 && !(fp64prod && fp64prod[0] == '1');
        if (!mfma32) {
            int rc = autotune_her2k(ctx, fp32, uw, sizeof(double2) * (size_t)uel, Vst, P,
                                    std::min<int64_t>(voff, (int64_t)m0 * b), zeros);
            if (rc) return rc;
        }
        HCK(bse_launch_zero_words(st, 128, zeros));
        HCK(hipEventRecord(tv[0], st));
        const int kind = mfma32 ? 3 : (ctx->her2k_valu[fp32] >= 0 ? ctx->her2k_valu[fp32] : 1);
        ctx->band_fp32prod = mfma32;
        hipEvent_t* pev = nullptr;
        if (ctx->prof) {
            while ((int64_t)ctx->prof_ev.size() < 4 * (int64_t)npan) {
                hipEvent_t ev;
                HCK(hipEventCreate(&ev));
                ctx->prof_ev.push_back(ev);
            }
            pev = ctx->prof_ev.data();
        }
        // Look-ahead (trailing-update kinds on 64 x 64 tiles): once Z of panel
        // p is known, the side stream updates the trailing block's first tile
        // column (the next panel's columns), factors panel p + 1 and applies
        // its reflectors to y = Q^H rho, while st updates every other tile;
        // st then waits for the factorisation before the next block product.
        // BSE_BAND_LOOKAHEAD=0 keeps everything on st.
        const char* lae = getenv("BSE_BAND_LOOKAHEAD");
        const bool look = (kind == 0 || kind == 1 || kind == 3) && !(lae && lae[0] == '0') && npan > 1;
        ctx->band_lookahead = look;
        // FP32 matrix-core update: its operands V and Z packed once per panel
        // in FP32 planes (BSE_HER2K_PACK=0 reads the FP64 panels per tile)
        const char* pke = getenv("BSE_HER2K_PACK");
        const bool pack = look && kind == 3 && !(pke && pke[0] == '0');
        ctx->band_pack = pack;
        hipStream_t ss = ctx->side;
        double2* qw = rho ? BUF(double2, S_W1, bse_apply_q_thin_workspace(n)) : nullptr;
        y = rho ? BUF(double2, S_Y, n) : nullptr;
        // First attempt: CholeskyQR2 panels only.  A panel whose Gram matrix is
        // not numerically positive definite sets *fail; the matrix, saved
        // before, is then restored and the reduction repeated with the TSQR
        // kernels behind every CholeskyQR2 panel (flag-gated, so only the
        // flagged panels run them).  One host read of *fail decides.
        int* qr_fail = BUF(int, S_BFAIL, 4);
        const size_t hbytes = (size_t)tiled_elems(n) * (fp32 ? sizeof(float2) : sizeof(double2));
        void* H0 = BUF(char, S_BH0, hbytes);
        HCK(hipMemcpyAsync(H0, H, hbytes, hipMemcpyDeviceToDevice, st));
        int tsqr = 0;
        // factor panel q on stream sq; apply its reflectors to y = Q^H rho
        auto factor = [&](hipStream_t sq, int q) -> int {
            const bse_ctx::BandPanel& bq = ctx->band_panels[q];
            int kq = 0;
            HCK(bse_band_gather(sq, fp32, n, q * b, b, H, P, AB, ldab));
            HCK(bse_band_panel_qr(sq, bq.m, b, P, Vst + bq.voff, Tall + (size_t)q * b * b, tw, AB, ldab, q * b, &kq,
                                  qr_fail, tsqr));
            if (kq != bq.k) return fail(ctx, BSE_ERR_STATE, "band panel sizes disagree");
            return BSE_OK;
        };
        auto apply_y = [&](hipStream_t sq, int q) -> int {
            const bse_ctx::BandPanel& bq = ctx->band_panels[q];
            if (rho)
                HCK(bse_apply_q_panel(sq, bq.m, bq.k, Vst + bq.voff, Tall + (size_t)q * b * b, b, 1, y + bq.s, n, 1, qw));
            return BSE_OK;
        };
        for (int attempt = 0; attempt < 2; ++attempt) {
            tsqr = attempt;
            if (attempt == 1) {
                HCK(hipMemcpyAsync(H, H0, hbytes, hipMemcpyDeviceToDevice, st));
                HCK(bse_launch_zero_words(st, 2 * (int64_t)n * ldab, AB));
            }
            HCK(bse_launch_zero_words(st, 2, qr_fail));
            if (rho) HCK(hipMemcpyAsync(y, rho, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
            if (int rc = factor(st, 0)) return rc;
            if (int rc = apply_y(st, 0)) return rc;
            for (int p = 0; p < npan; ++p) {
                const bse_ctx::BandPanel& bp = ctx->band_panels[p];
                double2* Vp = Vst + bp.voff;
                double2* Tp = Tall + (size_t)p * b * b;
                hipEvent_t* pe = pev ? pev + 4 * p : nullptr;
                if (!look || p + 1 == npan) {
                    HCK(bse_band_update(st, fp32, kind, n, bp.s, bp.k, b, H, Vp, Tp, uw, zeros, pe));
                    if (p + 1 < npan) {
                        if (int rc = factor(st, p + 1)) return rc;
                        if (int rc = apply_y(st, p + 1)) return rc;
                    }
                    continue;
                }
                double2* Z = nullptr;
                float* planes = nullptr;
                HCK(bse_band_product(st, fp32, kind, n, bp.s, bp.k, b, H, Vp, Tp, uw, pe, &Z,
                                     pack ? &planes : nullptr));
                // FP32 storage on matrix cores: both parts read V and Z from FP32 planes
                const int pk = pack ? 4 : kind;
                const double2* pv = pack ? reinterpret_cast<const double2*>(planes) : Vp;
                HCK(hipEventRecord(ctx->la_ev[0], st));
                HCK(hipStreamWaitEvent(ss, ctx->la_ev[0], 0));
                HCK(bse_launch_her2k_part(ss, fp32, pk, 1, n, bp.s, H, pv, Z, bp.m, bp.s, bp.k, zeros));
                if (int rc = factor(ss, p + 1)) return rc;
                HCK(hipEventRecord(ctx->la_ev[1], ss));
                // y's update runs on behind the factorisation, beside st's next
                // block product (it reads only this panel's V and T)
                if (int rc = apply_y(ss, p + 1)) return rc;
                HCK(bse_launch_her2k_part(st, fp32, pk, 2, n, bp.s, H, pv, Z, bp.m, bp.s, bp.k, zeros));
                if (pe) HCK(hipEventRecord(pe[3], st));
                HCK(hipStreamWaitEvent(st, ctx->la_ev[1], 0));
            }
            // join the side stream's last update of y
            if (look) {
                HCK(hipEventRecord(ctx->la_ev[0], ss));
                HCK(hipStreamWaitEvent(st, ctx->la_ev[0], 0));
            }
            int failed = 0;
            HCK(hipMemcpyAsync(&failed, qr_fail, sizeof(int), hipMemcpyDeviceToHost, st));
            HCK(hipStreamSynchronize(st));
            if (!failed) break;
        }
    }
    if (npan == 0) HCK(hipEventRecord(tv[0], st));
    HCK(bse_band_gather_tail(st, fp32, n, npan * b, H, AB, ldab));
    HCK(bse_launch_copy_words(st, 2 * (int64_t)n * ldab, AB, AB0));
    if (rho && npan == 0) {
        y = BUF(double2, S_Y, n);
        HCK(hipMemcpyAsync(y, rho, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
    }
    HCK(hipEventRecord(tv[1], st));
    const int64_t ncnt = bse_chase_counters(n);
    int* prog = BUF(int, S_BPROG, ncnt + 4);
    HCK(bse_launch_zero_words(st, (ncnt + 4) / 2, prog));
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) ncu = 64;
    // one workgroup per CU at most (all resident, so every waiter's
    // predecessor runs); BSE_CHASE_GRID sets the count for measurements
    int cgrid = std::max(1, ncu - 16);
    if (const char* g = getenv("BSE_CHASE_GRID")) cgrid = std::max(1, std::min(atoi(g), ncu));
    HCK(bse_launch_chase(st, n, b, AB, ldab, y, d, e, prog, prog + ncnt, cgrid));
    HCK(hipEventRecord(tv[2], st));
    int err = 0;
    HCK(hipMemcpyAsync(&err, prog + ncnt, sizeof(int), hipMemcpyDeviceToHost, st));
    HCK(hipStreamSynchronize(st));
    float ms1 = 0.f, ms2 = 0.f;
    HCK(hipEventElapsedTime(&ms1, tv[0], tv[1]));
    HCK(hipEventElapsedTime(&ms2, tv[1], tv[2]));
    ctx->band_ms = ms1;
    ctx->chase_ms = ms2;
    ctx->stage_ms[ST_BAND] = ms1;
    ctx->stage_ms[ST_CHASE] = ms2;
    if (ctx->prof && npan > 0) {
        // per-launch times of the two panel-update kernels; algorithmic
        // flops: X = A22 V costs 8 m^2 k, A22 -= V Z^H + Z V^H on the lower
        // triangle 8 m (m + 1) k (complex multiply-add = 8 real flops)
        double t2k = 0.0, f2k = 0.0, thm = 0.0, fhm = 0.0;
        for (int p = 0; p < npan; ++p) {
            const bse_ctx::BandPanel& bp = ctx->band_panels[p];
            float a = 0.f, c = 0.f;
            HCK(hipEventElapsedTime(&a, ctx->prof_ev[4 * p], ctx->prof_ev[4 * p + 1]));
            HCK(hipEventElapsedTime(&c, ctx->prof_ev[4 * p + 2], ctx->prof_ev[4 * p + 3]));
            thm += a;
            t2k += c;
            const double m = bp.m, k = bp.k;
            fhm += 8.0 * m * m * k;
            // with the look-ahead the timed launch skips the first tile column
            const double mr = (ctx->band_lookahead && p + 1 < npan) ? (double)std::max(0, n - 64 * ((bp.s >> 6) + 1)) : m;
            f2k += 8.0 * mr * (mr + 1.0) * k;
        }
        ctx->hemm_ms = thm;
        ctx->hemm_flops = fhm;
        ctx->b2k_ms = t2k;
        ctx->b2k_flops = f2k;
        ctx->band_launches = npan;
        ctx->prof_two_stage = true;
    }
    if (err) return fail(ctx, BSE_ERR_NUMERIC, "band to tridiagonal: a sweep waited too long on its predecessor");
    return BSE_OK;
}

// Eigenvectors of the lowest nstore eigenvalues (d_lam, ascending) after a
// two-stage reduction: inverse iteration on the band matrix, then Z = Q1 X.
// prep (may be null): recorded once bse_apply_q_grouped_prepare has formed
// the T of every four band panels in gws (panel offsets at voff_d); Z = Q1 X
// then runs four panels per pass, else one panel and 16 columns per pass.
static int twostage_vectors(bse_ctx* ctx, hipStream_t st, int n, const double* d_lam, int nstore, double tnorm,
                            double2* Z, hipEvent_t prep = nullptr, const int64_t* voff_d = nullptr,
                            double2* gws = nullptr, int64_t gws_el = 0) {
    const int b = kBand, ldab = 2 * b + 1;
    double2* ws = BUF(double2, S_BII, bse_band_ii_workspace(n, b, nstore));
    HCK(bse_launch_band_ii(st, n, b, ldab, (const double2*)ctx->bufs[S_BAB0].p, d_lam, nstore, tnorm, ws, Z, n));
    // the band vectors are formed: columns beyond the first 16 may be
    // back-transformed by another stream (twostage_vectors_rest)
    HCK(hipEventRecord(ctx->la_ev[1], st));
    const double2* Vst = (const double2*)ctx->bufs[S_VSTORE].p;
    const double2* Tall = (const double2*)ctx->bufs[S_TALL].p;
    if (prep) {
        HCK(hipStreamWaitEvent(st, prep, 0));
        HCK(bse_apply_q_grouped_run(st, n, (int)ctx->band_panels.size(), b, 0, Vst, voff_d, Z, n, nstore, gws,
                                    gws_el));
        return BSE_OK;
    }
    double2* qw = BUF(double2, S_W1, bse_apply_q_thin_workspace(n));
    // the first 16 columns here; the rest by twostage_vectors_rest on the
    // stream that joins this one (it would otherwise only wait)
    const int kc = std::min(16, nstore);
    for (int p = (int)ctx->band_panels.size() - 1; p >= 0; --p) {
        const bse_ctx::BandPanel& bp = ctx->band_panels[p];
        HCK(bse_apply_q_panel(st, bp.m, bp.k, Vst + bp.voff, Tall + (size_t)p * b * b, b, kc, Z + bp.s, n, 0, qw));
    }
    return BSE_OK;
}

// Z = Q1 X for the columns 16 .. nstore - 1 of the band vectors, on stream st
// once twostage_vectors' band inverse iteration is done (its own workspace,
// so the two streams' column groups run side by side; each group's result
// is the same as in one stream).
static int twostage_vectors_rest(bse_ctx* ctx, hipStream_t st, int n, int nstore, double2* Z) {
    if (nstore <= 16) return BSE_OK;
    const int b = kBand;
    HCK(hipStreamWaitEvent(st, ctx->la_ev[1], 0));
    double2* qw = BUF(double2, S_W2, bse_apply_q_thin_workspace(n));
    const double2* Vst = (const double2*)ctx->bufs[S_VSTORE].p;
    const double2* Tall = (const double2*)ctx->bufs[S_TALL].p;
    for (int p = (int)ctx->band_panels.size() - 1; p >= 0; --p) {
        const bse_ctx::BandPanel& bp = ctx->band_panels[p];
        for (int c0 = 16; c0 < nstore; c0 += 16) {
            const int kc = std::min(16, nstore - c0);
            HCK(bse_apply_q_panel(st, bp.m, bp.k, Vst + bp.voff, Tall + (size_t)p * b * b, b, kc,
                                  Z + bp.s + (int64_t)c0 * n, n, 0, qw));
        }
    }
    return BSE_OK;
}

// guard_tol > 0: eigenvalues within guard_tol above the last requested one
// are formed as well, up to max_vectors in total (ctx->nvec reports the count).
// vec_first > 0: the vectors formed are those of the sorted eigenvalues
// vec_first .. vec_first + n_vectors - 1 (one-stage reduction, no guard).
static int eig_pipeline(bse_ctx* ctx, int64_t n64, void* H, int fp32,
                        const double2* rho, int n_vectors, double guard_tol = 0.0,
                        int max_vectors = 0, int vec_first = 0)
{
    const int n = (int)n64;
    const int nb = kPanel;
    hipStream_t st = ctx->st;
    if (vec_first < 0 || vec_first > n) return fail(ctx, BSE_ERR_ARG, "first vector index out of range");
    const bool two = vec_first == 0 && use_two_stage(ctx, n, n_vectors, max_vectors, fp32);
    ctx->last_two_stage = two;
    ctx->stage_ms[ST_BAND] = ctx->stage_ms[ST_CHASE] = 0.0;
    if (ctx->prof) ctx->prof_two_stage = false;
    // ---------------- K5: tridiagonalisation
    int64_t vs_el = 1, yr_el = 1, yc_el = 1, part_d = 1;
    std::vector<int64_t> voff;
    int npan = 0;
    double2* Vst = nullptr;
    double2* Tall = nullptr;
    double2* skw = nullptr;
    int64_t sk_el = 0;
    double* d = nullptr;
    double* e = nullptr;
    if (two) {
        stage_begin(ctx, ST_TRD);
        int rc2 = twostage_tridiag(ctx, n, H, fp32, rho);
        if (rc2) return rc2;
        d = (double*)ctx->bufs[S_D].p;
        e = (double*)ctx->bufs[S_E].p;
        sk_el = (int64_t)64 * 1024 * 1024 / sizeof(double2);
        skw = BUF(double2, S_SPLITK, std::min<int64_t>(sk_el, std::max<int64_t>((int64_t)64 * nb * 64, (int64_t)nb * nb * 512)));
        sk_el = (int64_t)(ctx->bufs[S_SPLITK].cap / sizeof(double2));
        stage_end(ctx, ST_TRD);
    } else {
    bse_hetrd_workspace(n, nb, &vs_el, &yr_el, &yc_el, &part_d);
    bse_hetrd_workspace(n, nb, &vs_el, &yr_el, &yc_el, &part_d);
    {
        int64_t off = 0;
        for (int j0 = 0; j0 < n - 1; j0 += nb) { voff.push_back(off); off += (int64_t)(n - j0) * nb; }
    }
    npan = (int)voff.size();
    d = BUF(double, S_D, n);
    e = BUF(double, S_E, std::max(n - 1, 1));
    double2* tauh = BUF(double2, S_TAUH, std::max(n - 1, 1));
    Vst = BUF(double2, S_VSTORE, vs_el);
    double2* Wp = BUF(double2, S_WP, (size_t)n * nb);
    double2* acol = BUF(double2, S_ACOL, n);
    double2* Yr = BUF(double2, S_YR, yr_el);
    double2* Yc = BUF(double2, S_YC, yc_el);
    double2* t12 = BUF(double2, S_T12, 128);
    double* part = BUF(double, S_PART, part_d);
    double* scal = BUF(double, S_SCAL, 8);
    unsigned int* counter = BUF(unsigned int, S_COUNTER, 16);
    HCK(hipMemsetAsync(counter, 0, 16 * sizeof(unsigned int), st));
    stage_begin(ctx, ST_TRD);
    hipEvent_t* pev = nullptr;
    if (ctx->prof && n > 1) {
        while ((int64_t)ctx->prof_ev.size() < 2 * (int64_t)(n - 1)) {
            hipEvent_t ev;
            HCK(hipEventCreate(&ev));
            ctx->prof_ev.push_back(ev);
        }
        pev = ctx->prof_ev.data();
    }
    Tall = BUF(double2, S_TALL, (size_t)std::max(npan, 1) * nb * nb);
    double2* Gram = BUF(double2, S_GRAM, (size_t)std::max(npan, 1) * 64 * 64);   // one block per panel
    double2* alpha2 = BUF(double2, S_ALPHA2, 64);
    {
        int rc = autotune_hemv(ctx, fp32, n, H, acol, Yr, Yc);
        if (rc) return rc;
        rc = autotune_her2k(ctx, fp32, Yr, sizeof(double2) * (size_t)yr_el, Vst, Wp,
                            std::min<int64_t>(vs_el, (int64_t)n * nb), alpha2);
        if (rc) return rc;
    }
    double* hsys = BUF(double, S_HSYS, 32);   // one HetrdSys (device copy)
    HCK(bse_hetrd(st, fp32, n, H, nb, d, e, tauh, Vst, Tall, Wp, alpha2, Gram, acol, Yr, Yc, t12,
                  part, scal, counter, pev, ctx->hemv_ch[fp32], ctx->her2k_valu[fp32] >= 0 ? ctx->her2k_valu[fp32] : 1,
                  hsys));
    if (pev) {
        HCK(hipStreamSynchronize(st));
        double ms = 0.0, bytes = 0.0;
        const double es = fp32 ? 8.0 : 16.0;
        for (int c = 0; c < n - 1; ++c) {
            float t = 0.f;
            HCK(hipEventElapsedTime(&t, pev[2 * c], pev[2 * c + 1]));
            ms += t;
            const double nt = (double)(n - c - 1);
            bytes += es * nt * (nt + 1.0) * 0.5;   // lower triangle of the trailing block
        }
        ctx->hemv_ms = ms;
        ctx->hemv_bytes = bytes;
        ctx->hemv_launches = n - 1;
    }
    const int64_t skws = (int64_t)64 * 1024 * 1024 / sizeof(double2) * 8;   // 512 MiB split-K workspace cap
    sk_el = std::min<int64_t>(skws, std::max<int64_t>((int64_t)64 * nb * std::max(n_vectors, 64), (int64_t)nb * nb * 512));
    skw = BUF(double2, S_SPLITK, sk_el);
    stage_end(ctx, ST_TRD);
    }

    // ---------------- K6: tridiagonal eigenvalues (host bookkeeping + bisection)
    stage_begin(ctx, ST_TEVAL);
    std::vector<double> hd(n), he(std::max(n - 1, 0));
    HCK(hipMemcpyAsync(hd.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (n > 1) HCK(hipMemcpyAsync(he.data(), e, sizeof(double) * (n - 1), hipMemcpyDeviceToHost, st));
    HCK(hipStreamSynchronize(st));
    if (!finite_all(hd.data(), hd.size()) || !finite_all(he.data(), he.size()))
        return fail(ctx, BSE_ERR_NUMERIC, "non-finite tridiagonal (input matrix not finite?)");
    // ---------------- y = Q^H rho for the oscillator contraction, on the side
    // stream: it needs only the reflectors, so it overlaps the host's block
    // bookkeeping and the bisection; st waits for it before the first
    // contraction.  Nothing on st touches y, w1, w2 or skw in between.
    double2* y = nullptr;
    if (rho && two) {
        y = (double2*)ctx->bufs[S_Y].p;     // Q^H rho came out of the chase
    } else if (rho) {
        y = BUF(double2, S_Y, n);
        double2* w1 = BUF(double2, S_W1, (size_t)nb * 64);
        double2* w2 = BUF(double2, S_W2, (size_t)nb * 64);
        hipStream_t ss = ctx->side;
        HCK(hipEventRecord(ctx->fork_ev, st));
        HCK(hipStreamWaitEvent(ss, ctx->fork_ev, 0));
        HCK(hipMemcpyAsync(y, rho, sizeof(double2) * n, hipMemcpyDeviceToDevice, ss));
        if (bse_apply_q_thin_workspace(n) <= sk_el) {
            // one vector: the device-wide thin application (three short kernels per panel)
            HCK(bse_apply_q_thin(ss, n, nb, Vst, Tall, 1, y, n, 1, skw));
        } else for (int p = 0; p < npan; ++p) {
            const int j0 = p * nb, m = n - j0, nbp = std::min(nb, n - 1 - j0);
            const double2* Vp = Vst + voff[p];
            const double2* Tp = Tall + (size_t)p * nb * nb;
            HCK(bse_zgemm(ss, 1, 0, nbp, 1, m, 1, 0, Vp, m, y + j0, n, 0, 0, w1, nbp, skw, sk_el));
            HCK(bse_zgemm(ss, 1, 0, nbp, 1, nbp, 1, 0, Tp, nb, w1, nbp, 0, 0, w2, nbp, nullptr, 0));
            HCK(bse_zgemm(ss, 0, 0, m, 1, nbp, -1, 0, Vp, m, w2, nbp, 1, 0, y + j0, n, nullptr, 0));
        }
        HCK(hipEventRecord(ctx->join_ev, ss));
    }
    double tnorm = 0.0, emax2 = 0.0;
    for (int i = 0; i < n; ++i) {
        double r = fabs(hd[i]);
        if (i > 0) r += fabs(he[i - 1]);
        if (i < n - 1) r += fabs(he[i]);
        tnorm = std::max(tnorm, r);
        if (i < n - 1) emax2 = std::max(emax2, he[i] * he[i]);
    }
    if (tnorm == 0.0) tnorm = 1.0;
    const double pivmin = DBL_MIN * std::max(1.0, emax2);
    std::vector<double> he2(std::max(n - 1, 1), 0.0);
    std::vector<int> bstart;
    bstart.push_back(0);
    for (int i = 0; i < n - 1; ++i) {
        const double ee = he[i] * he[i];
        if (fabs(hd[i] * hd[i + 1]) * DBL_EPSILON * DBL_EPSILON + DBL_MIN >= ee) {
            he[i] = 0.0;
            bstart.push_back(i + 1);
        } else {
            he2[i] = ee;
        }
    }
    const int nblk = (int)bstart.size();
    std::vector<int4> items(n);
    std::vector<double2> bounds(nblk);
    for (int b = 0; b < nblk; ++b) {
        const int bs = bstart[b], be = (b + 1 < nblk) ? bstart[b + 1] : n;
        double gl = DBL_MAX, gu = -DBL_MAX;
        for (int i = bs; i < be; ++i) {
            double r = 0.0;
            if (i > bs) r += fabs(he[i - 1]);
            if (i < be - 1) r += fabs(he[i]);
            gl = std::min(gl, hd[i] - r);
            gu = std::max(gu, hd[i] + r);
        }
        const double fud = 2.1 * DBL_EPSILON * tnorm * (be - bs) + 4.2 * pivmin;
        bounds[b] = make_double2(gl - fud, gu + fud);
        for (int i = bs; i < be; ++i) items[i] = make_int4(bs, be - bs, i - bs, b);
    }
    double* d_e2 = BUF(double, S_E2, std::max(n - 1, 1));
    int4* d_items = BUF(int4, S_ITEMS, n);
    double2* d_bounds = BUF(double2, S_BOUNDS, nblk);
    double* d_lami = BUF(double, S_LAMI, n);
    HCK(hipMemcpyAsync(d_e2, he2.data(), sizeof(double) * he2.size(), hipMemcpyHostToDevice, st));
    HCK(hipMemcpyAsync(e, he.data(), sizeof(double) * std::max(n - 1, 0), hipMemcpyHostToDevice, st));
    HCK(hipMemcpyAsync(d_items, items.data(), sizeof(int4) * n, hipMemcpyHostToDevice, st));
    HCK(hipMemcpyAsync(d_bounds, bounds.data(), sizeof(double2) * nblk, hipMemcpyHostToDevice, st));
    HCK(bse_launch_bisect(st, n, d_items, d_bounds, d, d_e2, pivmin, tnorm, d_lami));
    std::vector<double> lami(n);
    HCK(hipMemcpyAsync(lami.data(), d_lami, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    HCK(hipStreamSynchronize(st));
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return lami[a] < lami[b]; });
    std::vector<double> lam(n);
    std::vector<int> bsv(n), nbv(n);
    for (int i = 0; i < n; ++i) {
        lam[i] = lami[ord[i]];
        bsv[i] = items[ord[i]].x;
        nbv[i] = items[ord[i]].y;
    }
    double* d_lam = BUF(double, S_LAM, n);
    int* d_bsv = BUF(int, S_BSV, n);
    int* d_nbv = BUF(int, S_NBV, n);
    HCK(hipMemcpyAsync(d_lam, lam.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
    HCK(hipMemcpyAsync(d_bsv, bsv.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    HCK(hipMemcpyAsync(d_nbv, nbv.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    stage_end(ctx, ST_TEVAL);

    // ---------------- K6: tridiagonal eigenvectors by batches
    stage_begin(ctx, ST_TVEC);
    int nstore = (n_vectors < 0) ? n - vec_first : std::min(n_vectors, n - vec_first);
    if (guard_tol < 0.0) guard_tol = 1e-4 * tnorm;
    if (vec_first == 0 && nstore > 0 && nstore < n && guard_tol > 0.0) {
        // guard vectors: eigenvalues just above the last requested one that
        // lie within guard_tol of it (a cluster straddling the cut) are formed
        // too, so that a refinement sees the whole cluster
        while (nstore < n && nstore < max_vectors && lam[nstore] - lam[nstore - 1] <= guard_tol) ++nstore;
    }
    ctx->hnorm = tnorm;
    const double tolc = 1e-6 * tnorm;
    // Clusters of close eigenvalues inside one unreduced block.  The global
    // ascending order interleaves the blocks, so each block's eigenvalues are
    // walked on their own: memb = sorted positions ordered by (block start,
    // position); a cluster is a run of memb with gaps <= tolc, given as
    // (offset into memb, size).  Its members span [memb[off], memb[off+size-1]]
    // in the global order.
    std::vector<int> memb(n);
    for (int i = 0; i < n; ++i) memb[i] = i;
    std::stable_sort(memb.begin(), memb.end(), [&](int a, int b) { return bsv[a] < bsv[b]; });
    std::vector<int2> clall;
    std::vector<int> nocut(n + 1, 0);   // difference array: cut points inside a cluster's span
    for (int q = 0; q < n;) {
        int j = q + 1;
        while (j < n && bsv[memb[j]] == bsv[memb[q]] && lam[memb[j]] - lam[memb[j - 1]] <= tolc) ++j;
        if (j - q > 1) {
            clall.push_back(make_int2(q, j - q));
            ++nocut[memb[q] + 1];
            --nocut[memb[j - 1] + 1];
        }
        q = j;
    }
    for (int i = 1; i <= n; ++i) nocut[i] += nocut[i - 1];
    // one thread per eigenvalue; as many per batch as 24 GB of scratch allows
    const int64_t cap = std::max<int64_t>(4096, (int64_t)(24e9 / (48.0 * n)));
    int bsz = (int)std::min<int64_t>(n, cap);
    const int64_t iscr = (int64_t)5 * n * bsz;
    double* scr = BUF(double, S_IISCR, iscr);
    double* xb = BUF(double, S_IIX, (size_t)n * bsz);
    double2* Z = nullptr;
    if (nstore > 0) Z = BUF(double2, S_Z, (size_t)n * nstore);
    double2* fvec = rho ? BUF(double2, S_F, n) : nullptr;
    int2* d_cl = BUF(int2, S_CLUST, std::max<size_t>(clall.size(), 1));
    int* d_memb = BUF(int, S_MEMB, n);
    HCK(hipMemcpyAsync(d_memb, memb.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    if (rho && !two) HCK(hipStreamWaitEvent(st, ctx->join_ev, 0));   // y = Q^H rho is ready
    // two-stage: the band inverse iteration and Z = Q1 X need only the
    // eigenvalues, so they run on the side stream beside the tridiagonal
    // vectors of the oscillator contraction; st joins before the refinement
    const bool vec_side = two && nstore > 0;
    std::vector<int64_t> bvoff;
    if (vec_side) {
        HCK(hipEventRecord(ctx->fork_ev, st));
        // Z = Q1 X one panel and 16 columns per pass (k_pq_*: about 4000
        // small launches that pipeline behind each other).  BSE_BT16_GROUPED=1:
        // four band panels per pass on FP64 matrix cores, the T of every group
        // formed on st and awaited by the side stream after the band inverse
        // iteration (measured 8 ms slower at N = 20736 with 20 vectors: a
        // V^H Z launch then has about 32 workgroups).
        const int npb = (int)ctx->band_panels.size();
        hipEvent_t prep = nullptr;
        int64_t* bvoff_d = nullptr;
        double2* gws = nullptr;
        int64_t gws_el = 0;
        if (npb > 0 && getenv("BSE_BT16_GROUPED")) {
            bvoff.resize(npb);
            for (int p = 0; p < npb; ++p) bvoff[p] = ctx->band_panels[p].voff;
            bvoff_d = BUF(int64_t, S_BTVOFF, npb);
            gws_el = bse_apply_q_grouped_workspace(n, npb, nstore);
            gws = BUF(double2, S_BTWS, gws_el);
            HCK(hipMemcpyAsync(bvoff_d, bvoff.data(), sizeof(int64_t) * npb, hipMemcpyHostToDevice, st));
            HCK(bse_apply_q_grouped_prepare(st, n, npb, kBand, 0, (const double2*)ctx->bufs[S_VSTORE].p, bvoff_d,
                                            (const double2*)ctx->bufs[S_TALL].p, gws));
            HCK(hipEventRecord(ctx->la_ev[0], st));
            prep = ctx->la_ev[0];
        }
        HCK(hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
        ctx->vec_grouped = prep != nullptr;
        int rc2 = twostage_vectors(ctx, ctx->side, n, d_lam, nstore, tnorm, Z, prep, bvoff_d, gws, gws_el);
        if (rc2) return rc2;
        HCK(hipEventRecord(ctx->join_ev, ctx->side));
    }
    int m0 = 0;
    // the two-stage path forms its vectors from the band matrix: the
    // tridiagonal vectors are needed only for the oscillator contraction
    if (two && !rho) m0 = n;
    size_t ci = 0;   // clusters are visited in order of their first member
    std::vector<int2> clall_sorted = clall;
    std::sort(clall_sorted.begin(), clall_sorted.end(),
              [&](int2 a, int2 b) { return memb[a.x] < memb[b.x]; });
    while (m0 < n) {
        int m1 = std::min(n, m0 + bsz);
        // do not split a cluster across batches
        while (m1 < n && m1 > m0 && nocut[m1] > 0) --m1;
        if (m1 == m0) return fail(ctx, BSE_ERR_UNSUPPORTED, "eigenvalue cluster larger than the inverse-iteration batch");
        const int cnt = m1 - m0;
        std::vector<int2> cls;
        while (ci < clall_sorted.size() && memb[clall_sorted[ci].x] < m1) cls.push_back(clall_sorted[ci++]);
        HCK(bse_launch_inverse_iter(st, bsz, cnt, m0, d_lam, d_bsv, d_nbv, d, e, tnorm, scr, n, xb));
        if (!cls.empty()) {
            HCK(hipMemcpyAsync(d_cl, cls.data(), sizeof(int2) * cls.size(), hipMemcpyHostToDevice, st));
            HCK(bse_launch_cluster_mgs(st, bsz, (int)cls.size(), d_cl, m0, d_bsv, d_nbv, d_memb, xb));
        }
        if (rho) {
            // the inverse-iteration scratch is idle until the next batch: it holds the segment partials
            if (bse_contract_split_workspace(cnt) <= iscr)
                HCK(bse_launch_contract_split(st, bsz, cnt, m0, d_bsv, d_nbv, xb, y, fvec, scr));
            else
                HCK(bse_launch_contract(st, bsz, cnt, m0, d_bsv, d_nbv, xb, y, fvec));
        }
        // commit the batch's members of [vec_first, vec_first + nstore)
        const int lo = std::max(m0, vec_first), hi = std::min(m1, vec_first + nstore);
        if (!two && hi > lo)
            HCK(bse_launch_commit(st, n, bsz, cnt - (lo - m0), hi - lo, lo, d_bsv, d_nbv, xb + (lo - m0), Z, n,
                                  lo - vec_first));
        HCK(hipStreamSynchronize(st));   // keeps d_cl and the host vectors alive
        m0 = m1;
    }
    stage_end(ctx, ST_TVEC);

    // ---------------- K7: back-transform Z <- Q Z
    stage_begin(ctx, ST_BT);
    if (vec_side) {
        if (!ctx->vec_grouped) {
            int rc3 = twostage_vectors_rest(ctx, st, n, nstore, Z);
            if (rc3) return rc3;
        }
        HCK(hipStreamWaitEvent(st, ctx->join_ev, 0));
    } else if (nstore > 0 && nstore <= 16 && bse_apply_q_thin_workspace(n) <= sk_el) {
        HCK(bse_apply_q_thin(st, n, nb, Vst, Tall, nstore, Z, n, 0, skw));
    } else if (nstore >= 2048 && nb == 32 && !getenv("BSE_BT_STRIPS")) {
        // many vectors: four panels per pass on FP64 matrix cores
        int64_t* voff_d = BUF(int64_t, S_BTVOFF, npan);
        HCK(hipMemcpyAsync(voff_d, voff.data(), sizeof(int64_t) * npan, hipMemcpyHostToDevice, st));
        const int64_t wel = bse_apply_q_grouped_workspace(n, npan, nstore);
        double2* bws = BUF(double2, S_BTWS, wel);
        HCK(bse_apply_q_grouped(st, n, nb, npan, Vst, voff_d, Tall, Z, n, nstore, bws, wel));
    } else if (nstore >= 2048) {
        // many vectors: one fused launch per panel over 32-column strips
        // (at least 64 strips, so the launch fills the chip)
        HCK(bse_apply_q_strips(st, n, nb, npan, Vst, voff.data(), Tall, Z, n, nstore));
    } else if (nstore > 0) {
        double2* w1 = BUF(double2, S_W1, (size_t)nb * nstore);
        double2* w2 = BUF(double2, S_W2, (size_t)nb * nstore);
        for (int p = npan - 1; p >= 0; --p) {
            const int j0 = p * nb, m = n - j0, nbp = std::min(nb, n - 1 - j0);
            const double2* Vp = Vst + voff[p];
            const double2* Tp = Tall + (size_t)p * nb * nb;
            HCK(bse_zgemm(st, 1, 0, nbp, nstore, m, 1, 0, Vp, m, Z + j0, n, 0, 0, w1, nbp, skw, sk_el));
            HCK(bse_zgemm(st, 0, 0, nbp, nstore, nbp, 1, 0, Tp, nb, w1, nbp, 0, 0, w2, nbp, nullptr, 0));
            HCK(bse_zgemm(st, 0, 0, m, nstore, nbp, -1, 0, Vp, m, w2, nbp, 1, 0, Z + j0, n, nullptr, 0));
        }
    }
    stage_end(ctx, ST_BT);
    double* ev = BUF(double, S_EVAL, n);   // eigenvalues live in the S_LAM slot; keep a stable copy
    HCK(hipMemcpyAsync(ev, d_lam, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    HCK(hipStreamSynchronize(st));
    ctx->n_eig = n;
    ctx->nvec = nstore;
    ctx->have_eig = true;
    ctx->have_f = (rho != nullptr);
    return BSE_OK;
}

// ---------------------------------------------------------------------------
// K8: FP64 Rayleigh-Ritz refinement of the lowest k eigenpairs against the
// FP64 matrix recomputed on the fly (no stored FP64 matrix).  Basis
// S = [X, R] with R the normalised residuals H X - X diag(x^H H x); Ritz
// pairs of S by the SVQB-stabilised small eigenproblem; iters rounds.
// ---------------------------------------------------------------------------
static int apply_H(bse_ctx* ctx, int k, const double2* X, int64_t ldx, double2* Y, int64_t ldy) {
    const bse_system& s = ctx->sys;
    const double2* coef = (const double2*)ctx->bufs[S_COEF].p;
    const double alpha = -s.direct_scale / (ctx->cell * ctx->nk);
    const double beta = s.f_s * s.exchange_scale / (ctx->cell * ctx->nk);
    if (!ctx->force_valu_assembly && bse_assemble_mfma_supported(s.nv, s.nc, ctx->norb)) {
        // matrix cores: tiles of H formed in LDS and multiplied at once, at
        // most 32 columns per launch (one H tile serves them all)
        double2* ws = BUF(double2, S_APPLYP, bse_apply_mfma_workspace(ctx->nk, s.nv, s.nc));
        for (int c0 = 0; c0 < k; c0 += 32) {
            const int kc = std::min(32, k - c0);
            HCK(bse_launch_apply_mfma(ctx->st, ctx->nk, s.n1, s.n2, ctx->n3, s.nv, s.nc, ctx->norb, coef,
                                      (const double2*)ctx->bufs[S_WTAB].p, (const double2*)ctx->bufs[S_PV].p,
                                      (const double2*)ctx->bufs[S_P].p, (const double*)ctx->bufs[S_DE].p,
                                      alpha, beta, kc, X + (int64_t)c0 * ldx, ldx, Y + (int64_t)c0 * ldy,
                                      ldy, ws));
        }
        return BSE_OK;
    }
    HCK(hipMemsetAsync(Y, 0, sizeof(double2) * ldy * k, ctx->st));
    HCK(bse_launch_apply(ctx->st, ctx->nk, s.n1, s.n2, ctx->n3, s.nv, s.nc, ctx->norb,
                         bse_apply_pick_tk(s.nv, s.nc, ctx->norb), coef, coef, coef, coef,
                         (const double2*)ctx->bufs[S_WTAB].p, (const double2*)ctx->bufs[S_PV].p,
                         (const double2*)ctx->bufs[S_P].p, (const double*)ctx->bufs[S_DE].p, alpha,
                         beta, k, X, ldx, Y, ldy, nullptr, 0, nullptr));
    return BSE_OK;
}

// Stopping rule of the refinement.  Each round is a Rayleigh-Ritz step on
// [X, R]; it stops when every requested vector's residual
// ||H x - theta x|| is below kRefineTol * ||H|| (||H|| bounded by the
// Gershgorin norm of the tridiagonal), when a round takes less than 20% off
// the largest residual (slow convergence, or the FP64 rounding floor of the
// matrix-free product), or after kRefineMaxRounds rounds.  The result is accepted if the largest
// residual is below kRefineAccept * ||H||: for a Hermitian H each refined
// value theta_j then lies within ||r_j|| of an eigenvalue.
static const double kRefineTol = 1e-11, kRefineAccept = 1e-7;
static const int kRefineMaxRounds = 10;

// k_req requested vectors, kw >= k_req formed ones (guard vectors of a cluster
// that straddles the cut); the block refined is all kw, 2 kw <= 40.
static int refine_lowest(bse_ctx* ctx, int k_req, int kw) {
    const int64_t N = ctx->N;
    const int k = kw, m = 2 * k;
    hipStream_t st = ctx->st;
    double2* Z = (double2*)ctx->bufs[S_Z].p;               // N x nvec, internal order
    double2* S = BUF(double2, S_RS, (size_t)N * m);        // [X | R]
    double2* HS = BUF(double2, S_RHS, (size_t)N * m);      // [HX | HR]
    double2* T = BUF(double2, S_RX, (size_t)N * 2 * k);    // new X, new HX
    double2* Gm = BUF(double2, S_RG, (size_t)m * m);
    double2* Hm = BUF(double2, S_RHG, (size_t)m * m);
    double2* Cm = BUF(double2, S_RC, (size_t)m * k);
    // small: [0, k) Ritz values, [k, 2k) residual norms, [2k] rank (int),
    // [2k+8, 3k+8) Rayleigh quotients
    double* small = BUF(double, S_RSMALL, 4 * (size_t)k + 16);
    double* rq = small + 2 * k + 8;
    double2* skw = (double2*)ctx->bufs[S_SPLITK].p;
    const int64_t sk_el = (int64_t)(ctx->bufs[S_SPLITK].cap / sizeof(double2));
    const double hn = ctx->hnorm > 0.0 ? ctx->hnorm : 1.0;
    HCK(hipMemcpyAsync(S, Z, sizeof(double2) * N * k, hipMemcpyDeviceToDevice, st));
    HCK(hipMemsetAsync(small + 2 * k, 0, sizeof(double) * 8, st));
    int rc = apply_H(ctx, k, S, N, HS, N);
    if (rc) return rc;
    std::vector<double> hsm(3 * (size_t)k + 8);
    double prev = 0.0;
    int rounds = 0;
    ctx->ref_status = BSE_OK;
    for (;; ++rounds) {
        // residuals and Rayleigh quotients of the current X; one host read per
        // round carries them and the rank kept by the previous Rayleigh-Ritz
        HCK(bse_launch_residual(st, N, k, S, HS, N, S + N * k, small + k, rq));
        HCK(hipMemcpyAsync(hsm.data(), small + k, sizeof(double) * (k + 8), hipMemcpyDeviceToHost, st));
        HCK(hipStreamSynchronize(st));
        int rank = 0;
        std::memcpy(&rank, &hsm[k], sizeof(int));
        if (rounds > 0 && rank < k)
            return fail(ctx, BSE_ERR_NUMERIC, "refinement basis lost rank (" + std::to_string(rank) +
                                                  " of " + std::to_string(k) + ")");
        double rmax = 0.0;
        for (int j = 0; j < k_req; ++j) {
            if (!std::isfinite(hsm[j])) return fail(ctx, BSE_ERR_NUMERIC, "non-finite refinement residual");
            rmax = std::max(rmax, hsm[j]);
        }
        ctx->ref_rnorm.assign(hsm.begin(), hsm.begin() + k);
        const bool done = rmax <= kRefineTol * hn || rounds >= kRefineMaxRounds ||
                          (rounds >= 2 && rmax > 0.8 * prev);
        prev = rmax;
        if (done) break;
        rc = apply_H(ctx, k, S + N * k, N, HS + N * k, N);
        if (rc) return rc;
        HCK(bse_zgemm(st, 1, 0, m, m, (int)N, 1, 0, S, N, S, N, 0, 0, Gm, m, skw, sk_el));
        HCK(bse_zgemm(st, 1, 0, m, m, (int)N, 1, 0, S, N, HS, N, 0, 0, Hm, m, skw, sk_el));
        HCK(bse_launch_small_rr(st, m, k, Gm, Hm, 1e-10, small, Cm, (int*)(small + 2 * k)));
        HCK(bse_zgemm(st, 0, 0, (int)N, k, m, 1, 0, S, N, Cm, m, 0, 0, T, N, nullptr, 0));
        HCK(bse_zgemm(st, 0, 0, (int)N, k, m, 1, 0, HS, N, Cm, m, 0, 0, T + N * k, N, nullptr, 0));
        HCK(hipMemcpyAsync(S, T, sizeof(double2) * N * k, hipMemcpyDeviceToDevice, st));
        HCK(hipMemcpyAsync(HS, T + N * k, sizeof(double2) * N * k, hipMemcpyDeviceToDevice, st));
    }
    ctx->ref_rounds = rounds;
    double rmax = 0.0;
    for (int j = 0; j < k_req; ++j) rmax = std::max(rmax, ctx->ref_rnorm[j]);
    HCK(hipMemcpyAsync(Z, S, sizeof(double2) * N * k, hipMemcpyDeviceToDevice, st));
    // refined eigenvalues: the Rayleigh quotients of the final vectors
    HCK(hipMemcpyAsync(ctx->bufs[S_EVAL].p, rq, sizeof(double) * k, hipMemcpyDeviceToDevice, st));
    // oscillator strengths of the refined vectors: f^T = rho^H X
    HCK(bse_zgemm(st, 1, 0, 1, k, (int)N, 1, 0, (const double2*)ctx->bufs[S_RHO].p, N, S, N, 0, 0,
                  (double2*)ctx->bufs[S_F].p, 1, skw, sk_el));
    HCK(hipStreamSynchronize(st));
    if (rmax > kRefineAccept * hn) {
        ctx->ref_status = BSE_ERR_NUMERIC;
        return fail(ctx, BSE_ERR_NUMERIC, "refinement did not converge: largest residual " +
                                              std::to_string(rmax) + " eV");
    }
    return BSE_OK;
}

// Matrix-free FP64 product Y = H X with the loaded problem (reference order).
extern "C" int bse_apply(bse_ctx* ctx, int k, const double* x, int64_t ldx, double* y, int64_t ldy) {
    if (!ctx || !x || !y || k <= 0) return BSE_ERR_ARG;
    if (!ctx->have_tables) return fail(ctx, BSE_ERR_STATE, "build the problem first");
    const int64_t N = ctx->N;
    if (ldx < N || ldy < N) return fail(ctx, BSE_ERR_ARG, "leading dimension < N");
    (void)hipSetDevice(ctx->device);
    double2* xin = BUF(double2, S_RX, (size_t)N * 2 * k);
    double2* tmp = xin + N * k;
    double2* yout = BUF(double2, S_RHS, (size_t)N * k);
    HCK(hipMemcpy2DAsync(tmp, sizeof(double2) * N, x, sizeof(double2) * ldx, sizeof(double2) * N, k,
                         hipMemcpyHostToDevice, ctx->st));
    // reference order -> internal order: xin[r'] = x[ref(r')]
    HCK(bse_launch_rows_from_reference(ctx->st, N, ctx->B, ctx->nk, k, tmp, N, xin, N, kmap_dev(ctx)));
    int rc = apply_H(ctx, k, xin, N, yout, N);
    if (rc) return rc;
    HCK(bse_launch_rows_to_reference(ctx->st, N, ctx->B, ctx->nk, k, yout, N, tmp, N, kmap_dev(ctx)));
    HCK(hipMemcpy2DAsync(y, sizeof(double2) * ldy, tmp, sizeof(double2) * N, sizeof(double2) * N, k,
                         hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    return BSE_OK;
}

// ---------------------------------------------------------------------------
// K10: Hellmann-Feynman gradients of the lowest exciton energies,
// dE_m/dtheta = A_m^H (dH/dtheta) A_m, with dH applied matrix-free.
// Parameters (column order): 0 hopping scale (all R != 0 blocks), 1 scissor,
// 2 r0, 3 eps_env, 4 exchange field h (h * diag(+1,-1,+1,-1,...) on H(R=0)).
// The TB-derived ones use first-order perturbation theory for the band
// energies and coefficients (parallel-transport gauge); the gauge part of dc
// drops out of A^H dH A.
// ---------------------------------------------------------------------------
static int grad_tb_apply(bse_ctx* ctx, int which, int k, const double2* A, double2* Y,
                         const double2* Bm = nullptr, double2* trp = nullptr) {
    const bse_system& s = ctx->sys;
    const int no = ctx->norb, nr = ctx->nr;
    std::vector<double> dh((size_t)nr * no * no * 2, 0.0);
    for (int r = 0; r < nr; ++r) {
        const double* R = &ctx->rcart_h[3 * r];
        const bool zero = fabs(R[0]) < 1e-12 && fabs(R[1]) < 1e-12 && fabs(R[2]) < 1e-12;
        for (int l = 0; l < no; ++l)
            for (int mm = 0; mm < no; ++mm) {
                const size_t e = ((size_t)r * no + l) * no + mm;
                if (which == 0 && !zero) {
                    dh[2 * e] = ctx->hr_h[2 * e];
                    dh[2 * e + 1] = ctx->hr_h[2 * e + 1];
                } else if (which == 4 && zero && l == mm) {
                    dh[2 * e] = (l % 2 == 0) ? 1.0 : -1.0;
                }
            }
    }
    double2* d_dh = BUF(double2, S_GDHR, (size_t)nr * no * no);
    double* deps = BUF(double, S_GDEPS, (size_t)ctx->nk * ctx->ns);
    double2* dcoef = BUF(double2, S_GDCOEF, (size_t)ctx->nk * ctx->ns * no);
    double2* dP = BUF(double2, S_GDP, (size_t)ctx->N * no);
    double2* dPV = BUF(double2, S_GDPV, (size_t)ctx->N * no);
    double2* drho = BUF(double2, S_GDRHO, (size_t)ctx->N);
    double* ddE = BUF(double, S_GDDE, (size_t)ctx->N);
    HCK(hipMemcpyAsync(d_dh, dh.data(), sizeof(double) * dh.size(), hipMemcpyHostToDevice, ctx->st));
    // TB derivatives (energies and coefficients of the same gauge-fixed states)
    double* eps2 = BUF(double, S_GDW, (size_t)ctx->nk * ctx->ns + (size_t)ctx->nk * ctx->ns * no * 2);
    double2* coef2 = (double2*)(eps2 + (size_t)ctx->nk * ctx->ns);
    HCK(bse_launch_tb_states(ctx->st, ctx->nk, no, nr, (const double*)ctx->bufs[S_RCART].p,
                             (const double2*)ctx->bufs[S_HR].p, (const double*)ctx->bufs[S_KPTS].p,
                             s.q, s.n_occ, s.nv, s.nc, s.scissor, d_dh, eps2, coef2, deps, dcoef));
    HCK(bse_launch_transitions(ctx->st, ctx->nk, s.nv, s.nc, no, (const double*)ctx->bufs[S_EPS].p,
                               (const double2*)ctx->bufs[S_COEF].p, deps, dcoef,
                               (const double*)ctx->bufs[S_TAU].p, ctx->qhat,
                               (const double2*)ctx->bufs[S_VX].p, dP, dPV, drho, ddE,
                               ctx->dip_nr > 0 ? (const double2*)ctx->bufs[S_DIPQ].p : nullptr));
    const double2* coef = (const double2*)ctx->bufs[S_COEF].p;
    const double2* wt = (const double2*)ctx->bufs[S_WTAB].p;
    const double alpha = -s.direct_scale / (ctx->cell * ctx->nk);
    const double beta = s.f_s * s.exchange_scale / (ctx->cell * ctx->nk);
    const int tk = bse_apply_pick_tk(s.nv, s.nc, no);
    if (!Bm) HCK(hipMemsetAsync(Y, 0, sizeof(double2) * ctx->N * k, ctx->st));
    const double2* fac[4][4] = {{dcoef, coef, coef, coef}, {coef, dcoef, coef, coef},
                                {coef, coef, dcoef, coef}, {coef, coef, coef, dcoef}};
    for (int t = 0; t < 4; ++t)
        HCK(bse_launch_apply(ctx->st, ctx->nk, s.n1, s.n2, ctx->n3, s.nv, s.nc, no, tk, fac[t][0], fac[t][1],
                             fac[t][2], fac[t][3], wt, nullptr, nullptr, nullptr, alpha, 0.0, k, A,
                             ctx->N, Y, ctx->N, Bm, ctx->N, trp));
    // exchange: beta (dP V P^H + P V dP^H); diagonal d(eps_c - eps_v)
    HCK(bse_launch_apply(ctx->st, ctx->nk, s.n1, s.n2, ctx->n3, s.nv, s.nc, no, tk, coef, coef, coef, coef, wt,
                         dPV, (const double2*)ctx->bufs[S_P].p, ddE, 0.0, beta, k, A, ctx->N, Y, ctx->N, Bm, ctx->N, trp));
    HCK(bse_launch_apply(ctx->st, ctx->nk, s.n1, s.n2, ctx->n3, s.nv, s.nc, no, tk, coef, coef, coef, coef, wt,
                         (const double2*)ctx->bufs[S_PV].p, dP, nullptr, 0.0, beta, k, A, ctx->N, Y,
                         ctx->N, Bm, ctx->N, trp));
    return BSE_OK;
}

static int grad_w_apply(bse_ctx* ctx, int which, int k, const double2* A, double2* Y,
                        const double2* Bm = nullptr, double2* trp = nullptr) {
    const bse_system& s = ctx->sys;
    const int no = ctx->norb;
    const int64_t nw = ktable_size(s.n1, s.n2, ctx->n3) * no * no;
    double2* dW = BUF(double2, S_GDW, nw);
    // derivative of the mini-BZ average (the 3D potential has no r0)
    const double da = 4.0 * M_PI * M_PI / (ctx->cell * ctx->nk);
    const double rad = sqrt(da / M_PI);
    double dwbar;
    if (which == 2 && ctx->dim == 3) {
        dwbar = 0.0;
    } else if (which == 2) {
        const double r0 = s.r0;
        dwbar = (r0 > 0.0) ? (M_PI * kC / (s.eps_env * da)) * (rad / ((1.0 + r0 * rad) * r0) - log1p(r0 * rad) / (r0 * r0))
                           : -(M_PI * kC / (s.eps_env * da)) * 0.5 * rad * rad;
    } else {
        dwbar = -ctx->wbar / s.eps_env;
    }
    HCK(bse_launch_w_table(ctx->st, s.n1, s.n2, ctx->n3, ctx->dim, no, ctx->ng,
                           (const double*)ctx->bufs[S_GVEC].p, ctx->recip,
                           (const double*)ctx->bufs[S_TAU].p, s.r0, s.eps_env, s.radius, kC, dwbar,
                           which == 2 ? 1 : 2, dW));
    const double2* coef = (const double2*)ctx->bufs[S_COEF].p;
    const double alpha = -s.direct_scale / (ctx->cell * ctx->nk);
    if (!Bm) HCK(hipMemsetAsync(Y, 0, sizeof(double2) * ctx->N * k, ctx->st));
    HCK(bse_launch_apply(ctx->st, ctx->nk, s.n1, s.n2, ctx->n3, s.nv, s.nc, no,
                         bse_apply_pick_tk(s.nv, s.nc, no), coef, coef, coef, coef, dW, nullptr,
                         nullptr, nullptr, alpha, 0.0, k, A, ctx->N, Y, ctx->N, Bm, ctx->N, trp));
    return BSE_OK;
}

extern "C" int bse_grad_energies(bse_ctx* ctx, int n_exc, const int* which, double* dE) {
    if (!ctx || !dE || n_exc <= 0) return BSE_ERR_ARG;
    if (!ctx->have_eig || !ctx->eig_on_bse) return fail(ctx, BSE_ERR_STATE, "run bse_eigh on a built system first");
    if (n_exc > ctx->nvec) return fail(ctx, BSE_ERR_STATE, "bse_eigh formed fewer eigenvectors than n_exc");
    if (ctx->sys.screening != 0 || ctx->full_eig)
        return fail(ctx, BSE_ERR_UNSUPPORTED, "gradients hold the model W fixed: TDA with model screening only");
    (void)hipSetDevice(ctx->device);
    const int64_t N = ctx->N;
    const double2* A = (const double2*)ctx->bufs[S_Z].p;
    double2* Y = BUF(double2, S_RHS, (size_t)N * n_exc);
    double* out = BUF(double, S_GOUT, n_exc);
    std::vector<double> col(n_exc);
    for (int t = 0; t < 5; ++t) {
        double* dst = dE + (size_t)t * n_exc;
        if (which && !which[t]) { for (int m = 0; m < n_exc; ++m) dst[m] = 0.0; continue; }
        if (t == 1) { for (int m = 0; m < n_exc; ++m) dst[m] = 1.0; continue; }   // dH = I
        int rc = (t == 0 || t == 4) ? grad_tb_apply(ctx, t, n_exc, A, Y) : grad_w_apply(ctx, t, n_exc, A, Y);
        if (rc) return rc;
        HCK(bse_launch_coldot(ctx->st, N, n_exc, A, Y, N, out));
        HCK(hipMemcpyAsync(dst, out, sizeof(double) * n_exc, hipMemcpyDeviceToHost, ctx->st));
        HCK(hipStreamSynchronize(ctx->st));
    }
    return BSE_OK;
}

// ---------------------------------------------------------------------------
// K10 (spectrum): gradient of the misfit L = sum_w (Im eps(w) - target(w))^2.
// With G(z) = (z - H)^-1 = A diag(g) A^H and Im eps = -pref Im(rho^H G rho):
//   dL/dtheta = -2 pref Im[ tr(dH Bm) + drho^H A sb + sa^T A^H drho ],
//   Bm = Xr U^H,  Xr = A Fbr,  U = A Fac   (factors from k_spec_factors),
// i.e. the sum over frequencies collapses into one N x N matrix Bm formed by
// three GEMMs, and tr(dH Bm) is accumulated over dH tiles recomputed on the
// fly (no stored derivative matrix).  Needs every eigenvector (bse_eigh with
// n_vectors = -1).
// ---------------------------------------------------------------------------
extern "C" int bse_grad_spectrum(bse_ctx* ctx, const double* omega, int n_omega, double eta,
                                 const double* target, const int* which, double* L, double* dL) {
    if (!ctx || !omega || !target || !dL || n_omega <= 0 || !(eta > 0.0)) return BSE_ERR_ARG;
    if (!ctx->have_eig || !ctx->eig_on_bse || !ctx->have_f)
        return fail(ctx, BSE_ERR_STATE, "run bse_eigh on a built system first");
    if (ctx->nvec != ctx->n_eig)
        return fail(ctx, BSE_ERR_STATE, "the spectrum gradient needs every eigenvector (bse_eigh(ctx, -1, 0))");
    if (ctx->dip_nr > 0)
        return fail(ctx, BSE_ERR_UNSUPPORTED, "spectrum gradients use point-orbital dipoles");
    if (ctx->sys.screening != 0 || ctx->full_eig)
        return fail(ctx, BSE_ERR_UNSUPPORTED, "gradients hold the model W fixed: TDA with model screening only");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->st;
    const int64_t N = ctx->N;
    const int nw = n_omega;
    const double2* A = (const double2*)ctx->bufs[S_Z].p;
    const double* E = (const double*)ctx->bufs[S_EVAL].p;
    const double2* f = (const double2*)ctx->bufs[S_F].p;
    // spectrum and misfit residual
    double* d_v = BUF(double, S_SGV, 3 * (size_t)nw + 8 * (size_t)N);
    double* d_om = d_v;
    double* d_re = d_v + nw;
    double* d_im = d_v + 2 * (size_t)nw;
    double2* sa = (double2*)(d_v + 3 * (size_t)nw);
    double2* sb = sa + N;
    double2* p1 = sb + N;
    double2* q = p1 + N;
    HCK(hipMemcpyAsync(d_om, omega, sizeof(double) * nw, hipMemcpyHostToDevice, st));
    const double pref = kC / (ctx->volume * ctx->nk);
    HCK(bse_launch_dielectric(st, nw, d_om, eta, (int)N, E, f, pref, d_re, d_im));
    std::vector<double> im(nw), r(nw);
    HCK(hipMemcpyAsync(im.data(), d_im, sizeof(double) * nw, hipMemcpyDeviceToHost, st));
    HCK(hipStreamSynchronize(st));
    double loss = 0.0;
    for (int w = 0; w < nw; ++w) {
        r[w] = im[w] - target[w];
        loss += r[w] * r[w];
    }
    if (L) *L = loss;
    double* d_r = d_re;   // reuse: the real part is not needed
    HCK(hipMemcpyAsync(d_r, r.data(), sizeof(double) * nw, hipMemcpyHostToDevice, st));
    double2* Fbr = BUF(double2, S_SGFB, (size_t)N * nw);
    double2* Fac = BUF(double2, S_SGFA, (size_t)N * nw);
    double2* Xr = BUF(double2, S_SGX, (size_t)N * nw);
    double2* U = BUF(double2, S_SGU, (size_t)N * nw);
    double2* Bm = BUF(double2, S_SGB, (size_t)N * N);
    HCK(bse_launch_spec_factors(st, (int)N, nw, d_om, eta, E, f, d_r, Fbr, Fac, sa, sb));
    HCK(bse_zgemm(st, 0, 0, (int)N, nw, (int)N, 1, 0, A, N, Fbr, N, 0, 0, Xr, N, nullptr, 0));
    HCK(bse_zgemm(st, 0, 0, (int)N, nw, (int)N, 1, 0, A, N, Fac, N, 0, 0, U, N, nullptr, 0));
    HCK(bse_zgemm(st, 0, 1, (int)N, (int)N, nw, 1, 0, Xr, N, U, N, 0, 0, Bm, N, nullptr, 0));
    HCK(bse_zgemm(st, 0, 0, (int)N, 1, (int)N, 1, 0, A, N, sb, N, 0, 0, p1, N, nullptr, 0));
    const int tk = bse_apply_pick_tk(ctx->sys.nv, ctx->sys.nc, ctx->norb);
    const int nblk = (int)ceil_div64(ctx->nk, tk);
    double2* part = BUF(double2, S_SGPART, std::max(nblk, 1));
    std::vector<double2> hpart(std::max(nblk, 1));
    std::vector<double2> hv(3 * (size_t)N);
    for (int t = 0; t < 5; ++t) {
        if (which && !which[t]) { dL[t] = 0.0; continue; }
        HCK(hipMemsetAsync(part, 0, sizeof(double2) * std::max(nblk, 1), st));
        int rc = BSE_OK;
        if (t == 1) HCK(bse_launch_trace(st, N, Bm, N, part));                       // dH = I
        else if (t == 0 || t == 4) rc = grad_tb_apply(ctx, t, 0, nullptr, nullptr, Bm, part);
        else rc = grad_w_apply(ctx, t, 0, nullptr, nullptr, Bm, part);
        if (rc) return rc;
        HCK(hipMemcpyAsync(hpart.data(), part, sizeof(double2) * std::max(nblk, 1), hipMemcpyDeviceToHost, st));
        HCK(hipStreamSynchronize(st));
        double2 term = cmk(0, 0);
        for (int b = 0; b < std::max(nblk, 1); ++b) term = cadd(term, hpart[b]);
        if (t == 0 || t == 4) {
            // dipole derivative terms: drho^H (A sb) + sa^T (A^H drho)
            const double2* drho = (const double2*)ctx->bufs[S_GDRHO].p;
            HCK(bse_zgemm(st, 1, 0, (int)N, 1, (int)N, 1, 0, A, N, drho, N, 0, 0, q, N, nullptr, 0));
            HCK(hipMemcpyAsync(hv.data(), drho, sizeof(double2) * N, hipMemcpyDeviceToHost, st));
            HCK(hipMemcpyAsync(hv.data() + N, p1, sizeof(double2) * N, hipMemcpyDeviceToHost, st));
            HCK(hipMemcpyAsync(hv.data() + 2 * N, q, sizeof(double2) * N, hipMemcpyDeviceToHost, st));
            std::vector<double2> hsa(N);
            HCK(hipMemcpyAsync(hsa.data(), sa, sizeof(double2) * N, hipMemcpyDeviceToHost, st));
            HCK(hipStreamSynchronize(st));
            for (int64_t i = 0; i < N; ++i) {
                cfmac(term, hv[i], hv[N + i]);              // conj(drho_i) (A sb)_i
                cfma(term, hsa[i], hv[2 * N + i]);          // sa_i (A^H drho)_i
            }
        }
        dL[t] = -2.0 * pref * term.y;
    }
    return BSE_OK;
}

extern "C" int bse_eigh(bse_ctx* ctx, int n_vectors, int refine) {
    if (!ctx) return BSE_ERR_ARG;
    if (!ctx->built) return fail(ctx, BSE_ERR_STATE, "matrix not built");
    if (ctx->have_eig) return fail(ctx, BSE_ERR_STATE, "matrix already reduced; build again");
    if (refine && (n_vectors <= 0 || n_vectors > 16))
        return fail(ctx, BSE_ERR_UNSUPPORTED, "refinement needs 1 <= n_vectors <= 16");
    if (refine && n_vectors > ctx->N)
        return fail(ctx, BSE_ERR_ARG, "refinement needs n_vectors <= N");
    (void)hipSetDevice(ctx->device);
    const int fp32 = ctx->precision != BSE_PREC_FP64;
    ctx->eig_on_bse = true;
    ctx->full_eig = false;
    ctx->ref_rnorm.clear();
    ctx->ref_rounds = 0;
    ctx->ref_status = BSE_OK;
    // with refinement, up to 4 guard vectors join a cluster that straddles
    // the cut (eigenvalues within 1e-4 ||T|| of the last requested one), as
    // long as the Rayleigh-Ritz basis [X, R] stays within 40 columns and N
    const int maxv = refine ? (int)std::min<int64_t>(std::min(n_vectors + 4, 20), ctx->N) : 0;
    int rc = eig_pipeline(ctx, ctx->N, ctx->bufs[S_H].p, fp32,
                          (const double2*)ctx->bufs[S_RHO].p, n_vectors,
                          refine ? -1.0 : 0.0, maxv);
    // an early error return may leave Q^H rho running on the side stream
    if (rc) (void)hipStreamSynchronize(ctx->side);
    if (rc || !refine) return rc;
    stage_begin(ctx, ST_REF);
    rc = refine_lowest(ctx, n_vectors, ctx->nvec);
    stage_end(ctx, ST_REF);
    return rc;
}

// ---------------------------------------------------------------------------
// Full BSE beyond the Tamm-Dancoff approximation: the coupling block
// (add_coupling_part, 2929-3040) and a structure-preserving solve of
// H = [[A, B], [-conj B, -conj A]] in place of eig_gen (3326): Cholesky of
// M = [[A, B], [conj B, conj A]] = L L^H, the Hermitian W = L^H s3 L through
// the ordinary eigen-pipeline with z = L^-1 rho~ as the contraction vector,
// then E > 0 with f = sqrt(E) z^H y and x = sqrt(E) L^-H y (full_bse.hip).
// ---------------------------------------------------------------------------
extern "C" int bse_build_full(bse_ctx* ctx) {
    if (!ctx) return BSE_ERR_ARG;
    if (!ctx->have_problem) return fail(ctx, BSE_ERR_STATE, "no problem loaded");
    {
        const double* q = ctx->sys.q;
        if (q[0] * q[0] + q[1] * q[1] + q[2] * q[2] > BSE_OPTICAL_Q * BSE_OPTICAL_Q)
            return fail(ctx, BSE_ERR_UNSUPPORTED, "the full BSE is formed in the optical limit |q| <= 1e-4 1/A");
    }
    int rc = bse_build(ctx, BSE_PREC_FP64);
    if (rc) return rc;
    const bse_system& s = ctx->sys;
    const int64_t N = ctx->N;
    hipStream_t st = ctx->st;
    double2* Bm = BUF(double2, S_BCPL, tiled_elems(N));
    HCK(hipMemsetAsync(Bm, 0, sizeof(double2) * tiled_elems(N), st));
    const double alpha = -s.direct_scale / (ctx->cell * ctx->nk);
    const double beta = s.f_s * s.exchange_scale / (ctx->cell * ctx->nk);
    const int tk = bse_assemble_pick_tk(s.nv, s.nc, ctx->norb);
    HCK(bse_launch_assemble_coupling(st, ctx->nk, s.n1, s.n2, ctx->n3, s.nv, s.nc, ctx->norb, tk,
                                     (const double2*)ctx->bufs[S_COEF].p, (const double2*)ctx->bufs[S_WTAB].p,
                                     (const double2*)ctx->bufs[S_PV].p, (const double2*)ctx->bufs[S_P].p, alpha,
                                     beta, Bm));
    double2* M = BUF(double2, S_MFULL, tiled_elems(2 * N));
    HCK(bse_launch_compose_full(st, N, (const double2*)ctx->bufs[S_H].p, Bm, M));
    HCK(hipStreamSynchronize(st));
    ctx->full_built = true;
    ctx->have_coupling = true;
    return BSE_OK;
}

extern "C" int bse_get_coupling(bse_ctx* ctx, double* dst, int64_t ld) {
    if (!ctx || !dst) return BSE_ERR_ARG;
    if (!ctx->have_coupling) return fail(ctx, BSE_ERR_STATE, "coupling block not built (bse_build_full)");
    const int64_t N = ctx->N;
    if (ld < N) return fail(ctx, BSE_ERR_ARG, "ld < N");
    (void)hipSetDevice(ctx->device);
    double2* tmp = BUF(double2, S_W2, (size_t)ld * N);
    HCK(bse_launch_expand_sym(ctx->st, N, ctx->B, ctx->nk, (const double2*)ctx->bufs[S_BCPL].p, tmp, ld, 1,
                              kinv_dev(ctx)));
    HCK(hipMemcpyAsync(dst, tmp, sizeof(double2) * (size_t)ld * N, hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    return BSE_OK;
}

extern "C" int bse_eigh_full(bse_ctx* ctx, int n_vectors) {
    if (!ctx) return BSE_ERR_ARG;
    if (!ctx->full_built) return fail(ctx, BSE_ERR_STATE, "full BSE matrix not built (bse_build_full)");
    const int64_t N = ctx->N;
    if (n_vectors < -1 || n_vectors > N) return fail(ctx, BSE_ERR_ARG, "n_vectors out of range");
    if (2 * N > (int64_t)1 << 30) return fail(ctx, BSE_ERR_UNSUPPORTED, "full BSE order too large");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->st;
    const int n2 = (int)(2 * N), nt2 = tile_count(n2);
    const int nvec = n_vectors < 0 ? (int)N : n_vectors;
    ctx->full_built = false;                 // M is overwritten by its Cholesky factor
    ctx->full_eig = false;
    ctx->have_eig = ctx->have_f = false;
    double2* M = (double2*)ctx->bufs[S_MFULL].p;
    double2* Linv = BUF(double2, S_LINV, (size_t)nt2 * 64 * 64);
    int* ferr = BUF(int, S_FERR, 4);
    HCK(bse_launch_zero_words(st, 2, ferr));
    const int64_t pm = std::max<int64_t>(n2 - 64, 1);
    double2* P = BUF(double2, S_FPAN, (size_t)pm * 64);
    double2* Ph = BUF(double2, S_FPANH, (size_t)pm * 64);
    double2* zeros = BUF(double2, S_BZERO, 64);
    HCK(bse_launch_zero_words(st, 128, zeros));
    const int kind = ctx->her2k_valu[0] >= 0 ? ctx->her2k_valu[0] : 1;
    HCK(hipEventRecord(ctx->tev[0], st));
    for (int J = 0; J < nt2; ++J) {
        HCK(bse_launch_potrf_tile(st, n2, J, M, Linv, ferr));
        HCK(bse_launch_trsm_panel(st, n2, J, M, Linv, P, Ph));
        const int s0 = (J + 1) * 64;
        if (s0 < n2) {
            const int m = n2 - s0;
            // A22 -= P P^H as two rank-64 halves of P Ph^H + Ph P^H, Ph = P / 2
            HCK(bse_launch_her2k(st, 0, kind, n2, s0, M, P, Ph, m, s0, 32, zeros));
            HCK(bse_launch_her2k(st, 0, kind, n2, s0, M, P + (int64_t)32 * m, Ph + (int64_t)32 * m, m, s0, 32,
                                 zeros));
        }
    }
    int herr = 0;
    HCK(hipMemcpyAsync(&herr, ferr, sizeof(int), hipMemcpyDeviceToHost, st));
    HCK(hipStreamSynchronize(st));
    if (herr)
        return fail(ctx, BSE_ERR_NUMERIC,
                    "full BSE: [[A, B], [conj B, conj A]] is not positive definite (pivot at row " +
                        std::to_string(herr) + "): the system is unstable beyond Tamm-Dancoff");
    double2* W = BUF(double2, S_WFULL, tiled_elems(n2));
    HCK(bse_launch_lsl(st, N, M, W));
    const int64_t npad = (int64_t)nt2 * 64;
    double2* rt = BUF(double2, S_RHOT, npad);
    double2* z = BUF(double2, S_ZT, npad);
    HCK(bse_launch_full_rho(st, N, (const double2*)ctx->bufs[S_RHO].p, rt, npad));
    HCK(bse_launch_lower_solve(st, n2, M, Linv, rt, z));
    HCK(hipEventRecord(ctx->tev[1], st));
    HCK(hipEventSynchronize(ctx->tev[1]));
    float ms = 0.f;
    HCK(hipEventElapsedTime(&ms, ctx->tev[0], ctx->tev[1]));
    ctx->eig_on_bse = false;
    int rc = eig_pipeline(ctx, n2, W, 0, z, nvec, 0.0, 0, (int)N);
    if (rc) {
        (void)hipStreamSynchronize(ctx->side);
        return rc;
    }
    ctx->stage_ms[ST_CHOL] = ms;
    // the spectrum of W is {-E, +E}: keep the positive half
    std::vector<double> lam(n2);
    std::vector<double2> fp(n2);
    HCK(hipMemcpyAsync(lam.data(), ctx->bufs[S_EVAL].p, sizeof(double) * n2, hipMemcpyDeviceToHost, st));
    HCK(hipMemcpyAsync(fp.data(), ctx->bufs[S_F].p, sizeof(double2) * n2, hipMemcpyDeviceToHost, st));
    HCK(hipStreamSynchronize(st));
    double emax = 0.0, pair = 0.0;
    for (int i = 0; i < n2; ++i) emax = std::max(emax, fabs(lam[i]));
    for (int64_t i = 0; i < N; ++i) pair = std::max(pair, fabs(lam[N + i] + lam[N - 1 - i]));
    if (!(lam[N] > 0.0) || !(lam[N - 1] < 0.0) || pair > 1e-8 * emax)
        return fail(ctx, BSE_ERR_NUMERIC, "full BSE: the spectrum of L^H s3 L is not paired +-E");
    std::vector<double> E(N), sc(std::max(nvec, 1));
    std::vector<double2> f(N);
    for (int64_t i = 0; i < N; ++i) {
        E[i] = lam[N + i];
        f[i] = cscale(fp[N + i], sqrt(E[i]));
    }
    for (int j = 0; j < nvec; ++j) sc[j] = sqrt(E[j]);
    HCK(hipMemcpyAsync(ctx->bufs[S_EVAL].p, E.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
    HCK(hipMemcpyAsync(ctx->bufs[S_F].p, f.data(), sizeof(double2) * N, hipMemcpyHostToDevice, st));
    if (nvec > 0) {
        double* dsc = BUF(double, S_FSCALE, nvec);
        double2* X = BUF(double2, S_XFULL, (size_t)npad * nvec);
        HCK(hipMemcpyAsync(dsc, sc.data(), sizeof(double) * nvec, hipMemcpyHostToDevice, st));
        HCK(bse_launch_upper_solve(st, n2, nvec, M, Linv, (const double2*)ctx->bufs[S_Z].p, n2, X, npad, dsc));
    }
    HCK(hipStreamSynchronize(st));
    ctx->full_ldx = npad;
    ctx->n_eig = N;
    ctx->nvec = nvec;
    ctx->have_eig = ctx->have_f = true;
    ctx->full_eig = true;
    ctx->eig_on_bse = true;
    return BSE_OK;
}

extern "C" int bse_refine_info(const bse_ctx* ctx, double* rnorm, int n, int* rounds) {
    if (!ctx || n < 0) return BSE_ERR_ARG;
    if (n > (int)ctx->ref_rnorm.size()) return BSE_ERR_ARG;
    for (int j = 0; j < n; ++j) rnorm[j] = ctx->ref_rnorm[j];
    if (rounds) *rounds = ctx->ref_rounds;
    return ctx->ref_status;
}

extern "C" int bse_get_eigenvalues(bse_ctx* ctx, double* evals) {
    if (!ctx || !evals) return BSE_ERR_ARG;
    if (!ctx->have_eig) return fail(ctx, BSE_ERR_STATE, "no eigendecomposition");
    HCK(hipMemcpy(evals, ctx->bufs[S_EVAL].p, sizeof(double) * ctx->n_eig, hipMemcpyDeviceToHost));
    return BSE_OK;
}

extern "C" int bse_get_eigenvectors(bse_ctx* ctx, double* z, int64_t ld, int nvec, int dst_on_device) {
    if (!ctx || !z) return BSE_ERR_ARG;
    if (!ctx->have_eig) return fail(ctx, BSE_ERR_STATE, "no eigendecomposition");
    if (nvec < 0 || nvec > ctx->nvec) return fail(ctx, BSE_ERR_ARG, "nvec exceeds the formed eigenvectors");
    if (ld < (ctx->full_eig ? 2 : 1) * ctx->n_eig) return fail(ctx, BSE_ERR_ARG, "ld too small");
    if (nvec == 0) return BSE_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->full_eig) {
        // (x; y) halves, each from the internal to the reference row order
        const int64_t N = ctx->N;
        const double2* X = (const double2*)ctx->bufs[S_XFULL].p;
        double2* dst = (double2*)z;
        if (!dst_on_device) dst = BUF(double2, S_W2, (size_t)ld * nvec);
        HCK(bse_launch_rows_to_reference(ctx->st, N, ctx->B, ctx->nk, nvec, X, ctx->full_ldx, dst, ld,
                                         kmap_dev(ctx)));
        HCK(bse_launch_rows_to_reference(ctx->st, N, ctx->B, ctx->nk, nvec, X + N, ctx->full_ldx, dst + N, ld,
                                         kmap_dev(ctx)));
        if (!dst_on_device)
            HCK(hipMemcpyAsync(z, dst, sizeof(double2) * ld * nvec, hipMemcpyDeviceToHost, ctx->st));
        HCK(hipStreamSynchronize(ctx->st));
        return BSE_OK;
    }
    const int64_t n = ctx->n_eig;
    const double2* Z = (const double2*)ctx->bufs[S_Z].p;
    const int B = ctx->eig_on_bse ? ctx->B : 1;
    const int nk = ctx->eig_on_bse ? ctx->nk : (int)n;
    double2* dst = (double2*)z;
    if (!dst_on_device) dst = BUF(double2, S_W2, (size_t)ld * nvec);
    HCK(bse_launch_rows_to_reference(ctx->st, n, B, nk, nvec, Z, n, dst, ld,
                                     ctx->eig_on_bse ? kmap_dev(ctx) : nullptr));
    if (!dst_on_device)
        HCK(hipMemcpyAsync(z, dst, sizeof(double2) * ld * nvec, hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    return BSE_OK;
}

extern "C" int bse_oscillator(bse_ctx* ctx, double* f) {
    if (!ctx || !f) return BSE_ERR_ARG;
    if (!ctx->have_f) return fail(ctx, BSE_ERR_STATE, "no oscillator strengths (run bse_eigh on a built matrix)");
    HCK(hipMemcpy(f, ctx->bufs[S_F].p, sizeof(double2) * ctx->n_eig, hipMemcpyDeviceToHost));
    return BSE_OK;
}

extern "C" int bse_spectrum(bse_ctx* ctx, const double* omega, int n_omega, double eta,
                            double* eps_re, double* eps_im) {
    if (!ctx || !omega || !eps_re || !eps_im || n_omega <= 0) return BSE_ERR_ARG;
    if (!ctx->have_f) return fail(ctx, BSE_ERR_STATE, "no oscillator strengths");
    (void)hipSetDevice(ctx->device);
    double* d_om = BUF(double, S_EPSOUT, 3 * (size_t)n_omega);
    HCK(hipMemcpyAsync(d_om, omega, sizeof(double) * n_omega, hipMemcpyHostToDevice, ctx->st));
    const double pref = kC / (ctx->volume * ctx->nk);
    stage_begin(ctx, ST_SPEC);
    HCK(bse_launch_dielectric(ctx->st, n_omega, d_om, eta, (int)ctx->n_eig,
                              (const double*)ctx->bufs[S_EVAL].p, (const double2*)ctx->bufs[S_F].p,
                              pref, d_om + n_omega, d_om + 2 * n_omega));
    std::vector<double> re2, im2;
    if (ctx->full_eig) {
        // beyond Tamm-Dancoff the anti-resonant branch adds
        // + pref sum |f|^2 / (w + E + i eta): the same sum at -E, subtracted
        const int64_t N = ctx->n_eig;
        double* eneg = BUF(double, S_EVNEG, N + 2 * (int64_t)n_omega);
        std::vector<double> e(N);
        HCK(hipMemcpyAsync(e.data(), ctx->bufs[S_EVAL].p, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->st));
        HCK(hipStreamSynchronize(ctx->st));
        for (double& x : e) x = -x;
        HCK(hipMemcpyAsync(eneg, e.data(), sizeof(double) * N, hipMemcpyHostToDevice, ctx->st));
        HCK(bse_launch_dielectric(ctx->st, n_omega, d_om, eta, (int)N, eneg, (const double2*)ctx->bufs[S_F].p,
                                  pref, eneg + N, eneg + N + n_omega));
        re2.resize(n_omega);
        im2.resize(n_omega);
        HCK(hipMemcpyAsync(re2.data(), eneg + N, sizeof(double) * n_omega, hipMemcpyDeviceToHost, ctx->st));
        HCK(hipMemcpyAsync(im2.data(), eneg + N + n_omega, sizeof(double) * n_omega, hipMemcpyDeviceToHost,
                           ctx->st));
    }
    stage_end(ctx, ST_SPEC);
    HCK(hipMemcpyAsync(eps_re, d_om + n_omega, sizeof(double) * n_omega, hipMemcpyDeviceToHost, ctx->st));
    HCK(hipMemcpyAsync(eps_im, d_om + 2 * n_omega, sizeof(double) * n_omega, hipMemcpyDeviceToHost, ctx->st));
    HCK(hipStreamSynchronize(ctx->st));
    if (ctx->full_eig) {
        for (int i = 0; i < n_omega; ++i) {
            eps_re[i] = eps_re[i] - re2[i] + 1.0;
            eps_im[i] = eps_im[i] - im2[i];
        }
    }
    return BSE_OK;
}

// Dense Hermitian eig of a caller matrix; outputs on the host or the device.
int eigh_dense_impl(bse_ctx* ctx, int n, const double* a, int64_t lda, int a_on_device, int precision,
                    int n_vectors, double* evals, int evals_on_device, double* z, int64_t ldz,
                    int z_on_device) {
    if (!ctx || !a || n <= 0 || lda < n) return BSE_ERR_ARG;
    if (precision != BSE_PREC_FP64 && precision != BSE_PREC_FP32)
        return fail(ctx, BSE_ERR_ARG, "dense eig precision must be FP64 or FP32 storage");
    (void)hipSetDevice(ctx->device);
    const int fp32 = precision == BSE_PREC_FP32;
    int rc = BSE_OK;
    // the lower triangle of the FP64 input goes to tiled storage (FP64 or FP32);
    // a device input is read in place
    const double2* src = (const double2*)a;
    int64_t lds = lda;
    if (!a_on_device) {
        double2* stage = BUF(double2, S_STAGE, (size_t)n * n);
        HCK(hipMemcpy2DAsync(stage, sizeof(double2) * n, a, sizeof(double2) * lda, sizeof(double2) * n,
                             n, hipMemcpyHostToDevice, ctx->st));
        src = stage;
        lds = n;
    }
    void* Hm = buf(ctx, S_H, (size_t)tiled_elems(n) * (fp32 ? sizeof(float2) : sizeof(double2)), &rc);
    if (!Hm) return rc;
    HCK(bse_launch_to_tiles(ctx->st, fp32, n, src, lds, Hm));
    ctx->built = false;
    ctx->have_tables = false;
    ctx->eig_on_bse = false;
    rc = eig_pipeline(ctx, n, Hm, fp32, nullptr, n_vectors);
    if (rc) return rc;
    if (evals)
        HCK(hipMemcpyAsync(evals, ctx->bufs[S_EVAL].p, sizeof(double) * n,
                           evals_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->st));
    if (z && ctx->nvec > 0) {
        if (ldz < n) return fail(ctx, BSE_ERR_ARG, "ldz < n");
        HCK(hipMemcpy2DAsync(z, sizeof(double2) * ldz, ctx->bufs[S_Z].p, sizeof(double2) * n,
                             sizeof(double2) * n, ctx->nvec,
                             z_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->st));
    }
    HCK(hipStreamSynchronize(ctx->st));
    return BSE_OK;
}

extern "C" int bse_eigh_dense(bse_ctx* ctx, int n, const double* a, int64_t lda, int a_on_device,
                              int precision, int n_vectors, double* evals, double* z, int64_t ldz) {
    return eigh_dense_impl(ctx, n, a, lda, a_on_device, precision, n_vectors, evals, 0, z, ldz, 0);
}
