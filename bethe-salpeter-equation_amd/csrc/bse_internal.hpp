// Internal state of a bse_ctx and the host helpers shared by the C ABI's
// translation units (single-system path and batched sweep).
#pragma once
#include "bse_common.hpp"
#include "../../include/bse_mi355x.h"

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

// ---- kernel launchers implemented in the other translation units ----------
extern "C" {
hipError_t bse_launch_tb_states(hipStream_t, int, int, int, const double*, const double2*,
                                const double*, const double*, int, int, int, double,
                                const double2*, double*, double2*, double*, double2*);
hipError_t bse_launch_w_table(hipStream_t, int, int, int, int, int, int, const double*,
                              const double*, const double*, double, double, double, double, double,
                              int, double2*);
hipError_t bse_launch_v_exchange(hipStream_t, int, int, int, const double*, const double*, double,
                                 double, const double*, double2*);
hipError_t bse_launch_transitions(hipStream_t, int, int, int, int, const double*, const double2*,
                                  const double*, const double2*, const double*, const double*,
                                  const double2*, double2*, double2*, double2*, double*,
                                  const double2* dq = nullptr);
hipError_t bse_launch_dipole_table(hipStream_t, int, int, int, const double*, const double*, const double*,
                                   const double*, double2*);
hipError_t bse_launch_wannier_integrals(hipStream_t, const int*, const int*, const double*, const double*, double,
                                        int, int, int, const int*, int, const double*, const double*, double2*,
                                        double*, double*);
size_t bse_assemble_lds_bytes(int, int, int, int);
int bse_assemble_pick_tk(int, int, int);
hipError_t bse_launch_assemble(hipStream_t, int, int, int, int, int, int, int, int, int, int64_t,
                               const double2*, const double2*, const double2*, const double2*,
                               const double2*, const double2*, const double2*, const double*,
                               double, double, void*, int);
hipError_t bse_launch_expand(hipStream_t, int, int64_t, int, int, const void*, void*, int64_t, int,
                             const int*);
hipError_t bse_launch_to_tiles(hipStream_t, int, int64_t, const double2*, int64_t, void*);
hipError_t bse_launch_tile_pad_zero(hipStream_t, int, int64_t, void*);
void bse_hetrd_workspace(int, int, int64_t*, int64_t*, int64_t*, int64_t*);
hipError_t bse_hetrd(hipStream_t, int, int, void*, int, double*, double*, double2*, double2*,
                     double2*, double2*, double2*, double2*, double2*, double2*, double2*, double2*,
                     double*, double*, unsigned int*, hipEvent_t*, const unsigned char*, int, void*);
hipError_t bse_launch_hemv(hipStream_t, int, int, int, int, const void*, const double2*, double2*,
                           double2*);
hipError_t bse_launch_last_diag(hipStream_t, int, int, const void*, double*);
hipError_t bse_hetrd_batch_column(hipStream_t, int, int, int, int, int, const void*, int);
hipError_t bse_hetrd_batch_panel(hipStream_t, int, int, int, int, const void*, int64_t, int);
hipError_t bse_launch_her2k(hipStream_t, int, int, int, int, void*, const double2*,
                            const double2*, int64_t, int, int, const double2*);
hipError_t bse_launch_tri_prepare(hipStream_t, int, const double*, double*, double*, int4*,
                                  double2*, double*);
hipError_t bse_launch_tri_sort(hipStream_t, int, const double*, const int4*, const double*,
                               double*, int*, int*, int2*, int*, int*);
hipError_t bse_launch_tri_vectors_dev(hipStream_t, int, const int4*, const double2*, const double*,
                                      const double*, const double*, const double*, double*,
                                      const double*, const int*, const int*, const int2*,
                                      const int*, const int*, double*, double*, int);
hipError_t bse_launch_apply_q(hipStream_t, int, int, const double2*, const double2*, int, double2*,
                              int64_t, int);
int64_t bse_apply_q_thin_workspace(int);
hipError_t bse_apply_q_thin(hipStream_t, int, int, const double2*, const double2*, int, double2*, int64_t,
                            int, double2*);
hipError_t bse_apply_q_panel(hipStream_t, int, int, const double2*, const double2*, int, int, double2*,
                             int64_t, int, double2*);
// two-stage reduction (band_kernels.hip, chase_kernels.hip)
hipError_t bse_band_gather(hipStream_t, int, int, int, int, const void*, double2*, double2*, int);
hipError_t bse_band_gather_tail(hipStream_t, int, int, int, const void*, double2*, int);
int64_t bse_tsqr_workspace(int, int);
hipError_t bse_band_panel_qr(hipStream_t, int, int, double2*, double2*, double2*, double2*, double2*, int,
                             int, int*, int*, int);
hipError_t bse_launch_assemble_coupling(hipStream_t, int, int, int, int, int, int, int, int, const double2*,
                                        const double2*, const double2*, const double2*, double, double,
                                        double2*);
hipError_t bse_launch_expand_sym(hipStream_t, int64_t, int, int, const double2*, double2*, int64_t, int,
                                 const int*);
hipError_t bse_launch_compose_full(hipStream_t, int64_t, const double2*, const double2*, double2*);
hipError_t bse_launch_potrf_tile(hipStream_t, int, int, double2*, double2*, int*);
hipError_t bse_launch_trsm_panel(hipStream_t, int, int, double2*, const double2*, double2*, double2*);
hipError_t bse_launch_lsl(hipStream_t, int64_t, const double2*, double2*);
hipError_t bse_launch_lower_solve(hipStream_t, int, const double2*, const double2*, const double2*, double2*);
hipError_t bse_launch_upper_solve(hipStream_t, int, int, const double2*, const double2*, const double2*,
                                  int64_t, double2*, int64_t, const double*);
hipError_t bse_launch_full_rho(hipStream_t, int64_t, const double2*, double2*, int64_t);
int64_t bse_rpa_pi_elems(int, int, int, int);
hipError_t bse_launch_rpa_table(hipStream_t, int, int, int, int, int, int, const double*, int, int, double, double,
                                double, double, const double*, const double2*, const double*, const double*,
                                double2*, double2*);
hipError_t bse_apply_q_strips(hipStream_t, int, int, int, const double2*, const int64_t*, const double2*,
                              double2*, int64_t, int);
int64_t bse_apply_q_grouped_workspace(int, int, int);
hipError_t bse_apply_q_grouped_prepare(hipStream_t, int, int, int, int, const double2*, const int64_t*,
                                       const double2*, double2*);
hipError_t bse_apply_q_grouped_run(hipStream_t, int, int, int, int, const double2*, const int64_t*, double2*,
                                   int64_t, int, double2*, int64_t);
hipError_t bse_apply_q_grouped(hipStream_t, int, int, int, const double2*, const int64_t*, const double2*, double2*,
                               int64_t, int, double2*, int64_t);
int64_t bse_band_update_workspace(int, int, int);
hipError_t bse_band_update(hipStream_t, int, int, int, int, int, int, void*, const double2*, const double2*,
                           double2*, const double2*, hipEvent_t*);
hipError_t bse_band_product(hipStream_t, int, int, int, int, int, int, const void*, const double2*,
                            const double2*, double2*, hipEvent_t*, double2**, float**);
hipError_t bse_launch_her2k_part(hipStream_t, int, int, int, int, int, void*, const double2*, const double2*,
                                 int64_t, int, int, const double2*);
int64_t bse_chase_counters(int);
hipError_t bse_launch_chase(hipStream_t, int, int, double2*, int, double2*, double*, double*, int*, int*, int);
int64_t bse_band_ii_workspace(int, int, int);
hipError_t bse_launch_band_ii(hipStream_t, int, int, int, const double2*, const double*, int, double,
                              double2*, double2*, int64_t);
hipError_t bse_launch_copy_words(hipStream_t, int64_t, const void*, void*);
hipError_t bse_launch_zero_words(hipStream_t, int64_t, void*);
hipError_t bse_launch_stamp(hipStream_t, unsigned long long*);
hipError_t bse_launch_spec_factors(hipStream_t, int, int, const double*, double, const double*,
                                   const double2*, const double*, double2*, double2*, double2*,
                                   double2*);
hipError_t bse_launch_trace(hipStream_t, int64_t, const double2*, int64_t, double2*);
hipError_t bse_launch_bisect(hipStream_t, int, const int4*, const double2*, const double*,
                             const double*, double, double, double*);
hipError_t bse_launch_inverse_iter(hipStream_t, int, int, int, const double*, const int*,
                                   const int*, const double*, const double*, double, double*, int,
                                   double*);
hipError_t bse_launch_cluster_mgs(hipStream_t, int, int, const int2*, int, const int*, const int*,
                                  const int*, double*);
hipError_t bse_launch_commit(hipStream_t, int, int, int, int, int, const int*, const int*,
                             const double*, double2*, int64_t, int);
hipError_t bse_launch_contract(hipStream_t, int, int, int, const int*, const int*, const double*,
                               const double2*, double2*);
hipError_t bse_launch_contract_split(hipStream_t, int, int, int, const int*, const int*, const double*,
                                     const double2*, double2*, double*);
int64_t bse_contract_split_workspace(int);
hipError_t bse_zgemm(hipStream_t, int, int, int, int, int, double, double, const double2*, int64_t,
                     const double2*, int64_t, double, double, double2*, int64_t, double2*, int64_t);
hipError_t bse_launch_larft(hipStream_t, int, const double2*, const double2*, double2*, int);
hipError_t bse_launch_dielectric(hipStream_t, int, const double*, double, int, const double*,
                                 const double2*, double, double*, double*);
hipError_t bse_launch_rows_to_reference(hipStream_t, int64_t, int, int, int, const double2*,
                                        int64_t, double2*, int64_t, const int*);
hipError_t bse_launch_convert_to_fp32(hipStream_t, int64_t, const double2*, float2*);
hipError_t bse_launch_rows_from_reference(hipStream_t, int64_t, int, int, int, const double2*,
                                          int64_t, double2*, int64_t, const int*);
int bse_apply_pick_tk(int, int, int);
int bse_assemble_mfma_supported(int, int, int);
int bse_apply_mfma_groups(int, int, int);
int64_t bse_apply_mfma_workspace(int, int, int);
hipError_t bse_launch_apply_mfma(hipStream_t, int, int, int, int, int, int, int, const double2*, const double2*,
                                 const double2*, const double2*, const double*, double, double, int,
                                 const double2*, int64_t, double2*, int64_t, double2*);
hipError_t bse_launch_assemble_mfma(hipStream_t, int, int, int, int, int, int, int, int, int64_t,
                                    const double2*, const double2*, const double2*, const double2*,
                                    const double*, double, double, void*);
hipError_t bse_launch_apply(hipStream_t, int, int, int, int, int, int, int, int, const double2*,
                            const double2*, const double2*, const double2*, const double2*,
                            const double2*, const double2*, const double*, double, double, int,
                            const double2*, int64_t, double2*, int64_t, const double2*, int64_t,
                            double2*);
hipError_t bse_launch_small_rr(hipStream_t, int, int, const double2*, const double2*, double,
                               double*, double2*, int*);
hipError_t bse_launch_residual(hipStream_t, int64_t, int, const double2*, const double2*, int64_t,
                               double2*, double*, double*);
hipError_t bse_launch_coldot(hipStream_t, int64_t, int, const double2*, const double2*, int64_t,
                             double*);
}

static const double kC = (1.602176634 * 1000.0) / 8.8541878176;   // e^2/eps0, eV*A
static const int kPanel = 32;                                      // hetrd panel width

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

enum Stage { ST_TB, ST_TABLES, ST_ASM, ST_TRD, ST_TEVAL, ST_TVEC, ST_BT, ST_REF, ST_SPEC, ST_N };
// stage_ms slots past the event-timed stages: the two halves of a two-stage
// tridiagonalisation (dense -> band with Q1^H rho; band -> tridiagonal)
enum { ST_BAND = ST_N, ST_CHASE = ST_N + 1, ST_CHOL = ST_N + 2 };

}  // namespace

struct bse_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;
    std::vector<DevBuf> bufs;   // indexed by slot id
    // problem
    bool have_problem = false;
    bse_system sys{};
    int norb = 0, nr = 0, nk = 0, ng = 0, B = 0, ns = 0, n3 = 1, dim = 2;
    int64_t N = 0;
    double recip[9] = {0}, area = 0, volume = 0, cell = 0, wbar = 0, vbar = 0, qhat[3] = {1, 0, 0};
    std::vector<double> kpts_h, gvec_h, hr_h, rcart_h;
    int dip_nr = 0;                 // lattice vectors of the Wannier position table (0: point orbitals)
    // state
    bool have_tb = false, have_tables = false, built = false, have_eig = false, have_f = false;
    int precision = BSE_PREC_FP64;
    int64_t n_eig = 0;          // size of the last eigenproblem
    int nvec = 0;               // eigenvectors formed
    bool eig_on_bse = false;    // last eig was of the assembled BSE matrix
    int TK = 1;
    hipEvent_t ev[ST_N + 1] = {};
    bool ev_used[ST_N] = {};
    double stage_ms[16] = {0};
    // optional per-launch timing of the tridiagonalisation's hemv kernel
    bool prof = false;
    std::vector<hipEvent_t> prof_ev;
    double hemv_ms = 0.0, hemv_bytes = 0.0;
    // two-stage profile: the band reduction's rank-2k update and Hermitian
    // block product (total ms, algorithmic flops, launches)
    double b2k_ms = 0.0, b2k_flops = 0.0, hemm_ms = 0.0, hemm_flops = 0.0;
    int64_t band_launches = 0;
    bool prof_two_stage = false;
    bool band_fp32prod = false;     // last band reduction ran on FP32 matrix cores
    bool band_lookahead = false;    // ... and updated each next panel's tile column first
    bool band_pack = false;         // ... with the update's operands packed in FP32 planes
    bool vec_grouped = false;       // the lowest vectors' Q1 back-transform ran four panels per pass
    bool force_valu_assembly = false;   // BSE_ASSEMBLY=valu selects the VALU tiles
    int asm_kernel = 0;                 // 0 VALU, 1 FP64 MFMA, 2 FP32 MFMA, 3 BF16x2 MFMA
    int64_t hemv_launches = 0;
    // autotuned hemv chunk length per floor(log2(trailing size)) bucket, one
    // table per storage precision (0 = not tuned yet)
    unsigned char hemv_ch[2][BSE_HEMV_BUCKETS] = {};
    // trailing-update kernel per storage precision: -1 not tuned, 0 FP64 MFMA
    // (64 x 64 tiles), 1 FP64 VALU, 2 FP64 MFMA (128 x 128 macro tiles)
    int her2k_valu[2] = {-1, -1};
    bool autotune = true;
    hipEvent_t tev[3] = {};     // autotuner timing pair (+1 for the two-stage split)
    // second stream (default priority): Q^H rho runs on it while the host
    // sorts the tridiagonal blocks and the bisection runs on st
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // band reduction look-ahead: product done (st -> side), next panel
    // factored (side -> st)
    hipEvent_t la_ev[2] = {};
    // caller's k order (bse_set_problem_klist): list position of each grid
    // index and its inverse; empty for the plain grid
    std::vector<int> kmap_h, kinv_h;
    // two-stage reduction: 0 auto (for at most 20 eigenvectors and n >= 256),
    // 1 whenever the eigenvectors asked for are at most 20, -1 never
    // (BSE_TWO_STAGE=1 / 0); band panels of the last two-stage reduction
    // (row start, rows, reflectors, offset into the reflector store)
    int two_stage = 0;
    bool last_two_stage = false;
    // full (non-Tamm-Dancoff) BSE state: M composed; last eig was of it
    bool full_built = false, full_eig = false, have_coupling = false;
    int64_t full_ldx = 0;
    struct BandPanel { int s, m, k; int64_t voff; };
    std::vector<BandPanel> band_panels;
    double chase_ms = 0.0, band_ms = 0.0;
    // allocation-free calls: while arena_base is set, buffers are carved from
    // the device scratch array (bump allocation, arena_off) instead of
    // hipMalloc; arena_bufs is the slot table of those calls
    char* arena_base = nullptr;
    size_t arena_cap = 0, arena_off = 0, arena_hwm = 0;
    std::vector<DevBuf> arena_bufs;
    // Gershgorin norm of the last tridiagonal (= an upper bound of ||H||)
    double hnorm = 0.0;
    // last refinement: residual norms ||H x_j - theta_j x_j|| (eV) of the
    // formed vectors, rounds run, and whether the stopping rule was met
    std::vector<double> ref_rnorm;
    int ref_rounds = 0, ref_status = 0;
    // batched sweep state (owned by batch_api.hip)
    struct BatchState* batch = nullptr;
};

namespace {

enum Slot {
    S_RCART, S_HR, S_TAU, S_KPTS, S_GVEC, S_Q, S_EPS, S_COEF, S_WTAB, S_VX, S_P, S_PV, S_RHO, S_DE,
    S_H, S_D, S_E, S_TAUH, S_VSTORE, S_WP, S_ACOL, S_YR, S_YC, S_T12, S_PART, S_SCAL, S_COUNTER,
    S_HSYS,
    S_TALL, S_GRAM, S_E2, S_ITEMS, S_BOUNDS, S_LAMI, S_LAM, S_BSV, S_NBV, S_CLUST, S_IISCR, S_IIX,
    S_Y, S_F, S_Z, S_W1, S_W2, S_SPLITK, S_EVAL, S_EPSOUT, S_STAGE, S_ALPHA2,
    S_RS, S_RHS, S_RX, S_RG, S_RHG, S_RC, S_RSMALL,
    S_GDHR, S_GDEPS, S_GDCOEF, S_GDP, S_GDPV, S_GDRHO, S_GDDE, S_GDW, S_GY, S_GOUT,
    S_SGFB, S_SGFA, S_SGX, S_SGU, S_SGB, S_SGV, S_SGPART, S_APPLYP, S_MEMB, S_KMAP, S_KINV, S_TBK, S_TBH, S_TBR, S_TBE, S_TBC,
    S_BAB, S_BAB0, S_BPAN, S_BTSQR, S_BUPD, S_BPROG, S_BII, S_BZERO, S_BH0, S_BFAIL,
    // grouped back-transform: panel offsets, workspace
    S_BTVOFF, S_BTWS,
    // full BSE: coupling block, M / its Cholesky factor, W = L^H s3 L, tile
    // inverses, panel, rho~, L^-1 rho~, eigenvectors, error flag, scales
    S_BCPL, S_MFULL, S_WFULL, S_LINV, S_FPAN, S_FPANH, S_RHOT, S_ZT, S_XFULL, S_FERR, S_FSCALE, S_EVNEG,
    // RPA screening: window states at the grid points, Pi(Q)
    S_RPAEPS, S_RPACOEF, S_RPAPI,
    // Wannier position table (R, D_ij(R), Dq_ij(k)) and the real-space
    // integrals' grid, translations, vectors and outputs
    S_DIPR, S_DIPD, S_DIPQ, S_WGRID, S_WR, S_WK, S_WA, S_WD, S_WS,
    S_NSLOTS
};

int fail(bse_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HCK(expr)                                                                       \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return fail(ctx, BSE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

void* buf(bse_ctx* ctx, int slot, size_t bytes, int* status) {
    DevBuf& b = ctx->bufs[slot];
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return b.p;
    if (ctx->arena_base) {
        // allocation-free mode: carve from the fixed scratch array, never free
        const size_t off = (ctx->arena_off + 255) & ~(size_t)255;
        if (off + bytes > ctx->arena_cap) {
            *status = fail(ctx, BSE_ERR_ARG, "workspace exceeds the fixed scratch array (" +
                                                 std::to_string(ctx->arena_cap) + " bytes)");
            return nullptr;
        }
        b.p = ctx->arena_base + off;
        b.cap = bytes;
        ctx->arena_off = off + bytes;
        ctx->arena_hwm = std::max(ctx->arena_hwm, ctx->arena_off);
        return b.p;
    }
    if (b.p) { (void)hipFree(b.p); b.p = nullptr; b.cap = 0; }
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
        b.p = nullptr;
        *status = fail(ctx, BSE_ERR_ALLOC, "hipMalloc failed for " + std::to_string(bytes) + " bytes");
        return nullptr;
    }
    b.cap = bytes;
    return b.p;
}

template <typename T>
T* bufT(bse_ctx* ctx, int slot, size_t count, int* status) {
    return (T*)buf(ctx, slot, count * sizeof(T), status);
}

#define BUF(T, slot, count)                                  \
    ({                                                        \
        int _s = BSE_OK;                                      \
        T* _p = bufT<T>(ctx, slot, (size_t)(count), &_s);     \
        if (!_p) return _s;                                   \
        _p;                                                   \
    })

void tick(bse_ctx* ctx, int stage) { (void)hipEventRecord(ctx->ev[stage], ctx->st); }

void stage_begin(bse_ctx* ctx, int stage) {
    (void)hipEventRecord(ctx->ev[stage], ctx->st);
    ctx->ev_used[stage] = true;
}
void stage_end(bse_ctx* ctx, int stage) {
    (void)hipEventRecord(ctx->ev[ST_N], ctx->st);
    (void)hipEventSynchronize(ctx->ev[ST_N]);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev[stage], ctx->ev[ST_N]);
    ctx->stage_ms[stage] = ms;
}

bool finite_all(const double* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

// Geometry and sizes of one problem, derived on the host from the model and
// the system parameters: Crystal_Lattice (bse.version13.cpp:85-124;
// b_i = 2 pi (a_j x a_k) / Omega), K_points (183-242; k = ((i+s1)/n1) b1 +
// ((j+s2)/n2) b2, index i*n2 + j), G_points (319-385; 2D parallelepiped,
// G = 0 at index Ng/2), and the q = 0 mini-BZ average of the Keldysh W over
// a disk of area (2 pi)^2 / (A Nk).
struct Geom {
    bse_system sys{};
    int norb = 0, nr = 0, nk = 0, B = 0, ns = 0, ng = 0;
    int n3 = 1, dim = 2;
    int64_t N = 0;
    // cell: the measure of the kernel prefactor 1/(cell Nk) -- the in-plane
    // area for 2D potentials, the volume for 3D ones
    double recip[9] = {0}, volume = 0, area = 0, cell = 0, wbar = 0, vbar = 0, qhat[3] = {1, 0, 0};
    std::vector<double> kpts, gvec;
};

int make_geom(const bse_tb_model* m, const bse_system* s, Geom& g, std::string& msg) {
    if (!m || !s || !m->r_cart || !m->h_r || !m->tau) { msg = "null argument"; return BSE_ERR_ARG; }
    if (m->norb < 1 || m->norb > BSE_MAXORB) { msg = "norb must be in [1, 16]"; return BSE_ERR_UNSUPPORTED; }
    if (m->nr < 1) { msg = "nr must be >= 1"; return BSE_ERR_ARG; }
    if (s->n1 < 1 || s->n2 < 1 || s->n3 < 0) { msg = "k grid must be at least 1 x 1"; return BSE_ERR_ARG; }
    if (s->dimension != 0 && s->dimension != 2 && s->dimension != 3) {
        msg = "dimension must be 2 or 3";
        return BSE_ERR_ARG;
    }
    const int dim = s->dimension == 3 ? 3 : 2;
    const int n3 = s->n3 > 0 ? s->n3 : 1;
    if (dim == 2 && n3 != 1) { msg = "a 2D system has n3 = 1"; return BSE_ERR_ARG; }
    if (!(s->radius >= 0.0)) { msg = "radius must be >= 0"; return BSE_ERR_ARG; }
    if (s->nv < 1 || s->nc < 1 || s->n_occ < s->nv || s->n_occ + s->nc > m->norb) {
        msg = "band window outside the Hamiltonian (nv <= n_occ, n_occ + nc <= norb)";
        return BSE_ERR_ARG;
    }
    if (!(s->eps_env > 0.0) || !(s->r0 >= 0.0) || !(s->g_cut >= 0.0)) {
        msg = "screening parameters must satisfy eps_env > 0, r0 >= 0, g_cut >= 0";
        return BSE_ERR_ARG;
    }
    if (s->screening != 0 && s->screening != 1) { msg = "screening must be 0 (model) or 1 (RPA)"; return BSE_ERR_ARG; }
    if (s->dipole_model != 0 && s->dipole_model != 1) { msg = "dipole_model must be 0 or 1"; return BSE_ERR_ARG; }
    if (s->dipole_model == 1 && (s->dipole_nr < 1 || !s->dipole_r || !s->dipole_d)) {
        msg = "the Wannier dipole model needs dipole_nr >= 1 lattice vectors and the position table";
        return BSE_ERR_ARG;
    }
    if (s->screening == 1 && (s->rpa_nv < 0 || s->rpa_nc < 0 || s->n_occ + 1 > m->norb)) {
        msg = "RPA screening needs rpa_nv, rpa_nc >= 0 and an empty band";
        return BSE_ERR_ARG;
    }
    const size_t nh = (size_t)m->nr * m->norb * m->norb * 2;
    if (!finite_all(m->h_r, nh) || !finite_all(m->r_cart, 3 * (size_t)m->nr) ||
        !finite_all(m->tau, 3 * (size_t)m->norb) || !finite_all(s->bravais, 9)) {
        msg = "non-finite model input";
        return BSE_ERR_NUMERIC;
    }
    g.sys = *s;
    g.sys.n3 = n3;
    g.sys.dimension = dim;
    g.n3 = n3;
    g.dim = dim;
    g.norb = m->norb;
    g.nr = m->nr;
    g.nk = s->n1 * s->n2 * n3;
    g.B = s->nv * s->nc;
    g.ns = s->nv + s->nc;
    g.N = (int64_t)g.nk * g.B;
    const double* a = s->bravais;
    auto cross = [](const double* x, const double* y, double* z) {
        z[0] = x[1] * y[2] - x[2] * y[1];
        z[1] = x[2] * y[0] - x[0] * y[2];
        z[2] = x[0] * y[1] - x[1] * y[0];
    };
    double c0[3], c1[3], c2[3];
    cross(a + 3, a + 6, c0);
    cross(a + 6, a + 0, c1);
    cross(a + 0, a + 3, c2);
    const double det = a[0] * c0[0] + a[1] * c0[1] + a[2] * c0[2];
    if (det == 0.0) { msg = "singular Bravais lattice"; return BSE_ERR_ARG; }
    g.volume = fabs(det);
    const double f = 2.0 * M_PI / g.volume;
    for (int i = 0; i < 3; ++i) {
        g.recip[i] = f * c0[i];
        g.recip[3 + i] = f * c1[i];
        g.recip[6 + i] = f * c2[i];
    }
    double axb[3];
    cross(a, a + 3, axb);
    g.area = sqrt(axb[0] * axb[0] + axb[1] * axb[1] + axb[2] * axb[2]);
    g.cell = dim == 3 ? g.volume : g.area;
    // k = ((i+s1)/n1) b1 + ((j+s2)/n2) b2 + ((l+s3)/n3) b3, index (i n2 + j) n3 + l
    g.kpts.assign(3 * (size_t)g.nk, 0.0);
    for (int i = 0; i < s->n1; ++i)
        for (int j = 0; j < s->n2; ++j)
            for (int l = 0; l < n3; ++l) {
                const double f1 = (i + s->shift[0]) / s->n1, f2 = (j + s->shift[1]) / s->n2,
                             f3 = dim == 3 ? (l + s->shift3) / n3 : 0.0;
                double* kp = &g.kpts[3 * (((size_t)i * s->n2 + j) * n3 + l)];
                for (int x = 0; x < 3; ++x)
                    kp[x] = f1 * g.recip[x] + f2 * g.recip[3 + x] + f3 * g.recip[6 + x];
            }
    // G-list: the parallelepiped |i| <= int(G_cut / |b_i|) over b1, b2 (and b3
    // in 3D), lexicographic, so G = 0 sits at index Ng/2 (334-351)
    int gm[3] = {0, 0, 0};
    for (int d = 0; d < dim; ++d) {
        const double* b = g.recip + 3 * d;
        gm[d] = (int)(s->g_cut / sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]));
    }
    g.gvec.clear();
    for (int i = -gm[0]; i <= gm[0]; ++i)
        for (int j = -gm[1]; j <= gm[1]; ++j)
            for (int l = -gm[2]; l <= gm[2]; ++l)
                for (int x = 0; x < 3; ++x)
                    g.gvec.push_back(i * g.recip[x] + j * g.recip[3 + x] + l * g.recip[6 + x]);
    g.ng = (int)(g.gvec.size() / 3);
    if (s->screening == 1 && (g.ng > 48 || m->norb > 16)) {
        msg = "RPA screening holds at most 48 G vectors and 16 orbitals";
        return BSE_ERR_UNSUPPORTED;
    }
    // q = 0 term: mean of W over the mini Brillouin zone around Q = 0, a disk
    // of area (2 pi)^2/(A Nk) (2D) or a ball of volume (2 pi)^3/(Omega Nk) (3D)
    if (dim == 2) {
        const double da = 4.0 * M_PI * M_PI / (g.area * g.nk);
        const double rad = sqrt(da / M_PI);
        g.wbar = (s->r0 > 0.0) ? (M_PI * kC / s->eps_env) * log1p(s->r0 * rad) / (s->r0 * da)
                               : (M_PI * kC / s->eps_env) * rad / da;
        g.vbar = (M_PI * kC) * rad / da;           // the bare potential's mean (RPA head)
    } else {
        const double vs = 8.0 * M_PI * M_PI * M_PI / (g.volume * g.nk);
        const double rad = cbrt(3.0 * vs / (4.0 * M_PI));
        const double r = s->radius;
        const double integ = (r > 0.0) ? rad - sqrt(r) * atan(rad / sqrt(r)) : rad;
        g.wbar = 4.0 * M_PI * kC * integ / (s->eps_env * vs);
        g.vbar = 4.0 * M_PI * kC * integ / vs;
    }
    double qh[3] = {s->dipole_dir[0], s->dipole_dir[1], s->dipole_dir[2]};
    const double qn = sqrt(qh[0] * qh[0] + qh[1] * qh[1] + qh[2] * qh[2]);
    if (!(qn > 0.0)) { qh[0] = 1.0; qh[1] = qh[2] = 0.0; }
    else for (double& x : qh) x /= qn;
    for (int x = 0; x < 3; ++x) g.qhat[x] = qh[x];
    return BSE_OK;
}

}  // namespace

// Dense eig with host or device outputs (bse_api.hip).
int eigh_dense_impl(bse_ctx* ctx, int n, const double* a, int64_t lda, int a_on_device, int precision,
                    int n_vectors, double* evals, int evals_on_device, double* z, int64_t ldz,
                    int z_on_device);

// Releases the batched-sweep state of a context (batch_api.hip).
void batch_destroy(bse_ctx* ctx);

// Autotuner of the tridiagonalisation's hemv chunk length (bse_api.hip).
int autotune_hemv(bse_ctx* ctx, int fp32, int n, const void* H, double2* v, double2* Yr,
                  double2* Yc);
