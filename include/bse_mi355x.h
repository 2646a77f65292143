/*
 * bse_mi355x.h -- C ABI of the MI355X excitonic hot path.
 *
 * Drop-in boundary for the build-and-eig calls of the tight-binding
 * Bethe-Salpeter solver (reference: bse.version13.cpp).  Every entry point
 * names the reference member function it replaces.  The reference seams are
 * C++ member functions that take and return Armadillo objects by value and
 * have no error returns; here every call returns an int status (0 = ok,
 * negative = error class) and buffers are caller-owned plain pointers.
 *
 * Conventions
 *   - complex numbers are interleaved (re, im) doubles or floats, layout
 *     compatible with std::complex / arma::cx_double;
 *   - matrices are column-major with an explicit leading dimension;
 *   - BSE basis index r = c*nv*Nk + v*Nk + k (k fastest), as the reference's
 *     energies_diff / rho layout (bse.version13.cpp:2004, 2022) and row index
 *     of the direct term (2876);
 *   - valence slot v = 0 is the highest valence band (812-813), conduction
 *     slot c = 0 the lowest conduction band (815-816); valence states are
 *     taken at k - q (1966, 2836);
 *   - eigenvalues ascending, eigenvectors of unit 2-norm (3334-3345);
 *   - textbook (Rohlfing-Louie) orientation of the direct term; the
 *     reference stores its complex conjugate (Armadillo .t() at 2878), which
 *     has the same eigenvalues and conjugated eigenvectors;
 *   - energies in eV, lengths in Angstrom, k and G in 1/Angstrom;
 *   - one context per host thread; calls on a context are serialised on the
 *     context's HIP stream.
 */
#ifndef BSE_MI355X_H
#define BSE_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSE_ABI_VERSION 3

/* status codes */
#define BSE_OK               0
#define BSE_ERR_ARG         -1   /* bad argument, size or null pointer        */
#define BSE_ERR_HIP         -2   /* HIP runtime failure                       */
#define BSE_ERR_ALLOC       -3   /* device or host allocation failed          */
#define BSE_ERR_STATE       -4   /* call made out of order                    */
#define BSE_ERR_NUMERIC     -5   /* non-finite input or failed convergence    */
#define BSE_ERR_UNSUPPORTED -6   /* shape or option not supported             */

/* precision of the assembled matrix and of the eigensolver path */
#define BSE_PREC_FP64   0  /* FP64 assembly, FP64 storage, FP64 reduction     */
#define BSE_PREC_FP32   1  /* FP32 MFMA assembly, FP32 storage; FP64 reflectors */
#define BSE_PREC_BF16X2 2  /* BF16 hi+lo split MFMA products, FP32 accumulation
                              and storage; FP64 reflectors                     */

/* matrix order for copies */
#define BSE_ORDER_REFERENCE 0  /* r = c*nv*Nk + v*Nk + k                       */
#define BSE_ORDER_KMAJOR    1  /* internal r' = k*(nc*nv) + c*nv + v           */

typedef struct bse_ctx bse_ctx;

/* Tight-binding model: H_lm(k) = sum_n H_lm(R_n) exp(i k.R_n), the reference's
 * Hamiltonian_TB::FFT (bse.version13.cpp:647-705) with ndegen weights already
 * applied (divided, the Wannier90 meaning). */
typedef struct {
    int32_t norb;          /* orbitals per cell incl. spin components (<= 16) */
    int32_t nr;            /* number of lattice vectors                       */
    const double* r_cart;  /* nr x 3 row-major Cartesian R, Angstrom          */
    const double* h_r;     /* complex nr x norb x norb, C order [n][l][m]     */
    const double* tau;     /* norb x 3 row-major orbital centres, Angstrom    */
} bse_tb_model;

/* One excitonic problem (q = 0 TDA).  Fields after dipole_dir are new in
 * ABI version 2 (dipole_model .. dipole_d: version 3); zero-initialised they
 * give a 2D system with model
 * screening. */
typedef struct {
    double bravais[9];     /* column-major 3x3, column i = a_i (93-96)        */
    int32_t n1, n2;        /* k grid, k = ((i+s1)/n1) b1 + ((j+s2)/n2) b2
                              (+ ((l+s3)/n3) b3 in 3D), k index
                              (i*n2 + j)*n3 + l (K_points, 183-242)           */
    double shift[2];
    int32_t n_occ;         /* occupied bands; conduction starts at n_occ       */
    int32_t nv, nc;        /* valence / conduction bands kept                 */
    double scissor;        /* eV added to conduction energies (816)           */
    double q[3];           /* exciton momentum; valence states at k - q.  Up to
                              |q| = 1e-4 1/A the optical limit q -> 0 (the
                              reference's minval, 29, lies inside it): the
                              exchange runs over v(|G|).  Beyond it a finite
                              momentum (adding_momentum, 2549): v(|q + G|),
                              G != 0 (2567, 2811-2816)                        */
    double r0, eps_env;    /* Rytova-Keldysh screening                        */
    double g_cut;          /* 2D G-list cutoff (319-385, parallelepiped)      */
    double f_s;            /* exchange spin factor, 2 - spinorial (2801)      */
    double exchange_scale; /* multiplies the exchange term                    */
    double direct_scale;   /* multiplies the screened direct term            */
    double dipole_dir[3];  /* direction qhat of the optical dipole (3562)     */
    int32_t dimension;     /* 2 (or 0): Rytova-Keldysh W and bare 2D v over the
                              in-plane G-list, kernel prefactor 1/(A Nk);
                              3: W = C/(eps_env (q^2 + radius)), v = C/(q^2 +
                              radius) (Coulomb_Potential::pull, 1281-1285)
                              over the 3D G-list (334-351), prefactor
                              1/(Omega Nk) (2846)                             */
    int32_t n3;            /* k grid along b3 (3D only; 0 or 1 in 2D)        */
    double shift3;         /* grid shift along b3, in grid steps             */
    double radius;         /* 3D regularisation of C/(q^2 + radius), 1/A^2   */
    int32_t screening;     /* 0: the model W above (r0, eps_env); 1: static
                              RPA from the TB states (Dielectric_Function::
                              pull_values, 2214-2303, order 1 at omega = 0;
                              W = eps^-1 v, 2575-2584)                        */
    int32_t rpa_nv, rpa_nc;/* RPA screening window: highest occupied / lowest
                              empty bands (0: all)                           */
    int32_t dipole_model;  /* 0: point orbitals, M_ii = i tau_i.qhat (1459-1501
                              in the small-q limit); 1: the Wannier position
                              matrix below, the real-space route's small-q
                              dipole (1535-1570, 1609-1650):
                              rho = i sum_ij conj(c_ck,i) Dq_ij(k) c_v,k-q,j,
                              Dq_ij(k) = sum_R e^{-ik.R} qhat.D_ij(R)         */
    int32_t dipole_nr;     /* lattice vectors of the position table          */
    const double* dipole_r;/* dipole_nr x 3 Cartesian R, Angstrom            */
    const double* dipole_d;/* real [R][i][j][3]: D_ij(R) = int w_i(r - R) r
                              w_j(r) dr (bse_wannier_integrals), norb x norb
                              per R in the model's orbital order              */
} bse_system;

/* Real-space Wannier functions on the grid of a supercell (Real_space_wannier,
 * bse.version13.cpp:1081-1212): n_sc[a] unit cells along a_a, p[a] points per
 * cell; point (cell c, sub-point s) lies at origin + sum_a (c_a + s_a/p_a) a_a.
 * Values: w[(spin*nw + f)*Npts + cell*P + point], cell = (c0 n1 + c1) n2 + c2,
 * point = (s0 p1 + s1) p2 + s2, P = p0 p1 p2 (the reference's
 * real_space_wannier_functions_list; the xsf reader in wannier.py fills it). */
typedef struct {
    int32_t nw;            /* Wannier functions per spin block                */
    int32_t nspin;         /* 1, or 2 (collinear: up block, then down block)  */
    int32_t n_sc[3];       /* unit cells of the supercell along a1, a2, a3    */
    int32_t p[3];          /* grid points per unit cell, at most 256 per
                              supercell axis (n_sc[a] p[a] <= 256)           */
    double origin[3];      /* Cartesian position of grid point 0, Angstrom    */
    double axis[9];        /* column-major, column a = unit-cell vector a_a   */
    const double* w;       /* host values, layout above                       */
} bse_wannier_grid;

/* ---- context ------------------------------------------------------------- */
int  bse_create(bse_ctx** ctx, int device);
int  bse_destroy(bse_ctx* ctx);
const char* bse_last_error(const bse_ctx* ctx);
int  bse_abi_version(void);

/* Upload a problem (model + parameters); computes nothing.  Replaces the
 * object construction of main (3686-3843): Crystal_Lattice, K_points,
 * G_points, Hamiltonian_TB, Coulomb_Potential. */
int  bse_set_problem(bse_ctx* ctx, const bse_tb_model* model, const bse_system* sys);

/* Same, with the k-points given as a list in the caller's order (crystal
 * coordinates, 3 x nk column-major, the k-list file of
 * K_points::push_k_points_list_values, 183-242, with its shift already
 * added).  The list must be a complete regular grid n1 x n2 (x n3 in 3D)
 * with one common offset, in any order and with any reciprocal-lattice
 * translation of each point; sys->n1, n2, n3 and the shifts are ignored and
 * taken from the list.  Anything else returns BSE_ERR_UNSUPPORTED.  The BSE
 * index k of every reference-order input and output (r = c*nv*Nk + v*Nk + k,
 * bse_tb_states' rows) is then the position in this list, as in the
 * reference. */
int  bse_set_problem_klist(bse_ctx* ctx, const bse_tb_model* model, const bse_system* sys,
                           const double* k_cryst, int nk);

/* Sizes of the loaded problem: out[0]=Nk, out[1]=N, out[2]=norb, out[3]=Ng. */
int  bse_problem_sizes(const bse_ctx* ctx, int64_t out[4]);

/* K1+K2: per-k H(k) and FP64 Hermitian eig for all grid k (conduction) and
 * k - q (valence).  Replaces Hamiltonian_TB::pull_ks_states_subset
 * (bse.version13.cpp:778-821) called for every k.
 *   eps  : Nk x (nv+nc) doubles, [k][slot], slots = valence(top-down) then
 *          conduction(bottom-up, + scissor); may be NULL
 *   coef : complex Nk x (nv+nc) x norb, [k][slot][orb]; may be NULL        */
int  bse_tb_states(bse_ctx* ctx, double* eps, double* coef);

/* K1+K2 at caller-given k: Hamiltonian_TB::pull_ks_states(k) (437, 706-777)
 * and pull_ks_states_subset(k, nv, nc) (438, 778-821) for a list of k.
 *   model : tight-binding model, or NULL for the loaded problem's model
 *   kpts  : 3 x nk Cartesian k (1/A), column-major
 *   nv = nc = 0 : every band, ascending (pull_ks_states); eps nk x norb,
 *                 coef complex nk x norb x norb ([k][band][orb])
 *   otherwise   : the window of pull_ks_states_subset -- valence bands
 *                 n_occ-1, n_occ-2, ... (nv of them), then conduction bands
 *                 n_occ, n_occ+1, ... (nc) with the scissor added; eps
 *                 nk x (nv+nc), coef complex nk x (nv+nc) x norb
 * Each vector is unit-norm with its largest component real and positive.
 * Used for band structures (pull_bands, 823-875) as well. */
int  bse_tb_states_at(bse_ctx* ctx, const bse_tb_model* model, const double* kpts, int nk,
                      int n_occ, int nv, int nc, double scissor, double* eps, double* coef);

/* Real-space Wannier integrals on the GPU (Dipole_Elements::
 * function_building_real_space_wannier_dipole_ij 1503-1533 and _small_q
 * 1535-1570; overlaps as in the orthogonality test 1165-1180), for every
 * same-spin pair (i, j) and lattice translation R = r_cells[r] (in unit
 * cells; the translated function's grid values come from cell c - R and are
 * zero outside the supercell, 1441-1456):
 *   A[s][i][j][r][q] = f sum_x w_si(x - R) e^{i K_q.x} w_sj(x)  (complex)
 *   D[s][i][j][r][a] = f sum_x w_si(x - R) x_a w_sj(x)
 *   S[s][i][j][r]    = f sum_x w_si(x - R) w_sj(x)
 * f = Omega_supercell / Npts; kvec: n_k x 3 Cartesian 1/A (n_k may be 0,
 * then A may be null).  Host arrays in and out.                             */
int  bse_wannier_integrals(bse_ctx* ctx, const bse_wannier_grid* grid, int n_r, const int32_t* r_cells,
                           int n_k, const double* kvec, double* A, double* D, double* S);

/* K3+K4: assemble the TDA matrix on the device in the given precision.
 * Replaces Excitonic_Hamiltonian::extract_hbse_and_rcv (3041-3047), i.e.
 * pull_coulomb_potentials (2547) + pull_resonant_part_and_rcv (2767). */
int  bse_build(bse_ctx* ctx, int precision);

/* Copy the assembled matrix out.  dst: complex N x N (ld >= N), doubles for
 * BSE_PREC_FP64 and floats otherwise; both triangles are written.           */
int  bse_get_matrix(bse_ctx* ctx, void* dst, int64_t ld, int dst_on_device, int order);

/* rho_r(G0) of the optical dipole (the rcv half of extract_hbse_and_rcv),
 * complex N, reference order. */
int  bse_get_dipoles(bse_ctx* ctx, double* rho);

/* K5-K8: eigendecomposition of the assembled matrix.  Replaces
 * Excitonic_Hamiltonian::common_diagonalization (3120-3368; eig_sym -> zheevd
 * at 3330).  All N eigenvalues are computed.
 *   n_vectors : eigenvectors formed for the lowest n_vectors (-1 = all)
 *   refine    : 1 = FP64 Rayleigh-Ritz refinement of the formed vectors
 *               against the FP64 matrix recomputed on the fly, iterated
 *               until the residuals stop decreasing (1 <= n_vectors <= 16,
 *               n_vectors <= N; BSE_ERR_NUMERIC if it does not converge) */
int  bse_eigh(bse_ctx* ctx, int n_vectors, int refine);
/* Outcome of the last refinement (new capability; the reference has none).
 * rnorm: residual norms ||H x_j - theta_j x_j|| in eV of the first n formed
 * vectors (n <= formed vectors, which include up to 4 guard vectors when a
 * cluster straddles the requested count); rounds: Rayleigh-Ritz rounds run.
 * Returns the refinement's status: BSE_OK when the largest residual of the
 * requested vectors is below 1e-7 ||H||, so that every refined value lies
 * within its residual of an eigenvalue; BSE_ERR_NUMERIC otherwise. */
int  bse_refine_info(const bse_ctx* ctx, double* rnorm, int n, int* rounds);
int  bse_get_eigenvalues(bse_ctx* ctx, double* evals);
/* complex N x nvec (ld >= N), reference order, lowest nvec columns. */
int  bse_get_eigenvectors(bse_ctx* ctx, double* z, int64_t ld, int nvec, int dst_on_device);

/* Matrix-free FP64 product Y = H X with the assembled system's matrix, every
 * element recomputed from the interaction tables (the kernel used by the
 * refinement and the gradients).  x, y: complex N x k host buffers,
 * reference order.  Requires bse_build. */
int  bse_apply(bse_ctx* ctx, int k, const double* x, int64_t ldx, double* y, int64_t ldy);

/* K9: oscillator strengths f_m = sum_r conj(rho_r) A_rm for all m (3486-3502)
 * and the TDA dielectric function (3621-3634):
 *   eps(w) = 1 - C/(Omega Nk) sum_m |f_m|^2 / (w - E_m + i eta).           */
int  bse_oscillator(bse_ctx* ctx, double* f);
int  bse_spectrum(bse_ctx* ctx, const double* omega, int n_omega, double eta,
                  double* eps_re, double* eps_im);

/* K10: Hellmann-Feynman gradients of the lowest n_exc exciton energies,
 * dE_m/dtheta = A_m^H (dH/dtheta) A_m, with dH recomputed on the fly (no
 * stored derivative matrix).  theta columns: 0 relative scale lambda of the
 * loaded model's R != 0 blocks (H(R) -> lambda H(R), at lambda = 1), 1 scissor, 2 Keldysh r0, 3 eps_env, 4 exchange field h
 * (h * diag(+1,-1,+1,...) added to H(R=0)).  which: 5 flags (NULL = all).
 * dE: n_exc x 5, column-major.  Requires bse_eigh with n_vectors >= n_exc.
 * No counterpart in the reference (new capability). */
int  bse_grad_energies(bse_ctx* ctx, int n_exc, const int* which, double* dE);

/* K10 (spectrum): L = sum_w (Im eps(w) - target(w))^2 and dL/dtheta for the
 * same five parameters, from the resolvent identity
 *   dL/dtheta = -2 pref Im[tr(dH B) + drho^H G rho + rho^H G drho],
 * B = A K^T A^H with the frequency sum folded in (three GEMMs), dH tiles
 * recomputed on the fly.  Requires bse_eigh(ctx, -1, 0) (all eigenvectors).
 * which: 5 flags (NULL = all); L (may be NULL); dL: 5 doubles.
 * No counterpart in the reference (new capability). */
int  bse_grad_spectrum(bse_ctx* ctx, const double* omega, int n_omega, double eta,
                       const double* target, const int* which, double* L, double* dL);

/* Dense Hermitian eig of any matrix (drop-in for arma::eig_sym, zheevd):
 * a: complex n x n (lower triangle read), doubles, host or device.
 * evals: n ascending; z: complex n x n_vectors (may be NULL).             */
int  bse_eigh_dense(bse_ctx* ctx, int n, const double* a, int64_t lda, int a_on_device,
                    int precision, int n_vectors, double* evals, double* z, int64_t ldz);

/* Stage timings of the last build/eigh (milliseconds, HIP events):
 * [0] tb, [1] tables, [2] assembly, [3] tridiagonalisation, [4] tridiagonal
 * eigenvalues, [5] tridiagonal vectors, [6] back-transform, [7] refinement,
 * [8] oscillator+spectrum; after a two-stage reduction (few vectors, large
 * N) [9] dense -> band with Q1^H rho and [10] band -> tridiagonal, the two
 * parts of [3] (0 after a one-stage reduction).  n <= 16. */
int  bse_stage_times(const bse_ctx* ctx, double* ms, int n);

/* Kernel-level profile of the tridiagonalisation's dominant kernels (HIP
 * events around every launch on the context stream).  enable switches
 * recording on for later calls; out (may be NULL, else 9 doubles) receives,
 * for the last bse_eigh:
 *   one-stage reduction: [0] total time in ms of the Hermitian matrix-vector
 *     kernel, [1] its algorithmic bytes (lower triangle of every trailing
 *     block), [2] launches, [4] = 0;
 *   two-stage reduction: [0] ms, [1] = 1 when the products ran on FP32
 *     matrix cores, [2] launches, [3] algorithmic flops of the panels'
 *     rank-2k update (8 m (m+1) k per panel), [4] = 1, and [5] ms, [6] flops
 *     (8 m^2 k), [7] launches of the Hermitian block product, [8] = 2; [8] = 3
 *     when the look-ahead updated each next panel's tile column on the second
 *     stream, and [0], [3] then cover the other tiles (8 m_r (m_r+1) k, m_r
 *     the order beyond that tile column); [8] = 4 as 3, with the update's
 *     operands packed once per panel in FP32 planes. */
int  bse_profile(bse_ctx* ctx, int enable, double* out);

/* ---- K11: batched sweep of many small systems ------------------------------
 * Many independent excitonic problems of uneven size (a parameter sweep) run
 * as ONE graph launch: per-system TB states, tables and assembly, one
 * lockstep Householder reduction whose every step serves all systems at
 * once, then per-system tridiagonal eigenpairs, oscillator strengths,
 * lowest-exciton vectors with optional FP64 refinement and the spectrum.
 * Tile shapes (hemv chunk length per trailing size, trailing-update kernel)
 * are picked by the library's autotuner on the batch itself.
 * No counterpart in the reference (which handles one system per run); each
 * system's results follow bse_build + bse_eigh + bse_spectrum.
 * Limits: 2 <= N <= 4096 per system, 1 <= n_lowest <= 16.               */
typedef struct {
    int32_t precision;     /* BSE_PREC_*: assembly and storage precision       */
    int32_t n_lowest;      /* excitons returned (and refined) per system       */
    int32_t refine;        /* 1: FP64 Rayleigh-Ritz of the lowest n_lowest      */
    int32_t n_omega;       /* spectrum points per system (0: no spectrum)      */
    const double* omega;   /* n_omega frequencies in eV (host)                 */
    double eta;            /* Lorentzian broadening in eV                      */
} bse_batch_options;

/* Upload all systems, size the device arena, autotune and record the graph.
 * models[i] / systems[i] describe system i (as bse_set_problem).          */
int bse_batch_prepare(bse_ctx* ctx, int n_systems, const bse_tb_model* models,
                      const bse_system* systems, const bse_batch_options* opt);
/* One launch of the recorded graph (inputs stay resident between runs).
 * total_ms: launch-to-completion time (HIP events); system_ms (may be NULL):
 * per system, time from the launch to that system's last kernel (device
 * wall clock), in the order given to bse_batch_prepare.                    */
int bse_batch_run(bse_ctx* ctx, double* total_ms, double* system_ms);
/* Results of the last run, in input order: evals n_systems x n_lowest
 * (row per system), eps_re / eps_im n_systems x n_omega (may be NULL).    */
int bse_batch_results(bse_ctx* ctx, double* evals, double* eps_re, double* eps_im);
/* info[0] graph nodes, [1] device arena bytes, [2] largest N, [3] systems,
 * [4] hemv variant chosen for the largest trailing size (chunk length in
 *     tile rows; plus 32 for the two-row-blocks-per-step form or 64 for the
 *     per-slab column-sum form; plus 128 for default-policy instead of
 *     non-temporal matrix loads),
 * [5] trailing-update kernel (0 FP64 MFMA on 64 x 64 tiles, 1 FP64 VALU,
 *     2 FP64 MFMA on 128 x 128 tiles).                                  */
int bse_batch_info(const bse_ctx* ctx, int64_t info[6]);

/* ---- full BSE beyond Tamm-Dancoff -------------------------------------------
 * bse_build_full assembles the resonant block A (as bse_build with
 * BSE_PREC_FP64) and the coupling block B (complex symmetric) and composes
 * M = [[A, B], [conj B, conj A]] (FP64, order 2N).  Replaces
 * extract_hbse_and_rcv with tamn_dancoff = 0 (add_coupling_part, 2929-3040).
 * The anti-resonant half of rho is rho_vc = -conj(rho_cv) (rows N..2N of
 * rho_q_diagk_cv, 2951). */
int bse_build_full(bse_ctx* ctx);
/* B, N x N complex (ld >= N), reference order, both triangles, host. */
int bse_get_coupling(bse_ctx* ctx, double* dst, int64_t ld);
/* Eigenpairs of H = [[A, B], [-conj B, -conj A]] by a structure-preserving
 * solve: Cholesky M = L L^H, the Hermitian W = L^H s3 L (s3 = diag(1, -1)),
 * its eigenvalues +-E and, for the lowest n_vectors positive ones (-1 =
 * all N), x = sqrt(E) L^-H y.  Replaces the eig_gen branch of
 * common_diagonalization (3326) and the augmentation matrix of the non-TDA
 * spectrum (pull_augmentation_matrix 3504-3522, sum 3607-3620), which the
 * pseudo-orthonormality x^H s3 x = 1 makes unnecessary.  Afterwards
 * bse_get_eigenvalues gives the N positive energies (ascending),
 * bse_oscillator f = rho~^H x (N), bse_get_eigenvectors the 2N-row vectors
 * (x; y) (ld >= 2N, each half in reference order) and bse_spectrum
 *   eps(w) = 1 - C/(Omega Nk) sum_E |f|^2 [1/(w - E + i eta) - 1/(w + E + i eta)].
 * BSE_ERR_NUMERIC when M is not positive definite (the system is unstable
 * beyond Tamm-Dancoff; the reference's eig_gen then returns complex
 * energies).  Requires bse_build_full; consumes M. */
int bse_eigh_full(bse_ctx* ctx, int n_vectors);

/* ---- allocation-free dense eig --------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory while they run (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): bse_eigh_dense with every input and output
 * on the device (a: complex n x n, lower triangle read; evals: n doubles; z:
 * complex n x n_vectors, ld ldz, may be NULL when n_vectors = 0) and every
 * work buffer cut from one fixed device array of 64 MiB that the library
 * holds per GPU until the process ends.  n <= 512 (BSE_ERR_ARG beyond);
 * precision and n_vectors as for bse_eigh_dense.  It calls no hipMalloc or
 * hipFree.  Calls on one GPU share that GPU's array and leave its contents
 * behind: calls that overlap on one GPU (two contexts on two host threads)
 * race.  The call leaves no eigen-results in the context (the getters then
 * report BSE_ERR_STATE). */
int bse_eigh_dense_noalloc(bse_ctx* ctx, int n, const double* a, int64_t lda, int precision,
                           int n_vectors, double* evals, double* z, int64_t ldz);
/* info[0] bytes of the fixed array, [1] largest n, [2] bytes the last
 * allocation-free call used. */
int bse_noalloc_info(const bse_ctx* ctx, int64_t info[3]);

/* ---- instruction-rate probe ----------------------------------------------
 * No counterpart in the reference; for the roofline report.  The rate in
 * TFLOP/s that one instruction form reaches on the current device with no
 * memory traffic (best of three launches of a few ms, on the null stream):
 * which = 0 v_mfma_f32_16x16x4 (the band reduction's products), 1
 * v_mfma_f64_16x16x4 (FP64 back-transform and matrix-free apply), 2 FP64
 * vector FMA.  Allocates and frees a small device buffer. */
int bse_probe_rate(int which, double* tflops);

#ifdef __cplusplus
}
#endif
#endif /* BSE_MI355X_H */
