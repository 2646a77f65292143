// Small dense FP64 linear algebra for the Rayleigh-Ritz refinement (K8):
// everything lives in one workgroup's LDS (n <= 64).
//
// * Hermitian eigensolver: cyclic Jacobi with the Brent-Luk parallel ordering
//   (n/2 disjoint rotations per step, n-1 steps per sweep), each rotation the
//   complex 2x2 Jacobi rotation of Golub & Van Loan sec. 8.5, applied as a
//   block J^H A J.
// * Rayleigh-Ritz with SVQB basis stabilisation (Stathopoulos & Wu 2002): for
//   a basis S (not orthonormal) with G = S^H S and Hs = S^H H S, eig G =
//   U diag(s) U^H, P = U diag(s^-1/2) (dropping s < tol s_max), eig of
//   P^H Hs P = Y diag(l) Y^H, coefficients C = P Y of the lowest k Ritz pairs.
#include "bse_common.hpp"

#define SD_MAX 40

// Jacobi on the LDS matrix A (ld SD_MAX), eigenvectors into V.  n <= SD_MAX.
// On exit the diagonal of A holds the eigenvalues (unsorted).  One thread per
// (pair, row) task: blockDim.x >= SD_JACOBI_THREADS.  A step reads A and
// writes the rotated columns into A2, then reads A2 and writes the rotated
// rows back into A, so no phase reads an entry another thread writes; every
// thread forms its pair's rotation from A itself and keeps it for both
// phases (two barriers per step).
#define SD_JACOBI_THREADS ((SD_MAX / 2) * SD_MAX)
__device__ void sd_jacobi(int n, double2 (*A)[SD_MAX + 1], double2 (*V)[SD_MAX + 1])
{
    __shared__ double2 A2[SD_MAX][SD_MAX + 1];
    __shared__ double sh[32];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int m = (n + 1) & ~1;   // even number of slots; slot n (if n odd) is a dummy
    for (int e = tid; e < SD_MAX * SD_MAX; e += nt) {
        const int i = e / SD_MAX, j = e % SD_MAX;
        V[i][j] = cmk(i == j ? 1.0 : 0.0, 0.0);
    }
    // round-robin order at step r of a sweep: slot 0 holds 0, slot i >= 1
    // holds ((i - 1 - r) mod (m - 1)) + 1 (slots 1.. shift right by one per
    // step; after m - 1 steps the order is the identity again)
    auto slot = [m](int i, int r) { return i == 0 ? 0 : ((i - 1 - r) % (m - 1) + (m - 1)) % (m - 1) + 1; };
    __syncthreads();
    double fro = 0.0;
    for (int e = tid; e < n * n; e += nt) fro += cabs2(A[e / n][e % n]);
    fro = block_sum(fro, sh);
    const bool act = tid < (m / 2) * n;
    const int t = act ? tid / n : 0, i = act ? tid % n : 0;
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0;
        for (int e = tid; e < n * n; e += nt) {
            const int r = e / n, j = e % n;
            if (r != j) off += cabs2(A[r][j]);
        }
        off = block_sum(off, sh);
        if (off <= 1e-30 * fro) break;
        for (int step = 0; step < m - 1; ++step) {
            // pair t of this step: (slot t, slot m-1-t); rotation from A
            int p = slot(t, step), q = slot(m - 1 - t, step);
            if (p > q) { const int x = p; p = q; q = x; }
            double c = 1.0, s = 0.0;
            double2 u = cmk(1.0, 0.0);
            if (act && q < n) {
                const double2 apq = A[p][q];
                const double absa = sqrt(cabs2(apq));
                if (absa > 0.0) {
                    const double app = A[p][p].x, aqq = A[q][q].x;
                    u = cmk(apq.x / absa, -apq.y / absa);
                    const double theta = (aqq - app) / (2.0 * absa);
                    double tt = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
                    if (theta < 0.0) tt = -tt;
                    c = 1.0 / sqrt(1.0 + tt * tt);
                    s = tt * c;
                }
            }
            const bool rot = act && q < n && s != 0.0;
            // columns: (A J)(i, {p, q}) into A2, V J in place, with
            // J = [[c, s], [-s u, c u]] on (p, q); other pairs copied
            if (act) {
                const double2 aip = A[i][p];
                if (rot) {
                    const double2 aiq = A[i][q];
                    const double2 su = cscale(u, s), cu = cscale(u, c);
                    A2[i][p] = csub(cscale(aip, c), cmul(aiq, su));
                    A2[i][q] = cadd(cscale(aip, s), cmul(aiq, cu));
                    const double2 vip = V[i][p], viq = V[i][q];
                    V[i][p] = csub(cscale(vip, c), cmul(viq, su));
                    V[i][q] = cadd(cscale(vip, s), cmul(viq, cu));
                } else {
                    A2[i][p] = aip;
                    if (q < n) A2[i][q] = A[i][q];
                }
            }
            __syncthreads();
            // rows: J^H (A J) back into A; the rotated pair's off-diagonal
            // entries are set to zero and its diagonal to real
            if (act) {
                const int j = i;
                const double2 bpj = A2[p][j];
                if (rot) {
                    const double2 bqj = A2[q][j];
                    const double2 uc = cconj(u);
                    double2 np = csub(cscale(bpj, c), cmul(bqj, cscale(uc, s)));
                    double2 nq = cadd(cscale(bpj, s), cmul(bqj, cscale(uc, c)));
                    if (j == q) { np = cmk(0, 0); nq.y = 0.0; }
                    if (j == p) { nq = cmk(0, 0); np.y = 0.0; }
                    A[p][j] = np;
                    A[q][j] = nq;
                } else {
                    A[p][j] = bpj;
                    if (q < n) A[q][j] = A2[q][j];
                }
            }
            __syncthreads();
        }
    }
}

// Sort the Jacobi result ascending: writes eigenvalues to w[0..n) and the
// permutation to perm[0..n) (shared).
__device__ void sd_sort(int n, double2 (*A)[SD_MAX + 1], double* w, int* perm)
{
    __shared__ double tmp[SD_MAX];
    if (threadIdx.x == 0) {
        for (int i = 0; i < n; ++i) { perm[i] = i; w[i] = A[i][i].x; }
        for (int i = 1; i < n; ++i) {
            const int pi = perm[i];
            const double x = w[pi];
            int j = i - 1;
            while (j >= 0 && w[perm[j]] > x) { perm[j + 1] = perm[j]; --j; }
            perm[j + 1] = pi;
        }
        for (int i = 0; i < n; ++i) tmp[i] = w[perm[i]];
        for (int i = 0; i < n; ++i) w[i] = tmp[i];
    }
    __syncthreads();
}

// Rayleigh-Ritz with SVQB: inputs G, Hs (m x m, ld m, device), outputs the
// lowest k Ritz values (evals) and the m x k coefficient matrix C (ld m).
// Returns the rank kept in info[0].
__global__ __launch_bounds__(SD_JACOBI_THREADS) void k_small_rr(int m, int k, const double2* __restrict__ G,
                                                 const double2* __restrict__ Hs, double tol,
                                                 double* __restrict__ evals,
                                                 double2* __restrict__ C, int* __restrict__ info)
{
    __shared__ double2 A[SD_MAX][SD_MAX + 1];
    __shared__ double2 V[SD_MAX][SD_MAX + 1];
    __shared__ double2 P[SD_MAX][SD_MAX + 1];
    __shared__ double w[SD_MAX];
    __shared__ int perm[SD_MAX];
    __shared__ int rank;
    const int tid = threadIdx.x, nt = blockDim.x;
    // 1. eig of G
    for (int e = tid; e < m * m; e += nt) A[e % m][e / m] = G[e];
    __syncthreads();
    sd_jacobi(m, A, V);
    sd_sort(m, A, w, perm);
    // 2. P = U diag(s^-1/2) over the kept (largest) eigenvalues, in descending order
    if (tid == 0) {
        const double smax = w[m - 1];
        int r = 0;
        for (int i = m - 1; i >= 0; --i)
            if (w[i] > tol * smax) ++r;
        rank = r;
    }
    __syncthreads();
    const int r = rank;
    for (int e = tid; e < m * r; e += nt) {
        const int i = e % m, j = e / m;            // column j of P <- eigenvector perm[m-1-j]
        const int src = perm[m - 1 - j];
        P[i][j] = cscale(V[i][src], 1.0 / sqrt(w[m - 1 - j]));
    }
    __syncthreads();
    // 3. Hr = P^H Hs P  (r x r) into A
    for (int e = tid; e < m * r; e += nt) {          // V <- Hs P  (m x r)
        const int i = e % m, j = e / m;
        double2 s = cmk(0, 0);
        for (int l = 0; l < m; ++l) cfma(s, Hs[i + l * m], P[l][j]);
        V[i][j] = s;
    }
    __syncthreads();
    for (int e = tid; e < r * r; e += nt) {
        const int i = e % r, j = e / r;
        double2 s = cmk(0, 0);
        for (int l = 0; l < m; ++l) cfmac(s, P[l][i], V[l][j]);
        A[i][j] = s;
    }
    __syncthreads();
    for (int e = tid; e < r * r; e += nt) {         // exact Hermitian symmetrisation
        const int i = e % r, j = e / r;
        if (i < j) {
            const double2 a = A[i][j], b = A[j][i];
            const double2 h = cmk(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
            A[i][j] = h;
            A[j][i] = cconj(h);
        } else if (i == j) {
            A[i][i].y = 0.0;
        }
    }
    __syncthreads();
    // 4. eig of Hr, lowest k
    sd_jacobi(r, A, V);
    sd_sort(r, A, w, perm);
    const int kk = k < r ? k : r;
    for (int e = tid; e < m * kk; e += nt) {         // C = P Y(:, lowest)
        const int i = e % m, j = e / m;
        const int src = perm[j];
        double2 s = cmk(0, 0);
        for (int l = 0; l < r; ++l) cfma(s, P[i][l], V[l][src]);
        C[i + j * m] = s;
    }
    for (int j = tid; j < kk; j += nt) evals[j] = w[j];
    if (tid == 0) info[0] = r;
}

extern "C" hipError_t bse_launch_small_rr(hipStream_t st, int m, int k, const double2* G,
                                          const double2* Hs, double tol, double* evals,
                                          double2* C, int* info)
{
    if (m > SD_MAX || k > m) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_small_rr, dim3(1), dim3(SD_JACOBI_THREADS), 0, st, m, k, G, Hs, tol, evals, C, info);
    return hipGetLastError();
}

// R = Y - X diag(lam) with lam_j = Re(x_j^H y_j) / (x_j^H x_j), then scale each
// residual column to unit norm (columns with zero residual stay zero).  One
// workgroup per column.  rnorm_j = ||y_j - lam_j x_j|| / ||x_j||; rq (may be
// null) receives lam_j.
__global__ __launch_bounds__(256) void k_residual(int64_t N, const double2* __restrict__ X,
                                                 const double2* __restrict__ Y, int64_t ld,
                                                 double2* __restrict__ R, double* __restrict__ rnorm,
                                                 double* __restrict__ rq)
{
    __shared__ double sh[32];
    const int j = blockIdx.x;
    const double2* x = X + (int64_t)j * ld;
    const double2* y = Y + (int64_t)j * ld;
    double2* r = R + (int64_t)j * ld;
    double xy = 0.0, xx = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
        xy += x[i].x * y[i].x + x[i].y * y[i].y;
        xx += cabs2(x[i]);
    }
    xy = block_sum(xy, sh);
    xx = block_sum(xx, sh);
    const double lam = xy / xx;
    double rr = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
        const double2 v = csub(y[i], cscale(x[i], lam));
        r[i] = v;
        rr += cabs2(v);
    }
    rr = block_sum(rr, sh);
    const double inv = (rr > 0.0) ? 1.0 / sqrt(rr) : 0.0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) r[i] = cscale(r[i], inv);
    if (threadIdx.x == 0) {
        rnorm[j] = sqrt(rr / xx);
        if (rq) rq[j] = lam;
    }
}

extern "C" hipError_t bse_launch_residual(hipStream_t st, int64_t N, int k, const double2* X,
                                          const double2* Y, int64_t ld, double2* R, double* rnorm,
                                          double* rq)
{
    hipLaunchKernelGGL(k_residual, dim3(k), dim3(256), 0, st, N, X, Y, ld, R, rnorm, rq);
    return hipGetLastError();
}

// out[j] = Re(x_j^H y_j) for k columns (one workgroup per column).
__global__ __launch_bounds__(256) void k_coldot(int64_t N, const double2* __restrict__ X,
                                               const double2* __restrict__ Y, int64_t ld,
                                               double* __restrict__ out)
{
    __shared__ double sh[32];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
        const double2 a = X[i + (int64_t)j * ld], b = Y[i + (int64_t)j * ld];
        s += a.x * b.x + a.y * b.y;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[j] = s;
}

extern "C" hipError_t bse_launch_coldot(hipStream_t st, int64_t N, int k, const double2* X,
                                        const double2* Y, int64_t ld, double* out)
{
    hipLaunchKernelGGL(k_coldot, dim3(k), dim3(256), 0, st, N, X, Y, ld, out);
    return hipGetLastError();
}
