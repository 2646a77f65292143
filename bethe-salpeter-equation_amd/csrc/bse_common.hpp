// Shared device/host helpers for the gfx950 excitonic kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

// Largest number of orbitals (spin components included) per cell.
#define BSE_MAXORB 16

// Number of floor(log2(size)) buckets of the per-device autotune tables.
#define BSE_HEMV_BUCKETS 24

// ---------------------------------------------------------------------------
// Complex arithmetic on HIP vector types (double2 = FP64 complex, float2 =
// FP32 complex); .x = real part, .y = imaginary part.
// ---------------------------------------------------------------------------
__host__ __device__ inline double2 cmk(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ inline double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
// acc += a * b
__host__ __device__ inline void cfma(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
// acc += conj(a) * b
__host__ __device__ inline void cfmac(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
}
__host__ __device__ inline double2 cdiv(double2 a, double2 b) {
    double s = fabs(b.x) + fabs(b.y);
    double br = b.x / s, bi = b.y / s;
    double ar = a.x / s, ai = a.y / s;
    double d = br * br + bi * bi;
    return make_double2((ar * br + ai * bi) / d, (ai * br - ar * bi) / d);
}

// Storage conversions between the matrix element types and FP64 complex.
template <typename T> struct cplx;
template <> struct cplx<double2> {
    __device__ static inline double2 load(const double2* p) { return *p; }
    __device__ static inline void store(double2* p, double2 v) { *p = v; }
};
template <> struct cplx<float2> {
    __device__ static inline double2 load(const float2* p) {
        float2 v = *p;
        return make_double2((double)v.x, (double)v.y);
    }
    __device__ static inline void store(float2* p, double2 v) {
        *p = make_float2((float)v.x, (float)v.y);
    }
};

// Streaming (non-temporal) loads of complex elements.
typedef float bse_f2v __attribute__((ext_vector_type(2)));
typedef double bse_d2v __attribute__((ext_vector_type(2)));
__device__ inline float2 nt_load(const float2* p) {
    const bse_f2v v = __builtin_nontemporal_load(reinterpret_cast<const bse_f2v*>(p));
    return make_float2(v.x, v.y);
}
__device__ inline double2 nt_load(const double2* p) {
    const bse_d2v v = __builtin_nontemporal_load(reinterpret_cast<const bse_d2v*>(p));
    return make_double2(v.x, v.y);
}

// Loads through the global address space (global_load_* rather than flat
// loads, which also count against the LDS wait counter).
typedef __attribute__((address_space(1))) const bse_f2v bse_gf2v;
typedef __attribute__((address_space(1))) const bse_d2v bse_gd2v;
__device__ inline float2 gload(const float2* p) {
    const bse_f2v v = *(bse_gf2v*)p;
    return make_float2(v.x, v.y);
}
__device__ inline double2 gload(const double2* p) {
    const bse_d2v v = *(bse_gd2v*)p;
    return make_double2(v.x, v.y);
}

// Buffer loads: a wave-uniform base in a buffer descriptor (SGPRs), a
// per-lane byte offset (one VGPR) and a uniform byte offset (SGPR), so a
// wave reading 16 columns of a column-major matrix keeps one address VGPR
// instead of sixteen 64-bit addresses.  Descriptor word 3 = 0x00020000
// (gfx9 family); loads past num_records return zero.
typedef int bse_i2v __attribute__((ext_vector_type(2)));
typedef int bse_i4v __attribute__((ext_vector_type(4)));
__device__ inline __amdgpu_buffer_rsrc_t bse_rsrc(const void* base, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, bytes, 0x00020000);
}
// Bit casts of single vector elements go through a scalar argument: a
// __builtin_bit_cast applied directly to an element access (v.y) was compiled
// as a cast of the vector's low 32 bits, dropping the imaginary part.
__device__ inline float bse_i2f(int x) { return __builtin_bit_cast(float, x); }
// Buffer loads of one complex element; AUX is the cache-policy operand of the
// load (BSE_LOAD_STREAM marks data that is read once: non-temporal).
#define BSE_LOAD_STREAM 2
template <typename T> struct bufld;
template <> struct bufld<float2> {
    template <int AUX = 0>
    __device__ static inline float2 ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
        const bse_i2v v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX);
        const int re = v[0], im = v[1];
        return make_float2(bse_i2f(re), bse_i2f(im));
    }
};
template <> struct bufld<double2> {
    template <int AUX = 0>
    __device__ static inline double2 ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
        const bse_i4v v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, AUX);
        const bse_i2v lo = {v.x, v.y}, hi = {v.z, v.w};
        return make_double2(__builtin_bit_cast(double, lo), __builtin_bit_cast(double, hi));
    }
};

// ---------------------------------------------------------------------------
// Wave-level reductions (64-lane wavefronts on CDNA4).
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The same sum through DPP lane moves (no LDS round trips): quad
// permutations, half-row and row mirrors give every lane its row's sum of 16;
// row_bcast15 and row_bcast31 carry the row sums up to lane 63, which is read
// back into every lane.  A fixed order, but not the butterfly's.
template <int CTRL, int ROWS>
__device__ inline double dpp_mov_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double wave_sum_dpp(double v) {
    v += dpp_mov_d<0xB1, 0xF>(v);     // quad_perm [1,0,3,2]
    v += dpp_mov_d<0x4E, 0xF>(v);     // quad_perm [2,3,0,1]
    v += dpp_mov_d<0x141, 0xF>(v);    // row_half_mirror
    v += dpp_mov_d<0x140, 0xF>(v);    // row_mirror
    v += dpp_mov_d<0x142, 0xA>(v);    // row_bcast15 into rows 1 and 3
    v += dpp_mov_d<0x143, 0xC>(v);    // row_bcast31 into rows 2 and 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}
// Wave maximum of a non-negative double by the same DPP steps; lanes that a
// row-masked step leaves out take max(v, 0), which changes nothing.
__device__ inline double wave_max_dpp(double v) {
    v = fmax(v, dpp_mov_d<0xB1, 0xF>(v));
    v = fmax(v, dpp_mov_d<0x4E, 0xF>(v));
    v = fmax(v, dpp_mov_d<0x141, 0xF>(v));
    v = fmax(v, dpp_mov_d<0x140, 0xF>(v));
    v = fmax(v, dpp_mov_d<0x142, 0xA>(v));
    v = fmax(v, dpp_mov_d<0x143, 0xC>(v));
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}
// 1 / a for complex a as conj(a) / |a|^2: one real division.
__host__ __device__ inline double2 crecip(double2 a) {
    const double r = 1.0 / (a.x * a.x + a.y * a.y);
    return make_double2(a.x * r, -a.y * r);
}
// Block reduction of a double over blockDim.x threads (multiple of 64, <= 1024).
// Every thread receives the total.  Deterministic order.
__device__ inline double block_sum(double v, double* sh /* >= 16 doubles */) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) t += sh[i];
    return t;
}

__device__ inline double2 block_sum2(double2 v, double* sh /* >= 32 doubles */) {
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { sh[2 * w] = v.x; sh[2 * w + 1] = v.y; }
    __syncthreads();
    double2 t = make_double2(0.0, 0.0);
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) { t.x += sh[2 * i]; t.y += sh[2 * i + 1]; }
    return t;
}

// "Last block done" ticket: returns true in exactly one block (the last to
// arrive) after every block has published its partial results.  The counter
// is a 32-bit device word that the last block resets to zero.
// Every thread releases its stores at agent scope; only the last block
// acquires, so the other blocks never invalidate their XCD's L2.
__device__ inline bool last_block_ticket(unsigned int* counter, unsigned int nblocks) {
    __shared__ bool is_last;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int prev =
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (prev == nblocks - 1);
    }
    __syncthreads();
    if (is_last) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (threadIdx.x == 0)
            __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return is_last;
}

// Output-row layout of the 16x16x4 MFMA tiles (measured on gfx950): lane l,
// accumulator register r holds D[row][l % 16] with
//   f32 (v_mfma_f32_16x16x4, and the bf16 16x16x32 variants): row = 4 (l/16) + r
//   f64 (v_mfma_f64_16x16x4):                                  row = (l/16) + 4 r
__device__ inline int mfma16_row_f32(int lane, int r) { return 4 * (lane >> 4) + r; }
__device__ inline int mfma16_row_f64(int lane, int r) { return (lane >> 4) + 4 * r; }

__host__ __device__ inline int64_t ceil_div64(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// k grid n1 x n2 x n3 (n3 = 1 for 2D systems), k index (i n2 + j) n3 + l.
// The interaction table holds W_ij(Q) for every grid difference
// Q = (di/n1) b1 + (dj/n2) b2 + (dl/n3) b3, |di| < n1, ..., ordered as a
// mixed-radix number with digits di + n1 - 1, dj + n2 - 1, dl + n3 - 1 of
// widths 2 n1 - 1, 2 n2 - 1, 2 n3 - 1.  kcode(k) writes (i, j, l) in that
// radix, so table_index(k, k') = kcode(k) - kcode(k') + kcode_origin is the
// entry of k - k' (each digit difference stays inside its width, so the
// subtraction never borrows).  For n3 = 1 this is (di + n1 - 1)(2 n2 - 1) +
// dj + n2 - 1.
// ---------------------------------------------------------------------------
__host__ __device__ inline int kcode(int k, int n2, int n3) {
    const int l = k % n3, ij = k / n3;
    return ((ij / n2) * (2 * n2 - 1) + ij % n2) * (2 * n3 - 1) + l;
}
__host__ __device__ inline int kcode_origin(int n1, int n2, int n3) {
    return ((n1 - 1) * (2 * n2 - 1) + (n2 - 1)) * (2 * n3 - 1) + (n3 - 1);
}
__host__ __device__ inline int ktable_index(int k, int kp, int n1, int n2, int n3) {
    return kcode(k, n2, n3) - kcode(kp, n2, n3) + kcode_origin(n1, n2, n3);
}
__host__ __device__ inline int64_t ktable_size(int n1, int n2, int n3) {
    return (int64_t)(2 * n1 - 1) * (2 * n2 - 1) * (2 * n3 - 1);
}

// ---------------------------------------------------------------------------
// Storage of the assembled N x N Hermitian matrix: lower triangle only, in
// 64 x 64 tiles that are contiguous in memory (column-major inside a tile),
// tiles ordered by tile column and, inside a column, by tile row from the
// diagonal down.  A tile column is one contiguous run, so a workgroup reading
// a tile reads 32 KB (FP32) or 64 KB (FP64) of consecutive memory, and the
// upper triangle is not stored (half the memory of a full square).
// ---------------------------------------------------------------------------
// Exciton momenta up to this length (1/Angstrom) are the optical limit
// q -> 0: the exchange then runs over v(|G|); the reference's minval = 1e-5
// (line 29) lies inside it.
#define BSE_OPTICAL_Q 1e-4

#define BSE_TILE 64
__host__ __device__ inline int tile_count(int64_t n) { return (int)ceil_div64(n, BSE_TILE); }
// Position of tile (ti, tj), ti >= tj, among nt x nt lower tiles.
__host__ __device__ inline int64_t tile_pos(int ti, int tj, int nt) {
    return (int64_t)tj * nt - (int64_t)tj * (tj - 1) / 2 + (ti - tj);
}
// Element offset of (r, c), r >= c (or any (r, c) inside a diagonal tile).
__host__ __device__ inline int64_t tix(int64_t r, int64_t c, int nt) {
    return tile_pos((int)(r >> 6), (int)(c >> 6), nt) * (BSE_TILE * BSE_TILE) + (r & 63) + ((c & 63) << 6);
}
// Elements of the tiled lower storage of an n x n matrix.
__host__ __device__ inline int64_t tiled_elems(int64_t n) {
    const int64_t nt = tile_count(n);
    return nt * (nt + 1) / 2 * BSE_TILE * BSE_TILE;
}
// Lower tile k of an m x m tile triangle, enumerated row by row: (ti, tj).
__host__ __device__ inline void tri_tile(int64_t k, int& ti, int& tj) {
    int64_t r = (int64_t)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > k) --r;
    while ((r + 1) * (r + 2) / 2 <= k) ++r;
    ti = (int)r;
    tj = (int)(k - r * (r + 1) / 2);
}

// Compact-WY factor of one Householder panel (LAPACK larft, forward,
// columnwise): Q = H_0 H_1 ... H_{nb-1} = I - V T V^H with T upper
// triangular, built column by column as
//   T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i),  T(i, i) = tau_i,
// from the Gram column G(l, i) = V(:, l)^H v_i (ld 64).  One block of >= nb
// threads (nb <= 64); col is a 64-entry shared scratch.
__device__ __attribute__((always_inline)) inline void larft_body(int nb, const double2* __restrict__ tau,
                                  const double2* __restrict__ G, double2* __restrict__ T, int ldt,
                                  double2* col)
{
    if (nb <= 32) {
        // the recurrence runs in LDS (global round trips per column would
        // dominate); T is written out once at the end
        __shared__ double2 sT[32][33], sG[32][33], stau[32];
        for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
            const int r = e & 31, c = e >> 5;
            sT[r][c] = cmk(0, 0);
            sG[r][c] = (r < nb && c < nb) ? G[r + c * 64] : cmk(0, 0);
        }
        if ((int)threadIdx.x < nb) stau[threadIdx.x] = tau[threadIdx.x];
        __syncthreads();
        for (int i = 0; i < nb; ++i) {
            const double2 ti = stau[i];
            const int r = threadIdx.x;
            double2 cv = cmk(0, 0);
            if (r < i) {
                double2 sacc = cmk(0, 0);
                for (int l = r; l < i; ++l) cfma(sacc, sT[r][l], sG[l][i]);
                cv = cscale(cmul(ti, sacc), -1.0);
            }
            __syncthreads();
            if (r < i) sT[r][i] = cv;
            if (r == 0) sT[i][i] = ti;
            __syncthreads();
        }
        for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) T[(e % nb) + (e / nb) * ldt] = sT[e % nb][e / nb];
        return;
    }
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) T[(e % nb) + (e / nb) * ldt] = cmk(0, 0);
    __syncthreads();
    for (int i = 0; i < nb; ++i) {
        const double2 ti = tau[i];
        const int r = threadIdx.x;
        if (r < i) {
            double2 s = cmk(0, 0);
            for (int l = r; l < i; ++l) cfma(s, T[r + l * ldt], G[l + i * 64]);
            col[r] = cscale(cmul(ti, s), -1.0);
        }
        __syncthreads();
        if (r < i) T[r + i * ldt] = col[r];
        if (r == 0) T[i + i * ldt] = ti;
        __syncthreads();
    }
}

// Per-matrix descriptor of the batched Householder reduction (storage and
// workspace of one matrix of an uneven batch; filled on the host).
struct HetrdSys {
    int n;
    // 0: one 64 x 64 Gram block (G) reused by every panel and the panel's
    // compact-WY factor formed when the panel closes; > 0: one Gram block per
    // panel at this stride (elements) and every factor formed after the
    // reduction (off the column-step chain)
    int g_stride;
    void* A;          // tiled lower storage

    double2* Vstore;
    double2* Tall;
    double2* Wp;
    double2* alpha2;
    double2* G;
    double2* acol;
    double* d;
    double* e;
    double2* tau;
    double* scal;
    double* part;
    unsigned int* counter;
    double2* Yr;
    double2* Yc;
    double2* t12;
};

// Hemv variant for trailing size nt from an autotune table indexed by
// floor(log2(nt)): code = ch + 32 u2 + 128 t (ch in {4, 8, 16}; u2 = 1: two
// row blocks per step, u2 = 2: per-slab cross-lane column sums, u2 = 3: waves
// owning rows, chunks of 2 ch tiles; t = 1: default cache policy instead of
// non-temporal loads); untuned entries give ch = 8, u2 = 2, t = 0.
inline int hemv_ch_table_lookup(const unsigned char* table, int nt) {
    if (!table || nt <= 0) return 8 + 64;
    int b = 0;
    while ((2 << b) <= nt && b < 30) ++b;
    const int code = table[b < BSE_HEMV_BUCKETS ? b : BSE_HEMV_BUCKETS - 1];
    const int ch = code & 31;
    return (ch == 4 || ch == 8 || ch == 16) ? code : 8 + 64;
}
