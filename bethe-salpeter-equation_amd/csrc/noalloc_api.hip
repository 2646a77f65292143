// Allocation-free variant of the dense Hermitian eigensolver entry point, for
// callers that cannot allocate device memory while they run it (code under
// HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// Its workspace is one file-scope device array, `scratch`, of
// BSE_NOALLOC_SCRATCH_BYTES bytes; every GPU the library runs on holds its
// own copy until the process ends.  The host never uses the array's name as
// an address: once per device and with that device current,
// hipGetSymbolAddress gives the device address of that device's
// copy, kept in scratch_ptr[device] (taken on the entry point's first call on
// that device, with the device current); the entry point hands that address to
// the library's buffer carving (the buffers of the ordinary entry point are
// cut from it in 256-byte aligned pieces instead of being hipMalloc'ed).
// The entry point calls no hipMalloc or hipFree; orders up to
// BSE_NOALLOC_MAX_N fit the array with all eigenvectors.  Every call on one
// GPU shares that GPU's array and leaves its contents behind: the pipeline
// clears what it needs cleared, and calls that overlap on one GPU (from two
// contexts on two host threads) race.
#include "bse_internal.hpp"

#include <mutex>

#define BSE_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define BSE_NOALLOC_MAX_N 512
#define BSE_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[BSE_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex scratch_mtx;
void* scratch_ptr[BSE_NOALLOC_MAX_DEVICES] = {};
hipError_t scratch_err[BSE_NOALLOC_MAX_DEVICES] = {};
bool scratch_done[BSE_NOALLOC_MAX_DEVICES] = {};
}  // namespace

// With `device` current: that device's address of the scratch array.
static int noalloc_init(int device) {
    if (device < 0 || device >= BSE_NOALLOC_MAX_DEVICES) return BSE_ERR_ARG;
    std::lock_guard<std::mutex> lk(scratch_mtx);
    if (!scratch_done[device]) {
        void* ptr = nullptr;
        scratch_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        scratch_ptr[device] = (scratch_err[device] == hipSuccess) ? ptr : nullptr;
        scratch_done[device] = true;
    }
    return scratch_err[device] == hipSuccess ? BSE_OK : BSE_ERR_HIP;
}

extern "C" int bse_eigh_dense_noalloc(bse_ctx* ctx, int n, const double* a, int64_t lda, int precision,
                                      int n_vectors, double* evals, double* z, int64_t ldz) {
    if (!ctx || !a || !evals || n <= 0 || lda < n) return BSE_ERR_ARG;
    if (n > BSE_NOALLOC_MAX_N) {
        ctx->err = "the allocation-free eigensolver takes n <= " + std::to_string(BSE_NOALLOC_MAX_N);
        return BSE_ERR_ARG;
    }
    if (n_vectors != 0 && !z) return BSE_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    void* base = nullptr;
    if (noalloc_init(ctx->device) == BSE_OK) {
        std::lock_guard<std::mutex> lk(scratch_mtx);
        base = scratch_ptr[ctx->device];
    }
    if (!base) {
        ctx->err = "the scratch array of the allocation-free eigensolver has no device address";
        return BSE_ERR_HIP;
    }
    // run the ordinary pipeline with its buffers carved from the scratch
    // array: an empty slot table for this call, swapped back afterwards
    if (ctx->arena_bufs.size() != ctx->bufs.size()) ctx->arena_bufs.assign(ctx->bufs.size(), DevBuf{});
    std::swap(ctx->bufs, ctx->arena_bufs);
    const bool prof = ctx->prof;
    ctx->prof = false;
    ctx->arena_base = (char*)base;
    ctx->arena_cap = BSE_NOALLOC_SCRATCH_BYTES;
    ctx->arena_off = 0;
    ctx->arena_hwm = 0;
    const int rc = eigh_dense_impl(ctx, n, a, lda, 1, precision, n_vectors, evals, 1, z, ldz, 1);
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamSynchronize(ctx->st);
    ctx->arena_base = nullptr;
    ctx->prof = prof;
    std::swap(ctx->bufs, ctx->arena_bufs);
    for (DevBuf& b : ctx->arena_bufs) b = DevBuf{};
    // nothing of this call stays addressable through the getters
    ctx->have_eig = ctx->have_f = ctx->built = false;
    return rc;
}

// info[0] scratch bytes, [1] largest order, [2] bytes the last call used.
extern "C" int bse_noalloc_info(const bse_ctx* ctx, int64_t info[3]) {
    if (!ctx || !info) return BSE_ERR_ARG;
    info[0] = (int64_t)BSE_NOALLOC_SCRATCH_BYTES;
    info[1] = BSE_NOALLOC_MAX_N;
    info[2] = (int64_t)ctx->arena_hwm;
    return BSE_OK;
}
