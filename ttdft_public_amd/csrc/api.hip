// C ABI of the MI355X ChFSI engine: contexts, Hamiltonian upload, and the
// host orchestration of the subspace iteration (filter, Cholesky-QR,
// Rayleigh-Ritz, rotation, Lanczos bounds, density).  All device work of a
// context is issued on its one stream; dense k x k factorisations use
// rocSOLVER in FP64 on that stream.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.h"

namespace ttdft {

int density_compute(ttdft_ctx ctx, const void *X, int64_t ldx, int dtype, int32_t k,
                    const double *occ, int32_t rx, int32_t ry, int32_t rz, const double *Bx,
                    const double *By, const double *Bz, int32_t px, int32_t py, int32_t pz,
                    double *rho);
int launch_initial_density(ttdft_ctx ctx, int32_t natoms, const double *d_xyz, int32_t nr,
                           const double *d_r, const double *d_coef, int32_t px, int32_t py,
                           int32_t pz, const double *gx, const double *gy, const double *gz,
                           double *rho);
int grid_reduce(ttdft_ctx ctx, const double *f, const double *g, int32_t px, int32_t py,
                int32_t pz, const double *wx, const double *wy, const double *wz, int squared,
                double *result);
int dev_dot(ttdft_ctx ctx, const double *a, const double *b, int64_t n, double *result);
int dev_axpby(ttdft_ctx ctx, double *y, double a, const double *x, double b, int64_t n);
int dev_scale(ttdft_ctx ctx, double *f, int64_t n, double s);
int dev_fill_uniform(ttdft_ctx ctx, double *f, int64_t n, uint64_t seed, double lo, double hi);
int launch_cm_to_bm(ttdft_ctx ctx, const double *src, int64_t lda, int64_t n, int32_t k,
                    void *dst, int64_t ldd, int dtype);
int launch_bm_to_cm(ttdft_ctx ctx, const void *src, int64_t lds, int dtype, int64_t n,
                    int32_t k, double *dst, int64_t lda);
int launch_lower_transpose(ttdft_ctx ctx, const double *L, double *LT, int32_t k);

// ---- errors ----------------------------------------------------------------
static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}
void clear_error() { g_err.clear(); }

// ---- scratch ---------------------------------------------------------------
int Scratch::ensure(size_t need) {
    if (need <= bytes) return TTDFT_OK;
    if (ptr) {
        // the old buffer may still be read by queued kernels
        hipError_t e = hipDeviceSynchronize();
        (void)e;
        e = hipFree(ptr);
        (void)e;
        ptr = nullptr;
        bytes = 0;
    }
    hipError_t e = hipMalloc(&ptr, need);
    if (e != hipSuccess && owner && !owner->blk_free.empty()) {
        dev_cache_clear(owner);
        e = hipMalloc(&ptr, need);
    }
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error("scratch allocation of %zu bytes failed: %s", need, hipGetErrorString(e));
        return TTDFT_ENOMEM;
    }
    bytes = need;
    return TTDFT_OK;
}
void Scratch::release() {
    if (ptr) {
        hipError_t e = hipFree(ptr);
        (void)e;
    }
    ptr = nullptr;
    bytes = 0;
}

// ---- timing ----------------------------------------------------------------
int timing_begin(ttdft_ctx ctx, int cls, PendingTiming *t) {
    t->cls = cls;
    t->start = t->stop = nullptr;
    if (!ctx->profiling) return TTDFT_OK;
    TT_HIP(hipEventCreate(&t->start));
    TT_HIP(hipEventCreate(&t->stop));
    TT_HIP(hipEventRecord(t->start, ctx->stream));
    return TTDFT_OK;
}
int timing_end(ttdft_ctx ctx, PendingTiming *t, double bytes, double flops) {
    if (!ctx->profiling) return TTDFT_OK;
    TT_HIP(hipEventRecord(t->stop, ctx->stream));
    t->bytes = bytes;
    t->flops = flops;
    ctx->pending->push_back(*t);
    if (ctx->pending->size() > 4096) return timing_flush(ctx);
    return TTDFT_OK;
}
int timing_flush(ttdft_ctx ctx) {
    if (!ctx->pending) return TTDFT_OK;
    for (auto &t : *ctx->pending) {
        TT_HIP(hipEventSynchronize(t.stop));
        float ms = 0.f;
        TT_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        KStat &s = ctx->stats[t.cls];
        s.launches++;
        s.ms += ms;
        s.bytes += t.bytes;
        s.flops += t.flops;
        TT_HIP(hipEventDestroy(t.start));
        TT_HIP(hipEventDestroy(t.stop));
    }
    ctx->pending->clear();
    return TTDFT_OK;
}

// ---- host random numbers (splitmix64 seeding + xoshiro256**) ---------------
namespace {
uint64_t splitmix64(uint64_t *s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t sm = seed;
        for (auto &v : s) v = splitmix64(&sm);
    }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};
}  // namespace

void rng_uniform(uint64_t seed, int64_t count, double *out) {
    Xoshiro g(seed);
    for (int64_t i = 0; i < count; ++i) out[i] = g.uniform();
}
void rng_normal(uint64_t seed, int64_t count, double *out) {
    Xoshiro g(seed);
    for (int64_t i = 0; i < count; i += 2) {
        const double u1 = 1.0 - g.uniform();
        const double u2 = g.uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        out[i] = r * std::cos(2.0 * M_PI * u2);
        if (i + 1 < count) out[i + 1] = r * std::sin(2.0 * M_PI * u2);
    }
}

// ---- host symmetric eigenvalues of a small matrix (cyclic Jacobi) ----------
int host_syev_small(int k, const double *A, double *evals) {
    std::vector<double> M((size_t)k * k);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
            const int r = i > j ? i : j, c = i > j ? j : i;
            M[i + (size_t)k * j] = A[r + (size_t)k * c];
        }
    double frob = 0.0;
    for (double v : M) frob += v * v;
    const double tol2 = frob * 1e-34;
    int sweep = 0;
    for (; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int q = 1; q < k; ++q)
            for (int p = 0; p < q; ++p) off += M[p + (size_t)k * q] * M[p + (size_t)k * q];
        if (off <= tol2) break;
        for (int p = 0; p < k - 1; ++p)
            for (int q = p + 1; q < k; ++q) {
                const double apq = M[p + (size_t)k * q];
                if (apq == 0.0) continue;
                const double theta = (M[q + (size_t)k * q] - M[p + (size_t)k * p]) / (2.0 * apq);
                const double t = std::fabs(theta) > 1e150
                                     ? 0.5 / theta
                                     : (theta >= 0 ? 1.0 : -1.0) /
                                           (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int r = 0; r < k; ++r) {
                    const double xp = M[r + (size_t)k * p], xq = M[r + (size_t)k * q];
                    M[r + (size_t)k * p] = c * xp - s * xq;
                    M[r + (size_t)k * q] = s * xp + c * xq;
                }
                for (int r = 0; r < k; ++r) {
                    const double xp = M[p + (size_t)k * r], xq = M[q + (size_t)k * r];
                    M[p + (size_t)k * r] = c * xp - s * xq;
                    M[q + (size_t)k * r] = s * xp + c * xq;
                }
                M[p + (size_t)k * q] = M[q + (size_t)k * p] = 0.0;
            }
    }
    for (int i = 0; i < k; ++i) evals[i] = M[i + (size_t)k * i];
    std::sort(evals, evals + k);
    return sweep < 100 ? 0 : -1;
}

// ---- clamped cubic spline (zero end slopes), second-derivative form --------
static void spline_build(int n, const double *x, const double *y, double *coef) {
    std::vector<double> h(n), d(n), rhs(n), M(n), up(n);
    for (int i = 0; i < n - 1; ++i) h[i] = x[i + 1] - x[i];
    d[0] = 2.0 * h[0];
    up[0] = h[0];
    rhs[0] = 6.0 * ((y[1] - y[0]) / h[0]);
    for (int i = 1; i < n - 1; ++i) {
        d[i] = 2.0 * (h[i - 1] + h[i]);
        up[i] = h[i];
        rhs[i] = 6.0 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1]);
    }
    d[n - 1] = 2.0 * h[n - 2];
    rhs[n - 1] = 6.0 * (0.0 - (y[n - 1] - y[n - 2]) / h[n - 2]);
    for (int i = 1; i < n; ++i) {
        const double w = up[i - 1] / d[i - 1];
        d[i] -= w * up[i - 1];
        rhs[i] -= w * rhs[i - 1];
    }
    M[n - 1] = rhs[n - 1] / d[n - 1];
    for (int i = n - 2; i >= 0; --i) M[i] = (rhs[i] - up[i] * M[i + 1]) / d[i];
    for (int i = 0; i < n - 1; ++i) {
        coef[4 * i + 0] = y[i];
        coef[4 * i + 1] = (y[i + 1] - y[i]) / h[i] - h[i] * (2.0 * M[i] + M[i + 1]) / 6.0;
        coef[4 * i + 2] = 0.5 * M[i];
        coef[4 * i + 3] = (M[i + 1] - M[i]) / (6.0 * h[i]);
    }
}

// ---- host-side small helpers -------------------------------------------------
static int upload(ttdft_ctx ctx, void **dptr, const void *src, size_t bytes) {
    *dptr = nullptr;
    if (bytes == 0) return TTDFT_OK;
    TT_TRY(dev_alloc(ctx, dptr, bytes));
    TT_HIP(hipMemcpyAsync(*dptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return TTDFT_OK;
}

// ---- allocation-free H application -------------------------------------------
// Work space of ttdft_ham_apply_noalloc: the nonlocal projections W = D C X
// (p x k) and their split partials.  Each GPU holds its own copy until the
// process ends.  In host code `scratch` names a host-side stand-in, not the
// device memory: the device address comes from hipGetSymbolAddress, taken
// once per device in ttdft_ctx_create with that device current.
constexpr size_t kNoAllocBytes = size_t(32) << 20;
constexpr int kNoAllocDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[kNoAllocDevices] = {};

// called with `device` current
static int noalloc_setup(int device) {
    TT_CHECK(device >= 0 && device < kNoAllocDevices, TTDFT_EBADARG,
             "device %d beyond the %d devices of the static work space table", device,
             kNoAllocDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return TTDFT_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    TT_CHECK(e == hipSuccess && ptr, TTDFT_EHIP,
             "hipGetSymbolAddress(static work space) on device %d: %s", device,
             hipGetErrorString(e));
    g_noalloc_ptr[device] = ptr;
    return TTDFT_OK;
}

static void *noalloc_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return device >= 0 && device < kNoAllocDevices ? g_noalloc_ptr[device] : nullptr;
}

static void dfree(void *p) {
    if (p) {
        hipError_t e = hipFree(p);
        (void)e;
    }
}

namespace {
constexpr size_t kBlockCacheCap = size_t(64) << 30;  // released bytes kept per context

// size class: 256-byte granules below 1 MiB, else rounded up to 1/8 of the
// power of two below the size (at most 12.5% slack)
size_t block_class(size_t bytes) {
    if (bytes < 8) bytes = 8;
    if (bytes <= (size_t(1) << 20)) return (bytes + 255) / 256 * 256;
    size_t top = size_t(1) << (63 - __builtin_clzll(bytes));
    const size_t g = top / 8;
    return (bytes + g - 1) / g * g;
}
}  // namespace

int dev_alloc(ttdft_ctx ctx, void **p, size_t bytes) {
    const size_t c = block_class(bytes);
    auto it = ctx->blk_free.find(c);
    if (it != ctx->blk_free.end()) {
        *p = it->second.second;
        ctx->blk_free.erase(it);
        ctx->blk_free_bytes -= c;
    } else {
        hipError_t e = hipMalloc(p, c);
        if (e != hipSuccess && !ctx->blk_free.empty()) {  // give the cached blocks back, retry
            dev_cache_clear(ctx);
            e = hipMalloc(p, c);
        }
        if (e != hipSuccess) {
            *p = nullptr;
            set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
            return TTDFT_ENOMEM;
        }
    }
    ctx->blk_size[*p] = c;
    return TTDFT_OK;
}

void dev_release(ttdft_ctx ctx, void *p) {
    if (!p) return;
    auto it = ctx->blk_size.find(p);
    if (it == ctx->blk_size.end()) {
        dfree(p);
        return;
    }
    const size_t c = it->second;
    ctx->blk_size.erase(it);
    if (c > kBlockCacheCap) {
        dfree(p);
        return;
    }
    // over the cap: give back the blocks released longest ago (an iterative
    // caller reuses what it released last; hipFree waits for pending work)
    while (ctx->blk_free_bytes + c > kBlockCacheCap && !ctx->blk_free.empty()) {
        auto oldest = ctx->blk_free.begin();
        for (auto b = ctx->blk_free.begin(); b != ctx->blk_free.end(); ++b)
            if (b->second.first < oldest->second.first) oldest = b;
        dfree(oldest->second.second);
        ctx->blk_free_bytes -= oldest->first;
        ctx->blk_free.erase(oldest);
    }
    ctx->blk_free.emplace(c, std::make_pair(++ctx->blk_seq, p));
    ctx->blk_free_bytes += c;
}

void dev_cache_clear(ttdft_ctx ctx) {
    if (!ctx->blk_free.empty()) {
        hipError_t e = hipStreamSynchronize(ctx->stream);  // pending work may still use them
        (void)e;
    }
    for (auto &b : ctx->blk_free) dfree(b.second.second);
    ctx->blk_free.clear();
    ctx->blk_free_bytes = 0;
}

int dev_malloc_retry(ttdft_ctx ctx, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 8);
    if (e != hipSuccess && ctx && !ctx->blk_free.empty()) {
        dev_cache_clear(ctx);
        e = hipMalloc(p, bytes ? bytes : 8);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return TTDFT_ENOMEM;
    }
    return TTDFT_OK;
}

void ctx_attach(ttdft_ctx ctx) { ctx->live_children++; }

void ctx_detach(ttdft_ctx ctx) {
    if (--ctx->live_children == 0 && ctx->destroy_pending) {
        ctx->destroy_pending = false;
        ttdft_ctx_destroy(ctx);
    }
}

static int orth_once(ttdft_ctx ctx, void *X, int64_t ldx, int64_t n, int32_t k, int dtype,
                     int mode) {
    const int64_t kk = (int64_t)k * k;
    TT_TRY(ctx->scratch[1].ensure((size_t)kk * 2 * sizeof(double)));
    double *S = (double *)ctx->scratch[1].ptr;
    double *QT = S + kk;
    TT_TRY(launch_gram(ctx, X, ldx, X, ldx, n, k, dtype, mode, S));
    if (small_dense_enabled(k)) {
        TT_TRY(small_chol_inv(ctx, S, k, QT));
        const size_t es = dtype_size(dtype);
        TT_TRY(ctx->scratch[0].ensure((size_t)n * k * es));
        void *Z = ctx->scratch[0].ptr;
        TT_TRY(launch_rotate(ctx, X, ldx, Z, k, n, k, QT, dtype, mode));
        TT_TRY(launch_copy_block(ctx, Z, k, dtype, X, ldx, dtype, n, k));
        return TTDFT_OK;
    }
    rocblas_status st = rocsolver_dpotrf(ctx->blas, rocblas_fill_lower, k, S, k, ctx->d_info);
    TT_CHECK(st == rocblas_status_success, TTDFT_ELIB, "rocsolver_dpotrf: status %d", (int)st);
    int info = 0;
    TT_HIP(hipMemcpyAsync(&info, ctx->d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    TT_CHECK(info == 0, TTDFT_ENOTSPD, "overlap matrix not positive definite (minor %d)", info);
    st = rocsolver_dtrtri(ctx->blas, rocblas_fill_lower, rocblas_diagonal_non_unit, k, S, k,
                          ctx->d_info);
    TT_CHECK(st == rocblas_status_success, TTDFT_ELIB, "rocsolver_dtrtri: status %d", (int)st);
    // Q = L^{-T} (column-major, strictly lower part of L^{-T} zero)
    TT_TRY(launch_lower_transpose(ctx, S, QT, k));
    const size_t es = dtype_size(dtype);
    TT_TRY(ctx->scratch[0].ensure((size_t)n * k * es));
    void *Z = ctx->scratch[0].ptr;
    TT_TRY(launch_rotate(ctx, X, ldx, Z, k, n, k, QT, dtype, mode));
    TT_TRY(launch_copy_block(ctx, Z, k, dtype, X, ldx, dtype, n, k));
    return TTDFT_OK;
}

}  // namespace ttdft

using namespace ttdft;

// =============================================================================
extern "C" {

const char *ttdft_last_error(void) { return g_err.c_str(); }

int ttdft_get_device_count(int *count) {
    TT_CHECK(count, TTDFT_EBADARG, "count is NULL");
    TT_HIP(hipGetDeviceCount(count));
    return TTDFT_OK;
}

int ttdft_ctx_create(int device, ttdft_ctx *out) {
    clear_error();
    TT_CHECK(out, TTDFT_EBADARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    TT_HIP(hipGetDeviceCount(&n));
    TT_CHECK(device >= 0 && device < n, TTDFT_EBADARG, "device %d not in [0,%d)", device, n);
    TT_HIP(hipSetDevice(device));
    TT_TRY(noalloc_setup(device));
    ttdft_ctx ctx = new ttdft_ctx_s();
    ctx->device = device;
    ctx->pending = new std::vector<PendingTiming>();
    for (auto &s : ctx->scratch) s.owner = ctx;
    for (auto &s : ctx->nn_scratch) s.owner = ctx;
    hipError_t e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx->pending;
        delete ctx;
        set_error("hipStreamCreate: %s", hipGetErrorString(e));
        return TTDFT_EHIP;
    }
    ctx->stream = ctx->own_stream;
    if (rocblas_create_handle(&ctx->blas) != rocblas_status_success ||
        rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success) {
        ttdft_ctx_destroy(ctx);
        set_error("rocblas_create_handle failed");
        return TTDFT_ELIB;
    }
    if (hipMalloc(&ctx->d_info, sizeof(int)) != hipSuccess) {
        ttdft_ctx_destroy(ctx);
        set_error("hipMalloc(info) failed");
        return TTDFT_ENOMEM;
    }
    *out = ctx;
    return TTDFT_OK;
}

int ttdft_ctx_destroy(ttdft_ctx ctx) {
    if (!ctx) return TTDFT_OK;
    if (ctx->destroy_pending) return TTDFT_OK;  // already handed to the last child
    hipError_t e = hipSetDevice(ctx->device);
    (void)e;
    if (ctx->stream) {
        e = hipStreamSynchronize(ctx->stream);
        (void)e;
    }
    timing_flush(ctx);
    if (ctx->live_children > 0) {
        // Hamiltonians / mixers still hold this context (their storage goes
        // back to its block cache): the last of them to be freed destroys it
        ctx->destroy_pending = true;
        return TTDFT_OK;
    }
    free_tucker_blocks(ctx->device, ctx->stream);
    dev_cache_clear(ctx);
    for (auto &s : ctx->scratch) s.release();
    for (auto &s : ctx->nn_scratch) s.release();
    dfree(ctx->d_info);
    if (ctx->blas) rocblas_destroy_handle(ctx->blas);
    if (ctx->own_stream) {
        e = hipStreamDestroy(ctx->own_stream);
        (void)e;
    }
    delete ctx->pending;
    delete ctx;
    return TTDFT_OK;
}

int ttdft_ctx_trim(ttdft_ctx ctx) {
    TT_CHECK(ctx, TTDFT_EBADARG, "ctx is NULL");
    TT_HIP(hipSetDevice(ctx->device));
    dev_cache_clear(ctx);  // waits for the stream before the blocks go back
    return TTDFT_OK;
}

int ttdft_ctx_set_stream(ttdft_ctx ctx, void *hip_stream) {
    TT_CHECK(ctx, TTDFT_EBADARG, "ctx is NULL");
    hipStream_t next = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    if (next == ctx->stream) return TTDFT_OK;
    // Released cache blocks and scratch buffers are reused in the order of
    // one stream: finish what the outgoing stream still has queued on them.
    TT_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stream = next;
    TT_CHECK(rocblas_set_stream(ctx->blas, ctx->stream) == rocblas_status_success, TTDFT_ELIB,
             "rocblas_set_stream failed");
    return TTDFT_OK;
}

int ttdft_ctx_synchronize(ttdft_ctx ctx) {
    TT_CHECK(ctx, TTDFT_EBADARG, "ctx is NULL");
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TTDFT_OK;
}

int ttdft_ctx_set_profiling(ttdft_ctx ctx, int enable) {
    TT_CHECK(ctx, TTDFT_EBADARG, "ctx is NULL");
    TT_TRY(timing_flush(ctx));
    ctx->profiling = enable ? 1 : 0;
    return TTDFT_OK;
}

int ttdft_ctx_kernel_stats(ttdft_ctx ctx, int cls, int64_t *launches, double *total_ms,
                           double *bytes, double *flops) {
    TT_CHECK(ctx && cls >= 0 && cls < TTDFT_KSTAT_COUNT, TTDFT_EBADARG, "bad kernel class");
    TT_TRY(timing_flush(ctx));
    const KStat &s = ctx->stats[cls];
    if (launches) *launches = s.launches;
    if (total_ms) *total_ms = s.ms;
    if (bytes) *bytes = s.bytes;
    if (flops) *flops = s.flops;
    return TTDFT_OK;
}

int ttdft_ctx_reset_stats(ttdft_ctx ctx) {
    TT_CHECK(ctx, TTDFT_EBADARG, "ctx is NULL");
    TT_TRY(timing_flush(ctx));
    for (auto &s : ctx->stats) s = KStat();
    return TTDFT_OK;
}

int ttdft_malloc(ttdft_ctx ctx, size_t bytes, void **dptr) {
    TT_CHECK(ctx && dptr, TTDFT_EBADARG, "null argument");
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    TT_CHECK(e == hipSuccess, TTDFT_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return TTDFT_OK;
}
int ttdft_free(ttdft_ctx ctx, void *dptr) {
    TT_CHECK(ctx, TTDFT_EBADARG, "ctx is NULL");
    if (dptr) TT_HIP(hipFree(dptr));
    return TTDFT_OK;
}
int ttdft_memcpy_h2d(ttdft_ctx ctx, void *dst, const void *src, size_t bytes) {
    TT_CHECK(ctx && (bytes == 0 || (dst && src)), TTDFT_EBADARG, "null argument");
    TT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TTDFT_OK;
}
int ttdft_memcpy_d2h(ttdft_ctx ctx, void *dst, const void *src, size_t bytes) {
    TT_CHECK(ctx && (bytes == 0 || (dst && src)), TTDFT_EBADARG, "null argument");
    TT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TTDFT_OK;
}
int ttdft_memcpy_d2d(ttdft_ctx ctx, void *dst, const void *src, size_t bytes) {
    TT_CHECK(ctx && (bytes == 0 || (dst && src)), TTDFT_EBADARG, "null argument");
    TT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return TTDFT_OK;
}

int ttdft_block_upload_colmajor(ttdft_ctx ctx, const double *host, int64_t lda, int64_t n,
                                int32_t k, void *dX, int64_t ldx, int dtype) {
    TT_CHECK(ctx && host && dX && n > 0 && k > 0 && lda >= n && ldx >= k, TTDFT_EBADARG,
             "block_upload: bad arguments");
    TT_TRY(ctx->scratch[7].ensure((size_t)lda * k * sizeof(double)));
    double *tmp = (double *)ctx->scratch[7].ptr;
    TT_HIP(hipMemcpyAsync(tmp, host, (size_t)lda * k * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
    TT_TRY(launch_cm_to_bm(ctx, tmp, lda, n, k, dX, ldx, dtype));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TTDFT_OK;
}

int ttdft_block_download_colmajor(ttdft_ctx ctx, const void *dX, int64_t ldx, int dtype,
                                  int64_t n, int32_t k, double *host, int64_t lda) {
    TT_CHECK(ctx && host && dX && n > 0 && k > 0 && lda >= n && ldx >= k, TTDFT_EBADARG,
             "block_download: bad arguments");
    TT_TRY(ctx->scratch[7].ensure((size_t)lda * k * sizeof(double)));
    double *tmp = (double *)ctx->scratch[7].ptr;
    TT_TRY(launch_bm_to_cm(ctx, dX, ldx, dtype, n, k, tmp, lda));
    TT_HIP(hipMemcpyAsync(host, tmp, (size_t)lda * k * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TTDFT_OK;
}

int ttdft_block_convert(ttdft_ctx ctx, const void *src, int64_t lds, int src_dtype, void *dst,
                        int64_t ldd, int dst_dtype, int64_t n, int32_t k) {
    TT_CHECK(ctx && src && dst && lds >= k && ldd >= k, TTDFT_EBADARG, "convert: bad arguments");
    return launch_copy_block(ctx, src, lds, src_dtype, dst, ldd, dst_dtype, n, k);
}

// ---- Hamiltonian ---------------------------------------------------------------
int ttdft_ham_upload(ttdft_ctx ctx, int64_t n, const int64_t *rowptr, const int32_t *colind,
                     const double *val, int32_t p, const int64_t *c_rowptr,
                     const int32_t *c_colind, const double *c_val, const double *inv_c_lm,
                     int copies, ttdft_ham *out) {
    clear_error();
    TT_CHECK(ctx && out && rowptr && n > 0, TTDFT_EBADARG, "ham_upload: bad arguments");
    TT_CHECK(copies & (TTDFT_HAM_F64 | TTDFT_HAM_F32), TTDFT_EBADARG,
             "ham_upload: no device copy requested");
    *out = nullptr;
    const int64_t nnz = rowptr[n] - rowptr[0];
    TT_CHECK(rowptr[0] == 0 && nnz >= 0, TTDFT_EBADARG, "ham_upload: rowptr must start at 0");
    TT_CHECK(nnz == 0 || (colind && val), TTDFT_EBADARG, "ham_upload: null CSR arrays");
    int32_t maxrow = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t len = rowptr[i + 1] - rowptr[i];
        TT_CHECK(len >= 0, TTDFT_EBADARG, "ham_upload: rowptr not monotone at row %lld",
                 (long long)i);
        if (len > maxrow) maxrow = (int32_t)len;
    }
    for (int64_t e = 0; e < nnz; ++e)
        TT_CHECK(colind[e] >= 0 && colind[e] < n, TTDFT_EBADARG,
                 "ham_upload: column index %d out of range at entry %lld", colind[e],
                 (long long)e);
    TT_CHECK(p >= 0 && (p == 0 || (c_rowptr && inv_c_lm)), TTDFT_EBADARG,
             "ham_upload: projector arrays missing");
    const int64_t nnz_c = p > 0 ? c_rowptr[p] - c_rowptr[0] : 0;
    if (p > 0) {
        TT_CHECK(c_rowptr[0] == 0, TTDFT_EBADARG, "ham_upload: c_rowptr must start at 0");
        TT_CHECK(nnz_c == 0 || (c_colind && c_val), TTDFT_EBADARG, "ham_upload: null C arrays");
        for (int64_t e = 0; e < nnz_c; ++e)
            TT_CHECK(c_colind[e] >= 0 && c_colind[e] < n, TTDFT_EBADARG,
                     "ham_upload: projector column %d out of range", c_colind[e]);
    }
    TT_HIP(hipSetDevice(ctx->device));
    ttdft_ham H = new ttdft_ham_s();
    H->ctx = ctx;
    ctx_attach(ctx);
    H->n = n;
    H->nnz = nnz;
    H->p = p;
    H->nnz_c = nnz_c;
    H->copies = copies;
    H->max_row_nnz = maxrow;
    int rc = TTDFT_OK;
#define UP(dst, src, bytes)                                                             \
    do {                                                                                \
        if (rc == TTDFT_OK) rc = upload(ctx, (void **)&(dst), (src), (bytes));          \
    } while (0)
    UP(H->rowptr, rowptr, sizeof(int64_t) * (n + 1));
    UP(H->col, colind, sizeof(int32_t) * nnz);
    if (copies & TTDFT_HAM_F64) UP(H->val64, val, sizeof(double) * nnz);
    std::vector<float> f32;
    if (copies & TTDFT_HAM_F32) {
        f32.resize(nnz);
        for (int64_t e = 0; e < nnz; ++e) f32[e] = (float)val[e];
        UP(H->val32, f32.data(), sizeof(float) * nnz);
    }
    // projectors: D C (row-scaled) and C^T (unscaled), ProjectHamiltonianSparse.cc:316-325
    std::vector<double> dc, ctv;
    std::vector<int64_t> ctp;
    std::vector<int32_t> ctc;
    std::vector<float> dc32, ct32;
    if (p > 0) {
        dc.resize(nnz_c);
        for (int r = 0; r < p; ++r)
            for (int64_t e = c_rowptr[r]; e < c_rowptr[r + 1]; ++e) dc[e] = c_val[e] * inv_c_lm[r];
        ctp.assign(n + 1, 0);
        for (int64_t e = 0; e < nnz_c; ++e) ctp[c_colind[e] + 1]++;
        for (int64_t i = 0; i < n; ++i) ctp[i + 1] += ctp[i];
        ctc.resize(nnz_c);
        ctv.resize(nnz_c);
        std::vector<int64_t> fill(ctp.begin(), ctp.end() - 1);
        for (int r = 0; r < p; ++r)  // rows of C in order => C^T rows sorted by projector
            for (int64_t e = c_rowptr[r]; e < c_rowptr[r + 1]; ++e) {
                const int64_t o = fill[c_colind[e]]++;
                ctc[o] = r;
                ctv[o] = c_val[e];
            }
        UP(H->c_rowptr, c_rowptr, sizeof(int64_t) * (p + 1));
        UP(H->c_col, c_colind, sizeof(int32_t) * nnz_c);
        UP(H->ct_rowptr, ctp.data(), sizeof(int64_t) * (n + 1));
        UP(H->ct_col, ctc.data(), sizeof(int32_t) * nnz_c);
        if (copies & TTDFT_HAM_F64) {
            UP(H->c_val64, dc.data(), sizeof(double) * nnz_c);
            UP(H->ct_val64, ctv.data(), sizeof(double) * nnz_c);
        }
        if (copies & TTDFT_HAM_F32) {
            dc32.resize(nnz_c);
            ct32.resize(nnz_c);
            for (int64_t e = 0; e < nnz_c; ++e) {
                dc32[e] = (float)dc[e];
                ct32[e] = (float)ctv[e];
            }
            UP(H->c_val32, dc32.data(), sizeof(float) * nnz_c);
            UP(H->ct_val32, ct32.data(), sizeof(float) * nnz_c);
        }
    }
#undef UP
    if (rc == TTDFT_OK) rc = maybe_build_dense(ctx, H, copies);
    if (rc == TTDFT_OK) rc = finish_dense(ctx, H, copies);
    if (rc == TTDFT_OK) {
        hipError_t e = hipStreamSynchronize(ctx->stream);  // host staging vectors go out of scope
        if (e != hipSuccess) {
            set_error("ham_upload: %s", hipGetErrorString(e));
            rc = TTDFT_EHIP;
        }
    }
    if (rc != TTDFT_OK) {
        std::string msg = g_err;
        ttdft_ham_free(H);
        g_err = msg;
        return rc;
    }
    *out = H;
    return TTDFT_OK;
}

int ttdft_ham_upload_tucker(ttdft_ctx ctx, int32_t rx, int32_t ry, int32_t rz,
                            const int64_t *rowptr, const int32_t *colind, const double *val,
                            int32_t p, const int64_t *c_rowptr, const int32_t *c_colind,
                            const double *c_val, const double *inv_c_lm, int copies,
                            ttdft_ham *out) {
    clear_error();
    TT_CHECK(rx > 0 && ry > 0 && rz > 0, TTDFT_EBADARG, "ham_upload_tucker: ranks must be > 0");
    const int64_t n = (int64_t)rx * ry * rz;
    TT_TRY(ttdft_ham_upload(ctx, n, rowptr, colind, val, p, c_rowptr, c_colind, c_val,
                            inv_c_lm, copies, out));
    int rc = build_band(ctx, *out, rx, ry, rz, rowptr, colind, val, copies);
    if (rc == TTDFT_OK && (*out)->format == 1) drop_dense(*out);
    if (rc != TTDFT_OK) {
        std::string msg = g_err;
        ttdft_ham_free(*out);
        *out = nullptr;
        g_err = msg;
    }
    return rc;
}

int ttdft_ham_format(ttdft_ham H, int32_t *format, int32_t *bands) {
    TT_CHECK(H, TTDFT_EBADARG, "ham is NULL");
    if (format) *format = H->format;
    if (bands) {
        bands[0] = H->bx;
        bands[1] = H->by;
        bands[2] = H->bz;
    }
    return TTDFT_OK;
}

int ttdft_ham_free(ttdft_ham H) {
    if (!H) return TTDFT_OK;
    hipError_t e = hipStreamSynchronize(H->ctx->stream);
    (void)e;
    dev_release(H->ctx, H->mf64);
    dev_release(H->ctx, H->mf32);
    dev_release(H->ctx, H->dense64);
    dev_release(H->ctx, H->dense32);
    dev_release(H->ctx, H->rowptr);
    dev_release(H->ctx, H->col);
    dev_release(H->ctx, H->val64);
    dev_release(H->ctx, H->val32);
    dev_release(H->ctx, H->c_rowptr);
    dev_release(H->ctx, H->c_col);
    dev_release(H->ctx, H->c_val64);
    dev_release(H->ctx, H->c_val32);
    dev_release(H->ctx, H->ct_rowptr);
    dev_release(H->ctx, H->ct_col);
    dev_release(H->ctx, H->ct_val64);
    dev_release(H->ctx, H->ct_val32);
    ttdft_ctx ctx = H->ctx;
    delete H;
    ctx_detach(ctx);
    return TTDFT_OK;
}

int ttdft_ham_info(ttdft_ham H, int64_t *n, int64_t *nnz, int32_t *p) {
    TT_CHECK(H, TTDFT_EBADARG, "ham is NULL");
    if (n) *n = H->n;
    if (nnz) *nnz = H->nnz;
    if (p) *p = H->p;
    return TTDFT_OK;
}

int ttdft_ham_apply(ttdft_ctx ctx, ttdft_ham H, const void *X, int64_t ldx, void *Y,
                    int64_t ldy, int32_t k, int dtype) {
    clear_error();
    TT_CHECK(ctx && H && X && Y && k >= 1 && k <= 1024 && ldx >= k && ldy >= k, TTDFT_EBADARG,
             "ham_apply: bad arguments (k=%d)", k);
    TT_CHECK(X != Y, TTDFT_EBADARG, "ham_apply: X and Y must not alias");
    return ham_step(ctx, H, X, ldx, Y, ldy, nullptr, 0, k, dtype, 1.0, 0.0, 0.0, 0, 0);
}

int ttdft_ham_apply_noalloc(ttdft_ctx ctx, ttdft_ham H, const void *X, int64_t ldx, void *Y,
                            int64_t ldy, int32_t k, int dtype) {
    clear_error();
    TT_CHECK(ctx && H && X && Y && k >= 1 && k <= 1024 && ldx >= k && ldy >= k, TTDFT_EBADARG,
             "ham_apply_noalloc: bad arguments (k=%d)", k);
    TT_CHECK(X != Y, TTDFT_EBADARG, "ham_apply_noalloc: X and Y must not alias");
    void *ws = noalloc_ptr(ctx->device);
    TT_CHECK(ws, TTDFT_EHIP, "ham_apply_noalloc: no static work space on device %d",
             ctx->device);
    // the same kernels as ttdft_ham_apply; not timed (timing creates events)
    const int prof = ctx->profiling;
    ctx->profiling = 0;
    const int rc = ham_step(ctx, H, X, ldx, Y, ldy, nullptr, 0, k, dtype, 1.0, 0.0, 0.0, 0, 0,
                            ws, kNoAllocBytes);
    ctx->profiling = prof;
    return rc;
}

int ttdft_cheb_filter(ttdft_ctx ctx, ttdft_ham H, void *X, int64_t ldx, int32_t k, int32_t m,
                      double a, double b, double a0, int32_t num_blocks, int dtype,
                      int degree_convention) {
    clear_error();
    TT_CHECK(ctx && H && X && k >= 1 && k <= 1024 && ldx >= k && m >= 1, TTDFT_EBADARG,
             "cheb_filter: bad arguments (k=%d, m=%d)", k, m);
    TT_CHECK(num_blocks >= 1 && num_blocks <= k, TTDFT_EBADARG, "cheb_filter: num_blocks=%d",
             num_blocks);
    TT_CHECK(degree_convention == TTDFT_DEGREE_REFERENCE || degree_convention == TTDFT_DEGREE_EXACT,
             TTDFT_EBADARG, "cheb_filter: degree convention %d", degree_convention);
    TT_CHECK(b > a, TTDFT_EBADARG, "cheb_filter: empty interval [%g, %g]", a, b);
    const int target = degree_convention == TTDFT_DEGREE_REFERENCE ? m - 1 : m;
    if (target == 0) return TTDFT_OK;  // Y_0 = X
    const size_t es = dtype_size(dtype);
    TT_TRY(ctx->scratch[0].ensure((size_t)H->n * ldx * es));
    char *B = (char *)ctx->scratch[0].ptr;
    char *A = (char *)X;
    int col0 = 0;
    for (int blk = 0; blk < num_blocks; ++blk) {
        // chunk sizes of ChebyShevFilter.cc:57-60
        const int cols = k / num_blocks + (blk < k % num_blocks ? 1 : 0);
        char *Ac = A + (size_t)col0 * es;
        char *Bc = B + (size_t)col0 * es;
        col0 += cols;
        const double e = (b - a) / 2, c = (b + a) / 2;
        double sigma = e / (a0 - c);
        const double sigma1 = sigma, gamma = 2.0 / sigma1;
        double alpha = sigma1 / e, beta = alpha * (-c);
        // Y_1 = alpha H X + beta X   -> B
        TT_TRY(ham_step(ctx, H, Ac, ldx, Bc, ldx, nullptr, 0, cols, dtype, alpha, beta, 0.0, 1,
                        0));
        char *cur = Bc, *prev = Ac;
        for (int it = 2; it <= target; ++it) {
            const double sigma2 = 1.0 / (gamma - sigma);
            alpha = 2.0 * sigma2 / e;
            beta = -c * alpha;
            const double delta = -sigma * sigma2;
            // Y_it = alpha H Y_{it-1} + beta Y_{it-1} + delta Y_{it-2}, written over Y_{it-2}
            TT_TRY(ham_step(ctx, H, cur, ldx, prev, ldx, prev, ldx, cols, dtype, alpha, beta,
                            delta, 1, 1));
            char *t = cur;
            cur = prev;
            prev = t;
            sigma = sigma2;
        }
        if (cur != Ac) TT_TRY(launch_copy_block(ctx, cur, ldx, dtype, Ac, ldx, dtype, H->n, cols));
    }
    return TTDFT_OK;
}

int ttdft_gram(ttdft_ctx ctx, const void *X, int64_t ldx, const void *Y, int64_t ldy, int64_t n,
               int32_t k, int dtype, int mode, double *dS) {
    clear_error();
    TT_CHECK(ctx && X && Y && dS && ldx >= k && ldy >= k, TTDFT_EBADARG, "gram: bad arguments");
    return launch_gram(ctx, X, ldx, Y, ldy, n, k, dtype, mode, dS);
}

int ttdft_orth(ttdft_ctx ctx, void *X, int64_t ldx, int64_t n, int32_t k, int dtype, int mode) {
    clear_error();
    TT_CHECK(ctx && X && n >= k && k >= 1 && ldx >= k, TTDFT_EBADARG, "orth: bad arguments");
    TT_TRY(orth_once(ctx, X, ldx, n, k, dtype, mode));
    // split-BF16 Gram: a second Cholesky-QR pass restores orthogonality to
    // FP32 level (CholQR2)
    if (mode == TTDFT_GEMM_BF16X3) TT_TRY(orth_once(ctx, X, ldx, n, k, dtype, mode));
    return TTDFT_OK;
}

int ttdft_rr_eig(ttdft_ctx ctx, const double *dA, int32_t k, double *d_evals, double *dQ) {
    clear_error();
    TT_CHECK(ctx && dA && d_evals && dQ && k >= 1, TTDFT_EBADARG, "rr_eig: bad arguments");
    TT_TRY(ctx->scratch[1].ensure((size_t)k * k * 2 * sizeof(double) + k * sizeof(double)));
    double *E = (double *)ctx->scratch[1].ptr + (size_t)k * k * 2;
    // the one-workgroup kernel reads A whole into LDS before it writes Q, so
    // dQ may alias dA
    if (small_dense_enabled(k)) return small_syev(ctx, dA, k, d_evals, dQ);
    if (dQ != dA)
        TT_HIP(hipMemcpyAsync(dQ, dA, sizeof(double) * k * k, hipMemcpyDeviceToDevice,
                              ctx->stream));
    rocblas_status st = rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_lower, k,
                                         dQ, k, d_evals, E, ctx->d_info);
    TT_CHECK(st == rocblas_status_success, TTDFT_ELIB, "rocsolver_dsyevd: status %d", (int)st);
    int info = 0;
    TT_HIP(hipMemcpyAsync(&info, ctx->d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    TT_CHECK(info == 0, TTDFT_ENOCONV, "rocsolver_dsyevd did not converge (info %d)", info);
    return TTDFT_OK;
}

int ttdft_rotate(ttdft_ctx ctx, void *X, int64_t ldx, int64_t n, int32_t k, const double *dQ,
                 int dtype, int mode) {
    clear_error();
    TT_CHECK(ctx && X && dQ && k >= 1 && ldx >= k, TTDFT_EBADARG, "rotate: bad arguments");
    const size_t es = dtype_size(dtype);
    TT_TRY(ctx->scratch[0].ensure((size_t)n * k * es));
    void *Z = ctx->scratch[0].ptr;
    TT_TRY(launch_rotate(ctx, X, ldx, Z, k, n, k, dQ, dtype, mode));
    TT_TRY(launch_copy_block(ctx, Z, k, dtype, X, ldx, dtype, n, k));
    return TTDFT_OK;
}

int ttdft_lanczos_bounds(ttdft_ctx ctx, ttdft_ham H, int32_t maxit, uint64_t seed, double *upper,
                         double *lower) {
    clear_error();
    TT_CHECK(ctx && H && upper && lower && maxit >= 1, TTDFT_EBADARG, "lanczos: bad arguments");
    TT_CHECK(H->val64, TTDFT_EBADARG, "lanczos needs the FP64 copy of H");
    const int64_t n = H->n;
    // EigenSolver.cc:186-388 with the start vector from the seeded generator
    std::vector<double> hv(n);
    rng_uniform(seed, n, hv.data());
    double nv = 0.0;
    for (double x : hv) nv += x * x;
    nv = std::sqrt(nv);
    for (double &x : hv) x *= 1.0 / nv;
    TT_TRY(ctx->scratch[7].ensure(sizeof(double) * 3 * n));
    double *v = (double *)ctx->scratch[7].ptr, *v0 = v + n, *f = v0 + n;
    TT_HIP(hipMemcpyAsync(v, hv.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    std::vector<double> T((size_t)maxit * maxit, 0.0);
    TT_TRY(ham_spmv(ctx, H, v, f));
    double alpha = 0.0, beta = 0.0;
    TT_TRY(dev_dot(ctx, f, v, n, &alpha));
    TT_TRY(dev_axpby(ctx, f, 1.0, v, -alpha, n));
    T[0] = alpha;
    for (int it = 1; it < maxit; ++it) {
        double ff = 0.0;
        TT_TRY(dev_dot(ctx, f, f, n, &ff));
        beta = std::sqrt(ff);
        TT_TRY(dev_axpby(ctx, f, 1.0 / beta, nullptr, 0.0, n));
        std::swap(v0, v);  // v0 <- v
        std::swap(v, f);   // v <- f (f's old storage is reused below)
        TT_TRY(ham_spmv(ctx, H, v, f));
        TT_TRY(dev_axpby(ctx, f, 1.0, v0, -beta, n));
        TT_TRY(dev_dot(ctx, v, f, n, &alpha));
        TT_TRY(dev_axpby(ctx, f, 1.0, v, -alpha, n));
        T[it + (size_t)maxit * (it - 1)] = beta;
        T[(it - 1) + (size_t)maxit * it] = beta;
        T[it + (size_t)maxit * it] = alpha;
    }
    double ff = 0.0;
    TT_TRY(dev_dot(ctx, f, f, n, &ff));
    beta = std::sqrt(ff);
    std::vector<double> ev(maxit);
    TT_CHECK(host_syev_small(maxit, T.data(), ev.data()) == 0, TTDFT_ENOCONV,
             "Lanczos tridiagonal eigensolve did not converge");
    *upper = beta + std::fabs(ev[maxit - 1]);
    *lower = ev[0];
    return TTDFT_OK;
}

int ttdft_fermi_occupancy(const double *evals, int32_t k, int32_t nelec, double kb, double T,
                          int32_t maxit, double *fermi, double *occ) {
    clear_error();
    TT_CHECK(evals && fermi && k >= 1 && nelec >= 1 && kb > 0 && T > 0, TTDFT_EBADARG,
             "fermi: bad arguments");
    int occNum = nelec / 2;
    if (nelec % 2 == 1) occNum += 1;
    TT_CHECK(occNum <= k, TTDFT_EBADARG, "fermi: %d electrons need more than %d states", nelec, k);
    // Newton iteration of KSDFTEnergyFunctional.cc:76-124; "s + a*b" is one
    // fused multiply-add as in the reference build
    double ef = evals[occNum - 1];
    for (int it = 0; it < maxit; ++it) {
        double func = 0, funcd = 0, sumfi = 0;
        for (int idx = 0; idx < k; ++idx) {
            const double arg = (evals[idx] - ef) / (kb * T);
            if (arg <= 0.0) {
                const double term = 1.0 / (1.0 + std::exp(arg));
                sumfi = std::fma(2.0, term, sumfi);
                funcd = std::fma(2.0 * (std::exp(arg) / (kb * T)) * term, term, funcd);
            } else {
                const double term = 1.0 / (1.0 + std::exp(-arg));
                sumfi = std::fma(2.0 * std::exp(-arg), term, sumfi);
                funcd = std::fma(2.0 * (std::exp(-arg) / (kb * T)) * term, term, funcd);
            }
        }
        func = sumfi - nelec;
        ef = ef - func / funcd;
        if (std::fabs(func) <= 1.0e-8) break;
    }
    *fermi = ef;
    if (occ) {
        for (int i = 0; i < k; ++i) {  // main.cc:1253-1261
            const double term = (evals[i] - ef) / (kb * T);
            occ[i] = term <= 0.0 ? 1.0 / (1.0 + std::exp(term))
                                 : std::exp(-term) / (1.0 + std::exp(-term));
        }
    }
    return TTDFT_OK;
}

double ttdft_band_energy(const double *evals, const double *occ, int32_t k) {
    double e = 0.0;
    for (int i = 0; i < k; ++i) e = std::fma(2 * occ[i], evals[i], e);
    return e;
}

int ttdft_density(ttdft_ctx ctx, const void *X, int64_t ldx, int dtype, int32_t k,
                  const double *occ, int32_t rx, int32_t ry, int32_t rz, const double *dBx,
                  const double *dBy, const double *dBz, int32_t px, int32_t py, int32_t pz,
                  double *d_rho) {
    clear_error();
    TT_CHECK(ctx && X && occ && dBx && dBy && dBz && d_rho && k >= 1 && ldx >= k && rx > 0 &&
                 ry > 0 && rz > 0 && px > 0 && py > 0 && pz > 0,
             TTDFT_EBADARG, "density: bad arguments");
    return density_compute(ctx, X, ldx, dtype, k, occ, rx, ry, rz, dBx, dBy, dBz, px, py, pz,
                           d_rho);
}

int ttdft_initial_density(ttdft_ctx ctx, int32_t natoms, const double *xyz, int32_t nr,
                          const double *r, const double *rho_r, int32_t px, int32_t py,
                          int32_t pz, const double *dgx, const double *dgy, const double *dgz,
                          double *d_rho) {
    clear_error();
    TT_CHECK(ctx && xyz && r && rho_r && dgx && dgy && dgz && d_rho && natoms >= 1 && nr >= 2,
             TTDFT_EBADARG, "initial_density: bad arguments");
    for (int i = 0; i + 1 < nr; ++i)
        TT_CHECK(r[i + 1] > r[i], TTDFT_EBADARG, "initial_density: radii not increasing");
    std::vector<double> host(3 * (size_t)natoms + nr + 4 * (size_t)(nr - 1));
    std::memcpy(host.data(), xyz, sizeof(double) * 3 * natoms);
    std::memcpy(host.data() + 3 * natoms, r, sizeof(double) * nr);
    spline_build(nr, r, rho_r, host.data() + 3 * natoms + nr);
    TT_TRY(ctx->scratch[6].ensure(sizeof(double) * (host.size() + 1024)));
    double *d = (double *)ctx->scratch[6].ptr;
    TT_HIP(hipMemcpyAsync(d, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    TT_TRY(launch_initial_density(ctx, natoms, d, nr, d + 3 * natoms, d + 3 * natoms + nr, px, py,
                                  pz, dgx, dgy, dgz, d_rho));
    TT_HIP(hipStreamSynchronize(ctx->stream));  // host staging goes out of scope
    return TTDFT_OK;
}

int ttdft_grid_integral(ttdft_ctx ctx, const double *d_f, int32_t px, int32_t py, int32_t pz,
                        const double *dwx, const double *dwy, const double *dwz,
                        double *result) {
    clear_error();
    TT_CHECK(ctx && d_f && dwx && dwy && dwz && result, TTDFT_EBADARG, "integral: bad arguments");
    return grid_reduce(ctx, d_f, nullptr, px, py, pz, dwx, dwy, dwz, 0, result);
}

int ttdft_grid_residual(ttdft_ctx ctx, const double *d_f, const double *d_g, int32_t px,
                        int32_t py, int32_t pz, const double *dwx, const double *dwy,
                        const double *dwz, double *result) {
    clear_error();
    TT_CHECK(ctx && d_f && d_g && dwx && dwy && dwz && result, TTDFT_EBADARG,
             "residual: bad arguments");
    return grid_reduce(ctx, d_f, d_g, px, py, pz, dwx, dwy, dwz, 1, result);
}

int ttdft_fill_uniform(ttdft_ctx ctx, double *d_f, int64_t count, uint64_t seed, double lo,
                       double hi) {
    clear_error();
    TT_CHECK(ctx && (d_f || count == 0) && count >= 0, TTDFT_EBADARG, "fill_uniform: bad arguments");
    if (count == 0) return TTDFT_OK;
    return dev_fill_uniform(ctx, d_f, count, seed, lo, hi);
}

int ttdft_grid_scale(ttdft_ctx ctx, double *d_f, int64_t count, double s) {
    clear_error();
    TT_CHECK(ctx && d_f && count >= 0, TTDFT_EBADARG, "scale: bad arguments");
    return dev_scale(ctx, d_f, count, s);
}

int ttdft_chfsi_state_init(ttdft_chfsi_state *st) {
    TT_CHECK(st, TTDFT_EBADARG, "state is NULL");
    st->upper = 1e16;
    st->lower_unwanted = -1e16;
    st->lower_wanted = 0.0;
    st->err_upper = 1e16;
    st->occ_energy = 1.0e9;
    st->passes = 0;
    st->lanczos_runs = 0;
    st->occ_tol = 1.0e-3;
    return TTDFT_OK;
}

int ttdft_chfsi_step(ttdft_ctx ctx, ttdft_ham H, void *X, int64_t ldx, int32_t k, int32_t m,
                     int32_t num_blocks, int32_t restart_times, int32_t occ_index,
                     int32_t scf_iter, int32_t start_iter, int32_t maxit_lanczos, uint64_t seed,
                     int degree_convention, int precision, double *evals,
                     ttdft_chfsi_state *st) {
    clear_error();
    TT_CHECK(ctx && H && X && evals && st && k >= 1 && k <= 1024 && ldx >= k, TTDFT_EBADARG,
             "chfsi_step: bad arguments");
    TT_CHECK(occ_index >= 0 && occ_index < k, TTDFT_EBADARG, "chfsi_step: occ_index %d",
             occ_index);
    TT_CHECK(precision == TTDFT_F64 || precision == TTDFT_F32, TTDFT_EBADARG,
             "chfsi_step: precision %d", precision);
    const int dtype = precision;
    const int mode = precision == TTDFT_F64 ? TTDFT_GEMM_F64 : TTDFT_GEMM_BF16X3;
    const int64_t n = H->n;
    st->passes = 0;
    st->lanczos_runs = 0;
    double temp_lower = 0.0;
    // bound update, main.cc:987-1003
    if (st->err_upper > 0.01) {
        double cur = 0.0;
        TT_TRY(ttdft_lanczos_bounds(ctx, H, maxit_lanczos, seed, &cur, &temp_lower));
        st->lanczos_runs++;
        st->err_upper = (st->upper - cur) / st->upper;
        st->upper = st->err_upper > 0.01 ? cur : cur * 1.05;
    }
    // main.cc:1007-1014
    if (scf_iter == 0) {
        st->lower_wanted = temp_lower;
        st->lower_unwanted = 0.0;
    } else if (scf_iter != start_iter) {
        st->lower_wanted = evals[0];
        st->lower_unwanted = evals[k - 1];
    }
    const int64_t kk = (int64_t)k * k;
    const size_t es = dtype_size(dtype);
    // per-call work (HX block, S, Q, eigenvalues) from the context's block
    // cache: after the first step of a shape no runtime allocation happens
    void *HX = nullptr;
    double *dS = nullptr;
    TT_TRY(dev_alloc(ctx, &HX, (size_t)n * k * es));
    if (dev_alloc(ctx, (void **)&dS, sizeof(double) * (2 * kk + k)) != TTDFT_OK) {
        dev_release(ctx, HX);
        return TTDFT_ENOMEM;
    }
    double *dQ = dS + kk, *dEv = dQ + kk;
    int rc = TTDFT_OK;
    int have_q = 0;
    std::vector<double> ev(k);
    for (int r = 0; r < restart_times && rc == TTDFT_OK; ++r) {
        rc = ttdft_cheb_filter(ctx, H, X, ldx, k, m, st->lower_unwanted, st->upper,
                               st->lower_wanted, num_blocks, dtype, degree_convention);
        if (rc != TTDFT_OK) break;
        rc = ttdft_orth(ctx, X, ldx, n, k, dtype, mode);
        if (rc != TTDFT_OK) break;
        rc = ham_step(ctx, H, X, ldx, HX, k, nullptr, 0, k, dtype, 1.0, 0.0, 0.0, 0, 0);
        if (rc != TTDFT_OK) break;
        rc = launch_gram(ctx, X, ldx, HX, k, n, k, dtype, mode, dS);
        if (rc != TTDFT_OK) break;
        rc = ttdft_rr_eig(ctx, dS, k, dEv, dQ);
        if (rc != TTDFT_OK) break;
        hipError_t ce = hipMemcpyAsync(ev.data(), dEv, sizeof(double) * k, hipMemcpyDeviceToHost,
                                       ctx->stream);
        if (ce == hipSuccess) ce = hipStreamSynchronize(ctx->stream);
        if (ce != hipSuccess) {
            set_error("chfsi_step: %s", hipGetErrorString(ce));
            rc = TTDFT_EHIP;
            break;
        }
        have_q = 1;
        st->passes++;
        for (int i = 0; i < k; ++i) evals[i] = ev[i];
        st->lower_wanted = evals[0];
        st->lower_unwanted = evals[k - 1];
        const double cur = evals[occ_index];
        const double err = std::fabs((cur - st->occ_energy) / st->occ_energy);
        st->occ_energy = cur;
        if (err < (st->occ_tol > 0.0 ? st->occ_tol : 1.0e-3)) break;  // main.cc:1130-1145
    }
    if (rc == TTDFT_OK && have_q) rc = ttdft_rotate(ctx, X, ldx, n, k, dQ, dtype, mode);
    if (rc == TTDFT_OK) {
        hipError_t se = hipStreamSynchronize(ctx->stream);
        if (se != hipSuccess) {
            set_error("chfsi_step: %s", hipGetErrorString(se));
            rc = TTDFT_EHIP;
        }
    }
    // back to the cache in stream order (an early error may leave work queued)
    dev_release(ctx, HX);
    dev_release(ctx, dS);
    return rc;
}

int ttdft_random_subspace(ttdft_ctx ctx, void *X, int64_t ldx, int64_t n, int32_t k,
                          uint64_t seed, int dtype) {
    clear_error();
    TT_CHECK(ctx && X && n >= k && k >= 1 && ldx >= k, TTDFT_EBADARG,
             "random_subspace: bad arguments");
    std::vector<double> h((size_t)n * k);
    rng_normal(seed, (int64_t)h.size(), h.data());
    TT_TRY(ctx->scratch[7].ensure(sizeof(double) * h.size()));
    double *tmp = (double *)ctx->scratch[7].ptr;
    TT_HIP(hipMemcpyAsync(tmp, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    TT_TRY(launch_copy_block(ctx, tmp, k, TTDFT_F64, X, ldx, dtype, n, k));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    const int mode = dtype == TTDFT_F64 ? TTDFT_GEMM_F64 : TTDFT_GEMM_BF16X3;
    TT_TRY(ttdft_orth(ctx, X, ldx, n, k, dtype, mode));
    TT_TRY(ttdft_orth(ctx, X, ldx, n, k, dtype, mode));
    return TTDFT_OK;
}

int ttdft_rng_uniform(uint64_t seed, int64_t count, double *out) {
    TT_CHECK(out && count >= 0, TTDFT_EBADARG, "rng: bad arguments");
    rng_uniform(seed, count, out);
    return TTDFT_OK;
}
int ttdft_rng_normal(uint64_t seed, int64_t count, double *out) {
    TT_CHECK(out && count >= 0, TTDFT_EBADARG, "rng: bad arguments");
    rng_normal(seed, count, out);
    return TTDFT_OK;
}

}  // extern "C"
