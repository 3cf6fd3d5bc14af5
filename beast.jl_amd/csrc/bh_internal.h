// Internal declarations shared by the HIP sources of libbeast_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <vector>

#include "../../include/beast_hip.h"

namespace bh {

// ---------------------------------------------------------------- errors
void set_error(const char *fmt, ...);
const char *get_error();

struct Error {
    int code;
};

#define BH_CHECK_HIP(expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            ::bh::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,               \
                            hipGetErrorString(e_));                                          \
            throw ::bh::Error{BH_ERR_HIP};                                                   \
        }                                                                                    \
    } while (0)

#define BH_FAIL(code, ...)                                                                   \
    do {                                                                                     \
        ::bh::set_error(__VA_ARGS__);                                                        \
        throw ::bh::Error{code};                                                             \
    } while (0)

#define BH_API_BEGIN                                                                         \
    (void)hipGetLastError(); /* drop errors left by earlier, unrelated calls */               \
    try {
#define BH_API_END                                                                           \
    }                                                                                        \
    catch (const ::bh::Error &e_) { return e_.code; }                                        \
    catch (const std::bad_alloc &) {                                                         \
        ::bh::set_error("host allocation failed");                                           \
        return BH_ERR_OOM;                                                                   \
    }                                                                                        \
    catch (...) {                                                                            \
        ::bh::set_error("unexpected exception");                                             \
        return BH_ERR_INVALID;                                                               \
    }                                                                                        \
    return BH_OK;

#define BH_REQUIRE(cond, ...)                                                                \
    do {                                                                                     \
        if (!(cond)) BH_FAIL(BH_ERR_INVALID, __VA_ARGS__);                                   \
    } while (0)

// ---------------------------------------------------------------- device memory
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
            throw Error{BH_ERR_OOM};
        }
        n = count;
    }
    void ensure(size_t count) {   // at least count elements; contents unspecified
        if (n < count) alloc(count);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        if (count > n) alloc(count);
        if (count) BH_CHECK_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    size_t bytes() const { return n * sizeof(T); }
    // ownership transfer to and from the context's buffer pool
    void adopt(T *q, size_t count) { release(); p = q; n = count; }
    T *detach() { T *q = p; p = nullptr; n = 0; return q; }
};

// ---------------------------------------------------------------- quadrature rules
// Dunavant triangle rules in chart parameters (u, v); weights sum to 1/2.
int triangle_rule(int degree, double *u, double *v, double *w);  // returns npts or -1
void legendre01(int n, double *x, double *w);                   // nodes ascending on [0,1]
constexpr int kMaxTriPts = 32;
constexpr int kMaxTriDegree = 10;   // Dunavant degrees 1..10 (25 points)
constexpr int kMaxGL = 16;

// ---------------------------------------------------------------- objects
struct Space;
struct Ctx;
struct Mesh;
// quadrature points of a triangle rule on all cells of a mesh (cached): [nf][Q][4]
const void *quad_points(Mesh &m, int degree, int *Q, bool fp32, hipStream_t s);
// dZ32 (optional, FP64 row-major assembly only): a row-major complex-float copy
// with leading dimension ld32 written by the same launch.  Returns whether it was.
bool assemble_device(Ctx &ctx, const bh_operator *op, const bh_quadstrat *qs, bool fp32, Space &X, Space &Y, void *dZ,
                     long long ld, bool rowmajor, bool accumulate, double sr, double si, void *dZ32 = nullptr,
                     long long ld32 = 0);

// Resident matrices of integral operators through the batched kernels (one request or many).
void assemble_resident_batch(bh_ctx *ctx, int64_t count, const bh_operator *ops, const bh_quadstrat *qs,
                             const bh_space *const *spaces, bh_matrix **out);

// Per-kernel timing with HIP events on the context stream (measurement only).
struct KernelStat {
    long long launches = 0;
    double work = 0.0;       // algorithmic flops or bytes, as the launch site defines
    double seconds = 0.0;    // resolved from events
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    std::string arch;
    bool profiling = false;
    std::map<std::string, KernelStat> kstats;
    std::vector<hipEvent_t> event_pool;   // timing events reused across launches
    std::string profile_only;             // when set, only this kernel name is timed
    // workspace of the FP32 factorisation kept between calls (split planes,
    // tournament buffers, the block-column stream), so that repeated
    // factorisations allocate and free nothing
    std::shared_ptr<void> lu_cache;
    // objects (meshes, matrices) created on this context; destruction of the
    // context is deferred until the last of them is gone
    std::atomic<int> refs{0};
    std::atomic<bool> closing{false};
    std::atomic<bool> freed{false};
    // device buffers of destroyed resident matrices, kept for the next matrix of
    // a similar size (a repeated assemble-factor-solve allocates nothing); at most
    // pool_limit bytes are kept, the oldest buffers are freed first
    std::mutex pool_mu;
    std::vector<std::pair<void *, size_t>> pool;
    size_t pool_bytes = 0;
    size_t pool_limit = size_t(64) << 30;
};
void ctx_retain(Ctx *c);
void ctx_release(Ctx *c);
// device address of the allocation-free solve's scratch array, looked up once per device
// (with that device current) when a context is created there
void scratch_register(int device);

// Pool of large device buffers (see Ctx::pool).  pool_alloc returns a pooled
// buffer of at least `bytes` and at most 5/4 of it, else a fresh allocation
// (after freeing the pool if the device is out of memory).  pool_free returns a
// buffer to the pool (or frees it when it is small or the pool is disabled).
// The caller orders the work: a buffer goes to the pool only once the kernels
// that use it are complete or ordered on the context stream.
void *pool_alloc(Ctx &ctx, size_t bytes);
void pool_free(Ctx &ctx, void *p, size_t bytes);
void pool_trim(Ctx &ctx, size_t keep_bytes);
template <typename T>
void pool_release(Ctx &ctx, DevBuf<T> &b)
{
    const size_t bytes = b.n * sizeof(T);
    if (b.p) pool_free(ctx, b.detach(), bytes);
}
template <typename T>
void pool_alloc(Ctx &ctx, DevBuf<T> &b, size_t count)
{
    if (b.n == count) return;
    pool_release(ctx, b);
    if (count) b.adopt((T *)pool_alloc(ctx, count * sizeof(T)), count);
}

// Brackets one launch on ctx.stream when profiling is on.
struct ProfScope {
    Ctx &ctx;
    const char *name;
    double work;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t stream;
    // brackets work on stream st (default: the context stream)
    ProfScope(Ctx &c, const char *n, double w, hipStream_t st = nullptr)
        : ctx(c), name(n), work(w), stream(st ? st : c.stream) {
        if (!ctx.profiling) return;
        if (!ctx.profile_only.empty() && ctx.profile_only != name) return;
        e0 = take();
        e1 = take();
        if (!e0 || !e1) return;
        (void)hipEventRecord(e0, stream);
    }
    hipEvent_t take() {
        hipEvent_t e = nullptr;
        if (!ctx.event_pool.empty()) {
            e = ctx.event_pool.back();
            ctx.event_pool.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
        }
        return e;
    }
    ~ProfScope() {
        if (!ctx.profiling || !e0 || !e1) return;
        (void)hipEventRecord(e1, stream);
        KernelStat &k = ctx.kstats[name];
        k.launches += 1;
        k.work += work;
        k.pending.emplace_back(e0, e1);
    }
};
void resolve_stats(Ctx &ctx);
// number of contexts alive in the process, and the lock that serialises
// cooperative (grid-barrier) launches between them
int live_contexts();
void context_created();
void context_destroyed();
std::mutex &coop_mutex();

struct Mesh {
    Ctx *ctx = nullptr;
    int64_t nv = 0, nf = 0;
    std::vector<double> V;      // [nv][3]
    std::vector<int32_t> C;     // [nf][3]
    std::vector<double> geo;    // per cell: vertices[9], centroid[3], area, normal[3]
    DevBuf<double> d_geo;       // same on device
    static constexpr int kGeo = 16;
    // quadrature points per triangle-rule degree: [nf][Q][4] (x - centroid, weight)
    std::map<int, DevBuf<double>> quad64;
    std::map<int, DevBuf<float>> quad32;
};

struct Space {
    Mesh *mesh = nullptr;
    int64_t ndofs = 0;
    std::vector<int32_t> ptr, cell, ref;
    std::vector<double> coeff;
    // element-to-dof map: per (cell, refid) up to kmax (dof, coeff) slots
    int kmax = 1;
    std::vector<int32_t> e2d;       // [nf][3][kmax], -1 = none
    std::vector<double> e2c;        // [nf][3][kmax]
    DevBuf<int32_t> d_e2d;
    DevBuf<double> d_e2c;
    DevBuf<int32_t> d_ptr, d_cell, d_ref;   // the shapes in CSR form, on the device
    DevBuf<double> d_coeff;
    // dof tiles for the far-field gather kernel
    bool tiles_built = false;
    int64_t ntiles = 0;
    DevBuf<int> t_ptr, t_cells, t_slot;
    DevBuf<double> t_coef;
};

}  // namespace bh

// Opaque C handles are the internal objects.
struct bh_ctx : bh::Ctx {};
struct bh_mesh : bh::Mesh {};
struct bh_space : bh::Space {};
