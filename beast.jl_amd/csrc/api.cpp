// C-ABI glue: errors, contexts, meshes and spaces.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <atomic>
#include <cstring>
#include <new>

#include "bh_internal.h"

namespace bh {

namespace {
thread_local char g_err[1024] = "";
}

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

const char *get_error() { return g_err; }

}  // namespace bh

using namespace bh;

extern "C" int bh_version(void) { return BH_VERSION; }

extern "C" const char *bh_last_error(void) { return bh::get_error(); }

extern "C" int bh_ctx_create(int device, bh_ctx **out)
{
    BH_API_BEGIN
    BH_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    int n = 0;
    BH_CHECK_HIP(hipGetDeviceCount(&n));
    BH_REQUIRE(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
    BH_CHECK_HIP(hipSetDevice(device));
    bh::scratch_register(device);
    hipDeviceProp_t prop;
    BH_CHECK_HIP(hipGetDeviceProperties(&prop, device));
    std::unique_ptr<bh_ctx> c(new bh_ctx());
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    c->arch = prop.gcnArchName;
    BH_CHECK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *out = c.release();
    bh::context_created();
    BH_API_END
}

namespace bh {
static void ctx_free(Ctx *c)
{
    bool expected = false;
    if (!c->freed.compare_exchange_strong(expected, true)) return;   // exactly once
    (void)hipSetDevice(c->device);
    resolve_stats(*c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    pool_trim(*c, 0);
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    c->event_pool.clear();
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete static_cast<bh_ctx *>(c);
    context_destroyed();
}

void ctx_retain(Ctx *c) { c->refs.fetch_add(1); }

// ---------------------------------------------------------------- buffer pool
namespace {
constexpr size_t kPoolMinBytes = size_t(1) << 20;   // smaller buffers are freed at once
}

void *pool_alloc(Ctx &ctx, size_t bytes)
{
    {
        std::lock_guard<std::mutex> g(ctx.pool_mu);
        int best = -1;
        for (int i = 0; i < (int)ctx.pool.size(); ++i) {
            const size_t b = ctx.pool[i].second;
            if (b >= bytes && b - bytes <= bytes / 4 && (best < 0 || b < ctx.pool[best].second)) best = i;
        }
        if (best >= 0) {
            void *p = ctx.pool[best].first;
            ctx.pool_bytes -= ctx.pool[best].second;
            ctx.pool.erase(ctx.pool.begin() + best);
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {   // pooled memory may be what is missing
        (void)hipGetLastError();
        pool_trim(ctx, 0);
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
        set_error("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
        throw Error{BH_ERR_OOM};
    }
    return p;
}

void pool_free(Ctx &ctx, void *p, size_t bytes)
{
    if (!p) return;
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> g(ctx.pool_mu);
        if (bytes < kPoolMinBytes || bytes > ctx.pool_limit) {
            drop.push_back(p);
        } else {
            ctx.pool.emplace_back(p, bytes);
            ctx.pool_bytes += bytes;
            while (ctx.pool_bytes > ctx.pool_limit) {   // oldest first
                drop.push_back(ctx.pool.front().first);
                ctx.pool_bytes -= ctx.pool.front().second;
                ctx.pool.erase(ctx.pool.begin());
            }
        }
    }
    for (void *q : drop) (void)hipFree(q);
}

void pool_trim(Ctx &ctx, size_t keep_bytes)
{
    std::vector<void *> drop;
    {
        std::lock_guard<std::mutex> g(ctx.pool_mu);
        while (ctx.pool_bytes > keep_bytes && !ctx.pool.empty()) {
            drop.push_back(ctx.pool.front().first);
            ctx.pool_bytes -= ctx.pool.front().second;
            ctx.pool.erase(ctx.pool.begin());
        }
    }
    for (void *q : drop) (void)hipFree(q);
}

void ctx_release(Ctx *c)
{
    if (c->refs.fetch_sub(1) == 1 && c->closing.load()) ctx_free(c);
}
}  // namespace bh

extern "C" int bh_ctx_destroy(bh_ctx *ctx)
{
    BH_API_BEGIN
    if (!ctx) return BH_OK;
    ctx->closing.store(true);
    if (ctx->refs.load() > 0) return BH_OK;   // freed with the last mesh or matrix created on it
    bh::ctx_free(ctx);
    BH_API_END
}

extern "C" int bh_ctx_pool(bh_ctx *ctx, int64_t limit_bytes, int64_t *pooled_bytes)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx != nullptr, "ctx is NULL");
    if (limit_bytes >= 0) {
        BH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        {
            std::lock_guard<std::mutex> g(ctx->pool_mu);
            ctx->pool_limit = (size_t)limit_bytes;
        }
        pool_trim(*ctx, (size_t)limit_bytes);
    }
    if (pooled_bytes) {
        std::lock_guard<std::mutex> g(ctx->pool_mu);
        *pooled_bytes = (int64_t)ctx->pool_bytes;
    }
    BH_API_END
}

extern "C" int bh_ctx_synchronize(bh_ctx *ctx)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx != nullptr, "ctx is NULL");
    BH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    BH_API_END
}

namespace bh {
namespace {
std::atomic<int> g_live{0};
}
int live_contexts() { return g_live.load(); }
void context_created() { g_live.fetch_add(1); }
void context_destroyed() { g_live.fetch_sub(1); }
std::mutex &coop_mutex()
{
    static std::mutex m;
    return m;
}

void resolve_stats(Ctx &ctx)
{
    (void)hipStreamSynchronize(ctx.stream);
    for (auto &kv : ctx.kstats) {
        for (auto &ev : kv.second.pending) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) kv.second.seconds += ms * 1e-3;
            ctx.event_pool.push_back(ev.first);
            ctx.event_pool.push_back(ev.second);
        }
        kv.second.pending.clear();
    }
}
}  // namespace bh

extern "C" int bh_ctx_profile(bh_ctx *ctx, int enable)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx != nullptr, "ctx is NULL");
    resolve_stats(*ctx);
    if (enable) ctx->kstats.clear();
    ctx->profiling = enable != 0;
    BH_API_END
}

extern "C" int bh_ctx_profile_only(bh_ctx *ctx, const char *kernel)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx != nullptr, "ctx is NULL");
    ctx->profile_only = kernel ? kernel : "";
    BH_API_END
}

extern "C" int bh_ctx_kernel_stats(bh_ctx *ctx, const char *kernel, int64_t *launches, double *seconds, double *work)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && kernel, "NULL argument");
    resolve_stats(*ctx);
    auto it = ctx->kstats.find(kernel);
    const KernelStat zero;
    const KernelStat &k = it == ctx->kstats.end() ? zero : it->second;
    if (launches) *launches = k.launches;
    if (seconds) *seconds = k.seconds;
    if (work) *work = k.work;
    BH_API_END
}

extern "C" int bh_mesh_create(bh_ctx *ctx, const double *vertices, int64_t nv, const int32_t *cells, int64_t nf,
                              bh_mesh **out)
{
    BH_API_BEGIN
    BH_REQUIRE(ctx && out, "NULL argument");
    *out = nullptr;
    BH_REQUIRE(nv >= 0 && nf >= 0 && nf < (int64_t(1) << 31), "bad mesh sizes");
    BH_REQUIRE(nv == 0 || vertices, "vertices is NULL");
    BH_REQUIRE(nf == 0 || cells, "cells is NULL");
    BH_CHECK_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<bh_mesh> m(new bh_mesh());
    m->ctx = ctx;
    m->nv = nv;
    m->nf = nf;
    m->V.assign(vertices, vertices + 3 * nv);
    m->C.assign(cells, cells + 3 * nf);
    for (int64_t k = 0; k < 3 * nv; ++k) BH_REQUIRE(std::isfinite(m->V[k]), "vertex coordinate not finite");
    m->geo.resize((size_t)nf * Mesh::kGeo);
    for (int64_t c = 0; c < nf; ++c) {
        double *g = &m->geo[(size_t)c * Mesh::kGeo];
        for (int a = 0; a < 3; ++a) {
            const int32_t v = m->C[3 * c + a];
            BH_REQUIRE(v >= 0 && v < nv, "cell %lld references vertex %d out of range", (long long)c, v);
            for (int k = 0; k < 3; ++k) g[3 * a + k] = m->V[3 * v + k];
        }
        for (int k = 0; k < 3; ++k) g[9 + k] = (g[k] + g[3 + k] + g[6 + k]) / 3.0;
        // area as 0.5 |(v1 - v3) x (v2 - v3)| (chart tangents, rtlocal.jl:8-27)
        const double t1[3] = {g[0] - g[6], g[1] - g[7], g[2] - g[8]};
        const double t2[3] = {g[3] - g[6], g[4] - g[7], g[5] - g[8]};
        const double n[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2],
                             t1[0] * t2[1] - t1[1] * t2[0]};
        const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        BH_REQUIRE(nn > 0, "cell %lld is degenerate (zero area)", (long long)c);
        g[12] = 0.5 * nn;
        for (int k = 0; k < 3; ++k) g[13 + k] = n[k] / nn;
    }
    m->d_geo.upload(m->geo.data(), m->geo.size(), ctx->stream);
    BH_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    ctx_retain(ctx);
    *out = m.release();
    BH_API_END
}

extern "C" int bh_mesh_destroy(bh_mesh *mesh)
{
    BH_API_BEGIN
    if (mesh) {
        Ctx *c = mesh->ctx;
        (void)hipSetDevice(c->device);
        delete mesh;
        ctx_release(c);
    }
    BH_API_END
}

extern "C" int bh_space_create(bh_mesh *mesh, int64_t ndofs, const int32_t *shape_ptr, const int32_t *shape_cell,
                               const int32_t *shape_ref, const double *shape_coeff, bh_space **out)
{
    BH_API_BEGIN
    BH_REQUIRE(mesh && out, "NULL argument");
    *out = nullptr;
    BH_REQUIRE(ndofs >= 0 && ndofs < (int64_t(1) << 31), "bad ndofs");
    BH_REQUIRE(ndofs == 0 || shape_ptr, "shape_ptr is NULL");
    std::unique_ptr<bh_space> s(new bh_space());
    s->mesh = mesh;
    s->ndofs = ndofs;
    s->ptr.assign(shape_ptr, shape_ptr + (ndofs + 1));
    BH_REQUIRE(ndofs == 0 || s->ptr[0] == 0, "shape_ptr[0] must be 0");
    for (int64_t b = 0; b < ndofs; ++b) BH_REQUIRE(s->ptr[b + 1] >= s->ptr[b], "shape_ptr must be non-decreasing");
    const int64_t ns = ndofs ? s->ptr[ndofs] : 0;
    if (ns) BH_REQUIRE(shape_cell && shape_ref && shape_coeff, "shape arrays are NULL");
    s->cell.assign(shape_cell, shape_cell + ns);
    s->ref.assign(shape_ref, shape_ref + ns);
    s->coeff.assign(shape_coeff, shape_coeff + ns);
    for (int64_t k = 0; k < ns; ++k) {
        BH_REQUIRE(s->cell[k] >= 0 && s->cell[k] < mesh->nf, "shape %lld: cell out of range", (long long)k);
        BH_REQUIRE(s->ref[k] >= 0 && s->ref[k] < 3, "shape %lld: refid out of range", (long long)k);
        BH_REQUIRE(std::isfinite(s->coeff[k]), "shape %lld: coefficient not finite", (long long)k);
    }
    // element-to-dof map (assemblydata, basis.jl:235-294)
    std::vector<int> cnt((size_t)mesh->nf * 3, 0);
    for (int64_t k = 0; k < ns; ++k) cnt[3 * s->cell[k] + s->ref[k]]++;
    int kmax = 1;
    for (int c : cnt) kmax = std::max(kmax, c);
    s->kmax = kmax;
    s->e2d.assign((size_t)mesh->nf * 3 * kmax, -1);
    s->e2c.assign((size_t)mesh->nf * 3 * kmax, 0.0);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int64_t b = 0; b < ndofs; ++b)
        for (int k = s->ptr[b]; k < s->ptr[b + 1]; ++k) {
            const int slot = 3 * s->cell[k] + s->ref[k];
            const size_t at = (size_t)slot * kmax + cnt[slot]++;
            s->e2d[at] = (int32_t)b;
            s->e2c[at] = s->coeff[k];
        }
    hipStream_t st = mesh->ctx->stream;
    s->d_e2d.upload(s->e2d.data(), s->e2d.size(), st);
    s->d_e2c.upload(s->e2c.data(), s->e2c.size(), st);
    s->d_ptr.upload(s->ptr.data(), s->ptr.size(), st);
    if (!s->cell.empty()) {
        s->d_cell.upload(s->cell.data(), s->cell.size(), st);
        s->d_ref.upload(s->ref.data(), s->ref.size(), st);
        s->d_coeff.upload(s->coeff.data(), s->coeff.size(), st);
    }
    BH_CHECK_HIP(hipStreamSynchronize(st));
    *out = s.release();
    BH_API_END
}

extern "C" int bh_space_destroy(bh_space *space)
{
    BH_API_BEGIN
    delete space;
    BH_API_END
}

extern "C" int bh_space_e2d(const bh_space *space, int32_t *kmax, int32_t *dof, double *coeff, int64_t cap)
{
    BH_API_BEGIN
    BH_REQUIRE(space && kmax, "NULL argument");
    *kmax = space->kmax;
    const int64_t need = (int64_t)space->e2d.size();
    if (dof && coeff) {
        BH_REQUIRE(cap >= need, "capacity %lld < %lld", (long long)cap, (long long)need);
        std::memcpy(dof, space->e2d.data(), sizeof(int32_t) * need);
        std::memcpy(coeff, space->e2c.data(), sizeof(double) * need);
    }
    BH_API_END
}
