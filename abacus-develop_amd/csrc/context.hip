// Context (device + stream), error reporting and the gemm_op mirror of the C ABI.
#include <cstring>

#include "apw_internal.hpp"

namespace apw {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

ProfEntry& Prof::entry(const char* name) {
    for (auto& e : entries)
        if (e.name == name) return e;
    entries.emplace_back();
    entries.back().name = name;
    return entries.back();
}

hipEvent_t Prof::event() {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e;
    APW_HIP(hipEventCreate(&e));
    return e;
}

void Prof::prefill(size_t n) {
    while (pool.size() < n) {
        hipEvent_t e;
        APW_HIP(hipEventCreate(&e));
        pool.push_back(e);
    }
}

void Prof::resolve() {
    for (auto& e : entries) {
        for (auto& pr : e.pending) {
            float ms = 0;
            APW_HIP(hipEventSynchronize(pr.second));
            APW_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
            e.ms += ms;
            pool.push_back(pr.first);
            pool.push_back(pr.second);
        }
        e.pending.clear();
    }
}

void Prof::scrub(hipStream_t s) {
    resolve();
    for (hipEvent_t ev : pool) APW_HIP(hipEventRecord(ev, s));
    APW_HIP(hipStreamSynchronize(s));
}

void Prof::reset() {
    resolve();
    entries.clear();
}

Prof::~Prof() {
    for (auto& e : entries)
        for (auto& pr : e.pending) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (auto ev : pool) (void)hipEventDestroy(ev);
}

ProfScope::ProfScope(Prof* prof, hipStream_t st, const char* name, double bytes, double flops, bool attached_) {
    if (!prof || !prof->on) return;
    p = prof;
    s = st;
    attached = attached_;
    e = &prof->entry(name);
    bytes_ = bytes;
    flops_ = flops;
    if (!attached) {
        e->launches += 1;
        e->bytes += bytes;
        e->flops += flops;
    }
    a = prof->event();
    b = prof->event();
    if (!attached) APW_HIP(hipEventRecord(a, s));
}

ProfScope::~ProfScope() {
    if (!p) return;
    if (attached) {
        if (used) {
            e->launches += 1;
            e->bytes += bytes_;
            e->flops += flops_;
            e->pending.emplace_back(a, b);
        } else {   // nothing was launched: the events go back to the pool
            p->pool.push_back(a);
            p->pool.push_back(b);
        }
        return;
    }
    if (hipEventRecord(b, s) == hipSuccess) e->pending.emplace_back(a, b);
}

}  // namespace apw

using apw::guarded;

extern "C" {

const char* apw_last_error(void) { return apw::g_err; }

int apw_ctx_create(apw_ctx** out, int device, void* hip_stream) {
    return guarded([&] {
        APW_CHECK(out, "apw_ctx_create: null out");
        int ndev = 0;
        APW_HIP(hipGetDeviceCount(&ndev));
        APW_CHECK(device >= 0 && device < ndev, "apw_ctx_create: device %d of %d", device, ndev);
        apw::DeviceGuard g(device);
        auto* c = new apw_ctx();
        c->device = device;
        if (hip_stream) {
            c->stream = reinterpret_cast<hipStream_t>(hip_stream);
            c->own_stream = false;
        } else {
            hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
            if (e != hipSuccess) {
                delete c;
                apw::set_error("hipStreamCreate: %s", hipGetErrorString(e));
                return (int)APW_EHIP;
            }
            c->own_stream = true;
        }
        hipDeviceProp_t prop;
        APW_HIP(hipGetDeviceProperties(&prop, device));
        c->num_cu = prop.multiProcessorCount;
        try {
            (void)apw::gemm_scratch_init(device);   // per device, once: apw_zgemm_noalloc's workspace
        } catch (...) {
            if (c->own_stream) (void)hipStreamDestroy(c->stream);
            delete c;
            throw;
        }
        *out = c;
        return (int)APW_OK;
    });
}

int apw_ctx_destroy(apw_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return (int)APW_OK;
        {
            apw::DeviceGuard g(ctx->device);
            (void)hipStreamSynchronize(ctx->stream);
            ctx->scratch.release();
            ctx->gws.part.release();
            if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
        }
        delete ctx;
        return (int)APW_OK;
    });
}

int apw_ctx_synchronize(apw_ctx* ctx) {
    return guarded([&] {
        APW_CHECK(ctx, "apw_ctx_synchronize: null ctx");
        apw::DeviceGuard g(ctx->device);
        APW_HIP(hipStreamSynchronize(ctx->stream));
        return (int)APW_OK;
    });
}

void* apw_ctx_stream(apw_ctx* ctx) { return ctx ? reinterpret_cast<void*>(ctx->stream) : nullptr; }

int apw_ctx_set_eig_single_wg(apw_ctx* ctx, int nmax) {
    return guarded([&] {
        APW_CHECK(ctx && nmax >= 0, "apw_ctx_set_eig_single_wg: null ctx or negative size");
        ctx->eig_single_max = nmax;
        return (int)APW_OK;
    });
}

int apw_ctx_set_profiling(apw_ctx* ctx, int on) {
    return guarded([&] {
        APW_CHECK(ctx, "apw_ctx_set_profiling: null ctx");
        apw::DeviceGuard g(ctx->device);
        ctx->prof.reset();
        ctx->prof.on = on != 0;
        // event creation costs tens of microseconds of host time: none inside the profiled code
        if (ctx->prof.on) ctx->prof.prefill(4096);
        return (int)APW_OK;
    });
}

int apw_ctx_profile_count(apw_ctx* ctx) {
    return guarded([&] {
        APW_CHECK(ctx, "apw_ctx_profile_count: null ctx");
        apw::DeviceGuard g(ctx->device);
        ctx->prof.resolve();
        return (int)ctx->prof.entries.size();
    });
}

int apw_ctx_profile_entry(apw_ctx* ctx, int i, char* name, int name_cap, long long* launches, double* ms,
                          double* bytes, double* flops) {
    return guarded([&] {
        APW_CHECK(ctx && name && name_cap > 0 && launches && ms && bytes && flops, "apw_ctx_profile_entry: null argument");
        APW_CHECK(i >= 0 && i < (int)ctx->prof.entries.size(), "apw_ctx_profile_entry: index %d out of range", i);
        const apw::ProfEntry& e = ctx->prof.entries[i];
        snprintf(name, name_cap, "%s", e.name.c_str());
        *launches = e.launches;
        *ms = e.ms;
        *bytes = e.bytes;
        *flops = e.flops;
        return (int)APW_OK;
    });
}

int apw_zgemm(apw_ctx* ctx, char transa, char transb, int m, int n, int k, const double* alpha, const double* a,
              int lda, const double* b, int ldb, const double* beta, double* c, int ldc) {
    return guarded([&] {
        APW_CHECK(ctx && alpha && beta && c, "apw_zgemm: null argument");
        APW_CHECK(k == 0 || (a && b), "apw_zgemm: null A/B");
        apw::DeviceGuard g(ctx->device);
        apw::gemm<double>(ctx->stream, ctx->gws, ctx->num_cu, transa, transb, m, n, k, alpha[0], alpha[1], a, lda, b, ldb,
                          beta[0], beta[1], c, ldc, &ctx->prof);
        APW_HIP(hipStreamSynchronize(ctx->stream));
        return (int)APW_OK;
    });
}

int apw_zgemm_x32(apw_ctx* ctx, char transa, char transb, int m, int n, int k, const double* alpha,
                  const double* a, int lda, const double* b, int ldb, const double* beta, double* c, int ldc) {
    return guarded([&] {
        APW_CHECK(ctx && alpha && beta && c, "apw_zgemm_x32: null argument");
        APW_CHECK(k == 0 || (a && b), "apw_zgemm_x32: null A/B");
        apw::DeviceGuard g(ctx->device);
        apw::gemm_x32(ctx->stream, ctx->gws, ctx->num_cu, transa, transb, m, n, k, alpha[0], alpha[1], a, lda, b, ldb,
                      beta[0], beta[1], c, ldc, &ctx->prof);
        APW_HIP(hipStreamSynchronize(ctx->stream));
        return (int)APW_OK;
    });
}

int apw_cgemm(apw_ctx* ctx, char transa, char transb, int m, int n, int k, const float* alpha, const float* a,
              int lda, const float* b, int ldb, const float* beta, float* c, int ldc) {
    return guarded([&] {
        APW_CHECK(ctx && alpha && beta && c, "apw_cgemm: null argument");
        APW_CHECK(k == 0 || (a && b), "apw_cgemm: null A/B");
        apw::DeviceGuard g(ctx->device);
        apw::gemm<float>(ctx->stream, ctx->gws, ctx->num_cu, transa, transb, m, n, k, alpha[0], alpha[1], a, lda, b, ldb,
                         beta[0], beta[1], c, ldc, &ctx->prof);
        APW_HIP(hipStreamSynchronize(ctx->stream));
        return (int)APW_OK;
    });
}

}  // extern "C"
