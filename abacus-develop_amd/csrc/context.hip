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

int apw_zgemm(apw_ctx* ctx, char transa, char transb, int m, int n, int k, const double* alpha, const double* a,
              int lda, const double* b, int ldb, const double* beta, double* c, int ldc) {
    return guarded([&] {
        APW_CHECK(ctx && alpha && beta && c, "apw_zgemm: null argument");
        APW_CHECK(k == 0 || (a && b), "apw_zgemm: null A/B");
        apw::DeviceGuard g(ctx->device);
        apw::gemm<double>(ctx->stream, ctx->gws, ctx->num_cu, transa, transb, m, n, k, alpha[0], alpha[1], a, lda, b, ldb,
                          beta[0], beta[1], c, ldc);
        APW_HIP(hipStreamSynchronize(ctx->stream));
        return (int)APW_OK;
    });
}

}  // extern "C"
