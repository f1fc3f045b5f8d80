// Hermitian-definite generalized eigenproblem H x = lambda S x on the GPU (lowest m pairs):
// the dngvx_op / dngvd_op path of Diago_DavSubspace (dngvd_op.cpp:82-330, LAPACK zhegvx /
// zhegvd: S = U^H U, C = U^-H H U^-1, eigenpairs of C, x = U^-1 y).
//
// U^-1 for any n: blocks of CHOL_MAX columns, right-looking.  For block j the register
// resident kernel factors the updated diagonal block and inverts its factor; the block row
// R_j,rest = U_jj^-H S'_j,rest and the trailing update S'_rest -= R_j,rest^H R_j,rest are
// GEMMs; the off-diagonal blocks of U^-1 follow by block back substitution,
// X[0:j, j] = -X[0:j, 0:j] R[0:j, j] X_jj.
#include <algorithm>

#include "apw_internal.hpp"

namespace apw {
namespace {

__global__ void k_gcopy(int n, const double2* __restrict__ a, int lda, double2* __restrict__ b, int ldb, int zero_only) {
    const size_t total = (size_t)n * n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % n), j = (int)(t / n);
        b[(size_t)j * ldb + i] = zero_only ? make_double2(0.0, 0.0) : a[(size_t)j * lda + i];
    }
}

int gblocks(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096)); }

}  // namespace

void chol_inv_blocked(hipStream_t s, GemmWorkspace& gws, GenWorkspace& w, int num_cu, int n, const double* g, int ldg,
                      double* x, int ldx, int* status, double min_norm, Prof* prof) {
    APW_CHECK(n >= 1, "chol_inv_blocked: n=%d", n);
    if (n <= CHOL_MAX) {
        chol_inv_upper(s, n, g, ldg, x, ldx, status, min_norm, prof);
        return;
    }
    const size_t nn = (size_t)n * n;
    w.a.reserve(2 * nn);
    w.r.reserve(2 * nn);
    w.t.reserve((size_t)2 * n * CHOL_MAX);
    double2* A = reinterpret_cast<double2*>(w.a.get());   // S', updated in place (ld n)
    double2* Rb = reinterpret_cast<double2*>(w.r.get());  // block rows R_j,rest (ld n)
    double2* X = reinterpret_cast<double2*>(x);
    hipLaunchKernelGGL(k_gcopy, dim3(gblocks(nn)), dim3(256), 0, s, n, reinterpret_cast<const double2*>(g), ldg, A, n, 0);
    hipLaunchKernelGGL(k_gcopy, dim3(gblocks(nn)), dim3(256), 0, s, n, (const double2*)nullptr, 0, X, ldx, 1);
    APW_LAUNCH_CHECK();
    auto dptr = [](double2* p) { return reinterpret_cast<double*>(p); };
    const int nblk = (n + CHOL_MAX - 1) / CHOL_MAX;
    const int bs = (n + nblk - 1) / nblk;
    for (int j0 = 0; j0 < n; j0 += bs) {
        const int j1 = std::min(n, j0 + bs), nb = j1 - j0, rest = n - j1;
        double2* Xjj = X + (size_t)j0 * ldx + j0;
        chol_inv_upper(s, nb, dptr(A + (size_t)j0 * n + j0), n, dptr(Xjj), ldx, status, min_norm, prof);
        if (rest > 0) {
            double2* Rjr = Rb + (size_t)j1 * n + j0;   // rows j0..j1, cols j1..n
            gemm<double>(s, gws, num_cu, 'C', 'N', nb, rest, nb, 1.0, 0.0, dptr(Xjj), ldx,
                         dptr(A + (size_t)j1 * n + j0), n, 0.0, 0.0, dptr(Rjr), n, prof);
            gemm<double>(s, gws, num_cu, 'C', 'N', rest, rest, nb, -1.0, 0.0, dptr(Rjr), n, dptr(Rjr), n, 1.0, 0.0,
                         dptr(A + (size_t)j1 * n + j1), n, prof);
        }
        if (j0 > 0) {
            // X[0:j0, j0:j1] = -X[0:j0, 0:j0] R[0:j0, j0:j1] X_jj
            double2* T = reinterpret_cast<double2*>(w.t.get());
            gemm<double>(s, gws, num_cu, 'N', 'N', j0, nb, j0, 1.0, 0.0, dptr(X), ldx, dptr(Rb + (size_t)j0 * n), n,
                         0.0, 0.0, dptr(T), n, prof);
            gemm<double>(s, gws, num_cu, 'N', 'N', j0, nb, nb, -1.0, 0.0, dptr(T), n, dptr(Xjj), ldx, 0.0, 0.0,
                         dptr(X + (size_t)j0 * ldx), ldx, prof);
        }
    }
}

void hegvx_lowest(hipStream_t s, GemmWorkspace& gws, EigWorkspace& ews, GenWorkspace& w, int num_cu, int n, int ldh,
                  const double* hcc, const double* scc, int nband, double* eig, double* vcc, int ldv, int* status,
                  Prof* prof) {
    APW_CHECK(n >= 1 && nband >= 1 && nband <= n && ldh >= n && ldv >= n, "hegvx: bad sizes n=%d nband=%d", n, nband);
    const size_t nn = (size_t)n * n;
    w.x.reserve(2 * nn);
    w.c.reserve(2 * nn);
    w.y.reserve((size_t)2 * n * nband);
    chol_inv_blocked(s, gws, w, num_cu, n, scc, ldh, w.x.get(), n, status, 0.0, prof);
    // C = X^H H X (X = U^-1), both triangles
    gemm<double>(s, gws, num_cu, 'N', 'N', n, n, n, 1.0, 0.0, hcc, ldh, w.x.get(), n, 0.0, 0.0, w.c.get(), n, prof);
    w.t2.reserve(2 * nn);
    gemm<double>(s, gws, num_cu, 'C', 'N', n, n, n, 1.0, 0.0, w.x.get(), n, w.c.get(), n, 0.0, 0.0, w.t2.get(), n, prof);
    heevx_lowest(s, ews, gws, num_cu, n, n, w.t2.get(), nband, eig, w.y.get(), n, prof);
    // x = U^-1 y
    gemm<double>(s, gws, num_cu, 'N', 'N', n, nband, n, 1.0, 0.0, w.x.get(), n, w.y.get(), n, 0.0, 0.0, vcc, ldv, prof);
}

}  // namespace apw

extern "C" {

int apw_dngvx_op(apw_ctx* ctx, int nstart, int ldh, const double* hcc, const double* scc, int m, double* eigenvalue,
                 double* vcc) {
    return apw::guarded([&] {
        APW_CHECK(ctx && hcc && scc && eigenvalue && vcc, "apw_dngvx_op: null argument");
        APW_CHECK(nstart >= 1 && nstart <= apw::HEEVX_MAX && m >= 1 && m <= nstart && ldh >= nstart,
                  "apw_dngvx_op: bad sizes nstart=%d m=%d ldh=%d", nstart, m, ldh);
        apw::DeviceGuard g(ctx->device);
        ctx->status.reserve(4);
        APW_HIP(hipMemsetAsync(ctx->status.get(), 0, 4 * sizeof(int), ctx->stream));
        ctx->ews.single_max = ctx->eig_single_max;
        apw::hegvx_lowest(ctx->stream, ctx->gws, ctx->ews, ctx->gen, ctx->num_cu, nstart, ldh, hcc, scc, m, eigenvalue,
                          vcc, ldh, ctx->status.get(), &ctx->prof);
        int st = 0;
        unsigned flag[2] = {0, 0};
        APW_HIP(hipMemcpyAsync(&st, ctx->status.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        if (nstart > 1)
            APW_HIP(hipMemcpyAsync(flag, ctx->ews.bar.get(), sizeof(flag), hipMemcpyDeviceToHost, ctx->stream));
        APW_HIP(hipStreamSynchronize(ctx->stream));
        APW_CHECK(st == 0, "apw_dngvx_op: B is not positive definite (block pivot %d)", st - 1);
        APW_CHECK(flag[1] == 0, "apw_dngvx_op: %s", apw::eig_fault_text(flag[1]));
        return APW_OK;
    });
}

int apw_dngvd_op(apw_ctx* ctx, int nstart, int ldh, const double* hcc, const double* scc, double* eigenvalue,
                 double* vcc) {
    return apw_dngvx_op(ctx, nstart, ldh, hcc, scc, nstart, eigenvalue, vcc);
}

}  // extern "C"
