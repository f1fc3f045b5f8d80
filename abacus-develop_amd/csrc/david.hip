// Block Davidson on the GPU: the DiagoDavid<complex<double>, DEVICE_GPU> call surface
// (diago_david.h:21-84) with the algorithm of diago_david.cpp:151-1240.
//
// The host drives the iteration; every vector and subspace operation runs on the device:
//   * the per-vector classical Gram-Schmidt of SchmidtOrth (diago_david.cpp:950-1075,
//     one host sync per vector) is done as one block: L = Q^H S V, G = V^H S V - L^H L,
//     G = R^H R (Cholesky), V <- (V - Q L) R^-1.  In exact arithmetic this is the same
//     vectors as the reference's Gram-Schmidt with Pythagorean norms (both are the QR
//     factor with a positive diagonal), and it breaks down (norm < 1e-12) in the same place;
//   * the subspace eigenproblem (dnevx_op, zheevx on the host in the ROCm reference) is the
//     GPU solver of eigh.hip;
//   * refresh() sets hcc/vcc on the device instead of a host round trip
//     (diago_david.cpp:882-920), keeping the reference's behaviour of clearing hcc and scc
//     but only setting the diagonal of vcc.
// One host sync per iteration remains: the nband eigenvalues come back for the
// convergence test |e_new - e_old| < thr (diago_david.cpp:270-280).
//
// FP32-mixed form: a first diag_once runs with the H|psi> chain in FP32 (subspace work in
// FP64) until the residual norms of the corrected bands drop below 1e-5 or no band is
// left unconverged; its Ritz vectors then restart the plain FP64 iteration, which alone
// decides convergence.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "apw_ham.hpp"
#include "apw_internal.hpp"

namespace apw {
namespace {

int gsz(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

// c128 <-> c64 column blocks (precision=single callers: c64 psi and callbacks)
__global__ void k_cols_narrow(int rows, int ncols, const double2* __restrict__ src, int lds, float2* __restrict__ dst,
                              int ldd) {
    const size_t total = (size_t)rows * ncols;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t % rows), c = (int)(t / rows);
        const double2 v = src[(size_t)c * lds + r];
        dst[(size_t)c * ldd + r] = make_float2((float)v.x, (float)v.y);
    }
}

__global__ void k_cols_widen(int rows, int ncols, const float2* __restrict__ src, int lds, double2* __restrict__ dst,
                             int ldd) {
    const size_t total = (size_t)rows * ncols;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t % rows), c = (int)(t / rows);
        const float2 v = src[(size_t)c * lds + r];
        dst[(size_t)c * ldd + r] = make_double2((double)v.x, (double)v.y);
    }
}

// vc[:, m] *= -eig[unconv[m]]                  (diago_david.cpp:449-475)
__global__ void k_scale_vc(int nbase, int nc, double2* __restrict__ vc, const double* __restrict__ eig,
                           const int* __restrict__ unconv) {
    const size_t total = (size_t)nbase * nc;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int m = (int)(t / nbase);
        const double f = -1.0 * eig[unconv[m]];
        double2 v = vc[t];
        v.x *= f;
        v.y *= f;
        vc[t] = v;
    }
}

// X[i, m] /= pre[i]                            (vector_div_vector_op, diago_david.cpp:505-532)
__global__ void k_precond(int dim, int nc, double2* __restrict__ x, const double* __restrict__ pre) {
    const size_t total = (size_t)dim * nc;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % dim);
        double2 v = x[t];
        v.x = v.x / pre[i];
        v.y = v.y / pre[i];
        x[t] = v;
    }
}

// out[m] = || x[:, m] - lam[m] * y[:, m] ||^2  (one workgroup of 1024 per column; y may be
// null).  Four elements per thread per step, so four loads are in flight per lane.
constexpr int CR_NT = 1024;
__global__ __launch_bounds__(CR_NT) void k_colres2(int rows, const double2* __restrict__ x, int ldx,
                                                   const double2* __restrict__ y, int ldy,
                                                   const double* __restrict__ lam, double* __restrict__ out) {
    __shared__ double red[CR_NT / 64];
    const int m = blockIdx.x;
    const double l = (y && lam) ? lam[m] : 0.0;
    const double2* xm = x + (size_t)m * ldx;
    const double2* ym = y ? y + (size_t)m * ldy : nullptr;
    double s[4] = {0, 0, 0, 0};
    for (int i0 = threadIdx.x; i0 < rows; i0 += 4 * CR_NT) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = i0 + u * CR_NT;
            if (i < rows) {
                double2 v = xm[i];
                if (ym) {
                    const double2 w = ym[i];
                    v.x -= l * w.x;
                    v.y -= l * w.y;
                }
                s[u] += v.x * v.x + v.y * v.y;
            }
        }
    }
    double t = (s[0] + s[1]) + (s[2] + s[3]);
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int w = 0; w < CR_NT / 64; ++w) r += red[w];
        out[m] = r;
    }
}

// refresh(): hcc = diag(eig[0..nband)), vcc[i][i] = 1 (off-diagonal entries of vcc kept)
__global__ void k_refresh_diag(int nband, double2* __restrict__ hcc, double2* __restrict__ vcc, int nbx,
                               const double* __restrict__ eig) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nband; i += gridDim.x * blockDim.x) {
        hcc[(size_t)i * nbx + i] = make_double2(eig[i], 0.0);
        vcc[(size_t)i * nbx + i] = make_double2(1.0, 0.0);
    }
}

// Stands in for the c128 callback while a c64 one is set (apply_h calls the c64 one).
int c64_placeholder(void*, const double*, double*, int, int, int, int) { return -1; }

}  // namespace

struct David {
    apw_ctx* ctx = nullptr;
    int nband = 0, dim = 0, ndim = 0, nbx = 0;
    DevBuf<double> pre, basis, hp, sp, hcc, vcc, eig, tmp, vc, lmat, resid;
    DevBuf<int> unconv_d, need_d, status;
    EigWorkspace ews;
    GemmWorkspace& gws() { return ctx->gws; }   // shared by the context: sized once

    // per call
    Ham* ham = nullptr;
    apw_hpsi_cb hcb = nullptr;
    apw_hpsi_cb_c64 hcb32 = nullptr;   // c64 callback: X narrowed into x32, HX widened from hx32
    apw_spsi_cb scb = nullptr;
    DevBuf<float> x32, hx32;
    DevBuf<double> psi64;
    void* user = nullptr;
    int prec = APW_FP64;
    int hpsi_calls = 0;
    long long hpsi_bands = 0;
    double last_resid = -1.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> hev;
    // pinned host block: [status 2 int][barrier words 2 uint][eig nband][resid nband][unconv nband int]
    // [need nband int]
    unsigned char* hbuf = nullptr;
    size_t hbuf_bytes() const { return 16 + (size_t)2 * nband * sizeof(double) + (size_t)2 * nband * sizeof(int); }
    int* h_status() { return reinterpret_cast<int*>(hbuf); }
    unsigned* h_bar() { return reinterpret_cast<unsigned*>(hbuf + 8); }
    double* h_eig() { return reinterpret_cast<double*>(hbuf + 16); }
    double* h_res() { return h_eig() + nband; }
    int* h_unconv() { return reinterpret_cast<int*>(h_res() + nband); }
    int* h_need() { return h_unconv() + nband; }
    ~David() {
        if (hbuf) (void)hipHostFree(hbuf);
    }

    hipStream_t st() const { return ctx->stream; }
    double2* B(double* p, int col) const { return reinterpret_cast<double2*>(p) + (size_t)col * dim; }
    double* Bd(double* p, int col) const { return p + (size_t)2 * col * dim; }

    void create(apw_ctx* c, const double* precondition, int nband_, int dim_, int ndim_) {
        ctx = c;
        nband = nband_;
        dim = dim_;
        ndim = ndim_;
        APW_CHECK(precondition, "DiagoDavid: null precondition");
        APW_CHECK(nband > 0 && dim > 0, "DiagoDavid: nband and dim must be positive");
        APW_CHECK(ndim > 1, "DiagoDavid: david_ndim must be > 1");                 // diago_david.cpp:60
        APW_CHECK((long)ndim * nband < (long)dim, "DiagoDavid: david_ndim*nband must be < dim");  // :61
        nbx = ndim * nband;
        APW_CHECK(nbx <= HEEVX_MAX, "DiagoDavid: david_ndim*nband = %d exceeds the subspace solver limit %d", nbx,
                  HEEVX_MAX);
        pre.reserve(dim);
        APW_HIP(hipMemcpy(pre.get(), precondition, dim * sizeof(double), hipMemcpyHostToDevice));
        basis.reserve((size_t)2 * dim * nbx);
        hp.reserve((size_t)2 * dim * nbx);
        hcc.reserve((size_t)2 * nbx * nbx);
        vcc.reserve((size_t)2 * nbx * nbx);
        eig.reserve(nbx);
        tmp.reserve((size_t)2 * dim * nband);
        vc.reserve((size_t)2 * nbx * nband);
        lmat.reserve((size_t)2 * nbx * nband);
        resid.reserve(nbx);
        unconv_d.reserve(nband);
        need_d.reserve(nband);
        status.reserve(4);
        heevx_reserve(ews, nbx, nband);
        gemm_reserve(ctx->gws, ctx->num_cu);
        // pinned host words of the per-iteration reads (pageable copies go through staging)
        APW_HIP(hipHostMalloc(reinterpret_cast<void**>(&hbuf), hbuf_bytes(), hipHostMallocDefault));
        zero(st(), (size_t)2 * nbx * nbx, vcc.get());
        zero(st(), (size_t)2 * nbx * nbx, hcc.get());
        APW_LAUNCH_CHECK();
        APW_HIP(hipStreamSynchronize(st()));
    }

    double* S() { return scb ? sp.get() : basis.get(); }

    void gemm(char ta, char tb, int m, int n, int k, double ar, const double* A, int lda, const double* Bm, int ldb,
              double br, double* C, int ldc) {
        apw::gemm<double>(st(), gws(), ctx->num_cu, ta, tb, m, n, k, ar, 0.0, A, lda, Bm, ldb, br, 0.0, C, ldc, &ctx->prof);
    }
    // the dim-long products of cal_grad and cal_elem: in the FP32 pre-phase on the FP32
    // matrix cores (c128 operands rounded to FP32 as they are staged, FP64 sums), else FP64.
    // Orthonormalisation, the restart vectors and refresh stay FP64.
    bool sub_x32 = false;
    void gemm_sub(char ta, char tb, int m, int n, int k, double ar, const double* A, int lda, const double* Bm,
                  int ldb, double br, double* C, int ldc) {
        if (sub_x32)
            apw::gemm_x32(st(), gws(), ctx->num_cu, ta, tb, m, n, k, ar, 0.0, A, lda, Bm, ldb, br, 0.0, C, ldc,
                          &ctx->prof);
        else
            gemm(ta, tb, m, n, k, ar, A, lda, Bm, ldb, br, C, ldc);
    }

    // HX = H X[:, i0..i1]  (hpsi_func semantics: X = head of basis, HX = exact destination)
    void apply_h(double* HX, int i0, int i1, int p) {
        const int nb = i1 - i0 + 1;
        hipEvent_t e0 = ctx->prof.event(), e1 = ctx->prof.event();   // pooled: returned after the solve
        APW_HIP(hipEventRecord(e0, st()));
        if (ham) {
            ham->hpsi(Bd(basis.get(), i0), dim, HX, dim, nb, p);
        } else if (hcb32) {
            // the callback sees a c64 copy of the basis columns it reads (same offsets from
            // the head, leading dimension dim) and writes a c64 block that is widened into HX
            float2* xs = reinterpret_cast<float2*>(x32.get());
            hipLaunchKernelGGL(k_cols_narrow, dim3(gsz((size_t)dim * nb)), dim3(256), 0, st(), dim, nb,
                               reinterpret_cast<const double2*>(Bd(basis.get(), i0)), dim, xs + (size_t)i0 * dim, dim);
            APW_LAUNCH_CHECK();
            APW_HIP(hipStreamSynchronize(st()));
            const int rc = hcb32(user, x32.get(), hx32.get(), nbx, dim, i0, i1);
            if (rc != 0) {
                set_error("DiagoDavid: hpsi callback returned %d", rc);
                throw Error{APW_ECALLBACK};
            }
            hipLaunchKernelGGL(k_cols_widen, dim3(gsz((size_t)dim * nb)), dim3(256), 0, st(), dim, nb,
                               reinterpret_cast<const float2*>(hx32.get()), dim, reinterpret_cast<double2*>(HX), dim);
            APW_LAUNCH_CHECK();
        } else {
            APW_HIP(hipStreamSynchronize(st()));
            const int rc = hcb(user, basis.get(), HX, nbx, dim, i0, i1);
            if (rc != 0) {
                set_error("DiagoDavid: hpsi callback returned %d", rc);
                throw Error{APW_ECALLBACK};
            }
        }
        APW_HIP(hipEventRecord(e1, st()));
        hev.emplace_back(e0, e1);
        ++hpsi_calls;
        hpsi_bands += nb;
    }

    void apply_s(int c0, int nc) {
        if (!scb) return;
        APW_HIP(hipStreamSynchronize(st()));
        for (int m = 0; m < nc; ++m) {   // called as spsi(in, out, dim, dim, 1) per band (diago_david.cpp:189-228)
            const int rc = scb(user, Bd(basis.get(), c0 + m), Bd(sp.get(), c0 + m), dim, dim, 1);
            if (rc != 0) {
                set_error("DiagoDavid: spsi callback returned %d", rc);
                throw Error{APW_ECALLBACK};
            }
        }
    }

    // S-orthonormalise basis[:, c0 .. c0+nc) against basis[:, 0 .. c0) and among themselves:
    // block classical Gram-Schmidt with Pythagorean norms (SchmidtOrth, diago_david.cpp:950-1075),
    // in blocks of at most CHOL_MAX columns, each against everything before it.
    void orthonormalize(int c0, int nc) {
        zero(st(), (size_t)2, reinterpret_cast<double*>(status.get()));
        APW_LAUNCH_CHECK();
        // equal blocks rounded up to whole 32-column GEMM tiles (two blocks of 80 left 16
        // columns of every third tile idle in the block's products; 96 + 64 fill them)
        const int nblk = (nc + CHOL_MAX - 1) / CHOL_MAX;
        const int bs = std::min(CHOL_MAX, ((nc + nblk - 1) / nblk + 31) / 32 * 32);
        for (int b0 = 0; b0 < nc; b0 += bs) orthonormalize_block(c0 + b0, std::min(bs, nc - b0));
    }

    // With Q = basis[:, 0..c0), V = basis[:, c0..c0+nc), W = [Q V]^H S V = [L; G0]:
    // G = G0 - L^H L = R^H R,  V <- (V - Q L) R^-1 = [Q V] [-L R^-1; R^-1].
    void orthonormalize_block(int c0, int nc) {
        double* V = Bd(basis.get(), c0);
        apply_s(c0, nc);
        double* SV = Bd(S(), c0);
        const int ld = c0 + nc;
        double* W = lmat.get();                                        // ld x nc: [L; G]
        double* G = reinterpret_cast<double*>(reinterpret_cast<double2*>(W) + c0);
        double* X = vc.get();                                          // ld x nc: [-L R^-1; R^-1]
        double* Rinv = reinterpret_cast<double*>(reinterpret_cast<double2*>(X) + c0);
        gemm('C', 'N', ld, nc, dim, 1.0, basis.get(), dim, SV, dim, 0.0, W, ld);
        if (c0 > 0) gemm('C', 'N', nc, nc, c0, -1.0, W, ld, W, ld, 1.0, G, ld);
        chol_inv_upper(st(), nc, G, ld, Rinv, ld, status.get(), 1.0e-12, &ctx->prof);
        if (c0 > 0) gemm('N', 'N', c0, nc, nc, -1.0, W, ld, Rinv, ld, 0.0, X, ld);
        gemm('N', 'N', dim, nc, ld, 1.0, basis.get(), dim, X, ld, 0.0, tmp.get(), dim);
        copy_cols(st(), dim, nc,
                  reinterpret_cast<const double2*>(tmp.get()), dim, reinterpret_cast<double2*>(V), dim);
        APW_LAUNCH_CHECK();
        apply_s(c0, nc);
    }

    // the breakdown and time-out words, the eigenvalues and (nres > 0) residual norms in one
    // host round trip; throws on a breakdown or a time-out
    void fetch_iteration(int nres) {
        APW_HIP(hipMemcpyAsync(h_status(), status.get(), 2 * sizeof(int), hipMemcpyDeviceToHost, st()));
        h_bar()[0] = h_bar()[1] = 0;
        if (ews.bar.get()) APW_HIP(hipMemcpyAsync(h_bar(), ews.bar.get(), 2 * sizeof(unsigned), hipMemcpyDeviceToHost, st()));
        APW_HIP(hipMemcpyAsync(h_eig(), eig.get(), nband * sizeof(double), hipMemcpyDeviceToHost, st()));
        if (nres > 0) APW_HIP(hipMemcpyAsync(h_res(), resid.get(), nres * sizeof(double), hipMemcpyDeviceToHost, st()));
        APW_HIP(hipStreamSynchronize(st()));
        check_words(h_status(), h_bar());
    }

    void check_status() {
        fetch_iteration(0);
    }

    void check_words(const int* h, const unsigned* b) {
        if (h[0] != 0) {
            set_error("DiagoDavid::SchmidtOrth: psi_norm < 1.0e-12 (vector %d of the block)", h[0] - 1);
            throw Error{APW_EBREAKDOWN};
        }
        if (b[1] != 0) {
            set_error("DiagoDavid: %s", eig_fault_text(b[1]));
            throw Error{APW_EEIG};
        }
    }

    void solve_subspace(int nbase) {
        ews.single_max = ctx->eig_single_max;
        heevx_lowest(st(), ews, gws(), ctx->num_cu, nbase, nbx, hcc.get(), nband, eig.get(), vcc.get(), nbx,
                     &ctx->prof);
    }
    // the eigenvalues only; the eigenvectors follow from solve_vectors once the host knows
    // which bands are unconverged
    void solve_values(int nbase) {
        ews.single_max = ctx->eig_single_max;
        heevx_values(st(), ews, ctx->num_cu, nbase, nbx, hcc.get(), nband, eig.get(), &ctx->prof);
    }
    // all eigenvectors (every column of vcc is read: psi = basis vcc and refresh), or only those
    // of the unconverged bands (cal_grad reads vcc[:, unconv] alone)
    void solve_vectors(int nbase, bool all, const std::vector<int>& unconv, int notconv) {
        const int* need = nullptr;
        if (!all) {
            // the pinned need block is rewritten only after the next fetch_iteration has
            // synchronised the stream, so this copy has completed by then
            int* hn = h_need();
            std::fill(hn, hn + nband, 0);
            for (int m = 0; m < notconv; ++m) hn[unconv[m]] = 1;
            APW_HIP(hipMemcpyAsync(need_d.get(), hn, nband * sizeof(int), hipMemcpyHostToDevice, st()));
            need = need_d.get();
        }
        heevx_vectors(st(), ews, gws(), ctx->num_cu, nbase, nband, eig.get(), vcc.get(), nbx, need, all ? nband : notconv,
                      &ctx->prof);
    }

    // hcc[nbase .. nbase+nc, 0 .. nbase+nc) = basis_new^H hpsi   (cal_elem, diago_david.cpp:626-717)
    void cal_elem(int nbase, int nc) {
        gemm_sub('C', 'N', nc, nbase + nc, dim, 1.0, Bd(basis.get(), nbase), dim, hp.get(), dim, 0.0,
                 reinterpret_cast<double*>(reinterpret_cast<double2*>(hcc.get()) + nbase), nbx);
    }

    // one diag_once (diago_david.cpp:151-341).  switch_thr > 0: FP32 pre-phase that stops once
    // the corrected bands' residual norms are all below switch_thr (or nothing is unconverged).
    int diag_once(int ldpsi, double* psi, double* eig_host, double thr, int maxiter, int& notconv_out,
                  double switch_thr) {
        std::vector<int> unconv(nband);
        for (int m = 0; m < nband; ++m) unconv[m] = m;
        int notconv = nband;
        // initial orthonormalisation of the start vectors (:186-228)
        copy_cols(st(), dim, nband,
                  reinterpret_cast<const double2*>(psi), ldpsi, reinterpret_cast<double2*>(basis.get()), dim);
        APW_LAUNCH_CHECK();
        orthonormalize(0, nband);
        apply_h(hp.get(), 0, nband - 1, prec);                                    // :233
        cal_elem(0, nband);
        int nbase = nband;
        solve_subspace(nbase);
        fetch_iteration(0);
        std::copy(h_eig(), h_eig() + nband, eig_host);
        int it = 0;
        while (true) {
            ++it;
            // ---- cal_grad (:359-611) ----
            // (the pinned unconv block is not rewritten before the stream has consumed it: the
            // next write follows the synchronisation of fetch_iteration below)
            std::copy(unconv.begin(), unconv.begin() + notconv, h_unconv());
            APW_HIP(hipMemcpyAsync(unconv_d.get(), h_unconv(), notconv * sizeof(int), hipMemcpyHostToDevice, st()));
            double2* vcp = reinterpret_cast<double2*>(vc.get());
            gather_cols(st(), nbase, notconv,
                        reinterpret_cast<const double2*>(vcc.get()), nbx, unconv_d.get(), vcp);
            double* NB = Bd(basis.get(), nbase);
            gemm_sub('N', 'N', dim, notconv, nbase, 1.0, hp.get(), dim, vc.get(), nbase, 0.0, NB, dim);
            hipLaunchKernelGGL(k_scale_vc, dim3(gsz((size_t)nbase * notconv)), dim3(256), 0, st(), nbase, notconv, vcp,
                               eig.get(), unconv_d.get());
            gemm_sub('N', 'N', dim, notconv, nbase, 1.0, S(), dim, vc.get(), nbase, 1.0, NB, dim);
            if (switch_thr > 0) {
                hipLaunchKernelGGL(k_colres2, dim3(notconv), dim3(CR_NT), 0, st(), dim,
                                   reinterpret_cast<const double2*>(NB), dim, (const double2*)nullptr, 0,
                                   (const double*)nullptr, resid.get());
            }
            hipLaunchKernelGGL(k_precond, dim3(gsz((size_t)dim * notconv)), dim3(256), 0, st(), dim, notconv,
                               reinterpret_cast<double2*>(NB), pre.get());
            APW_LAUNCH_CHECK();
            orthonormalize(nbase, notconv);
            apply_h(Bd(hp.get(), nbase), nbase, nbase + notconv - 1, prec);      // :604
            // ---- cal_elem, diag_zhegvx ----
            cal_elem(nbase, notconv);
            const int nc_grad = notconv;
            nbase += notconv;
            solve_values(nbase);
            fetch_iteration(switch_thr > 0 ? nc_grad : 0);
            const double* ehost = h_eig();
            double rmax = 0;
            for (int m = 0; m < nc_grad && switch_thr > 0; ++m) rmax = std::max(rmax, std::sqrt(h_res()[m]));
            // ---- convergence (:270-280) ----
            notconv = 0;
            for (int m = 0; m < nband; ++m) {
                const bool conv = std::abs(ehost[m] - eig_host[m]) < thr;
                if (!conv) unconv[notconv++] = m;
                eig_host[m] = ehost[m];
            }
            const bool leave_mixed = switch_thr > 0 && (rmax < switch_thr || notconv == 0);
            const bool full = !notconv || nbase + notconv > nbx || it == maxiter || leave_mixed;
            solve_vectors(nbase, full, unconv, notconv);
            if (full) {
                // psi = basis vcc (:292-306)
                gemm('N', 'N', dim, nband, nbase, 1.0, basis.get(), dim, vcc.get(), nbx, 0.0, psi, ldpsi);
                if (ldpsi > dim) {
                    zero_rows(st(), dim,
                              ldpsi, nband, reinterpret_cast<double2*>(psi), ldpsi);
                    APW_LAUNCH_CHECK();
                }
                if (!notconv || it == maxiter || leave_mixed) {
                    // residual norms ||H psi_m - e_m psi_m|| of the returned vectors, from the
                    // subspace (hpsi vcc - e psi): one GEMM, no extra H application
                    gemm('N', 'N', dim, nband, nbase, 1.0, hp.get(), dim, vcc.get(), nbx, 0.0, tmp.get(), dim);
                    hipLaunchKernelGGL(k_colres2, dim3(nband), dim3(CR_NT), 0, st(), dim,
                                       reinterpret_cast<const double2*>(tmp.get()), dim,
                                       reinterpret_cast<const double2*>(psi), ldpsi, eig.get(), resid.get());
                    APW_LAUNCH_CHECK();
                    std::vector<double> r2(nband);
                    APW_HIP(hipMemcpyAsync(r2.data(), resid.get(), nband * sizeof(double), hipMemcpyDeviceToHost, st()));
                    APW_HIP(hipStreamSynchronize(st()));
                    last_resid = 0;
                    for (double v : r2) last_resid = std::max(last_resid, std::sqrt(v));
                    break;
                }
                // refresh (:797-934)
                gemm('N', 'N', dim, nband, nbase, 1.0, hp.get(), dim, vcc.get(), nbx, 0.0, tmp.get(), dim);
                copy_cols(st(), dim, nband,
                          reinterpret_cast<const double2*>(tmp.get()), dim,
                          reinterpret_cast<double2*>(hp.get()), dim);
                if (scb) {
                    gemm('N', 'N', dim, nband, nbase, 1.0, sp.get(), dim, vcc.get(), nbx, 0.0, tmp.get(), dim);
                    copy_cols(st(), dim, nband,
                              reinterpret_cast<const double2*>(tmp.get()), dim,
                              reinterpret_cast<double2*>(sp.get()), dim);
                }
                copy_cols(st(), dim, nband,
                          reinterpret_cast<const double2*>(psi), ldpsi,
                          reinterpret_cast<double2*>(basis.get()), dim);
                zero(st(), (size_t)2 * nbx * nbx,
                     hcc.get());
                hipLaunchKernelGGL(k_refresh_diag, dim3(1), dim3(256), 0, st(), nband,
                                   reinterpret_cast<double2*>(hcc.get()), reinterpret_cast<double2*>(vcc.get()), nbx,
                                   eig.get());
                APW_LAUNCH_CHECK();
                nbase = nband;
            }
        }
        notconv_out = notconv;
        return it;
    }

};

}  // namespace apw

struct apw_david {
    apw::David d;
};

using apw::guarded;

extern "C" {

int apw_david_create(apw_ctx* ctx, const double* precondition, int nband, int dim, int david_ndim, apw_david** out) {
    return guarded([&] {
        APW_CHECK(ctx && out, "apw_david_create: null argument");
        apw::DeviceGuard g(ctx->device);
        auto* p = new apw_david();
        try {
            p->d.create(ctx, precondition, nband, dim, david_ndim);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
        return APW_OK;
    });
}

int apw_david_destroy(apw_david* dav) {
    return guarded([&] {
        if (!dav) return APW_OK;
        apw::DeviceGuard g(dav->d.ctx->device);
        (void)hipStreamSynchronize(dav->d.ctx->stream);
        dav->d.ctx->prof.scrub(dav->d.ctx->stream);   // profiled launches must not pin the freed buffers
        delete dav;
        return APW_OK;
    });
}

int apw_david_diag(apw_david* dav, apw_ham* ham, apw_hpsi_cb hpsi_cb, apw_spsi_cb spsi_cb, void* user, int ldpsi,
                   double* psi, double* eig, double david_diag_thr, int david_maxiter, int ntry_max, int notconv_max,
                   int precision, apw_diag_stats* stats) {
    return guarded([&] {
        APW_CHECK(dav && psi && eig, "apw_david_diag: null argument");
        apw::David& D = dav->d;
        APW_CHECK(ham || hpsi_cb, "apw_david_diag: need a Hamiltonian or an hpsi callback");
        APW_CHECK(ldpsi >= D.dim, "apw_david_diag: ldpsi < dim");
        APW_CHECK(precision == APW_FP64 || precision == APW_FP32_MIXED || precision == APW_FP32_MIXED_X32,
                  "apw_david_diag: bad precision");
        APW_CHECK(david_maxiter >= 1, "apw_david_diag: david_maxiter < 1");
        if (ham) APW_CHECK(ham->h.npw == D.dim, "apw_david_diag: ham npw %d != dim %d", ham->h.npw, D.dim);
        apw::DeviceGuard g(D.ctx->device);
        D.ham = ham ? &ham->h : nullptr;
        D.hcb = hpsi_cb;
        D.scb = ham ? nullptr : spsi_cb;
        D.user = user;
        D.hpsi_calls = 0;
        D.hpsi_bands = 0;
        D.hev.clear();
        if (D.scb) D.sp.reserve((size_t)2 * D.dim * D.nbx);
        const auto t0 = std::chrono::steady_clock::now();
        int total = 0, ntry = 0, notconv = 0, switch_iter = -1;
        int rc = APW_OK;
        try {
            D.sub_x32 = false;
            if ((precision == APW_FP32_MIXED || precision == APW_FP32_MIXED_X32) && ham) {
                D.prec = APW_FP32_MIXED;
                D.sub_x32 = precision == APW_FP32_MIXED_X32;
                total += D.diag_once(ldpsi, psi, eig, david_diag_thr, david_maxiter, notconv, 1.0e-5);
                switch_iter = total;
            }
            D.sub_x32 = false;
            D.prec = APW_FP64;
            do {
                total += D.diag_once(ldpsi, psi, eig, david_diag_thr, david_maxiter, notconv, 0.0);
                ++ntry;
            } while (!(ntry > ntry_max || notconv <= notconv_max));   // check_block_conv (:1223-1240)
            APW_HIP(hipStreamSynchronize(D.ctx->stream));
        } catch (const apw::Error& e) {
            rc = e.code;
        }
        const auto t1 = std::chrono::steady_clock::now();
        double th = 0;
        for (auto& pr : D.hev) {
            float ms = 0;
            if (rc == APW_OK && hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) th += ms * 1e-3;
            D.ctx->prof.pool.push_back(pr.first);
            D.ctx->prof.pool.push_back(pr.second);
        }
        D.hev.clear();
        if (rc != APW_OK) return rc;
        if (stats) {
            stats->iterations = total;
            stats->ntry = ntry;
            stats->notconv = notconv;
            stats->hpsi_calls = D.hpsi_calls;
            stats->hpsi_bands = D.hpsi_bands;
            stats->fp64_switch_iter = switch_iter;
            stats->max_residual = D.last_resid;
            stats->t_hpsi_s = th;
            stats->t_total_s = std::chrono::duration<double>(t1 - t0).count();
        }
        return total;
    });
}

int apw_david_diag_c64(apw_david* dav, apw_hpsi_cb_c64 hpsi_cb, void* user, int ldpsi, float* psi, double* eig,
                       double david_diag_thr, int david_maxiter, int ntry_max, int notconv_max,
                       apw_diag_stats* stats) {
    return guarded([&] {
        APW_CHECK(dav && psi && eig && hpsi_cb, "apw_david_diag_c64: null argument");
        apw::David& D = dav->d;
        APW_CHECK(ldpsi >= D.dim, "apw_david_diag_c64: ldpsi < dim");
        apw::DeviceGuard g(D.ctx->device);
        // c128 shadow of psi, and the callback's c64 operand / result blocks
        D.psi64.reserve((size_t)2 * ldpsi * D.nband);
        D.x32.reserve((size_t)2 * D.dim * D.nbx);
        D.hx32.reserve((size_t)2 * D.dim * D.nband);
        hipLaunchKernelGGL(apw::k_cols_widen, dim3(apw::gsz((size_t)ldpsi * D.nband)), dim3(256), 0, D.ctx->stream,
                           ldpsi, D.nband, reinterpret_cast<const float2*>(psi), ldpsi,
                           reinterpret_cast<double2*>(D.psi64.get()), ldpsi);
        APW_LAUNCH_CHECK();
        D.hcb32 = hpsi_cb;
        int rc = APW_OK;
        try {
            rc = apw_david_diag(dav, nullptr, &apw::c64_placeholder, nullptr, user, ldpsi, D.psi64.get(), eig,
                                david_diag_thr, david_maxiter, ntry_max, notconv_max, APW_FP64, stats);
        } catch (...) {
            D.hcb32 = nullptr;
            throw;
        }
        D.hcb32 = nullptr;
        if (rc < 0) return rc;
        hipLaunchKernelGGL(apw::k_cols_narrow, dim3(apw::gsz((size_t)ldpsi * D.nband)), dim3(256), 0, D.ctx->stream,
                           ldpsi, D.nband, reinterpret_cast<const double2*>(D.psi64.get()), ldpsi,
                           reinterpret_cast<float2*>(psi), ldpsi);
        APW_LAUNCH_CHECK();
        APW_HIP(hipStreamSynchronize(D.ctx->stream));
        return rc;
    });
}

int apw_david_diag_batched(int nunits, apw_david* const* dav, apw_ham* const* ham, const int* ldpsi,
                           double* const* psi, double* const* eig, double david_diag_thr, int david_maxiter,
                           int ntry_max, int notconv_max, int precision, apw_diag_stats* stats, int* status) {
    return guarded([&] {
        APW_CHECK(nunits >= 0, "apw_david_diag_batched: nunits=%d", nunits);
        if (nunits == 0) return (int)APW_OK;
        APW_CHECK(dav && ham && ldpsi && psi && eig && status, "apw_david_diag_batched: null argument");
        // units of one context share its stream and workspaces: they run in turn on one host
        // thread; units of different contexts run concurrently, one host thread each
        std::map<apw_ctx*, std::vector<int>> groups;
        for (int u = 0; u < nunits; ++u) {
            APW_CHECK(dav[u] && ham[u], "apw_david_diag_batched: unit %d has no solver or Hamiltonian", u);
            APW_CHECK(ham[u]->h.ctx == dav[u]->d.ctx,
                      "apw_david_diag_batched: unit %d: solver and Hamiltonian on different contexts", u);
            groups[dav[u]->d.ctx].push_back(u);
            status[u] = APW_OK;
        }
        std::vector<std::string> err(nunits);
        auto run = [&](const std::vector<int>& units) {
            for (int u : units) {
                status[u] = apw_david_diag(dav[u], ham[u], nullptr, nullptr, nullptr, ldpsi[u], psi[u], eig[u],
                                           david_diag_thr, david_maxiter, ntry_max, notconv_max, precision,
                                           stats ? &stats[u] : nullptr);
                if (status[u] < 0) err[u] = apw_last_error();
            }
        };
        std::vector<std::thread> threads;
        for (auto& g : groups) threads.emplace_back(run, std::cref(g.second));
        for (auto& t : threads) t.join();
        for (int u = 0; u < nunits; ++u)
            if (status[u] < 0) {
                apw::set_error("apw_david_diag_batched: unit %d: %s", u, err[u].c_str());
                return status[u];
            }
        return (int)APW_OK;
    });
}

}  // extern "C"
