// Internal declarations of the apw library (HIP, gfx950).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <functional>
#include <string>
#include <vector>

#include "apw.h"

namespace apw {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);

struct Error {
    int code;
};

#define APW_HIP(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::apw::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            throw ::apw::Error{APW_EHIP};                                                     \
        }                                                                                     \
    } while (0)

#define APW_CHECK(cond, ...)                      \
    do {                                          \
        if (!(cond)) {                            \
            ::apw::set_error(__VA_ARGS__);        \
            throw ::apw::Error{APW_EINVAL};       \
        }                                         \
    } while (0)

#define APW_LAUNCH_CHECK() APW_HIP(hipGetLastError())

// Runs f() translating exceptions into status codes (for the extern "C" layer).
template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const Error& e) {
        return e.code;
    } catch (...) {
        set_error("unexpected C++ exception");
        return APW_EINVAL;
    }
}

// ---------------------------------------------------------------- device memory
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // grow-only; contents are not preserved
    void reserve(size_t count) {
        if (count <= n) return;
        release();
        hipError_t e = hipMalloc(&p, count * sizeof(T) + 256);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
            throw Error{APW_ENOMEM};
        }
        n = count;
    }
    T* get() const { return p; }
};

struct GemmWorkspace {
    DevBuf<double> part;
    // When ext is set, the split-K partials go there (ext_n doubles, not owned) and a
    // product that needs more is an error: gemm then allocates nothing.
    double* ext = nullptr;
    size_t ext_n = 0;
    DevBuf<double> tune_c;   // the autotuner's scratch C (only when it measures)
};

// ---------------------------------------------------------------- kernel profiler
// When enabled, instrumented launches are bracketed by HIP events recorded on the launch
// stream; per kernel name the launch count, the device time between the events and the
// algorithmic bytes / flops of each launch are accumulated (resolved lazily).
struct ProfEntry {
    std::string name;
    long long launches = 0;
    double bytes = 0, flops = 0, ms = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};

struct Prof {
    bool on = false;
    std::deque<ProfEntry> entries;   // deque: references stay valid while entries are added
    std::vector<hipEvent_t> pool;
    ProfEntry& entry(const char* name);
    hipEvent_t event();                 // from the pool (created only when it is empty)
    void prefill(size_t n);             // create events up front, so timed code creates none
    void resolve();
    void reset();
    // A recorded event keeps the launch it was attached to alive, and with it that launch's
    // device buffers: memory freed meanwhile is not returned until the event is recorded
    // again or destroyed.  Before buffers that profiled launches used are freed, scrub
    // resolves the pending timings and re-records every pooled event as a plain marker on s.
    void scrub(hipStream_t s);
    ~Prof();
};

// attached = false: events recorded on the stream around whatever the scope launches.
// attached = true: the scope's one kernel is launched by APW_PLAUNCH, which hands the two
// events to the dispatch itself (no marker packets, no extra host work per launch).
struct ProfScope {
    Prof* p = nullptr;
    hipStream_t s = nullptr;
    ProfEntry* e = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    bool attached = false, used = false;
    double bytes_ = 0, flops_ = 0;
    ProfScope(Prof* prof, hipStream_t st, const char* name, double bytes, double flops, bool attached_ = false);
    ~ProfScope();
};

#define APW_PLAUNCH(ps, K, G, B, L, S, ...)                                                            \
    do {                                                                                               \
        if ((ps).attached && (ps).e) {                                                                 \
            hipExtLaunchKernelGGL(K, G, B, (std::uint32_t)(L), S, (ps).a, (ps).b, 0u, __VA_ARGS__);   \
            (ps).used = true;                                                                          \
        } else {                                                                                       \
            hipLaunchKernelGGL(K, G, B, L, S, __VA_ARGS__);                                            \
        }                                                                                              \
    } while (0)

struct EigWorkspace {
    DevBuf<double> a;      // n x n c128 working matrix
    DevBuf<double> v;      // n x n c128 Householder vectors
    DevBuf<double> tau;    // n c128
    DevBuf<double> d, e;   // tridiagonal
    DevBuf<double> ring;   // 2n^2 c128 exchange ring of the reduction
    DevBuf<double> z;      // n x nband real eigenvectors of T
    // blocked back-transform: the needed columns as c128 (n x nband), the Gram partials,
    // triangular factors and two nb x nband products of the WY blocks, the column list
    DevBuf<double> zc, wy;
    DevBuf<int> cols;
    DevBuf<unsigned> bar;  // [1]: abort word of the exchange polls
    // n <= single_max: the reduction runs in one workgroup (ordinary launch, not ordered
    // against other streams); 0 = always the multi-workgroup cooperative forms
    int single_max = 0;
};

// work arrays of the generalized (pencil) solver
struct GenWorkspace {
    DevBuf<double> a, r, t, x, c, t2, y;
};

}  // namespace apw

struct apw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cu = 0;
    // subspace matrices up to this size are reduced by one workgroup (apw_ctx_set_eig_single_wg)
    int eig_single_max = 0;
    // scratch shared by the functor-level entry points
    apw::DevBuf<double> scratch;
    apw::GemmWorkspace gws;
    apw::EigWorkspace ews;   // subspace eigensolver workspace of apw_dnevx_op
    apw::GenWorkspace gen;   // and of apw_dngvx_op / apw_dngvd_op
    apw::DevBuf<int> status;
    apw::DevBuf<double> vnl_tau;   // atom positions of apw_getvnl
    apw::DevBuf<double> io_stage;  // bands of a wavefunction file, file order
    apw::DevBuf<int> io_perm;      // file position of each plane wave of the basis
    apw::Prof prof;
};

namespace apw {

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev) {
        (void)hipGetDevice(&prev);
        if (prev != dev) APW_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard() {
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (cur != prev) (void)hipSetDevice(prev);
    }
};

// ---------------------------------------------------------------- autotuner (tune.hip)
// Process-wide table of measured choices (two ints per key).  tune_mode(): 0 = fixed rules,
// 1 = the table, fixed rules on a miss, 2 = as 1 and a miss is measured and recorded.
int tune_mode();
bool tune_find(const std::string& key, int& a, int& b);
void tune_store(const std::string& key, int a, int b, double us);
// Fastest of reps timed runs of run() on stream s (after one untimed run), in microseconds.
double time_on_stream(hipStream_t s, int reps, const std::function<void()>& run);

// ---------------------------------------------------------------- GEMM
// Column-major complex GEMM, C = alpha op(A) op(B) + beta C.  op: 'N', 'T', 'C'.
// R = double (c128 data) or float (c64 data).  ws: scratch for split-K partials.
template <class R>
void gemm(hipStream_t s, GemmWorkspace& ws, int num_cu, char ta, char tb, int m, int n, int k,
          double alpha_re, double alpha_im, const R* A, int lda, const R* B, int ldb,
          double beta_re, double beta_im, R* C, int ldc, Prof* prof = nullptr);
// Same product of c128 data on the FP32 matrix cores: operands rounded to FP32 on their way
// into LDS, FP32 sums folded into FP64 ones every 32-64 k values, C written in FP64.  The
// error is that of the FP32 operands (relative ~1e-7 of sum |a||b|), not of a long FP32 sum.
void gemm_x32(hipStream_t s, GemmWorkspace& ws, int num_cu, char ta, char tb, int m, int n, int k,
              double alpha_re, double alpha_im, const double* A, int lda, const double* B, int ldb,
              double beta_re, double beta_im, double* C, int ldc, Prof* prof = nullptr);
// Sizes the split-K partials for every call on this GPU (num_cu CUs), so gemm allocates
// nothing afterwards.
void gemm_reserve(GemmWorkspace& ws, int num_cu);
// Device address of this GPU's copy of the fixed split-K workspace of apw_zgemm_noalloc
// (looked up once per device; call with that device current).
double* gemm_scratch_init(int device);

// ---------------------------------------------------------------- column blocks (colops.hip)
// dst[:, c] = src[:, c] for rows x ncols column-major c128 blocks
void copy_cols(hipStream_t s, int rows, int ncols, const double2* src, int lds, double2* dst, int ldd);
// x[r0:r1, 0:ncols] = 0
void zero_rows(hipStream_t s, int r0, int r1, int ncols, double2* x, int ld);
// x[0:n] = 0 (doubles)
void zero(hipStream_t s, size_t n, double* x);
// vc[i + m nbase] = vcc[i + sel[m] nbx]: the columns sel[0..nc) of vcc, rows 0..nbase, packed
void gather_cols(hipStream_t s, int nbase, int nc, const double2* vcc, int nbx, const int* sel, double2* vc);

// ---------------------------------------------------------------- eigensolver
// Lowest nband eigenpairs of Hermitian hcc (n x n, lower triangle, ld ldh), n <= HEEVX_MAX.
constexpr int HEEVX_MAX = 2000;
// eig: device double[nband]; vcc: device c128 (ld ldv), columns 0..nband-1 rows 0..n-1.
// gws: split-K partials of the back-transform's products (large n).
void heevx_lowest(hipStream_t s, EigWorkspace& ws, GemmWorkspace& gws, int num_cu, int n, int ldh,
                  const double* hcc, int nband, double* eig, double* vcc, int ldv, Prof* prof = nullptr);
// The same in two parts: the eigenvalues (reduction and bisection), then the eigenvectors
// (inverse iteration and back-transform) of the bands j with need[j] != 0 (need: device int
// [nband], nullptr = all, with nneed the number of non-zero flags; a cluster of close
// eigenvalues is computed whole when any of its bands is needed).  Columns not computed keep
// their previous contents.
void heevx_values(hipStream_t s, EigWorkspace& ws, int num_cu, int n, int ldh, const double* hcc, int nband,
                  double* eig, Prof* prof = nullptr);
void heevx_vectors(hipStream_t s, EigWorkspace& ws, GemmWorkspace& gws, int num_cu, int n, int nband,
                   const double* eig, double* vcc, int ldv, const int* need, int nneed, Prof* prof = nullptr);
// Sizes ws for every n <= nmax, so heevx_lowest allocates nothing (a growing workspace
// frees and allocates, and hipFree waits for the device).
void heevx_reserve(EigWorkspace& ws, int nmax, int nband);
// Text of a non-zero fault word of the eigensolver (EigWorkspace::bar[1]): bit 1, an exchange
// poll of the reduction timed out; bit 2, d or e of the tridiagonal matrix is not finite.
const char* eig_fault_text(unsigned word);

// ---------------------------------------------------------------- small dense helpers
// Cholesky of Hermitian positive definite G (n x n, ld ldg, upper triangle read) and inverse
// of the upper factor: G = R^H R, R^-1 written to rinv (n x n, ld ldr, zeros below the
// diagonal).  status[0] = j + 1 if pivot j is not positive or its square root is below
// min_norm (then rinv is not written).  n <= CHOL_MAX (one register-resident workgroup).
constexpr int CHOL_MAX = 128;
void chol_inv_upper(hipStream_t s, int n, const double* g, int ldg, double* rinv, int ldr, int* status,
                    double min_norm, Prof* prof = nullptr);

// ---------------------------------------------------------------- generalized problem
// The same factor and inverse for any n (blocks of CHOL_MAX, GEMM updates); x gets U^-1
// (n x n, ld ldx, zeros below).  status[0] != 0 when a diagonal block breaks down.
void chol_inv_blocked(hipStream_t s, GemmWorkspace& gws, GenWorkspace& w, int num_cu, int n, const double* g,
                      int ldg, double* x, int ldx, int* status, double min_norm, Prof* prof = nullptr);
// Lowest nband pairs of H x = lambda S x (hcc, scc: n x n Hermitian, both triangles, ld
// ldh; S positive definite): eig device double[nband], vcc device c128 (ld ldv), x^H S x = I.
void hegvx_lowest(hipStream_t s, GemmWorkspace& gws, EigWorkspace& ews, GenWorkspace& w, int num_cu, int n, int ldh,
                  const double* hcc, const double* scc, int nband, double* eig, double* vcc, int ldv, int* status,
                  Prof* prof = nullptr);

}  // namespace apw
