// apw_abacus.hpp -- header-only C++ adapter: the shapes ABACUS's HSolverPW calls, over the C ABI
// of apw.h, so the GPU device path can be swapped in without changing the caller.
//
//   apw::abacus::DiagoDavid<std::complex<double>>
//       == hsolver::DiagoDavid<std::complex<double>, base_device::DEVICE_GPU>
//          (source/module_hsolver/diago_david.h:21-84): same constructor arguments
//          (precondition on the host, nband, dim, david_ndim, use_paw, diag_comm_info) and the
//          same diag(hpsi_func, spsi_func, ldPsi, psi, eigenvalue, thr, maxiter, ntry_max,
//          notconv_max) with the reference's HPsiFunc / SPsiFunc std::function types; psi is
//          device memory, eigenvalue host memory, the return value the iteration count.
//   apw::abacus::DiagoDavid<std::complex<float>>
//       == hsolver::DiagoDavid<std::complex<float>, base_device::DEVICE_GPU>
//          (diago_david.cpp:1246, ABACUS's precision=single, esolver.cpp:137-141): c64 psi
//          and callbacks; the solver itself runs in FP64 on c128 shadows of the blocks.
//   apw::abacus::HamiltPW
//       == the HamiltPW operator chain: hPsi (operator.h:48-49, operator.cpp:39-88, overwrite)
//          and the accumulating act() of Ekinetic / Veff / Nonlocal (operator.h:60-65), plus
//          hpsi_func(), the HPsiFunc HSolverPW builds from hm->ops->hPsi (hsolver_pw.cpp:488-500).
//
// Errors: the C ABI's status codes become apw::abacus::Error exceptions (the reference
// aborts).  Contexts: like the reference's process-wide hipBLAS handle, one context (one
// stream) per device is created on first use and shared by the adapters of that device.
// Dependencies: apw.h and the C++ standard library only.
#pragma once

#include <complex>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "apw.h"

namespace apw {
namespace abacus {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline int check(int rc) {
    if (rc < 0) throw Error(rc, apw_last_error());
    return rc;
}

// The shared context of `device` (created on first use, destroyed at process exit).
inline apw_ctx* default_context(int device = 0) {
    struct Slot {
        apw_ctx* ctx = nullptr;
        ~Slot() {
            if (ctx) (void)apw_ctx_destroy(ctx);
        }
    };
    static std::mutex m;
    static Slot slots[64];
    if (device < 0 || device >= 64) throw Error(APW_EINVAL, "apw::abacus: device index out of range");
    std::lock_guard<std::mutex> lock(m);
    if (!slots[device].ctx) check(apw_ctx_create(&slots[device].ctx, device, nullptr));
    return slots[device].ctx;
}

// ------------------------------------------------------------------ DiagoDavid
template <typename T>
class DiagoDavid;

namespace detail {

template <typename T>
struct Callbacks {
    using HPsiFunc = std::function<void(T*, T*, const int, const int, const int, const int)>;
    using SPsiFunc = std::function<void(T*, T*, const int, const int, const int)>;
    const HPsiFunc* h = nullptr;
    const SPsiFunc* s = nullptr;
    std::string error;
};

// C-ABI trampolines: a callback that throws reports non-zero, which the solver turns into
// APW_ECALLBACK; the exception text is kept for the rethrow.
template <typename T, typename S = typename T::value_type>
int hpsi_tramp(void* user, const S* X, S* HX, int nvec, int dim, int i0, int i1) {
    auto* cb = static_cast<Callbacks<T>*>(user);
    try {
        (*cb->h)(reinterpret_cast<T*>(const_cast<S*>(X)), reinterpret_cast<T*>(HX), nvec, dim, i0, i1);
        return 0;
    } catch (const std::exception& e) {
        cb->error = e.what();
        return 1;
    }
}

template <typename T>
int spsi_tramp(void* user, const double* X, double* SX, int nrow, int npw, int nbands) {
    auto* cb = static_cast<Callbacks<T>*>(user);
    try {
        (*cb->s)(reinterpret_cast<T*>(const_cast<double*>(X)), reinterpret_cast<T*>(SX), nrow, npw, nbands);
        return 0;
    } catch (const std::exception& e) {
        cb->error = e.what();
        return 1;
    }
}

}  // namespace detail

template <>
class DiagoDavid<std::complex<double>> {
  public:
    using T = std::complex<double>;
    using Real = double;
    using HPsiFunc = std::function<void(T*, T*, const int, const int, const int, const int)>;
    using SPsiFunc = std::function<void(T*, T*, const int, const int, const int)>;

    // diag_comm_info of the reference: one GPU, one rank (no MPI inside this path).
    template <typename Comm>
    DiagoDavid(const Real* precondition_in, const int nband_in, const int dim_in, const int david_ndim_in,
               const bool use_paw_in, const Comm& /*diag_comm_in*/, int device = 0)
        : ctx_(default_context(device)), nband_(nband_in), dim_(dim_in) {
        if (use_paw_in) throw Error(APW_EINVAL, "apw::abacus::DiagoDavid: PAW is not supported");
        check(apw_david_create(ctx_, precondition_in, nband_in, dim_in, david_ndim_in, &dav_));
    }
    ~DiagoDavid() {
        if (dav_) (void)apw_david_destroy(dav_);
    }
    DiagoDavid(const DiagoDavid&) = delete;
    DiagoDavid& operator=(const DiagoDavid&) = delete;

    // DiagoDavid::diag (diago_david.cpp:1173-1200).  spsi_func may be empty (S = I).
    int diag(const HPsiFunc& hpsi_func, const SPsiFunc& spsi_func, const int ldPsi, T* psi_in, Real* eigenvalue_in,
             const Real david_diag_thr, const int david_maxiter, const int ntry_max = 5, const int notconv_max = 0) {
        detail::Callbacks<T> cb;
        cb.h = &hpsi_func;
        cb.s = spsi_func ? &spsi_func : nullptr;
        const int rc = apw_david_diag(dav_, nullptr, &detail::hpsi_tramp<T>, cb.s ? &detail::spsi_tramp<T> : nullptr,
                                      &cb, ldPsi, reinterpret_cast<double*>(psi_in), eigenvalue_in, david_diag_thr,
                                      david_maxiter, ntry_max, notconv_max, APW_FP64, &stats_);
        if (rc == APW_ECALLBACK && !cb.error.empty()) throw Error(rc, cb.error);
        return check(rc);
    }

    // Fast path: H|psi> applied by the library's own Hamiltonian (no host callback per
    // block); precision APW_FP64 or APW_FP32_MIXED.
    int diag(apw_ham* ham, const int ldPsi, T* psi_in, Real* eigenvalue_in, const Real david_diag_thr,
             const int david_maxiter, const int ntry_max = 5, const int notconv_max = 0, int precision = APW_FP64) {
        return check(apw_david_diag(dav_, ham, nullptr, nullptr, nullptr, ldPsi, reinterpret_cast<double*>(psi_in),
                                    eigenvalue_in, david_diag_thr, david_maxiter, ntry_max, notconv_max, precision,
                                    &stats_));
    }

    const apw_diag_stats& stats() const { return stats_; }
    apw_ctx* context() const { return ctx_; }

  private:
    apw_ctx* ctx_ = nullptr;
    apw_david* dav_ = nullptr;
    int nband_ = 0, dim_ = 0;
    apw_diag_stats stats_{};
};

// precision=single: c64 psi and callbacks.  The blocks are widened to c128 on the device on
// their way into the solver (and narrowed on the way to the callbacks and back to psi), so
// the Davidson arithmetic is the FP64 form's; eigenvalues come back as float.
template <>
class DiagoDavid<std::complex<float>> {
  public:
    using T = std::complex<float>;
    using Real = float;
    using HPsiFunc = std::function<void(T*, T*, const int, const int, const int, const int)>;
    using SPsiFunc = std::function<void(T*, T*, const int, const int, const int)>;

    template <typename Comm>
    DiagoDavid(const Real* precondition_in, const int nband_in, const int dim_in, const int david_ndim_in,
               const bool use_paw_in, const Comm& /*diag_comm_in*/, int device = 0)
        : ctx_(default_context(device)), nband_(nband_in), dim_(dim_in) {
        if (use_paw_in) throw Error(APW_EINVAL, "apw::abacus::DiagoDavid: PAW is not supported");
        std::vector<double> pre(precondition_in, precondition_in + dim_in);
        check(apw_david_create(ctx_, pre.data(), nband_in, dim_in, david_ndim_in, &dav_));
    }
    ~DiagoDavid() {
        if (dav_) (void)apw_david_destroy(dav_);
    }
    DiagoDavid(const DiagoDavid&) = delete;
    DiagoDavid& operator=(const DiagoDavid&) = delete;

    int diag(const HPsiFunc& hpsi_func, const SPsiFunc& spsi_func, const int ldPsi, T* psi_in, Real* eigenvalue_in,
             const Real david_diag_thr, const int david_maxiter, const int ntry_max = 5, const int notconv_max = 0) {
        if (spsi_func) throw Error(APW_EINVAL, "apw::abacus::DiagoDavid<complex<float>>: only S = I is supported");
        detail::Callbacks<T> cb;
        cb.h = &hpsi_func;
        std::vector<double> eig(nband_);
        const int rc = apw_david_diag_c64(dav_, &detail::hpsi_tramp<T>, &cb, ldPsi, reinterpret_cast<float*>(psi_in),
                                          eig.data(), david_diag_thr, david_maxiter, ntry_max, notconv_max, &stats_);
        if (rc == APW_ECALLBACK && !cb.error.empty()) throw Error(rc, cb.error);
        check(rc);
        for (int i = 0; i < nband_; ++i) eigenvalue_in[i] = static_cast<float>(eig[i]);
        return rc;
    }

    const apw_diag_stats& stats() const { return stats_; }

  private:
    apw_ctx* ctx_ = nullptr;
    apw_david* dav_ = nullptr;
    int nband_ = 0, dim_ = 0;
    apw_diag_stats stats_{};
};

// ------------------------------------------------------------------ HamiltPW
// The operator chain of one k-point.  desc.precision picks the FP64 or FP32 H|psi> arithmetic.
class HamiltPW {
  public:
    using T = std::complex<double>;
    using HPsiFunc = DiagoDavid<T>::HPsiFunc;

    explicit HamiltPW(const apw_ham_desc& desc, int device = 0) : ctx_(default_context(device)) {
        check(apw_ham_create(ctx_, &desc, &ham_));
        npwx_ = desc.npwx;
    }
    ~HamiltPW() {
        if (ham_) (void)apw_ham_destroy(ham_);
    }
    HamiltPW(const HamiltPW&) = delete;
    HamiltPW& operator=(const HamiltPW&) = delete;

    // Operator::hPsi on a band range: hpsi[:, b] = H psi[:, b], b < nbands (overwrite).
    void hPsi(const T* psi, int ldpsi, T* hpsi, int ldhpsi, int nbands) const {
        check(apw_hpsi(ham_, reinterpret_cast<const double*>(psi), ldpsi, reinterpret_cast<double*>(hpsi), ldhpsi,
                       nbands));
    }
    // Operator::act of the three operators, each accumulating into tmhpsi (npol = 1).
    void act_ekinetic(int nbands, int nbasis, int npol, const T* psi, T* hpsi) const {
        require_npol1(npol);
        check(apw_act_ekinetic(ham_, nbands, nbasis, reinterpret_cast<const double*>(psi),
                               reinterpret_cast<double*>(hpsi)));
    }
    void act_veff(int nbands, int nbasis, int npol, const T* psi, T* hpsi) const {
        require_npol1(npol);
        check(apw_act_veff(ham_, nbands, nbasis, reinterpret_cast<const double*>(psi), reinterpret_cast<double*>(hpsi)));
    }
    void act_nonlocal(int nbands, int nbasis, int npol, const T* psi, T* hpsi) const {
        require_npol1(npol);
        check(apw_act_nonlocal(ham_, nbands, nbasis, reinterpret_cast<const double*>(psi),
                               reinterpret_cast<double*>(hpsi)));
    }
    // The HPsiFunc of HSolverPW::hamiltSolvePsiK (hsolver_pw.cpp:488-500): H applied to
    // X[:, id_start..id_end] (X = head of the block, leading dimension dim) into HX.
    HPsiFunc hpsi_func() const {
        apw_ham* h = ham_;
        return [h](T* X, T* HX, const int /*nvec*/, const int dim, const int id_start, const int id_end) {
            check(apw_hpsi(h, reinterpret_cast<const double*>(X + (size_t)id_start * dim), dim,
                           reinterpret_cast<double*>(HX), dim, id_end - id_start + 1));
        };
    }
    apw_ham* handle() const { return ham_; }
    void synchronize() const { check(apw_ctx_synchronize(ctx_)); }

  private:
    static void require_npol1(int npol) {
        if (npol != 1) throw Error(APW_EINVAL, "apw::abacus::HamiltPW: npol must be 1");
    }
    apw_ctx* ctx_ = nullptr;
    apw_ham* ham_ = nullptr;
    int npwx_ = 0;
};

}  // namespace abacus
}  // namespace apw
