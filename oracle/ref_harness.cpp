// TEST INFRASTRUCTURE ONLY -- drives the reference's own CPU DiagoDavid.
//
// Compiled with the reference sources (oracle/Makefile) into oracle/_ref/libref_david.so;
// used by tests/ to pin the numpy oracle and by bench.py's cpu_baseline leg.  The H|psi>
// callback reproduces HSolverPW's hpsi_func (hsolver_pw.cpp:496-514) and the Operator
// chain (operator.cpp:39-88): hpsi = 0; Ekinetic::act (ekinetic_pw.cpp:34-58); Veff::act per
// band with the PW_Basis_K transforms (veff_pw.cpp:38-105, pw_transform_k.cpp CPU path:
// box scatter, backward FFT, veff, forward FFT, gather with 1/nxyz); Nonlocal::act
// (nonlocal_pw.cpp:211-304).  All operator arithmetic is the reference's own functors
// (ekinetic_pw_op, veff_pw_op, nonlocal_pw_op, set_3d_fft_box_op,
// set_real_to_recip_output_op, gemm_op on BLAS).  The 3D FFT is the FFTW3 API exported by
// the image's MKL runtime (the reference links FFTW3 the same way, fft.cpp); sPsi is the
// identity copy of HamiltPW::sPsi for norm-conserving pseudopotentials (hamilt_pw.cpp:227).
#include <complex>
#include <cstring>
#include <vector>

#include "module_basis/module_pw/kernels/pw_op.h"
#include "module_hamilt_pw/hamilt_pwdft/kernels/ekinetic_op.h"
#include "module_hamilt_pw/hamilt_pwdft/kernels/nonlocal_op.h"
#include "module_hamilt_pw/hamilt_pwdft/kernels/veff_op.h"
#include "module_hsolver/diago_david.h"
#include "module_hsolver/diago_dav_subspace.h"
#include "module_hsolver/kernels/math_kernel_op.h"

// FFTW3 API (www.fftw.org/fftw3_doc), resolved from MKL's libmkl_rt.
extern "C" {
typedef double fftw_complex[2];
typedef struct fftw_plan_s* fftw_plan;
fftw_plan fftw_plan_dft_3d(int n0, int n1, int n2, fftw_complex* in, fftw_complex* out, int sign, unsigned flags);
void fftw_execute_dft(const fftw_plan p, fftw_complex* in, fftw_complex* out);
void fftw_destroy_plan(fftw_plan p);
}
static const int kFftwForward = -1;
static const int kFftwBackward = +1;
static const unsigned kFftwEstimate = 1U << 6;

using T = std::complex<double>;
using CPU = base_device::DEVICE_CPU;

namespace {

struct RefHam {
    int nx, ny, nz, nxyz, npw, npwx, nkb, ntype, nhm, nat;
    double tpiba2;
    const int* ig2ixyz;
    const double* gk2;
    const double* veff;
    const T* vkb;          // col-major npwx x nkb
    const int* na;
    const int* nh;
    const double* deeq;    // [nat][nhm][nhm]
    std::vector<T> box, becp, ps;
    fftw_plan pb = nullptr, pf = nullptr;
    const CPU* ctx = nullptr;
    long hpsi_calls = 0, hpsi_bands = 0;
    std::vector<int> trace;   // bands per H|psi> call
    std::vector<double> etrace;   // eigenvalue_in seen at each call
    const double* eig_watch = nullptr;
    int neig = 0;
    int dump_call = -1;                 // copy the block head X[:, 0..id_end] of this call
    std::vector<T> dump;

    void init() {
        nxyz = nx * ny * nz;
        box.assign(nxyz, T(0, 0));
        auto* b = reinterpret_cast<fftw_complex*>(box.data());
        pb = fftw_plan_dft_3d(nx, ny, nz, b, b, kFftwBackward, kFftwEstimate);
        pf = fftw_plan_dft_3d(nx, ny, nz, b, b, kFftwForward, kFftwEstimate);
    }
    ~RefHam() {
        if (pb) fftw_destroy_plan(pb);
        if (pf) fftw_destroy_plan(pf);
    }

    // Operator::hPsi on nb bands of leading dimension ld
    void hpsi(const T* psi, T* hp, int nb, int ld) {
        ++hpsi_calls;
        hpsi_bands += nb;
        trace.push_back(nb);
        for (int i = 0; eig_watch && i < neig; ++i) etrace.push_back(eig_watch[i]);
        std::memset(static_cast<void*>(hp), 0, sizeof(T) * (size_t)nb * ld);
        hamilt::ekinetic_pw_op<double, CPU>()(ctx, nb, npw, ld, tpiba2, gk2, hp, psi);
        auto* b = reinterpret_cast<fftw_complex*>(box.data());
        for (int ib = 0; ib < nb; ++ib) {
            std::fill(box.begin(), box.end(), T(0, 0));
            ModulePW::set_3d_fft_box_op<double, CPU>()(ctx, npw, ig2ixyz, psi + (size_t)ib * ld, box.data());
            fftw_execute_dft(pb, b, b);
            hamilt::veff_pw_op<double, CPU>()(ctx, nxyz, box.data(), veff);
            fftw_execute_dft(pf, b, b);
            ModulePW::set_real_to_recip_output_op<double, CPU>()(ctx, npw, nxyz, true, 1.0, ig2ixyz, box.data(),
                                                                 hp + (size_t)ib * ld);
        }
        if (nkb > 0) {
            const T one(1, 0), zero(0, 0);
            becp.assign((size_t)nkb * nb, zero);
            ps.assign((size_t)nkb * nb, zero);
            hsolver::gemm_op<T, CPU>()(ctx, 'C', 'N', nkb, nb, npw, &one, vkb, npwx, psi, ld, &zero, becp.data(), nkb);
            int sum = 0, iat = 0;
            for (int it = 0; it < ntype; ++it) {
                hamilt::nonlocal_pw_op<double, CPU>()(ctx, na[it], nb, nh[it], sum, iat, 0, nkb, nat, nhm, nhm, deeq,
                                                      ps.data(), becp.data());
            }
            hsolver::gemm_op<T, CPU>()(ctx, 'N', 'T', npw, nb, nkb, &one, vkb, npwx, ps.data(), nb, &one, hp, ld);
        }
    }
};

}  // namespace

extern "C" {

// Runs the reference DiagoDavid<complex<double>, CPU>::diag from psi (nband x ldpsi,
// row = band) with H of the given arrays.  Returns the iteration count; eig[nband] and psi
// are overwritten; counts[0..1] = hpsi calls, hpsi bands; trace[i] = bands of call i.
int ref_david_diag(int nx, int ny, int nz, int npw, int npwx, const int* ig2ixyz, const double* gk2, double tpiba2,
                   const double* veff, int nkb, const double* vkb, int ntype, const int* na, const int* nh, int nhm,
                   const double* deeq, const double* precondition, int nband, int david_ndim, int ldpsi, double* psi,
                   double* eig, double thr, int maxiter, int ntry_max, int notconv_max, long* counts,
                   int* trace, int trace_cap, double* etrace, int dump_call, double* dump_out) {
    RefHam h;
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.npw = npw;
    h.npwx = npwx;
    h.nkb = nkb;
    h.ntype = ntype;
    h.nhm = nhm;
    h.tpiba2 = tpiba2;
    h.ig2ixyz = ig2ixyz;
    h.gk2 = gk2;
    h.veff = veff;
    h.vkb = reinterpret_cast<const T*>(vkb);
    h.na = na;
    h.nh = nh;
    h.deeq = deeq;
    h.nat = 0;
    for (int it = 0; it < ntype; ++it) h.nat += na[it];
    h.init();
    h.eig_watch = eig;
    h.dump_call = dump_call;
    h.neig = nband;
    const int dim = npw;
    auto hpsi_func = [&h](T* X, T* HX, const int nvec, const int d, const int i0, const int i1) {
        (void)nvec;
        if ((long)h.dump_call == h.hpsi_calls) h.dump.assign(X, X + (size_t)(i1 + 1) * d);
        h.hpsi(X + (size_t)i0 * d, HX, i1 - i0 + 1, d);
    };
    auto spsi_func = [](const T* X, T* SX, const int nrow, const int npw_, const int nbands) {
        (void)npw_;
        std::memcpy(static_cast<void*>(SX), X, sizeof(T) * (size_t)nrow * nbands);
    };
    hsolver::diag_comm_info comm(0, 1);
    hsolver::DiagoDavid<T, CPU> david(precondition, nband, dim, david_ndim, false, comm);
    const int it = david.diag(hpsi_func, spsi_func, ldpsi, reinterpret_cast<T*>(psi), eig, thr, maxiter, ntry_max,
                              notconv_max);
    if (counts) {
        counts[0] = h.hpsi_calls;
        counts[1] = h.hpsi_bands;
    }
    for (int i = 0; trace && i < (int)h.trace.size() && i < trace_cap; ++i) trace[i] = h.trace[i];
    for (size_t i = 0; etrace && i < h.etrace.size() && i < (size_t)trace_cap * nband; ++i) etrace[i] = h.etrace[i];
    if (dump_out && !h.dump.empty()) std::memcpy(static_cast<void*>(dump_out), h.dump.data(), h.dump.size() * sizeof(T));
    return it;
}

// Runs the reference Diago_DavSubspace<complex<double>, CPU>::diag from psi (nband x
// ldpsi, band-major c128), H applied by the reference functors.  Returns the iteration
// count; eig receives the eigenvalues (in: used only as the previous values of the first
// check, as in HSolverPW).
int ref_dav_subspace_diag(int nx, int ny, int nz, int npw, int npwx, const int* ig2ixyz, const double* gk2,
                          double tpiba2, const double* veff, int nkb, const double* vkb, int ntype, const int* na,
                          const int* nh, int nhm, const double* deeq, const double* precondition, int nband,
                          int david_ndim, int ldpsi, double* psi, double* eig, double thr, int maxiter,
                          int need_subspace, int scf, long* counts) {
    RefHam h;
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.npw = npw;
    h.npwx = npwx;
    h.nkb = nkb;
    h.ntype = ntype;
    h.nhm = nhm;
    h.tpiba2 = tpiba2;
    h.ig2ixyz = ig2ixyz;
    h.gk2 = gk2;
    h.veff = veff;
    h.vkb = reinterpret_cast<const T*>(vkb);
    h.na = na;
    h.nh = nh;
    h.deeq = deeq;
    h.nat = 0;
    for (int it = 0; it < ntype; ++it) h.nat += na[it];
    h.init();
    h.neig = nband;
    const int dim = npw;
    auto hpsi_func = [&h](T* X, T* HX, const int nvec, const int d, const int i0, const int i1) {
        (void)nvec;
        h.hpsi(X + (size_t)i0 * d, HX, i1 - i0 + 1, d);
    };
    std::vector<double> pre(precondition, precondition + npwx);
    hsolver::diag_comm_info comm(0, 1);
    hsolver::Diago_DavSubspace<T, CPU> dav(pre, nband, dim, david_ndim, thr, maxiter, need_subspace != 0, comm);
    const std::vector<bool> occ(nband, true);
    const int it = dav.diag(hpsi_func, reinterpret_cast<T*>(psi), ldpsi, eig, occ, scf != 0);
    if (counts) {
        counts[0] = h.hpsi_calls;
        counts[1] = h.hpsi_bands;
    }
    return it;
}

// One Operator::hPsi application (for checking the harness itself).
void ref_hpsi(int nx, int ny, int nz, int npw, int npwx, const int* ig2ixyz, const double* gk2, double tpiba2,
              const double* veff, int nkb, const double* vkb, int ntype, const int* na, const int* nh, int nhm,
              const double* deeq, const double* psi, double* hpsi, int nb, int ld) {
    RefHam h;
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.npw = npw;
    h.npwx = npwx;
    h.nkb = nkb;
    h.ntype = ntype;
    h.nhm = nhm;
    h.tpiba2 = tpiba2;
    h.ig2ixyz = ig2ixyz;
    h.gk2 = gk2;
    h.veff = veff;
    h.vkb = reinterpret_cast<const T*>(vkb);
    h.na = na;
    h.nh = nh;
    h.deeq = deeq;
    h.nat = 0;
    for (int it = 0; it < ntype; ++it) h.nat += na[it];
    h.init();
    h.hpsi(reinterpret_cast<const T*>(psi), reinterpret_cast<T*>(hpsi), nb, ld);
}

}  // extern "C"
