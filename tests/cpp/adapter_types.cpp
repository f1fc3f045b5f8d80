// Compile-only check (g++ -fsyntax-only): the adapter's DiagoDavid and HamiltPW take exactly
// the types HSolverPW hands to the reference's DiagoDavid<T, DEVICE_GPU>
// (source/module_hsolver/diago_david.h:21-84), for T = complex<double> and complex<float>.
// Built with -I <reference>/source -I <reference>/source/module_base/module_container.
#include <type_traits>

#include "apw_abacus.hpp"
#include "module_hsolver/diag_comm_info.h"
#include "module_hsolver/diago_david.h"

template <typename T>
void check_one() {
    using Ref = hsolver::DiagoDavid<T, base_device::DEVICE_GPU>;
    using Ours = apw::abacus::DiagoDavid<T>;
    static_assert(std::is_same<typename Ref::HPsiFunc, typename Ours::HPsiFunc>::value, "HPsiFunc differs");
    static_assert(std::is_same<typename Ref::SPsiFunc, typename Ours::SPsiFunc>::value, "SPsiFunc differs");
    using Real = typename Ours::Real;
    // diag(hpsi_func, spsi_func, ldPsi, psi, eigenvalue, thr, maxiter, ntry_max, notconv_max)
    using RefDiag = int (Ref::*)(const typename Ref::HPsiFunc&, const typename Ref::SPsiFunc&, const int, T*, Real*,
                                 const Real, const int, const int, const int);
    using OurDiag = int (Ours::*)(const typename Ours::HPsiFunc&, const typename Ours::SPsiFunc&, const int, T*, Real*,
                                  const Real, const int, const int, const int);
    RefDiag r = &Ref::diag;
    OurDiag o = &Ours::diag;
    (void)r;
    (void)o;
    // the constructor HSolverPW calls: (precondition, nband, dim, david_ndim, use_paw, comm_info)
    static_assert(std::is_constructible<Ours, const Real*, int, int, int, bool, const hsolver::diag_comm_info&>::value,
                  "constructor differs");
    static_assert(std::is_constructible<Ref, const Real*, int, int, int, bool, const hsolver::diag_comm_info&>::value,
                  "reference constructor changed");
}

int main() {
    check_one<std::complex<double>>();
    check_one<std::complex<float>>();
    return 0;
}
