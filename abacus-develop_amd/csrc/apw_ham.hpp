// Hamiltonian of one k-point on the device (HamiltPW + PW_Basis_K + pseudopot data).
#pragma once

#include "apw_internal.hpp"
#include "pw_fft.hpp"

namespace apw {

struct Ham {
    apw_ctx* ctx = nullptr;
    PwGrid grid;
    int npw = 0, npwx = 0, nkb = 0, nhm = 0, nat = 0;
    double tpiba2 = 0;
    int precision = APW_FP64;
    int chunk = 1;                 // bands per local-pass launch (FP64)
    int chunk32 = 2;               // and FP32
    DevBuf<double> gk2, veff, vkb, deeq;
    DevBuf<float> veff32, vkb32;
    DevBuf<int> kb_iat, kb_off, kb_nh;
    // work buffers (grow-only)
    DevBuf<double> S, X, becp, ps, psi32, h32, wts;
    GemmWorkspace gws;

    void create(apw_ctx* c, const apw_ham_desc* d);
    // hpsi = H psi (overwrite), nb bands, c128 in/out
    void hpsi(const double* psi, int ld, double* out, int ldh, int nb, int prec);
    void local(const double* psi, int ld, double* out, int ldh, int nb, GatherMode mode, int prec);
    void nonlocal(const double* psi, int ld, double* out, int ldh, int nb, int prec);
    void act_ekinetic(int nb, int nbasis, const double* psi, double* out);
    void tune_local();   // autotuner of the local pass: chunk and pass schedule per precision
};

}  // namespace apw

struct apw_ham {
    apw::Ham h;
};
