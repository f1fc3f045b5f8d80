// TEST INFRASTRUCTURE ONLY -- deterministic eigenvector phases for the reference build.
//
// The reference's DiagoDavid keeps the off-diagonal entries of vcc across refresh()
// (diago_david.cpp:866-870 clear hcc and scc but only set vcc's diagonal), so its
// trajectory depends on the phases zheevx gives the subspace eigenvectors.  LAPACK's
// zheevx fixes them by its inverse-iteration convention (zhetrd 'L' reflectors, then the
// real tridiagonal eigenvector scaled so its largest component is positive) whenever fewer
// than all eigenpairs are requested, but when all are requested (IL = 1, IU = N: the first
// Rayleigh-Ritz step of every diag_once) it takes a QR-iteration path whose signs differ
// between LAPACK implementations.
//
// Linked into oracle/_ref/libref_david_detsign.so only: this zheevx_ calls the library's
// zheevx and, on that all-eigenpair path, re-phases each eigenvector v to the
// inverse-iteration convention: z = Q^H v (Q from zhetrd 'L'), v <- v * conj(z_i)/|z_i| for
// the first i of largest |z_i|.  The reference sources are compiled unchanged.
#include <complex>
#include <cmath>
#include <vector>

using cplx = std::complex<double>;

extern "C" {
// LAPACK (www.netlib.org/lapack), resolved from the runtime's non-underscored aliases
void zheevx(const char* jobz, const char* range, const char* uplo, const int* n, cplx* a, const int* lda,
            const double* vl, const double* vu, const int* il, const int* iu, const double* abstol, int* m, double* w,
            cplx* z, const int* ldz, cplx* work, const int* lwork, double* rwork, int* iwork, int* ifail, int* info);
void zhetrd(const char* uplo, const int* n, cplx* a, const int* lda, double* d, double* e, cplx* tau, cplx* work,
            const int* lwork, int* info);
void zunmtr(const char* side, const char* uplo, const char* trans, const int* m, const int* n, const cplx* a,
            const int* lda, const cplx* tau, cplx* c, const int* ldc, cplx* work, const int* lwork, int* info);

void zheevx_(const char* jobz, const char* range, const char* uplo, const int* n, cplx* a, const int* lda,
             const double* vl, const double* vu, const int* il, const int* iu, const double* abstol, int* m, double* w,
             cplx* z, const int* ldz, cplx* work, const int* lwork, double* rwork, int* iwork, int* ifail, int* info) {
    const int N = *n;
    const bool all = (*jobz == 'V' || *jobz == 'v') && (*range == 'I' || *range == 'i') && *il == 1 && *iu == N &&
                     *lwork != -1 && N > 1 && (*uplo == 'L' || *uplo == 'l');
    std::vector<cplx> acopy;
    if (all) {
        acopy.resize((size_t)N * N);
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) acopy[(size_t)j * N + i] = a[(size_t)j * (*lda) + i];
    }
    zheevx(jobz, range, uplo, n, a, lda, vl, vu, il, iu, abstol, m, w, z, ldz, work, lwork, rwork, iwork, ifail, info);
    if (!all || *info != 0) return;
    const int M = *m;
    std::vector<double> d(N), e(N);
    std::vector<cplx> tau(N), zq((size_t)N * M);
    int lw = -1, inf = 0;
    cplx q;
    zhetrd("L", n, acopy.data(), n, d.data(), e.data(), tau.data(), &q, &lw, &inf);
    lw = std::max(1, (int)q.real());
    {
        const int lw2 = -1;
        cplx q2;
        zunmtr("L", "L", "C", n, &M, acopy.data(), n, tau.data(), zq.data(), n, &q2, &lw2, &inf);
        lw = std::max(lw, (int)q2.real());
    }
    std::vector<cplx> wk(lw);
    zhetrd("L", n, acopy.data(), n, d.data(), e.data(), tau.data(), wk.data(), &lw, &inf);
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < N; ++i) zq[(size_t)j * N + i] = z[(size_t)j * (*ldz) + i];
    zunmtr("L", "L", "C", n, &M, acopy.data(), n, tau.data(), zq.data(), n, wk.data(), &lw, &inf);
    for (int j = 0; j < M; ++j) {
        int imax = 0;
        double vmax = -1.0;
        for (int i = 0; i < N; ++i) {
            const double v = std::abs(zq[(size_t)j * N + i]);
            if (v > vmax) {
                vmax = v;
                imax = i;
            }
        }
        const cplx ph = zq[(size_t)j * N + imax] / vmax;
        for (int i = 0; i < N; ++i) z[(size_t)j * (*ldz) + i] *= std::conj(ph);
    }
}
}
