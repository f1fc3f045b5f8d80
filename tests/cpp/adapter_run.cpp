// Runs the header-only adapter (include/apw_abacus.hpp) the way HSolverPW drives the
// reference's DiagoDavid<T, DEVICE_GPU>: a DiagoDavid object built from (precondition, nband,
// dim, david_ndim, use_paw, comm_info) and diag() with an HPsiFunc std::function.  H is a
// dense Hermitian matrix read from a file and applied on the device by the library's GEMM.
//
//   adapter_run <dir> <double|float>
// <dir>/h.bin: n x n complex<double> column-major; <dir>/psi0.bin: nband x n start vectors
// (row = band); <dir>/meta.txt: "n nband david_ndim thr".  Writes <dir>/eig_<type>.txt.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "apw_abacus.hpp"

namespace {

struct CommInfo {   // the reference's diag_comm_info without MPI: (rank, nproc)
    int rank = 0, nproc = 1;
};

template <typename T>
std::vector<T> read_bin(const std::string& path, size_t count) {
    std::vector<std::complex<double>> raw(count);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)(count * sizeof(raw[0])));
    if (!f) throw std::runtime_error("cannot read " + path);
    std::vector<T> out(count);
    for (size_t i = 0; i < count; ++i) out[i] = T(raw[i].real(), raw[i].imag());
    return out;
}

template <typename T>
T* to_device(const std::vector<T>& h) {
    T* d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(T)) != hipSuccess) throw std::runtime_error("hipMalloc");
    if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("hipMemcpy");
    return d;
}

int gemm(apw_ctx* ctx, int m, int nb, int k, const std::complex<double>* A, const std::complex<double>* B,
         std::complex<double>* C) {
    const double one[2] = {1.0, 0.0}, zero[2] = {0.0, 0.0};
    return apw_zgemm(ctx, 'N', 'N', m, nb, k, one, reinterpret_cast<const double*>(A), m,
                     reinterpret_cast<const double*>(B), k, zero, reinterpret_cast<double*>(C), m);
}

int gemm(apw_ctx* ctx, int m, int nb, int k, const std::complex<float>* A, const std::complex<float>* B,
         std::complex<float>* C) {
    const float one[2] = {1.0f, 0.0f}, zero[2] = {0.0f, 0.0f};
    return apw_cgemm(ctx, 'N', 'N', m, nb, k, one, reinterpret_cast<const float*>(A), m,
                     reinterpret_cast<const float*>(B), k, zero, reinterpret_cast<float*>(C), m);
}

template <typename T>
int run(const std::string& dir, const char* tag) {
    using Real = typename T::value_type;
    int n = 0, nband = 0, ndim = 0;
    double thr = 0;
    {
        std::ifstream m(dir + "/meta.txt");
        m >> n >> nband >> ndim >> thr;
        if (!m) throw std::runtime_error("bad meta.txt");
    }
    const std::vector<T> h = read_bin<T>(dir + "/h.bin", (size_t)n * n);
    const std::vector<T> psi0 = read_bin<T>(dir + "/psi0.bin", (size_t)n * nband);
    std::vector<Real> pre(n);
    for (int i = 0; i < n; ++i) {
        const Real d = h[(size_t)i * n + i].real();
        pre[i] = (d > -1 && d < 1) ? Real(1) : d;
    }
    T* dH = to_device(h);
    T* dpsi = to_device(psi0);
    CommInfo comm;
    apw::abacus::DiagoDavid<T> dav(pre.data(), nband, n, ndim, false, comm);
    apw_ctx* ctx = apw::abacus::default_context(0);
    // HPsiFunc: HX = H X[:, id_start..id_end], X the head of the block (hsolver_pw.cpp:488-500)
    typename apw::abacus::DiagoDavid<T>::HPsiFunc hpsi = [&](T* X, T* HX, const int, const int dim,
                                                             const int i0, const int i1) {
        apw::abacus::check(gemm(ctx, dim, i1 - i0 + 1, dim, dH, X + (size_t)i0 * dim, HX));
    };
    std::vector<Real> eig(nband);
    const int it = dav.diag(hpsi, {}, n, dpsi, eig.data(), (Real)thr, 300);
    FILE* f = std::fopen((dir + "/eig_" + tag + ".txt").c_str(), "w");
    std::fprintf(f, "%d\n", it);
    for (int i = 0; i < nband; ++i) std::fprintf(f, "%.17g\n", (double)eig[i]);
    std::fclose(f);
    // an HPsiFunc that throws comes back as an apw::abacus::Error carrying its message
    typename apw::abacus::DiagoDavid<T>::HPsiFunc bad = [](T*, T*, const int, const int, const int, const int) {
        throw std::runtime_error("hpsi failed on purpose");
    };
    bool caught = false;
    try {
        dav.diag(bad, {}, n, dpsi, eig.data(), (Real)thr, 300);
    } catch (const apw::abacus::Error& e) {
        caught = e.code() == APW_ECALLBACK && std::strstr(e.what(), "on purpose") != nullptr;
    }
    (void)hipFree(dH);
    (void)hipFree(dpsi);
    std::printf("%s: %d iterations, callback error %s\n", tag, it, caught ? "reported" : "NOT reported");
    return caught ? 0 : 3;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: adapter_run <dir> <double|float>\n");
        return 2;
    }
    try {
        return std::string(argv[2]) == "float" ? run<std::complex<float>>(argv[1], "float")
                                               : run<std::complex<double>>(argv[1], "double");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "adapter_run: %s\n", e.what());
        return 1;
    }
}
