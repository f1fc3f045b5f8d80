// Column-block helpers shared by the Davidson solvers (DiagoDavid and Diago_DavSubspace):
// copies, zero fills and the gather of selected eigenvector columns, on column-major c128
// blocks.  Grid-stride loops, one element per thread per step.
#include <algorithm>

#include "apw_internal.hpp"

namespace apw {
namespace {

int grid_of(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

__global__ void k_copy_cols(int rows, int ncols, const double2* __restrict__ src, int lds, double2* __restrict__ dst,
                            int ldd) {
    const size_t total = (size_t)rows * ncols;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t % rows), c = (int)(t / rows);
        dst[(size_t)c * ldd + r] = src[(size_t)c * lds + r];
    }
}

__global__ void k_zero_rows(int r0, int r1, int ncols, double2* __restrict__ x, int ld) {
    const int rows = r1 - r0;
    const size_t total = (size_t)rows * ncols;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int r = r0 + (int)(t % rows), c = (int)(t / rows);
        x[(size_t)c * ld + r] = make_double2(0.0, 0.0);
    }
}

__global__ void k_zero(size_t n, double* __restrict__ x) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) x[t] = 0.0;
}

__global__ void k_gather_cols(int nbase, int nc, const double2* __restrict__ vcc, int nbx, const int* __restrict__ sel,
                              double2* __restrict__ vc) {
    const size_t total = (size_t)nbase * nc;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nbase), m = (int)(t / nbase);
        vc[t] = vcc[(size_t)sel[m] * nbx + i];
    }
}

}  // namespace

void copy_cols(hipStream_t s, int rows, int ncols, const double2* src, int lds, double2* dst, int ldd) {
    if ((size_t)rows * ncols == 0) return;
    hipLaunchKernelGGL(k_copy_cols, dim3(grid_of((size_t)rows * ncols)), dim3(256), 0, s, rows, ncols, src, lds, dst,
                       ldd);
    APW_LAUNCH_CHECK();
}

void zero_rows(hipStream_t s, int r0, int r1, int ncols, double2* x, int ld) {
    if ((size_t)(r1 - r0) * ncols == 0) return;
    hipLaunchKernelGGL(k_zero_rows, dim3(grid_of((size_t)(r1 - r0) * ncols)), dim3(256), 0, s, r0, r1, ncols, x, ld);
    APW_LAUNCH_CHECK();
}

void zero(hipStream_t s, size_t n, double* x) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_zero, dim3(grid_of(n)), dim3(256), 0, s, n, x);
    APW_LAUNCH_CHECK();
}

void gather_cols(hipStream_t s, int nbase, int nc, const double2* vcc, int nbx, const int* sel, double2* vc) {
    if ((size_t)nbase * nc == 0) return;
    hipLaunchKernelGGL(k_gather_cols, dim3(grid_of((size_t)nbase * nc)), dim3(256), 0, s, nbase, nc, vcc, nbx, sel,
                       vc);
    APW_LAUNCH_CHECK();
}

}  // namespace apw
