// Complex GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).
//
// Replaces gemm_op<std::complex<T>, DEVICE_GPU> (hipBLAS/cuBLAS Zgemm/Cgemm in the
// reference, rocm/math_kernel_op.hip.cu:798) for the shapes of the path:
//   * inner-product shapes  C(m x n) = A^H B with k = npw (up to ~2.3e5) and m, n <= ~2e3
//     (becp = vkb^H psi, cal_elem, Gram matrices)  -> split-K with a deterministic reduction;
//   * update shapes         C(npw x n) += A(npw x k) B(k x n) with k <= ~2e3
//     (back-projection, cal_grad, refresh, final psi) -> plain 2D tiling.
//
// Tile: 64 x 64, 128 x 64 or 64 x 128 (m x n) per workgroup, k step 8 (16 for the wide
// tile) through two LDS stages, 4 waves in a 2 x 2 arrangement, each wave owning TM x TN MFMA 16x16 sub-tiles
// with separate accumulators (a complex product = 3 real MFMAs; 4 in the x32 form).  Operand roles are swapped
// (MFMA "A" <- op(B) indexed by n, MFMA "B" <- op(A) indexed by m) so that the 16 lanes of
// a group hold 16 consecutive rows m of one column n: the column-major stores of C are then
// 256-byte coalesced runs.  Accumulator register r of lane l holds D row drow(l, r) and
// column l % 16; drow is (l / 16) + 4 r for the f64 instruction and 4 (l / 16) + r for the
// f32 one (checked on the device by the layout probe below and its unit test).
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

#include "apw_internal.hpp"

namespace apw {
namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <class R>
struct Mfma;
template <>
struct Mfma<double> {
    using acc_t = d4_t;
    using c_t = double2;
    __device__ static inline acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of D held in accumulator register r of a lane (column = lane % 16)
    __device__ static inline int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
    using acc_t = f4_t;
    using c_t = float2;
    __device__ static inline acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int BKMAX = 16, NT = 256;

// op codes: 0 = N, 1 = T, 2 = C.  Each of the 4 waves (2 x 2) owns (16 TM) x (16 TN) of the
// (32 TM) x (32 TN) tile.  Two LDS stages: while the MFMAs consume stage s, the k-tile
// after the next is loaded into registers and the next one is written to stage s^1, so a
// k step costs one workgroup barrier.  LDS rows are padded (16 / BK elements) so that the
// k-fastest operand stores fall in distinct banks.
// S: element type of A, B, C and the split-K partials.  S = R is the plain form.  R = float
// with S = double is the FP32-MFMA form of an FP64 product: operands rounded to FP32 as
// they are staged in LDS, FP32 accumulators folded into FP64 ones every XFOLD k steps
// (32 or 64 k values), so the rounding is that of the FP32 operands, not of a long FP32
// sum; C and the partials stay FP64.
constexpr int XFOLD = 4;
template <class R, class S, int TA, int TB, int TM, int TN, int BK, int MAP>
__global__ __launch_bounds__(NT) void zgemm_kernel(int M, int N, int K, int kchunk,
                                                   const typename Mfma<S>::c_t* __restrict__ A, int lda,
                                                   const typename Mfma<S>::c_t* __restrict__ B, int ldb,
                                                   typename Mfma<S>::c_t* __restrict__ C, int ldc,
                                                   double ar, double ai, double br, double bi,
                                                   typename Mfma<S>::c_t* __restrict__ part) {
    using c_t = typename Mfma<S>::c_t;
    using acc_t = typename Mfma<R>::acc_t;
    constexpr bool X32 = !std::is_same<R, S>::value;
    static_assert(!X32 || (std::is_same<R, float>::value && std::is_same<S, double>::value), "x32: f32 on f64");
    // three real products per complex one (P1 = re re, P2 = im im, P3 = (re + im)(re + im);
    // re = P1 - P2, im = P3 - P1 - P2), 3 MFMAs instead of 4; the error stays normwise, a few
    // ulps of |a||b| per term.  The x32 form keeps four (its FP32 sums are folded into FP64)
    constexpr bool M3 = !X32;
    constexpr int BM = 32 * TM, BN = 32 * TN;
    constexpr int NA = BM * BK / NT, NB = BN * BK / NT;
    // row padding 16 / BK elements: a k-fastest staging store puts 16 lanes on BK consecutive
    // k rows of 16 / BK columns, which then fall on 16 distinct bank pairs (with one element of
    // padding and BK = 8, lanes 8..15 hit the banks of lanes 0..7); the m- or n-fastest stores
    // and the MFMA operand reads stay conflict-free
    constexpr int LA = BM + 16 / BK, LB = BN + 16 / BK;
    __shared__ R As[2][2][BK][LA];   // [stage][re, im][k][m]
    __shared__ R Bs[2][2][BK][LB];   // [stage][re, im][k][n]

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    // tile of this workgroup.  MAP 0: blockIdx.x walks m; MAP 1: it walks n.  MAP 2 (XCD-aware,
    // 1-D grid in x; measured slower for split-K, used for the tall updates only):
    // workgroup L runs on XCD L % 8, so each XCD takes whole m-tiles and walks their n-tiles
    // back to back; the A tile they share stays in that XCD's L2.  The grid is padded to a
    // multiple of 8 m-tiles; padding workgroups leave before any barrier.
    // MAP 3 (XCD-aware split-K, 1-D grid padded to a multiple of 8): the workgroups of one
    // k split run on one or two XCDs, so the A and B slices that all tiles of the split read
    // at the same k come from that XCD's L2 instead of once per tile from HBM.  XCD x takes
    // the x-th eighth of the (split, tile) sequence, tiles of a split consecutive.
    int bm = blockIdx.x, bn = blockIdx.y, sp = blockIdx.z;
    if (MAP == 1) {   // blockIdx.x walks n
        bm = blockIdx.y;
        bn = blockIdx.x;
    } else if (MAP == 2) {
        const int L = blockIdx.x;
        const int ntn = (N + BN - 1) / BN, tmt = (M + BM - 1) / BM;
        bm = (L % 8) + 8 * (L / (8 * ntn));
        bn = (L / 8) % ntn;
        if (bm >= tmt) return;
    } else if (MAP == 3) {
        const int tmt = (M + BM - 1) / BM, tiles = tmt * ((N + BN - 1) / BN);
        const int per = (int)gridDim.x / 8;
        const int g = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
        sp = g / tiles;
        if (sp * kchunk >= K) return;   // padding workgroups: no barrier has been reached
        const int t = g - sp * tiles;
        bm = t % tmt;
        bn = t / tmt;
    }
    const int m0 = bm * BM, n0 = bn * BN;
    const int kbeg = sp * kchunk;
    const int kend = min(K, kbeg + kchunk);

    acc_t acc_re[TN][TM], acc_im[TN][TM];   // 4M: re, im; 3M: P1, P3
    acc_t acc_p2[M3 ? TN : 1][M3 ? TM : 1];   // 3M: P2
    d4_t wide_re[X32 ? TN : 1][X32 ? TM : 1], wide_im[X32 ? TN : 1][X32 ? TM : 1];   // X32: FP64 sums
#pragma unroll
    for (int a = 0; a < TN; ++a)
#pragma unroll
        for (int b = 0; b < TM; ++b) {
            acc_re[a][b] = acc_t{0, 0, 0, 0};
            acc_im[a][b] = acc_t{0, 0, 0, 0};
            if constexpr (M3) acc_p2[a][b] = acc_t{0, 0, 0, 0};
            if constexpr (X32) {
                wide_re[a][b] = d4_t{0, 0, 0, 0};
                wide_im[a][b] = d4_t{0, 0, 0, 0};
            }
        }
    auto fold = [&]() {
        if constexpr (X32) {
#pragma unroll
            for (int a = 0; a < TN; ++a)
#pragma unroll
                for (int b = 0; b < TM; ++b) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        wide_re[a][b][r] += (double)acc_re[a][b][r];
                        wide_im[a][b][r] += (double)acc_im[a][b][r];
                    }
                    acc_re[a][b] = acc_t{0, 0, 0, 0};
                    acc_im[a][b] = acc_t{0, 0, 0, 0};
                }
        }
    };

    // Prefetch registers.  The loads are unconditional (an out-of-range element reads
    // element 0 instead, which is valid for k > 0) so the compiler never has to wait on
    // one inside a branch; range masking and the conjugation happen at the LDS store.
    c_t ra[NA], rb[NB];
    unsigned oka = 0, okb = 0;
    auto fetch = [&](int k0) {
        oka = 0;
        okb = 0;
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = tid + NT * i;
            int mm, kk;
            if (TA == 0) { mm = e % BM; kk = e / BM; } else { kk = e % BK; mm = e / BK; }
            const int gm = m0 + mm, gk = k0 + kk;
            const bool ok = gm < M && gk < kend;
            const size_t off = (TA == 0) ? (size_t)gm + (size_t)gk * lda : (size_t)gk + (size_t)gm * lda;
            ra[i] = A[ok ? off : 0];
            oka |= (unsigned)ok << i;
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = tid + NT * i;
            int nn, kk;
            if (TB == 0) { kk = e % BK; nn = e / BK; } else { nn = e % BN; kk = e / BN; }
            const int gn = n0 + nn, gk = k0 + kk;
            const bool ok = gn < N && gk < kend;
            const size_t off = (TB == 0) ? (size_t)gk + (size_t)gn * ldb : (size_t)gn + (size_t)gk * ldb;
            rb[i] = B[ok ? off : 0];
            okb |= (unsigned)ok << i;
        }
    };
    // LDS store of the prefetched elements: range mask and conjugation as bit operations
    // (no branches), so the stores can be scheduled among the MFMAs of the k step
    using bits_t = typename std::conditional<sizeof(R) == 8, unsigned long long, unsigned>::type;
    constexpr bits_t SIGN = (bits_t)1 << (8 * sizeof(R) - 1);
    auto masked = [](R v, bool ok, bool neg) {
        bits_t b = __builtin_bit_cast(bits_t, v);
        if (neg) b ^= SIGN;
        return __builtin_bit_cast(R, b & ((bits_t)0 - (bits_t)ok));
    };
    auto stage = [&](int st) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = tid + NT * i;
            int mm, kk;
            if (TA == 0) { mm = e % BM; kk = e / BM; } else { kk = e % BK; mm = e / BK; }
            const bool ok = (oka >> i) & 1u;
            As[st][0][kk][mm] = masked((R)ra[i].x, ok, false);
            As[st][1][kk][mm] = masked((R)ra[i].y, ok, TA == 2);
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = tid + NT * i;
            int nn, kk;
            if (TB == 0) { kk = e % BK; nn = e / BK; } else { nn = e % BN; kk = e / BN; }
            const bool ok = (okb >> i) & 1u;
            Bs[st][0][kk][nn] = masked((R)rb[i].x, ok, false);
            Bs[st][1][kk][nn] = masked((R)rb[i].y, ok, TB == 2);
        }
    };

    if (kbeg < kend) {
        fetch(kbeg);
        stage(0);
        fetch(kbeg + BK);
    }
    __syncthreads();
    int st = 0, nfold = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            R xr[TN], xi[TN], yr[TM], yi[TM];
#pragma unroll
            for (int t = 0; t < TN; ++t) {
                const int nn = wn * 16 * TN + t * 16 + (lane & 15);
                xr[t] = Bs[st][0][kk][nn];
                xi[t] = Bs[st][1][kk][nn];
            }
#pragma unroll
            for (int t = 0; t < TM; ++t) {
                const int mm = wm * 16 * TM + t * 16 + (lane & 15);
                yr[t] = As[st][0][kk][mm];
                yi[t] = As[st][1][kk][mm];
            }
            if constexpr (M3) {
                R xs[TN], ys[TM];
#pragma unroll
                for (int t = 0; t < TN; ++t) xs[t] = xr[t] + xi[t];
#pragma unroll
                for (int t = 0; t < TM; ++t) ys[t] = yr[t] + yi[t];
#pragma unroll
                for (int tn = 0; tn < TN; ++tn)
#pragma unroll
                    for (int tm = 0; tm < TM; ++tm) {
                        acc_re[tn][tm] = Mfma<R>::mma(xr[tn], yr[tm], acc_re[tn][tm]);
                        acc_p2[tn][tm] = Mfma<R>::mma(xi[tn], yi[tm], acc_p2[tn][tm]);
                        acc_im[tn][tm] = Mfma<R>::mma(xs[tn], ys[tm], acc_im[tn][tm]);
                    }
            } else {
#pragma unroll
                for (int tn = 0; tn < TN; ++tn)
#pragma unroll
                    for (int tm = 0; tm < TM; ++tm) {
                        acc_re[tn][tm] = Mfma<R>::mma(xr[tn], yr[tm], acc_re[tn][tm]);
                        acc_re[tn][tm] = Mfma<R>::mma(-xi[tn], yi[tm], acc_re[tn][tm]);
                        acc_im[tn][tm] = Mfma<R>::mma(xr[tn], yi[tm], acc_im[tn][tm]);
                        acc_im[tn][tm] = Mfma<R>::mma(xi[tn], yr[tm], acc_im[tn][tm]);
                    }
            }
        }
        // unconditional (one basic block with the MFMAs): past the end the masks are zero and
        // the loads read element 0
        stage(st ^ 1);
        fetch(k0 + 2 * BK);
        if constexpr (X32) {
            if (++nfold == XFOLD) {
                fold();
                nfold = 0;
            }
        }
        __syncthreads();
        st ^= 1;
    }
    fold();

    // epilogue: lane holds D[n = drow(lane, r)][m = lane%16] of each sub-tile
#pragma unroll
    for (int tn = 0; tn < TN; ++tn)
#pragma unroll
        for (int tm = 0; tm < TM; ++tm)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gm = m0 + wm * 16 * TM + tm * 16 + (lane & 15);
                const int gn = n0 + wn * 16 * TN + tn * 16 + Mfma<R>::drow(lane, r);
                if (gm >= M || gn >= N) continue;
                double xr, xi;
                if constexpr (X32) {
                    xr = wide_re[tn][tm][r];
                    xi = wide_im[tn][tm][r];
                } else if constexpr (M3) {
                    const double p1 = acc_re[tn][tm][r], p2 = acc_p2[tn][tm][r];
                    xr = p1 - p2;
                    xi = (double)acc_im[tn][tm][r] - (p1 + p2);
                } else {
                    xr = (double)acc_re[tn][tm][r];
                    xi = (double)acc_im[tn][tm][r];
                }
                if (part) {
                    c_t v;
                    v.x = (S)xr;
                    v.y = (S)xi;
                    part[(size_t)sp * M * N + gm + (size_t)gn * M] = v;
                } else {
                    double cr = ar * xr - ai * xi;
                    double ci = ar * xi + ai * xr;
                    c_t* cp = C + gm + (size_t)gn * ldc;
                    if (br != 0.0 || bi != 0.0) {
                        const c_t c = *cp;
                        cr += br * (double)c.x - bi * (double)c.y;
                        ci += br * (double)c.y + bi * (double)c.x;
                    }
                    c_t v;
                    v.x = (S)cr;
                    v.y = (S)ci;
                    *cp = v;
                }
            }
}

// C = alpha * sum_s part[s] + beta * C  (fixed summation order: deterministic)
template <class R>
__global__ void splitk_reduce(int M, int N, int S, const typename Mfma<R>::c_t* __restrict__ part,
                              typename Mfma<R>::c_t* __restrict__ C, int ldc, double ar, double ai,
                              double br, double bi) {
    using c_t = typename Mfma<R>::c_t;
    const size_t total = (size_t)M * N;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        R sr = 0, si = 0;
        for (int s = 0; s < S; ++s) {
            const c_t v = part[(size_t)s * total + idx];
            sr += v.x;
            si += v.y;
        }
        const int m = (int)(idx % M);
        const int n = (int)(idx / M);
        double cr = ar * (double)sr - ai * (double)si;
        double ci = ar * (double)si + ai * (double)sr;
        c_t* cp = C + m + (size_t)n * ldc;
        if (br != 0.0 || bi != 0.0) {
            const c_t c = *cp;
            cr += br * (double)c.x - bi * (double)c.y;
            ci += br * (double)c.y + bi * (double)c.x;
        }
        c_t v;
        v.x = (R)cr;
        v.y = (R)ci;
        *cp = v;
    }
}

template <class R>
__global__ void scale_c(int M, int N, typename Mfma<R>::c_t* __restrict__ C, int ldc, double br, double bi) {
    using c_t = typename Mfma<R>::c_t;
    const size_t total = (size_t)M * N;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const int m = (int)(idx % M), n = (int)(idx / M);
        c_t* cp = C + m + (size_t)n * ldc;
        c_t v;
        if (br == 0.0 && bi == 0.0) {
            v.x = 0;
            v.y = 0;
        } else {
            const c_t c = *cp;
            v.x = (R)(br * c.x - bi * c.y);
            v.y = (R)(br * c.y + bi * c.x);
        }
        *cp = v;
    }
}

int opcode(char t) {
    switch (t) {
        case 'N': case 'n': return 0;
        case 'T': case 't': return 1;
        case 'C': case 'c': return 2;
        default: return -1;
    }
}

template <class R, class S, int TA, int TB, int TM, int TN, int BK, int MAP>
void launch(ProfScope& ps, hipStream_t s, int splits, int M, int N, int K, int kchunk, const S* A, int lda,
            const S* B, int ldb, S* C, int ldc, double ar, double ai, double br, double bi, S* part) {
    using c_t = typename Mfma<S>::c_t;
    const int tm = (M + 32 * TM - 1) / (32 * TM), tn = (N + 32 * TN - 1) / (32 * TN);
    const dim3 grid = MAP == 3   ? dim3((unsigned)((tm * tn * splits + 7) / 8 * 8), 1, 1)
                      : MAP == 2 ? dim3((unsigned)(((tm + 7) / 8) * 8 * tn), 1, splits)
                                 : (MAP == 1 ? dim3(tn, tm, splits) : dim3(tm, tn, splits));
    APW_PLAUNCH(ps, (zgemm_kernel<R, S, TA, TB, TM, TN, BK, MAP>), grid, dim3(NT), 0, s, M, N, K, kchunk,
                       reinterpret_cast<const c_t*>(A), lda, reinterpret_cast<const c_t*>(B), ldb,
                       reinterpret_cast<c_t*>(C), ldc, ar, ai, br, bi, reinterpret_cast<c_t*>(part));
}

// R: MFMA type; S: storage type of A, B, C (S = R, or R = float on S = double, see zgemm_kernel)
template <class R, class S>
void gemm_impl(hipStream_t s, GemmWorkspace& ws, int num_cu, char ta, char tb, int m, int n, int k,
               double alpha_re, double alpha_im, const S* A, int lda, const S* B, int ldb,
               double beta_re, double beta_im, S* C, int ldc, Prof* prof) {
    using c_t = typename Mfma<S>::c_t;
    constexpr bool X32 = !std::is_same<R, S>::value;
    const int oa = opcode(ta), ob = opcode(tb);
    APW_CHECK(oa >= 0 && ob >= 0, "gemm: bad trans '%c' '%c'", ta, tb);
    APW_CHECK(m >= 0 && n >= 0 && k >= 0, "gemm: negative size");
    if (m == 0 || n == 0) return;
    APW_CHECK(ldc >= m, "gemm: ldc < m");
    APW_CHECK(lda >= (oa == 0 ? m : k) || k == 0, "gemm: lda too small");
    APW_CHECK(ldb >= (ob == 0 ? k : n) || k == 0, "gemm: ldb too small");
    if (k == 0 || (alpha_re == 0.0 && alpha_im == 0.0)) {
        const size_t total = (size_t)m * n;
        const int blocks = (int)std::min<size_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(scale_c<S>, dim3(blocks), dim3(256), 0, s, m, n, reinterpret_cast<c_t*>(C), ldc,
                           beta_re, beta_im);
        APW_LAUNCH_CHECK();
        return;
    }
    // Plans: a tile variant and a split rule.  Variants (TM x TN MFMA sub-tiles per wave, map):
    // 0: 64 x 64, m-major; 1: 64 x 32, m-major; 2: 128 x 64, n-major; 3: 128 x 64 and 4: 128 x 32,
    // XCD-aware (each XCD takes whole m-tiles); 5: 64 x 128 with k step 16; 6: 64 x 64 and
    // 7: 64 x 32, XCD-aware split-K (the tiles of one k split share one XCD's L2); 8: 128 x 64
    // and 9: 128 x 32 with whole m-tiles per XCD and a k step of 16 (half the barriers); 10: 64 x 64
    // with a k step of 16 and 11: the same XCD-aware (split-K candidates: each staged column
    // of a k-contiguous operand is 256 bytes instead of 128).
    // occ > 0 splits k so that every workgroup of the launch is resident at once (occ
    // workgroups per CU: floor of the resident slots over the tiles, at least 64 of k per
    // split; a launch with a partial last round of workgroups ran up to 1.5x longer); occ = 0
    // does not split.
    static const int VBM[12] = {64, 64, 128, 128, 128, 64, 64, 64, 128, 128, 64, 64};
    static const int VBN[12] = {64, 32, 64, 64, 32, 128, 64, 32, 64, 32, 64, 64};
    static const int VOCC[12] = {3, 6, 3, 3, 3, 2, 3, 6, 1, 1, 2, 2};   // resident workgroups per CU
    const int cus = std::max(num_cu, 64);
    auto tiles_of = [&](int bm, int bn) { return (long)((m + bm - 1) / bm) * ((n + bn - 1) / bn); };
    auto splits_of = [&](int v, int occ) {
        if (occ <= 0) return 1;
        const long tiles = tiles_of(VBM[v], VBN[v]), slots = (long)occ * cus;
        int sp = tiles < slots ? (int)(slots / tiles) : 1;
        sp = std::max(1, std::min(std::min(sp, std::max(1, k / 64)), 256));
        int kc = (k + sp - 1) / sp;
        kc = (kc + BKMAX - 1) / BKMAX * BKMAX;
        return (k + kc - 1) / kc;
    };
    const bool tall = m >= 4 * n && m >= 256, wide = n >= 4 * m && n >= 256;
    const bool inner = tiles_of(64, 64) < 2 * cus && k / 64 > 1;
    // Fixed rule (measured on the Davidson shapes): inner products (k >> m n) split over k
    // with 64 x 32 tiles when they cover n with fewer columns than 64 x 64 ones (n <= 32,
    // 65..96), else 64 x 64; tall updates 128 x 32 in the XCD-aware order (with three
    // accumulator sets the 128 x 64 tile runs one wave per SIMD and lost: 7.1 vs 6.6 ms on
    // 228487 x 256 x 1024; 64 x 64 made the 216-atom solve 3 % slower); wide 64 x 128; else
    // 64 x 64.
    int variant, occ = 0;
    if (inner) variant = (n <= 32 || (n > 64 && n <= 96)) ? 1 : 0;
    else if (tall) variant = 4;
    else variant = wide ? 5 : 0;
    if (inner) occ = VOCC[variant];
    const char* prec_tag = X32 ? "x32" : (std::is_same<R, double>::value ? "f64" : "f32");
    auto need_of = [&](int v, int oc) {
        const int sp = splits_of(v, oc);
        return sp > 1 ? (size_t)sp * m * n * 2 * sizeof(S) / sizeof(double) + 2 : (size_t)0;
    };
    // runs a plan: C = alpha op(A) op(B) + beta C through the split partials when it splits
    auto run = [&](int v, int oc, S* Cp, int ldcp, double b_re, double b_im, Prof* pf) {
        const int splits = splits_of(v, oc);
        int kchunk = (k + splits - 1) / splits;
        kchunk = (kchunk + BKMAX - 1) / BKMAX * BKMAX;
        S* part = nullptr;
        if (splits > 1) {
            const size_t need = need_of(v, oc);
            if (ws.ext) {
                APW_CHECK(need <= ws.ext_n,
                          "gemm: %c%c m=%d n=%d k=%d splits over k into %d partial products of %zu bytes, more than "
                          "the %zu-byte fixed workspace", ta, tb, m, n, k, splits, need * sizeof(double),
                          ws.ext_n * sizeof(double));
                part = reinterpret_cast<S*>(ws.ext);
            } else {
                ws.part.reserve(need);
                part = reinterpret_cast<S*>(ws.part.get());
            }
        }
        const double esz = 2.0 * sizeof(S);
        const bool beta_nz = (b_re != 0.0 || b_im != 0.0);
        const char* pname = X32 ? (splits > 1 ? "zgemm_x32_splitk" : "zgemm_x32_tiled")
                                : std::is_same<R, double>::value ? (splits > 1 ? "zgemm_f64_splitk" : "zgemm_f64_tiled")
                                                                 : (splits > 1 ? "cgemm_f32_splitk" : "cgemm_f32_tiled");
        {
            ProfScope ps(pf, s, pname,
                         esz * ((double)m * k + (double)k * n + (double)m * n * (splits > 1 ? splits : (beta_nz ? 2 : 1))),
                         8.0 * m * n * (double)k, true);
#define APW_GEMM_CASE(X, Y)                                                                                          \
    if (oa == X && ob == Y) {                                                                                       \
        switch (v) {                                                                                                \
            case 1: launch<R, S, X, Y, 2, 1, 8, 0>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                               alpha_im, b_re, b_im, part); break;                                  \
            case 2: launch<R, S, X, Y, 4, 2, 8, 1>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                               alpha_im, b_re, b_im, part); break;                                  \
            case 3: launch<R, S, X, Y, 4, 2, 8, 2>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                               alpha_im, b_re, b_im, part); break;                                  \
            case 4: launch<R, S, X, Y, 4, 1, 8, 2>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                               alpha_im, b_re, b_im, part); break;                                  \
            case 5: launch<R, S, X, Y, 2, 4, 16, 0>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                                alpha_im, b_re, b_im, part); break;                                 \
            case 6: launch<R, S, X, Y, 2, 2, 8, 3>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                               alpha_im, b_re, b_im, part); break;                                  \
            case 7: launch<R, S, X, Y, 2, 1, 8, 3>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                               alpha_im, b_re, b_im, part); break;                                  \
            case 8: launch<R, S, X, Y, 4, 2, 16, 2>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                                alpha_im, b_re, b_im, part); break;                                 \
            case 9: launch<R, S, X, Y, 4, 1, 16, 2>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                                alpha_im, b_re, b_im, part); break;                                 \
            case 10: launch<R, S, X, Y, 2, 2, 16, 0>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                                 alpha_im, b_re, b_im, part); break;                                \
            case 11: launch<R, S, X, Y, 2, 2, 16, 3>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                                 alpha_im, b_re, b_im, part); break;                                \
            default: launch<R, S, X, Y, 2, 2, 8, 0>(ps, s, splits, m, n, k, kchunk, A, lda, B, ldb, Cp, ldcp, alpha_re, \
                                                alpha_im, b_re, b_im, part); break;                                 \
        }                                                                                                           \
    }
            APW_GEMM_CASE(0, 0) APW_GEMM_CASE(0, 1) APW_GEMM_CASE(0, 2)
            APW_GEMM_CASE(1, 0) APW_GEMM_CASE(1, 1) APW_GEMM_CASE(1, 2)
            APW_GEMM_CASE(2, 0) APW_GEMM_CASE(2, 1) APW_GEMM_CASE(2, 2)
#undef APW_GEMM_CASE
            APW_LAUNCH_CHECK();
        }
        if (splits > 1) {
            ProfScope pr(pf, s, "gemm_splitk_reduce", esz * (double)m * n * (splits + (beta_nz ? 2 : 1)),
                         2.0 * m * n * (double)splits, true);
            const size_t total = (size_t)m * n;
            const int blocks = (int)std::min<size_t>((total + 255) / 256, 65536);
            APW_PLAUNCH(pr, splitk_reduce<S>, dim3(blocks), dim3(256), 0, s, m, n, splits,
                        reinterpret_cast<const c_t*>(part), reinterpret_cast<c_t*>(Cp), ldcp, alpha_re, alpha_im,
                        b_re, b_im);
            APW_LAUNCH_CHECK();
        }
    };
    // autotuned plan (the fixed-workspace entry point keeps the fixed rule, whose partials
    // it is sized for)
    const int tmode = ws.ext ? 0 : tune_mode();
    if (tmode > 0) {
        auto bucket = [](int x) { return x <= 2048 ? (x + 31) / 32 * 32 : (x + 1023) / 1024 * 1024; };
        char key[160];
        std::snprintf(key, sizeof(key), "gemm %s %c%c m=%d n=%d k=%d cu=%d", prec_tag, "NTC"[oa], "NTC"[ob],
                      bucket(m), bucket(n), bucket(k), cus);
        int a = 0, b = 0;
        if (tune_find(key, a, b) && a >= 0 && a < 12 && b >= 0 && b <= 6 && !(X32 && (a == 2 || a == 3 || a == 8))) {
            variant = a;
            occ = b;
        } else if (tmode == 2) {
            // candidates: every tile/map of the shape's kind, and for inner products every
            // split rule; timed into a scratch C (beta 0) so the caller's C is untouched
            std::vector<std::pair<int, int>> cand;
            if (inner) {
                for (int v : {0, 1, 6, 7, 10, 11})
                    for (int oc : {1, 2, 3, 4, 6}) cand.emplace_back(v, oc);
            } else if (tall) {
                for (int v : {3, 4, 0, 2, 8, 9}) cand.emplace_back(v, 0);
            } else {
                for (int v : {0, 1, 5}) cand.emplace_back(v, 0);
            }
            ws.tune_c.reserve((size_t)m * n * 2 * sizeof(S) / sizeof(double) + 2);
            S* Ct = reinterpret_cast<S*>(ws.tune_c.get());
            double best = 1e300;
            std::vector<std::pair<int, int>> seen;   // (variant, splits) already timed
            for (auto& c : cand) {
                if (X32 && (c.first == 2 || c.first == 3 || c.first == 8)) continue;
                const std::pair<int, int> vs(c.first, splits_of(c.first, c.second));
                if (std::find(seen.begin(), seen.end(), vs) != seen.end()) continue;
                seen.push_back(vs);
                const double us = time_on_stream(s, 2, [&] { run(c.first, c.second, Ct, m, 0.0, 0.0, nullptr); });
                if (us < best) {
                    best = us;
                    variant = c.first;
                    occ = c.second;
                }
            }
            tune_store(key, variant, occ, best);
        }
    }
    static const bool trace = std::getenv("APW_GEMM_TRACE") != nullptr;   // debugging aid: one line per call
    if (trace)
        std::fprintf(stderr, "apw_gemm %s %c%c m=%d n=%d k=%d variant=%d splits=%d\n", prec_tag, ta, tb, m, n, k,
                     variant, splits_of(variant, occ));
    run(variant, occ, C, ldc, beta_re, beta_im, prof);
}

}  // namespace

void gemm_reserve(GemmWorkspace& ws, int num_cu) {
    // splits x tiles <= resident slots, so splits m n <= slots x BM x BN; the largest product
    // of the split variants is 3 x 128 x 64 (c128: two doubles an element)
    ws.part.reserve((size_t)3 * std::max(num_cu, 64) * 128 * 64 * 2 + 2);
}

template <class R>
void gemm(hipStream_t s, GemmWorkspace& ws, int num_cu, char ta, char tb, int m, int n, int k,
          double alpha_re, double alpha_im, const R* A, int lda, const R* B, int ldb,
          double beta_re, double beta_im, R* C, int ldc, Prof* prof) {
    gemm_impl<R, R>(s, ws, num_cu, ta, tb, m, n, k, alpha_re, alpha_im, A, lda, B, ldb, beta_re, beta_im, C, ldc,
                    prof);
}

void gemm_x32(hipStream_t s, GemmWorkspace& ws, int num_cu, char ta, char tb, int m, int n, int k,
              double alpha_re, double alpha_im, const double* A, int lda, const double* B, int ldb,
              double beta_re, double beta_im, double* C, int ldc, Prof* prof) {
    gemm_impl<float, double>(s, ws, num_cu, ta, tb, m, n, k, alpha_re, alpha_im, A, lda, B, ldb, beta_re, beta_im,
                             C, ldc, prof);
}

template void gemm<double>(hipStream_t, GemmWorkspace&, int, char, char, int, int, int, double, double,
                           const double*, int, const double*, int, double, double, double*, int, Prof*);
template void gemm<float>(hipStream_t, GemmWorkspace&, int, char, char, int, int, int, double, double,
                          const float*, int, const float*, int, double, double, float*, int, Prof*);

// ---------------------------------------------------------------- layout probe
// Records, for every (lane, accumulator register) of one v_mfma_f64_16x16x4f64 and one
// v_mfma_f32_16x16x4f32, the (row, column) of D it holds.  Used by the unit tests to
// check the operand/accumulator layout the GEMM relies on.
__global__ void probe_mfma_kernel(double* out) {
    const int l = threadIdx.x;
    const double row_sel = (l >> 4) == 0 ? 1.0 : 0.0;
    d4_t z{0, 0, 0, 0};
    // D[i][j] = A[i][0]: A = row index, B[0][j] = 1
    d4_t r = __builtin_amdgcn_mfma_f64_16x16x4f64((double)(l & 15), row_sel, z, 0, 0, 0);
    // D[i][j] = B[0][j]: A[i][0] = 1, B = column index
    d4_t c = __builtin_amdgcn_mfma_f64_16x16x4f64(row_sel, (double)(l & 15), z, 0, 0, 0);
    f4_t zf{0, 0, 0, 0};
    f4_t rf = __builtin_amdgcn_mfma_f32_16x16x4f32((float)(l & 15), (float)row_sel, zf, 0, 0, 0);
    f4_t cf = __builtin_amdgcn_mfma_f32_16x16x4f32((float)row_sel, (float)(l & 15), zf, 0, 0, 0);
    for (int q = 0; q < 4; ++q) {
        out[(0 * 64 + l) * 4 + q] = r[q];
        out[(1 * 64 + l) * 4 + q] = c[q];
        out[(2 * 64 + l) * 4 + q] = rf[q];
        out[(3 * 64 + l) * 4 + q] = cf[q];
    }
}

// ---------------------------------------------------------------- fixed split-K workspace
// The split-K partials of apw_zgemm_noalloc live in this file-scope device array instead of
// memory the library allocates.  The default plan keeps every launch resident at once, so
// splits x m x n <= (resident workgroups per CU) x CUs x tile = 3 x 256 x 64 x 64 complex
// values for every shape on a 256-CU GPU: 48 MiB of c128 covers them all.  Each GPU holds
// its own copy from the time the library is loaded until the process ends.
constexpr size_t SCRATCH_DOUBLES = (size_t)3 * 256 * 64 * 64 * 2;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

// The device address of this GPU's copy of scratch; `scratch` itself names only a host-side
// stand-in.  Looked up once per device, with that device current (apw_ctx_create).
double* gemm_scratch_init(int device) {
    APW_CHECK(device >= 0 && device < MAX_DEVICES, "gemm scratch: device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        APW_HIP(hipGetDevice(&cur));
        APW_CHECK(cur == device, "gemm scratch: device %d is not current (%d)", device, cur);
        void* ptr = nullptr;
        APW_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        APW_CHECK(ptr != nullptr, "gemm scratch: no device address on device %d", device);
        g_scratch_ptr[device] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[device];
}

}  // namespace apw

extern "C" int apw_zgemm_noalloc(apw_ctx* ctx, char transa, char transb, int m, int n, int k, const double* alpha,
                                 const double* a, int lda, const double* b, int ldb, const double* beta, double* c,
                                 int ldc) {
    return apw::guarded([&] {
        APW_CHECK(ctx && alpha && beta && c, "apw_zgemm_noalloc: null argument");
        APW_CHECK(k == 0 || (a && b), "apw_zgemm_noalloc: null A/B");
        apw::DeviceGuard g(ctx->device);
        apw::GemmWorkspace ws;   // owns nothing: the partials go to the fixed array
        ws.ext = apw::gemm_scratch_init(ctx->device);   // looked up when the context was created
        ws.ext_n = apw::SCRATCH_DOUBLES;
        apw::gemm<double>(ctx->stream, ws, ctx->num_cu, transa, transb, m, n, k, alpha[0], alpha[1], a, lda, b, ldb,
                          beta[0], beta[1], c, ldc, nullptr);
        return (int)APW_OK;
    });
}

extern "C" int apw_probe_mfma_layout(apw_ctx* ctx, double* out_dev) {
    return apw::guarded([&] {
        APW_CHECK(ctx && out_dev, "apw_probe_mfma_layout: null argument");
        apw::DeviceGuard g(ctx->device);
        hipLaunchKernelGGL(apw::probe_mfma_kernel, dim3(1), dim3(64), 0, ctx->stream, out_dev);
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}
