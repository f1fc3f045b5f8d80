// FP64 Hermitian subspace eigensolver on the GPU (lowest nband eigenpairs).
//
// Replaces dnevx_op<complex<double>, DEVICE_GPU> (dngvd_op.h:85-107), which on ROCm copies
// hcc to the host and calls LAPACK zheevx every Davidson iteration
// (rocm/dngvd_op.hip.cu:146-170).  Same algorithm family as zheevx, all on the device:
//   1. Householder reduction of the lower triangle to a real symmetric tridiagonal T
//      (zhetrd 'L' form, unblocked) in one kernel: the trailing matrix is partitioned by
//      columns over workgroups that are all resident at once, one exchange through a ring
//      per column.  n <= 704: columns in LDS, vectors in registers; up to 2048: columns in
//      registers, rows split over the waves of a workgroup; global memory only where neither
//      fits the GPU (or in one workgroup, for many small concurrent solves);
//   2. bisection with Sturm counts for eigenvalues 0..nband-1 of T (dstebz form);
//   3. inverse iteration with partial-pivoting tridiagonal LU and Gram-Schmidt inside
//      clusters of close eigenvalues (dstein form);
//   4. back-transformation z -> Q z with the stored reflectors (zunmtr form): one wave per
//      vector for n <= 512, else blocks of 64 reflectors in compact WY form (zlarft), two
//      matrix-core products per block.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <deque>
#include <mutex>
#include <vector>

#include "apw_internal.hpp"

namespace apw {
namespace {

constexpr int NT = 256;
constexpr unsigned SPIN_LIMIT = 1u << 24;
constexpr size_t TRIDIAG_LDS_BYTES = 160 * 1024 - 512;   // LDS per workgroup on gfx950, less static use

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cconjmul(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Cross-lane moves through DPP (VALU latency, a few cycles) rather than LDS-routed
// permutes (~100 cycles each): the sequential chains of the reduction and the
// back-transform are made of these.
template <int CTRL, int ROWMASK>
__device__ inline double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// value of lane `lane` (wave-uniform) in every lane
__device__ inline double lane_bcast(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// sum over the 64 lanes, in every lane (fixed order: deterministic).  quad_perm [1,0,3,2]
// and [2,3,0,1], half-row and row mirrors leave every lane of a 16-lane row with the row
// sum; row_bcast15 (rows 1, 3) and row_bcast31 (rows 2, 3) carry it to lane 63.  Call from
// wave-uniform control flow only.
__device__ inline double wave_sum(double v) {
    v += dpp_d<0xB1, 0xF>(v);
    v += dpp_d<0x4E, 0xF>(v);
    v += dpp_d<0x141, 0xF>(v);
    v += dpp_d<0x140, 0xF>(v);
    v += dpp_d<0x142, 0xA>(v);
    v += dpp_d<0x143, 0xC>(v);
    return lane_bcast(v, 63);
}

// block-wide sum of two doubles (deterministic tree); red: LDS scratch of >= 2*NT/64 doubles
__device__ inline double2 block_sum2(double a, double b, double* red) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int nw = blockDim.x >> 6;
    __syncthreads();
    if (l == 0) {
        red[w] = a;
        red[nw + w] = b;
    }
    __syncthreads();
    double sa = 0, sb = 0;
    for (int i = 0; i < nw; ++i) {
        sa += red[i];
        sb += red[nw + i];
    }
    __syncthreads();
    return make_double2(sa, sb);
}

__global__ void k_copy_herm(int n, const double2* __restrict__ h, int ldh, double2* __restrict__ a) {
    const size_t total = (size_t)n * n;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % n), j = (int)(t / n);
        double2 v;
        if (i > j) {
            v = h[i + (size_t)j * ldh];
        } else if (i < j) {
            const double2 u = h[j + (size_t)i * ldh];
            v = make_double2(u.x, -u.y);
        } else {
            v = make_double2(h[i + (size_t)i * ldh].x, 0.0);
        }
        a[t] = v;
    }
}

__global__ void k_zero_u32(unsigned* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// Reflector (zlarfg convention) of x[0..mx): P^H x = beta e_1, P = I - tau v v^H, v[0] = 1.
// x is read from xs (any memory), v written to vo (LDS, may alias xs shifted); every thread
// of the block calls it and gets tau and beta.
__device__ void make_reflector(const double2* xs, int mx, double2* vo, double2& tau, double& beta, double* red) {
    double s2 = 0;
    for (int i = 1 + threadIdx.x; i < mx; i += NT) {
        const double2 x = xs[i];
        s2 += x.x * x.x + x.y * x.y;
    }
    const double xn2 = block_sum2(s2, 0.0, red).x;
    const double2 alpha = xs[0];
    double2 sc;
    if (xn2 == 0.0 && alpha.y == 0.0) {
        tau = make_double2(0.0, 0.0);
        beta = alpha.x;
        sc = make_double2(0.0, 0.0);
    } else {
        const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
        beta = alpha.x >= 0 ? -nrm : nrm;
        tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
        const double2 den = make_double2(alpha.x - beta, alpha.y);   // sc = 1 / (alpha - beta)
        const double dd = den.x * den.x + den.y * den.y;
        sc = make_double2(den.x / dd, -den.y / dd);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < mx; i += NT) vo[i] = (i == 0) ? make_double2(1.0, 0.0) : cmul(xs[i], sc);
    __syncthreads();
}

// Cross-workgroup exchange of the reduction: a ring with one slot per column step,
// pre-filled with a NaN bit pattern no arithmetic produces.  Producers store each value with
// device-scope stores (they bypass the non-coherent per-XCD caches); consumers poll the
// slots they need with device-scope loads until no word holds the pattern.  Every value
// validates itself, so there is no flag, no fence and no grid barrier: one memory round trip
// per column.  The check assumes a 16-byte load returns each 8-byte half whole (one
// request); a torn half could only pass as a wrong value, which k_bisect's finiteness test
// catches when it is not finite.
constexpr unsigned long long RING_EMPTY = 0x7ff4dead5eed0001ull;

__global__ void k_fill_ring(double2* ring, size_t count) {
    const double e = __longlong_as_double((long long)RING_EMPTY);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        ring[i] = make_double2(e, e);
}

// 16-byte forms: one vector buffer load / store per value with the device-scope cache
// policy (sc1, the bits of the 8-byte device-scope atomics).  Each 8-byte half
// still validates itself, so a load that sees one half written and the other not simply
// polls again.  r spans one step's slot (2n values); i is the value's index in it.
typedef unsigned ring_u4 __attribute__((ext_vector_type(4)));
constexpr int RING_SC1 = 16;
__device__ inline __amdgpu_buffer_rsrc_t ring_rsrc(const double2* slot, int count) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double2*>(slot), 0, count * (int)sizeof(double2), 0x00020000);
}
__device__ inline void ring_put16(__amdgpu_buffer_rsrc_t r, int i, double2 v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(ring_u4, v), r, (unsigned)i * 16u, 0, RING_SC1);
}
__device__ inline bool ring_get16(__amdgpu_buffer_rsrc_t r, int i, double2& v) {
    v = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (unsigned)i * 16u, 0, RING_SC1));
    return (unsigned long long)__double_as_longlong(v.x) != RING_EMPTY &&
           (unsigned long long)__double_as_longlong(v.y) != RING_EMPTY;
}

// Householder tridiagonalisation (lower, zhetrd 'L' steps).  Column c belongs to workgroup
// c % nwg for the whole reduction, in its LDS (GCOL = false) or in A itself (GCOL = true: the
// copy of hcc, updated in place).  Per column step k, what crosses workgroups is
// p_k = A_k^H v_k on each workgroup's columns and column k+1 of A_k; both go through the
// ring (slot k * 2n + i holds p_i, k * 2n + n + i the column).  Every workgroup applies the
// rank-2 correction of step k to that column and forms the next reflector v_{k+1} itself
// (redundantly) before it touches its own columns, so one pass over each own column both
// applies the rank-2 update of step k and accumulates p_{k+1} = A_{k+1}^H v_{k+1}: a trailing
// entry is read and written once per step.  The owner of column k+2 publishes it updated.
// abort_flag: raised when a poll exceeds its spin bound.
// LDS: [ncol x n columns], v_k and v_{k+1}, y, the gathered column (n each), scratch.
template <bool GCOL>
__global__ __launch_bounds__(NT) void k_tridiag_fused(double2* __restrict__ A, int n, int ncol,
                                                      double2* __restrict__ V, double2* __restrict__ tau,
                                                      double* __restrict__ d, double* __restrict__ e,
                                                      double2* __restrict__ ring, unsigned* abort_flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* Al = reinterpret_cast<double2*>(smem);
    double2* vb = Al + (GCOL ? 0 : (size_t)ncol * n);
    double2* y = vb + 2 * n;
    double2* cn = y + n;
    double* red = reinterpret_cast<double*>(cn + n);
    const int nwg = (int)gridDim.x, bid = (int)blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int NW = NT / 64;
    auto colp = [&](int lc) -> double2* {
        return GCOL ? A + (size_t)(bid + lc * nwg) * n : Al + (size_t)lc * n;
    };
    // first own column slot with c >= x
    auto first_slot = [&](int x) { return x > bid ? (x - bid + nwg - 1) / nwg : 0; };
    if (!GCOL) {
        for (int lc = 0; lc < ncol; ++lc) {
            const int c = bid + lc * nwg;
            if (c >= n) break;
            for (int i = threadIdx.x; i < n; i += NT) Al[(size_t)lc * n + i] = A[(size_t)c * n + i];
        }
        __syncthreads();
    }
    double2* v = vb;        // v_k: rows k+1 .. n-1 (v[0] = 1 at row k+1)
    double2* vn = vb + n;   // v_{k+1}
    double2 t;
    double beta;
    make_reflector(A + 1, n - 1, v, t, beta, red);
    if (bid == 0) {
        for (int i = threadIdx.x; i < n - 1; i += NT) V[1 + i] = v[i];
        if (threadIdx.x == 0) {
            tau[0] = t;
            d[0] = A[0].x;
            e[0] = beta;
        }
    }
    // p_0 = A^H v_0 on own columns c >= 1, and column 1 (as loaded), into slot 0
    {
        const __amdgpu_buffer_rsrc_t rs = ring_rsrc(ring, 2 * n);
        const int m = n - 1;
        for (int lc = first_slot(1) + wave; lc < ncol; lc += NW) {
            const int c = bid + lc * nwg;
            if (c >= n) break;
            const double2* col = colp(lc) + 1;
            double pr = 0, pi = 0;
            for (int i = lane; i < m; i += 64) {
                const double2 q = cconjmul(col[i], v[i]);
                pr += q.x;
                pi += q.y;
            }
            pr = wave_sum(pr);
            pi = wave_sum(pi);
            if (lane == 0) ring_put16(rs, c - 1, make_double2(pr, pi));
            if (c == 1)
                for (int i = lane; i < m; i += 64) ring_put16(rs, n + i, col[i]);
        }
    }
    for (int k = 0; k < n - 1; ++k) {
        const int m = n - k - 1;   // trailing rows/cols k+1 .. n-1
        // ---- gather p_k -> y and column k+1 of A_k -> cn (each wave polls on its own,
        // skipping the values it already has; one barrier publishes them) ----
        {
            const __amdgpu_buffer_rsrc_t rs = ring_rsrc(ring + (size_t)k * 2 * n, 2 * n);
            unsigned spins = 0, got = 0;
            while (true) {
                bool mine = true;
                int q = 0;
                for (int i = threadIdx.x; i < 2 * m; i += NT, ++q) {
                    if ((got >> q) & 1u) continue;
                    const bool isp = i < m;
                    const int ii = isp ? i : i - m;
                    double2 val;
                    if (ring_get16(rs, isp ? ii : n + ii, val)) {
                        (isp ? y : cn)[ii] = val;
                        got |= 1u << q;
                    } else {
                        mine = false;
                    }
                }
                if (__all(mine)) break;
                if ((++spins & 31u) == 0) {
                    int st = 0;
                    if (lane == 0) {
                        st = __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
                        if (!st && spins > SPIN_LIMIT) {
                            atomicExch(abort_flag, 1u);
                            st = 1;
                        }
                    }
                    if (__shfl(st, 0, 64)) break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __syncthreads();
        }
        // ---- y = tau p - (|tau|^2 mu / 2) v,  mu = v^H p ----
        double mr = 0;
        for (int i = threadIdx.x; i < m; i += NT) mr += cconjmul(v[i], y[i]).x;
        const double mu = block_sum2(mr, 0.0, red).x;
        const double c2 = 0.5 * (t.x * t.x + t.y * t.y) * mu;
        for (int i = threadIdx.x; i < m; i += NT) {
            const double2 tp = cmul(t, y[i]);
            y[i] = make_double2(tp.x - c2 * v[i].x, tp.y - c2 * v[i].y);
        }
        __syncthreads();
        // ---- column k+1 of A_{k+1} ----
        {
            const double2 v0 = v[0], y0 = y[0];
            for (int i = threadIdx.x; i < m; i += NT) {
                const double2 a1 = cconjmul(v0, y[i]);
                const double2 a2 = cconjmul(y0, v[i]);
                cn[i].x -= a1.x + a2.x;
                cn[i].y -= a1.y + a2.y;
            }
        }
        __syncthreads();
        const double dk1 = cn[0].x;
        if (m >= 2) {
            // ---- v_{k+1} from rows k+2 .. n-1 of that column (every workgroup) ----
            double2 tn;
            double bn;
            make_reflector(cn + 1, m - 1, vn, tn, bn, red);
            if (bid == 0) {
                for (int i = threadIdx.x; i < m - 1; i += NT) V[(size_t)(k + 1) * n + (k + 2) + i] = vn[i];
                if (threadIdx.x == 0) {
                    tau[k + 1] = tn;
                    d[k + 1] = dk1;
                    e[k + 1] = bn;
                }
            }
            // ---- rank-2 update of own columns c >= k+2 (rows k+2 ..) and p_{k+1} ----
            const __amdgpu_buffer_rsrc_t rs1 = ring_rsrc(ring + (size_t)(k + 1) * 2 * n, 2 * n);
            for (int lc = first_slot(k + 2) + wave; lc < ncol; lc += NW) {
                const int c = bid + lc * nwg;
                if (c >= n) break;
                double2* col = colp(lc) + (k + 1);   // relative row i = row k+1+i
                const int jj = c - (k + 1);
                const double2 vj = v[jj], yj = y[jj];
                double pr = 0, pi = 0;
                for (int i = 1 + lane; i < m; i += 64) {
                    const double2 a1 = cconjmul(vj, y[i]);
                    const double2 a2 = cconjmul(yj, v[i]);
                    double2 a = col[i];
                    a.x -= a1.x + a2.x;
                    a.y -= a1.y + a2.y;
                    col[i] = a;
                    const double2 q = cconjmul(a, vn[i - 1]);
                    pr += q.x;
                    pi += q.y;
                }
                pr = wave_sum(pr);
                pi = wave_sum(pi);
                if (lane == 0) ring_put16(rs1, c - (k + 2), make_double2(pr, pi));
                if (c == k + 2)   // owner: column k+2 of A_{k+1}, rows k+2 .. n-1 (its own writes)
                    for (int i = 1 + lane; i < m; i += 64) ring_put16(rs1, n + (i - 1), col[i]);
            }
            t = tn;
            double2* sw = v;
            v = vn;
            vn = sw;
        } else if (bid == 0 && threadIdx.x == 0) {
            d[k + 1] = dk1;
        }
        __syncthreads();
    }
}

// Sum over each 16-lane row, in every lane of the row (quad_perm [1,0,3,2], [2,3,0,1], half-row
// and row mirrors: every source lane is valid, so the moves need no zeroed old value).
template <int CTRL>
__device__ inline double dppm_d(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ inline double row_sum16(double v) {
    v += dppm_d<0xB1>(v);
    v += dppm_d<0x4E>(v);
    v += dppm_d<0x141>(v);
    v += dppm_d<0x140>(v);
    return v;
}

// sum over the 64 lanes, in every lane: the row tree above, then the four row sums read out
// (fixed order: deterministic)
__device__ inline double wave_sum_rows(double v) {
    v = row_sum16(v);
    return (lane_bcast(v, 0) + lane_bcast(v, 16)) + (lane_bcast(v, 32) + lane_bcast(v, 48));
}

// Value of a per-lane register array at absolute row i (lane i % 64, slot i / 64),
// broadcast to the whole wave.
template <int R>
__device__ inline double2 row_bcast(const double2 (&x)[R], int i) {
    double2 t = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (r == (i >> 6)) t = x[r];
    return make_double2(lane_bcast(t.x, i & 63), lane_bcast(t.y, i & 63));
}
// Householder tridiagonalisation (lower), LDS-resident columns, vectors in registers.
// Same ownership and exchange as k_tridiag_fused; what changes is where the per-column work
// happens.  Every wave holds the full vectors v_k, y_k and the next reflector v_{k+1} in
// registers (lane l: rows l + 64 r, r < R), so y, the rank-2-corrected column and the next
// reflector are formed with wave reductions only, and one pass over each own column both
// applies the rank-2 update of step k and accumulates p_{k+1} = A_{k+1}^H v_{k+1} (the
// column is read and written once per step).  Workgroup barriers: the gather's only.
template <int R, int NTR>
__global__ __launch_bounds__(NTR) void k_tridiag_reg(const double2* __restrict__ A, int n, int ncol,
                                                    double2* __restrict__ V, double2* __restrict__ tau,
                                                    double* __restrict__ d, double* __restrict__ e,
                                                    double2* __restrict__ ring, unsigned* abort_flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* Al = reinterpret_cast<double2*>(smem);   // Al[lc * n + i] = A[i][c], c = blockIdx.x + lc * nwg
    double2* gs = Al + (size_t)ncol * n;                // gathered p and column, two buffers of 2n
    const int nwg = (int)gridDim.x, bid = (int)blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int NW = NTR / 64;
    for (int lc = 0; lc < ncol; ++lc) {
        const int c = bid + lc * nwg;
        if (c >= n) break;
        for (int i = threadIdx.x; i < n; i += NTR) Al[(size_t)lc * n + i] = A[(size_t)c * n + i];
    }
    __syncthreads();
    double2 v[R], y[R], w[R];   // v_k, y_k, and the corrected column / v_{k+1}
    // reflector of w on rows > k0 (alpha at row k0 + 1, the norm over rows > k0 + 1):
    // overwrites w with it (1 at row k0 + 1, 0 at rows <= k0), returns tau and beta
    auto reflector = [&](int k0, double2& tv, double& beta) {
        double s2 = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane + 64 * r;
            if (i > k0 + 1 && i < n) s2 += w[r].x * w[r].x + w[r].y * w[r].y;
        }
        const double xn2 = wave_sum(s2);
        const double2 alpha = row_bcast<R>(w, k0 + 1);
        double2 sc;
        if (xn2 == 0.0 && alpha.y == 0.0) {
            tv = make_double2(0.0, 0.0);
            beta = alpha.x;
            sc = make_double2(0.0, 0.0);
        } else {
            const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
            beta = alpha.x >= 0 ? -nrm : nrm;
            tv = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
            const double2 den = make_double2(alpha.x - beta, alpha.y);
            const double dd = den.x * den.x + den.y * den.y;
            sc = make_double2(den.x / dd, -den.y / dd);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane + 64 * r;
            w[r] = (i == k0 + 1) ? make_double2(1.0, 0.0) : ((i > k0 + 1 && i < n) ? cmul(w[r], sc) : make_double2(0.0, 0.0));
        }
    };
    // p_{k+1}[c] = sum_{i > k+1} conj(A[i][c]) w_i over own columns c > k + 1 (k = -1 for
    // the first reflector), optionally after the rank-2 update with (v, y) of step k;
    // published into the ring slot of step k + 1 with the owner's copy of column k + 2.
    auto update_and_project = [&](int k, bool update) {
        const __amdgpu_buffer_rsrc_t rs = ring_rsrc(ring + (size_t)(k + 1) * 2 * n, 2 * n);
        const int lc1 = (k + 2 - bid + nwg - 1) / nwg;
        for (int lc = lc1 + wave; lc < ncol; lc += NW) {
            const int c = bid + lc * nwg;
            if (c >= n) break;
            double2* col = Al + (size_t)lc * n;
            double2 yc = make_double2(0.0, 0.0), vc = make_double2(0.0, 0.0);
            if (update) {
                yc = row_bcast<R>(y, c);
                vc = row_bcast<R>(v, c);
            }
            double pr = 0, pi = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = lane + 64 * r;
                if (i > k + 1 && i < n) {
                    // a -= v_i conj(y_c) + y_i conj(v_c);  p += conj(a) w_i  (as FMA chains)
                    double2 a = col[i];
                    if (update) {
                        a.x = fma(-yc.x, v[r].x, a.x);
                        a.x = fma(-yc.y, v[r].y, a.x);
                        a.x = fma(-vc.x, y[r].x, a.x);
                        a.x = fma(-vc.y, y[r].y, a.x);
                        a.y = fma(-yc.x, v[r].y, a.y);
                        a.y = fma(yc.y, v[r].x, a.y);
                        a.y = fma(-vc.x, y[r].y, a.y);
                        a.y = fma(vc.y, y[r].x, a.y);
                        col[i] = a;
                    }
                    pr = fma(a.x, w[r].x, pr);
                    pr = fma(a.y, w[r].y, pr);
                    pi = fma(a.x, w[r].y, pi);
                    pi = fma(-a.y, w[r].x, pi);
                }
            }
            pr = wave_sum_rows(pr);
            pi = wave_sum_rows(pi);
            if (k + 2 < n && lane == 0) ring_put16(rs, c - (k + 2), make_double2(pr, pi));
            if (c == k + 2) {
                // owner: publish the updated column k + 2 (rows k+2 .. n-1)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int i = lane + 64 * r;
                    if (i >= k + 2 && i < n) ring_put16(rs, n + (i - (k + 2)), col[i]);
                }
            }
        }
    };
    // first reflector from column 0 (rows 1 .. n-1)
    double2 t;
    double beta;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = lane + 64 * r;
        w[r] = (i >= 1 && i < n) ? A[i] : make_double2(0.0, 0.0);
    }
    reflector(0, t, beta);
    if (bid == 0 && wave == 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane + 64 * r;
            if (i >= 1 && i < n) V[i] = w[r];
        }
        if (lane == 0) {
            tau[0] = t;
            d[0] = A[0].x;
            e[0] = beta;
        }
    }
    // p_0 and column 1 (its owner publishes it as loaded)
    update_and_project(-1, false);
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = w[r];
    for (int k = 0; k < n - 1; ++k) {
        const int m = n - k - 1;
        const __amdgpu_buffer_rsrc_t rs = ring_rsrc(ring + (size_t)k * 2 * n, 2 * n);
        double2* g = gs + (size_t)(k & 1) * 2 * n;
        // ---- gather p_k and column k+1 ----
        // (a thread's q-th value, once valid, is not polled again: bit q of got; at most
        // ceil(2n / NTR) <= 8 values a thread for the sizes this form takes).  Each wave
        // polls on its own until its values are in (no workgroup barrier per poll round);
        // one barrier then publishes the gathered values to the whole workgroup.
        {
            unsigned spins = 0, got = 0;
            while (true) {
                bool mine = true;
                int q = 0;
                for (int i = threadIdx.x; i < 2 * m; i += NTR, ++q) {
                    if ((got >> q) & 1u) continue;
                    const int ii = i < m ? i : n + (i - m);
                    double2 val;
                    if (ring_get16(rs, ii, val)) {
                        g[ii] = val;
                        got |= 1u << q;
                    } else {
                        mine = false;
                    }
                }
                if (__all(mine)) break;
                // the abort word and the spin bound every 32 failed polls of this wave: its
                // load is a memory round trip that would otherwise lengthen every wait.  Every
                // wave leaves: its values arrive, it sees the abort word, or it raises it.
                if ((++spins & 31u) == 0) {
                    int st = 0;
                    if (lane == 0) {
                        st = __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
                        if (!st && spins > SPIN_LIMIT) {
                            atomicExch(abort_flag, 1u);
                            st = 1;
                        }
                    }
                    if (__shfl(st, 0, 64)) break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __syncthreads();
        }
        // ---- y = tau p - (|tau|^2 mu / 2) v ----
        double mr = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = lane + 64 * r;
            const bool in = i > k && i < n;
            y[r] = in ? g[i - (k + 1)] : make_double2(0.0, 0.0);
            w[r] = in ? g[n + (i - (k + 1))] : make_double2(0.0, 0.0);
            mr += cconjmul(v[r], y[r]).x;
        }
        const double mu = wave_sum(mr);
        const double c2 = 0.5 * (t.x * t.x + t.y * t.y) * mu;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double2 tp = cmul(t, y[r]);
            y[r] = make_double2(tp.x - c2 * v[r].x, tp.y - c2 * v[r].y);
        }
        // ---- column k+1 after the rank-2 update, then the next reflector ----
        {
            const double2 v0 = row_bcast<R>(v, k + 1), y0 = row_bcast<R>(y, k + 1);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double2 a1 = cconjmul(v0, y[r]);
                const double2 a2 = cconjmul(y0, v[r]);
                w[r].x -= a1.x + a2.x;
                w[r].y -= a1.y + a2.y;
            }
        }
        const double dk1 = row_bcast<R>(w, k + 1).x;
        if (m >= 2) {
            double2 tn;
            double bn;
            reflector(k + 1, tn, bn);
            if (bid == 0 && wave == 0) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int i = lane + 64 * r;
                    if (i >= k + 2 && i < n) V[(size_t)(k + 1) * n + i] = w[r];
                }
                if (lane == 0) {
                    tau[k + 1] = tn;
                    d[k + 1] = dk1;
                    e[k + 1] = bn;
                }
            }
            update_and_project(k, true);
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = w[r];
            t = tn;
        } else if (bid == 0 && threadIdx.x == 0) {
            d[k + 1] = dk1;
        }
    }
}

// Householder tridiagonalisation (lower) for large n (up to 64 * NW * CH rows): the own
// columns stay in registers for the whole reduction, their rows split over the NW waves of
// the workgroup (wave w, lane l holds rows 64 (w + NW ch) + l, ch < CH, of every own column),
// so the trailing matrix never leaves the register files of the grid (at n = 1920 it is
// 59 MB, more than all LDS together).  Same ownership and ring as k_tridiag_reg; each lane
// polls only its own rows of p_k and of column k + 1.  The sums over rows v^H p and the
// reflector norm are wave sums combined in LDS in a fixed order; the projections onto
// v_{k+1} are summed over 16-lane rows only and the 4 NW row partials of a column combined
// by wave 0.  Three workgroup barriers per column step.
// LDS: y_k and v_k by row (2n), 4 NC NW projection partials, broadcast values, 2 NW sums.
template <int CH, int NC, int NW>
__global__ __launch_bounds__(NW * 64) void k_tridiag_rows(const double2* __restrict__ A, int n, int ncol,
                                                         double2* __restrict__ V, double2* __restrict__ tau,
                                                         double* __restrict__ d, double* __restrict__ e,
                                                         double2* __restrict__ ring, unsigned* abort_flag) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* yl = reinterpret_cast<double2*>(smem);   // y_k by row
    double2* vl = yl + n;                              // v_k by row
    double2* pp = vl + n;                              // pp[(lc * NW + wave) * 4 + row]: partial projections
    double2* bc = pp + 4 * NC * NW;                    // v_k, p_k, column k+1 at row k+1; alpha
    double* red = reinterpret_cast<double*>(bc + 4);   // [NW] v^H p partials, [NW] norm partials
    const int nwg = (int)gridDim.x, bid = (int)blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double2 cols[NC][CH];
#pragma unroll
    for (int lc = 0; lc < NC; ++lc) {
        const int c = bid + lc * nwg;
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) {
            const int i = 64 * (wave + NW * ch) + lane;
            cols[lc][ch] = (lc < ncol && c < n && i < n) ? A[(size_t)c * n + i] : make_double2(0.0, 0.0);
        }
    }
    double2 v[CH], y[CH], w[CH];   // own rows of v_k, y_k, and the column / v_{k+1}
    // the owning lane of row i stores x at row i into *dst
    auto own_put = [&](int i, const double2 (&x)[CH], double2* dst) {
        const int j = i >> 6;
        if (i < n && j % NW == wave && lane == (i & 63)) {
#pragma unroll
            for (int ch = 0; ch < CH; ++ch)
                if (ch == j / NW) *dst = x[ch];
        }
    };
    // sum of the NW wave partials r[0..NW): one LDS read a lane and the same row tree in every
    // wave and every workgroup, so all of them get the same bits
    auto waves_sum = [&](const double* r) { return row_sum16((lane & 15) < NW ? r[lane & 15] : 0.0); };
    // reflector of w on rows > k0 (as in k_tridiag_reg); one workgroup barrier
    auto reflector = [&](int k0, double2& tv, double& beta) {
        double s2 = 0;
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) {
            const int i = 64 * (wave + NW * ch) + lane;
            if (i > k0 + 1 && i < n) s2 += w[ch].x * w[ch].x + w[ch].y * w[ch].y;
        }
        s2 = wave_sum(s2);
        if (lane == 0) red[NW + wave] = s2;
        own_put(k0 + 1, w, bc + 3);
        __syncthreads();
        const double xn2 = waves_sum(red + NW);
        const double2 alpha = bc[3];
        double2 sc;
        if (xn2 == 0.0 && alpha.y == 0.0) {
            tv = make_double2(0.0, 0.0);
            beta = alpha.x;
            sc = make_double2(0.0, 0.0);
        } else {
            const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
            beta = alpha.x >= 0 ? -nrm : nrm;
            const double rb = 1.0 / beta;
            tv = make_double2((beta - alpha.x) * rb, -alpha.y * rb);
            const double2 den = make_double2(alpha.x - beta, alpha.y);
            const double rd = 1.0 / (den.x * den.x + den.y * den.y);
            sc = make_double2(den.x * rd, -den.y * rd);
        }
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) {
            const int i = 64 * (wave + NW * ch) + lane;
            w[ch] = (i == k0 + 1) ? make_double2(1.0, 0.0)
                                  : ((i > k0 + 1 && i < n) ? cmul(w[ch], sc) : make_double2(0.0, 0.0));
        }
    };
    // as in k_tridiag_reg, with (y_c, v_c) read from LDS; one workgroup barrier before the
    // per-column sums over the waves are published
    auto update_and_project = [&](int k, bool update) {
        const __amdgpu_buffer_rsrc_t rs = ring_rsrc(ring + (size_t)(k + 1) * 2 * n, 2 * n);
#pragma unroll
        for (int lc = 0; lc < NC; ++lc) {
            const int c = bid + lc * nwg;
            if (lc < ncol && c > k + 1 && c < n) {   // uniform over the workgroup
                double2 yc = make_double2(0.0, 0.0), vc = make_double2(0.0, 0.0);
                if (update) {
                    yc = yl[c];
                    vc = vl[c];
                }
                // every row, unpredicated: v_k and y_k are zero on rows <= k and past n, so
                // the update changes nothing there but row k + 1, which no later step reads;
                // v_{k+1} (in w) is zero on rows <= k + 1 and past n, so those rows add nothing
                // to the projection
                // a -= v_i conj(y_c) + y_i conj(v_c);  p += conj(a) w_i  (as FMA chains)
                double pr = 0, pi = 0;
#pragma unroll
                for (int ch = 0; ch < CH; ++ch) {
                    double2 a = cols[lc][ch];
                    if (update) {
                        a.x = fma(-yc.x, v[ch].x, a.x);
                        a.x = fma(-yc.y, v[ch].y, a.x);
                        a.x = fma(-vc.x, y[ch].x, a.x);
                        a.x = fma(-vc.y, y[ch].y, a.x);
                        a.y = fma(-yc.x, v[ch].y, a.y);
                        a.y = fma(yc.y, v[ch].x, a.y);
                        a.y = fma(-vc.x, y[ch].y, a.y);
                        a.y = fma(vc.y, y[ch].x, a.y);
                        cols[lc][ch] = a;
                    }
                    pr = fma(a.x, w[ch].x, pr);
                    pr = fma(a.y, w[ch].y, pr);
                    pi = fma(a.x, w[ch].y, pi);
                    pi = fma(-a.y, w[ch].x, pi);
                }
                // row sums only: the 4 rows of every wave go to LDS
                pr = row_sum16(pr);
                pi = row_sum16(pi);
                if ((lane & 15) == 0) pp[(lc * NW + wave) * 4 + (lane >> 4)] = make_double2(pr, pi);
                if (c == k + 2) {
#pragma unroll
                    for (int ch = 0; ch < CH; ++ch) {
                        const int i = 64 * (wave + NW * ch) + lane;
                        if (i >= k + 2 && i < n) ring_put16(rs, n + (i - (k + 2)), cols[lc][ch]);
                    }
                }
            }
        }
        __syncthreads();
        // wave 0, lane 8 lc + g: partials g, g + 8, ... of column lc, then a tree over g
        if (wave == 0) {
            // the lane id read afresh (volatile: not hoisted out of the column loop, where a
            // long-lived per-lane column index was spilled to scratch)
            int l;
            asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
            const int lc = l >> 3, g = l & 7, c = bid + lc * nwg;
            double re = 0, im = 0;
#pragma unroll
            for (int q = g; q < 4 * NW; q += 8) {
                const double2 t = pp[lc * NW * 4 + q];
                re += t.x;
                im += t.y;
            }
            re += dppm_d<0xB1>(re);
            im += dppm_d<0xB1>(im);
            re += dppm_d<0x4E>(re);
            im += dppm_d<0x4E>(im);
            re += dppm_d<0x141>(re);
            im += dppm_d<0x141>(im);
            if (g == 0 && lc < ncol && c > k + 1 && c < n) ring_put16(rs, c - (k + 2), make_double2(re, im));
        }
    };
    // first reflector from column 0 (rows 1 .. n-1)
    double2 t;
    double beta;
#pragma unroll
    for (int ch = 0; ch < CH; ++ch) {
        const int i = 64 * (wave + NW * ch) + lane;
        w[ch] = (i >= 1 && i < n) ? A[i] : make_double2(0.0, 0.0);
    }
    reflector(0, t, beta);
    if (bid == 0) {
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) {
            const int i = 64 * (wave + NW * ch) + lane;
            if (i >= 1 && i < n) V[i] = w[ch];
        }
        if (threadIdx.x == 0) {
            tau[0] = t;
            d[0] = A[0].x;
            e[0] = beta;
        }
    }
    update_and_project(-1, false);
#pragma unroll
    for (int ch = 0; ch < CH; ++ch) v[ch] = w[ch];
    for (int k = 0; k < n - 1; ++k) {
        const int m = n - k - 1;
        const __amdgpu_buffer_rsrc_t rs = ring_rsrc(ring + (size_t)k * 2 * n, 2 * n);
        // ---- own rows of p_k (-> y) and of column k+1 (-> w) ----
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) {
            y[ch] = make_double2(0.0, 0.0);
            w[ch] = make_double2(0.0, 0.0);
        }
        {
            unsigned spins = 0, got = 0;
            while (true) {
                bool mine = true;
#pragma unroll
                for (int ch = 0; ch < CH; ++ch) {
                    const int i = 64 * (wave + NW * ch) + lane;
                    if (i > k && i < n) {
                        double2 val;
                        if (!((got >> (2 * ch)) & 1u)) {
                            if (ring_get16(rs, i - (k + 1), val)) {
                                y[ch] = val;
                                got |= 1u << (2 * ch);
                            } else {
                                mine = false;
                            }
                        }
                        if (!((got >> (2 * ch + 1)) & 1u)) {
                            if (ring_get16(rs, n + (i - (k + 1)), val)) {
                                w[ch] = val;
                                got |= 1u << (2 * ch + 1);
                            } else {
                                mine = false;
                            }
                        }
                    }
                }
                if (__all(mine)) break;
                if ((++spins & 31u) == 0) {
                    int st = 0;
                    if (lane == 0) {
                        st = __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
                        if (!st && spins > SPIN_LIMIT) {
                            atomicExch(abort_flag, 1u);
                            st = 1;
                        }
                    }
                    if (__shfl(st, 0, 64)) break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        // ---- mu = v^H p over the workgroup; y = tau p - (|tau|^2 mu / 2) v ----
        double mr = 0;
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) mr += cconjmul(v[ch], y[ch]).x;
        mr = wave_sum(mr);
        if (lane == 0) red[wave] = mr;
        own_put(k + 1, v, bc + 0);
        own_put(k + 1, y, bc + 1);
        own_put(k + 1, w, bc + 2);
        __syncthreads();
        const double mu = waves_sum(red);
        const double c2 = 0.5 * (t.x * t.x + t.y * t.y) * mu;
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) {
            const double2 tp = cmul(t, y[ch]);
            y[ch] = make_double2(tp.x - c2 * v[ch].x, tp.y - c2 * v[ch].y);
        }
        // ---- column k+1 after the rank-2 update ----
        const double2 v0 = bc[0];
        double2 y0 = cmul(t, bc[1]);
        y0 = make_double2(y0.x - c2 * v0.x, y0.y - c2 * v0.y);
#pragma unroll
        for (int ch = 0; ch < CH; ++ch) {
            const double2 a1 = cconjmul(v0, y[ch]);
            const double2 a2 = cconjmul(y0, v[ch]);
            w[ch].x -= a1.x + a2.x;
            w[ch].y -= a1.y + a2.y;
        }
        double dk1;
        {
            const double2 a1 = cconjmul(v0, y0);
            const double2 a2 = cconjmul(y0, v0);
            dk1 = bc[2].x - (a1.x + a2.x);
        }
        if (m >= 2) {
            // y_k and v_k by row, for the (y_c, v_c) of the own columns' updates
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) {
                const int i = 64 * (wave + NW * ch) + lane;
                if (i > k && i < n) {
                    yl[i] = y[ch];
                    vl[i] = v[ch];
                }
            }
            double2 tn;
            double bn;
            reflector(k + 1, tn, bn);
            if (bid == 0) {
#pragma unroll
                for (int ch = 0; ch < CH; ++ch) {
                    const int i = 64 * (wave + NW * ch) + lane;
                    if (i >= k + 2 && i < n) V[(size_t)(k + 1) * n + i] = w[ch];
                }
                if (threadIdx.x == 0) {
                    tau[k + 1] = tn;
                    d[k + 1] = dk1;
                    e[k + 1] = bn;
                }
            }
            update_and_project(k, true);
#pragma unroll
            for (int ch = 0; ch < CH; ++ch) v[ch] = w[ch];
            t = tn;
        } else if (bid == 0 && threadIdx.x == 0) {
            d[k + 1] = dk1;
        }
    }
}

// Sturm count: number of eigenvalues of T strictly below x, from the sign changes of the
// characteristic polynomials of the leading blocks, p_i = (d_i - x) p_{i-1} - e_{i-1}^2 p_{i-2}
// (division-free; the ratio q_i = p_i / p_{i-1} of the LAPACK form is never formed).  A zero
// p_i counts as a negative ratio and takes the sign opposite to p_{i-1}, as LAPACK's
// replacement of q = 0 by -pivmin does.  p and p_{i-1} are rescaled by a power of two every
// 8 steps (the recurrence is homogeneous).  d and e2 (squared off-diagonals) in LDS; every
// lane of a wave reads the same address.
__device__ int sturm(const double* d, const double* e2, int n, double x) {
    double pm = 1.0, p = d[0] - x;
    // the last p negative, zeros resolved as above: a per-lane boolean, so the sign logic
    // is a few lane-mask operations beside the p chain rather than selects of doubles
    bool neg = p == 0.0 || p < 0.0;
    int c = neg;
    auto step = [&](double di, double fi) {
        const double pn = fma(di - x, p, -fi * pm);
        pm = p;
        p = pn;
        const bool z = p == 0.0, lt = p < 0.0;
        c += z | (lt ^ neg);              // a zero always changes the sign
        neg = (z & !neg) | (!z & lt);     // bitwise: lane-mask logic, no branches
    };
    int i = 1;
    for (; i + 7 < n; i += 8) {
        double dd[8], ff[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            dd[q] = d[i + q];
            ff[q] = e2[i + q - 1];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) step(dd[q], ff[q]);
        const int ex = __builtin_amdgcn_frexp_exp(p != 0.0 ? p : pm);
        p = __builtin_amdgcn_ldexp(p, -ex);
        pm = __builtin_amdgcn_ldexp(pm, -ex);
    }
    for (; i < n; ++i) step(d[i], e2[i - 1]);
    return c;
}

// Multisection: one wave per eigenvalue index j; each step the 64 lanes evaluate the Sturm
// count at 64 interior points of [lo, hi] and the wave keeps the sub-interval holding the
// j-th eigenvalue (a 65x shrink per step, ~9 steps to full precision).
// A d or e that is not finite (a corrupted reduction) raises bit 2 of *bad (when given), so the
// caller reports an eigensolver error instead of returning meaningless pairs.
__global__ __launch_bounds__(256) void k_bisect(const double* __restrict__ d, const double* __restrict__ e, int n,
                                                int nband, double* __restrict__ w, unsigned* __restrict__ bad) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* ds = reinterpret_cast<double*>(smem);   // d, then e^2
    double* e2 = ds + n;
    bool finite = true;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        ds[i] = d[i];
        finite = finite && isfinite(ds[i]);
        if (i < n - 1) {
            e2[i] = e[i] * e[i];
            finite = finite && isfinite(e2[i]);
        }
    }
    if (!finite && bad && blockIdx.x == 0) atomicOr(bad, 2u);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= nband) return;   // whole wave; no barrier follows
    double lo = 1e300, hi = -1e300, emax2 = 0;
    for (int i = lane; i < n; i += 64) {
        const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        lo = fmin(lo, ds[i] - r);
        hi = fmax(hi, ds[i] + r);
        if (i < n - 1) emax2 = fmax(emax2, e2[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, o, 64));
        hi = fmax(hi, __shfl_xor(hi, o, 64));
        emax2 = fmax(emax2, __shfl_xor(emax2, o, 64));
    }
    const double pivmin = DBL_MIN * fmax(1.0, emax2);
    const double tnorm = fmax(fabs(lo), fabs(hi));
    lo -= 2.0 * DBL_EPSILON * tnorm * n + 2 * pivmin;
    hi += 2.0 * DBL_EPSILON * tnorm * n + 2 * pivmin;
    // invariant: sturm(lo) <= j < sturm(hi)
    for (int it = 0; it < 64; ++it) {
        if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin) break;
        const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
        const int c = sturm(ds, e2, n, x);
        const unsigned long long above = __ballot(c > j);
        const double x63 = lane_bcast(x, 63);
        if (above == 0ull) {
            lo = x63;
        } else {
            const int l = __ffsll((long long)above) - 1;
            const double xl = lane_bcast(x, l);
            const double xlm = lane_bcast(x, l > 0 ? l - 1 : 0);
            if (l > 0) lo = xlm;
            hi = xl;
        }
    }
    if (lane == 0) w[j] = 0.5 * (lo + hi);
}

__device__ inline unsigned hash_u32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Inverse iteration, one workgroup per cluster of close eigenvalues (the workgroup of the
// cluster's first index does the whole cluster; the others exit).  The LU factorisation
// and the two triangular solves are sequential (thread 0); their recurrences carry the
// running values in registers and read the next operands from LDS one step ahead, so the
// chain waits on arithmetic, not on LDS round trips.
// LDS: d and e, which the LU overwrites in place with IA = 1 / pivot and B (step k writes
// slot k after its reads of slots k + 1 and k + 2; reloaded for the next member of a
// cluster), the iterate xv, the LU factors C and DD (n doubles each), reduction scratch and
// the pivots (n bytes): under 80 KB up to n = 1980, two workgroups per CU.
__global__ __launch_bounds__(NT) void k_invit(const double* __restrict__ d, const double* __restrict__ e, int n,
                                              int nband, const double* __restrict__ w, double* __restrict__ Z,
                                              double ctol_rel, const int* __restrict__ need) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* ds = reinterpret_cast<double*>(smem);
    double* es = ds + n;
    double* xv = es + n;
    double* C = xv + n;
    double* DD = C + n;
    double* red = DD + n;
    unsigned char* piv = reinterpret_cast<unsigned char*>(red + 64);
    double* IA = ds;   // written over d by the LU
    double* B = es;    // written over e by the LU
    const int j0 = blockIdx.x;
    // ||T||_1 bound (Gershgorin), block-parallel
    double tn = 0;
    for (int i = threadIdx.x; i < n; i += NT) {
        const double di = d[i];
        const double ei = i < n - 1 ? e[i] : 0.0;
        ds[i] = di;
        es[i] = ei;
        tn = fmax(tn, fabs(di) + (i > 0 ? fabs(e[i - 1]) : 0.0) + fabs(ei));
    }
    for (int o = 32; o > 0; o >>= 1) tn = fmax(tn, __shfl_xor(tn, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = tn;
    __syncthreads();
    tn = DBL_MIN;
    for (int q = 0; q < NT / 64; ++q) tn = fmax(tn, red[q]);
    __syncthreads();
    const double ctol = ctol_rel * tn;
    if (j0 > 0 && w[j0] - w[j0 - 1] <= ctol) return;   // not a cluster head (uniform)
    int j1 = j0;
    while (j1 + 1 < nband && w[j1 + 1] - w[j1] <= ctol) ++j1;
    if (need) {   // the cluster is computed whole when any of its bands is needed (uniform)
        bool any = false;
        for (int j = j0; j <= j1; ++j) any = any || need[j] != 0;
        if (!any) return;
    }
    const double eps = DBL_EPSILON;
    const double pertol = 10.0 * eps * tn;
    const double tiny = eps * tn;
    double lam_prev = -1e300;
    for (int j = j0; j <= j1; ++j) {
        double lam = w[j];
        if (j > j0 && lam - lam_prev < pertol) lam = lam_prev + pertol;
        lam_prev = lam;
        if (j > j0) {   // the previous member's LU overwrote d and e (uniform)
            for (int i = threadIdx.x; i < n; i += NT) {
                ds[i] = d[i];
                es[i] = i < n - 1 ? e[i] : 0.0;
            }
            __syncthreads();
        }
        // LU of T - lam I with partial pivoting (thread 0, sequential); ak, bk carry the
        // running diagonal and superdiagonal of the active row
        if (threadIdx.x == 0) {
            double ak = ds[0] - lam, bk = n > 1 ? es[0] : 0.0;
            double an = n > 1 ? ds[1] - lam : 0.0, ck = es[0], bn = n > 2 ? es[1] : 0.0;
            for (int k = 0; k < n - 1; ++k) {
                // operands of step k+1, independent of this step: unconditional loads at
                // clamped indices (e[n-1] is 0, and d past n - 1 only feeds the unused
                // value after the last step), shift by lam applied where it is consumed
                const double dn2 = ds[min(k + 2, n - 1)];
                const double ck2 = es[min(k + 1, n - 1)];
                const double bn2 = es[min(k + 2, n - 1)];
                // branch-free (selects keep the exec mask whole on the sequential chain):
                // no interchange when |c_k| <= |a_k| (a tiny a_k is replaced by +-tiny), else
                // rows k and k+1 are swapped
                const bool keep = fabs(ck) <= fabs(ak);
                const double akk = (keep && fabs(ak) < tiny) ? (ak >= 0 ? tiny : -tiny) : ak;
                const double pv = keep ? akk : ck;
                const double ia = 1.0 / pv;
                const double mult = (keep ? ck : akk) * ia;
                piv[k] = keep ? 0 : 1;
                IA[k] = ia;
                B[k] = keep ? bk : an;
                C[k] = mult;
                DD[k] = keep ? 0.0 : bn;
                const double akn = keep ? fma(-mult, bk, an) : fma(-mult, an, bk);
                bk = keep ? bn : -mult * bn;
                ak = akn;
                an = dn2 - lam;
                ck = ck2;
                bn = bn2;
            }
            if (fabs(ak) < tiny) ak = (ak >= 0 ? tiny : -tiny);
            IA[n - 1] = 1.0 / ak;
        }
        // start vector
        for (int i = threadIdx.x; i < n; i += NT) {
            const unsigned h = hash_u32((unsigned)(i * 2654435761u) ^ (unsigned)(j * 40503u + 17u));
            xv[i] = ((double)h / 4294967296.0) - 0.5;
        }
        __syncthreads();
        for (int it = 0; it < 3; ++it) {
            if (threadIdx.x == 0) {
                // forward: L (row interchanges per piv), carry the running entry; the
                // operands of 8 steps are loaded before their dependent updates (each read
                // of xv precedes the write of its slot, as in the step-by-step order)
                double xk = xv[0];
                double nx = n > 1 ? xv[1] : 0.0;
                auto fwd = [&](double ck, int pk, double nx2, int k) {
                    const bool sw = pk != 0;
                    xv[k] = sw ? nx : xk;
                    xk = sw ? fma(-ck, nx, xk) : fma(-ck, xk, nx);
                    nx = nx2;
                };
                int k = 0;
                for (; k + 8 <= n - 1; k += 8) {
                    double c8[8], x8[8];
                    int p8[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        c8[q] = C[k + q];
                        p8[q] = piv[k + q];
                        x8[q] = xv[min(k + q + 2, n - 1)];   // past n - 1: fed to no step
                    }
#pragma unroll
                    for (int q = 0; q < 8; ++q) fwd(c8[q], p8[q], x8[q], k + q);
                }
                for (; k < n - 1; ++k) fwd(C[k], piv[k], k + 2 < n ? xv[k + 2] : 0.0, k);
                xv[n - 1] = xk;
                // backward: U (diagonal A, superdiagonals B, DD), carry x[k+1], x[k+2]
                double x1 = xk * IA[n - 1], x2 = 0.0;
                xv[n - 1] = x1;
                auto bwd = [&](double xx, double bb, double dd, double ia, int kk) {
                    const double x0 = (xx - bb * x1 - dd * x2) * ia;
                    xv[kk] = x0;
                    x2 = x1;
                    x1 = x0;
                };
                k = n - 2;
                for (; k - 7 >= 0; k -= 8) {
                    double xx[8], bb[8], dd[8], ia[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) {
                        const int kk = k - q;
                        xx[q] = xv[kk];
                        bb[q] = B[kk];
                        dd[q] = DD[kk];   // DD[n - 2] = 0: the LU's last step has no second superdiagonal
                        ia[q] = IA[kk];
                    }
#pragma unroll
                    for (int q = 0; q < 8; ++q) bwd(xx[q], bb[q], dd[q], ia[q], k - q);
                }
                for (; k >= 0; --k) bwd(xv[k], B[k], k < n - 2 ? DD[k] : 0.0, IA[k], k);
            }
            __syncthreads();
            // scale to unit max-norm first (avoid overflow), then Gram-Schmidt against the
            // cluster vectors already done, then normalise
            double mx = 0;
            for (int i = threadIdx.x; i < n; i += NT) mx = fmax(mx, fabs(xv[i]));
            for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
            __syncthreads();
            mx = 0;
            for (int q = 0; q < NT / 64; ++q) mx = fmax(mx, red[q]);
            __syncthreads();
            const double inv = mx > 0 ? 1.0 / mx : 1.0;
            for (int i = threadIdx.x; i < n; i += NT) xv[i] *= inv;
            __syncthreads();
            for (int q = j0; q < j; ++q) {
                const double* zq = Z + (size_t)q * n;
                double s = 0;
                for (int i = threadIdx.x; i < n; i += NT) s += zq[i] * xv[i];
                s = block_sum2(s, 0.0, red).x;
                for (int i = threadIdx.x; i < n; i += NT) xv[i] -= s * zq[i];
                __syncthreads();
            }
            double s2 = 0;
            for (int i = threadIdx.x; i < n; i += NT) s2 += xv[i] * xv[i];
            s2 = block_sum2(s2, 0.0, red).x;
            const double rn = s2 > 0 ? 1.0 / sqrt(s2) : 0.0;
            for (int i = threadIdx.x; i < n; i += NT) xv[i] *= rn;
            __syncthreads();
        }
        // sign convention of LAPACK's inverse iteration (dstein): the component of largest
        // magnitude (first one on ties) is positive
        {
            double bv = -1.0;
            int bi = n;
            for (int i = threadIdx.x; i < n; i += NT) {
                const double a = fabs(xv[i]);
                if (a > bv) {
                    bv = a;
                    bi = i;
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_xor(bv, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                if (ov > bv || (ov == bv && oi < bi)) {
                    bv = ov;
                    bi = oi;
                }
            }
            __shared__ double sv[NT / 64];
            __shared__ int si[NT / 64];
            if ((threadIdx.x & 63) == 0) {
                sv[threadIdx.x >> 6] = bv;
                si[threadIdx.x >> 6] = bi;
            }
            __syncthreads();
            bv = sv[0];
            bi = si[0];
            for (int q = 1; q < NT / 64; ++q)
                if (sv[q] > bv || (sv[q] == bv && si[q] < bi)) {
                    bv = sv[q];
                    bi = si[q];
                }
            const double sgn = xv[bi] < 0 ? -1.0 : 1.0;
            __syncthreads();
            double* zj = Z + (size_t)j * n;
            for (int i = threadIdx.x; i < n; i += NT) zj[i] = sgn * xv[i];
            __syncthreads();
        }
    }
}

// vcc[:, j] = Q z_j with one wave per column, no workgroup barriers: lane l holds rows
// l + 64 r (r < RPL) of z in registers; reflectors are read straight from V into a ring of
// D register sets, D - 1 reflectors ahead of their use, so each load has D - 1 reflector
// applications to arrive.  Rows below k + 1 are skipped a register slot at a time.
template <int RPL, int D>
__global__ __launch_bounds__(NT) void k_backtransform_w(int n, int nband, const double2* __restrict__ V,
                                                        const double2* __restrict__ tau,
                                                        const double* __restrict__ Z, double2* __restrict__ vcc,
                                                        int ldv, const int* __restrict__ need) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (j >= nband || (need && !need[j])) return;   // whole wave; the kernel has no workgroup barrier
    double2 z[RPL], vb[D][RPL], tb[D];
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const int i = lane + 64 * r;
        z[r] = make_double2(i < n ? Z[(size_t)j * n + i] : 0.0, 0.0);
    }
    // reflector k occupies rows k+1 .. n-1 of column k of V; k < 0 loads nothing useful
    auto load = [&](int k, double2 (&dst)[RPL], double2& t) {
        const int kk = k < 0 ? 0 : k;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int i = lane + 64 * r;
            dst[r] = V[(size_t)kk * n + (i < n ? i : 0)];
        }
        t = k < 0 ? make_double2(0.0, 0.0) : tau[kk];
    };
    auto apply = [&](int k, const double2 (&vc)[RPL], double2 tc) {
        const int rmin = (k + 1) >> 6;
        double sr = 0, si = 0;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            if (r < rmin) continue;
            const int i = lane + 64 * r;
            if (i > k && i < n) {
                const double2 q = cconjmul(vc[r], z[r]);
                sr += q.x;
                si += q.y;
            }
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        const double2 f = cmul(tc, make_double2(sr, si));
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            if (r < rmin) continue;
            const int i = lane + 64 * r;
            if (i > k && i < n) {
                const double2 q = cmul(vc[r], f);
                z[r].x -= q.x;
                z[r].y -= q.y;
            }
        }
    };
#pragma unroll
    for (int q = 0; q < D; ++q) load(n - 2 - q, vb[q], tb[q]);
    for (int k0 = n - 2; k0 >= 0; k0 -= D) {
#pragma unroll
        for (int q = 0; q < D; ++q) {
            const int k = k0 - q;
            if (k >= 0) {
                apply(k, vb[q], tb[q]);
                load(k - D, vb[q], tb[q]);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const int i = lane + 64 * r;
        if (i < n) vcc[(size_t)j * ldv + i] = z[r];
    }
}

// ---- blocked back-transform (zunmtr with compact WY blocks of zlarft's 'F','C' form) ----
// Q = P_0 P_1 ... P_{n-2} is taken WY_NB reflectors at a time.  Block b holds reflectors
// k0 .. k0 + nbk - 1 (k0 = b WY_NB): P_k0 ... P_{k0+nbk-1} = I - V_b T_b V_b^H, where V_b is
// rows k0 + 1 .. n - 1 of those columns of V (zero above each reflector's first row) and
// T_b is upper triangular:
//   T(c, c) = tau_c,   T(r, c) = -tau_c sum_{q = r}^{c - 1} T(r, q) G(q, c)   (r < c),
// with G = V_b^H V_b.  Q Z is Z <- Z - V_b (T_b (V_b^H Z)) for b from the last block down to
// 0: two matrix-core products per block (with U_b = V_b T_b^H, W = U_b^H Z, Z -= V_b W)
// instead of a chain of n dependent reflections.
constexpr int WY_NB = 64;
constexpr int WY_ROWS = 256;   // rows of V per workgroup of the Gram partials
constexpr int WY_MIN_N = 513;   // from this n the blocked form is faster than one wave per column

// cols[0 .. nneed) = the indices j with need[j] != 0, in order (need == nullptr: all);
// entries up to mc that no flag fills are set to -1 (one wave).
__global__ __launch_bounds__(64) void k_wy_cols(const int* __restrict__ need, int nband, int mc,
                                                int* __restrict__ cols) {
    const int lane = threadIdx.x;
    int base = 0;
    for (int j0 = 0; j0 < nband; j0 += 64) {
        const int j = j0 + lane;
        const bool f = j < nband && (need == nullptr || need[j] != 0);
        const unsigned long long m = __ballot(f);
        const int pos = base + (int)__popcll(m & ((1ull << lane) - 1ull));
        if (f && pos < mc) cols[pos] = j;
        base += (int)__popcll(m);
    }
    for (int c = base + lane; c < mc; c += 64) cols[c] = -1;
}

// zc[:, c] = (Z[:, cols[c]], 0)
__global__ void k_wy_gather(int n, const int* __restrict__ cols, const double* __restrict__ Z,
                            double2* __restrict__ zc) {
    const int c = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = cols[c];
    if (i < n) zc[(size_t)c * n + i] = make_double2(j >= 0 ? Z[(size_t)j * n + i] : 0.0, 0.0);
}

// vcc[:, cols[c]] = zc[:, c]
__global__ void k_wy_scatter(int n, const int* __restrict__ cols, const double2* __restrict__ zc,
                             double2* __restrict__ vcc, int ldv) {
    const int c = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = cols[c];
    if (i < n && j >= 0) vcc[(size_t)j * ldv + i] = zc[(size_t)c * n + i];
}

// rows k0 + 1 .. k of column k of V (above the reflector's first row, inside its block's
// panel) set to zero; one workgroup per column k = 0 .. n - 2
__global__ __launch_bounds__(WY_NB) void k_wy_zero(int n, double2* __restrict__ V) {
    const int k = blockIdx.x;
    const int i = k / WY_NB * WY_NB + 1 + (int)threadIdx.x;
    if (i <= k) V[(size_t)k * n + i] = make_double2(0.0, 0.0);
}

// partial Gram matrices of the blocks: gp[b S + s] = sum over rows r0 + s WY_ROWS ... of
// conj(V_b(i, q)) V_b(i, c), NB x NB row-major.  Thread (tq, tc) of a 16 x 16 grid keeps the
// 4 x 4 entries (4 tq + q, 4 tc + c); 32 rows of the panel at a time are staged in LDS.
template <int NB>
__global__ __launch_bounds__(256) void k_wy_gram(int n, int S, const double2* __restrict__ V,
                                                 double2* __restrict__ gp) {
    static_assert(NB == 64, "the 16 x 16 thread grid covers 64 x 64");
    __shared__ double2 tile[32][NB + 1];
    const int b = blockIdx.x, s = blockIdx.y;
    const int k0 = b * NB, nbk = min(NB, n - 1 - k0);
    const int rs = k0 + 1 + s * WY_ROWS, re = min(n, rs + WY_ROWS);
    const int tq = threadIdx.x >> 4, tc = threadIdx.x & 15;
    double2 acc[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[q][c] = make_double2(0.0, 0.0);
    for (int i0 = rs; i0 < re; i0 += 32) {
        for (int e = threadIdx.x; e < 32 * NB; e += 256) {
            const int c = e >> 5, rr = e & 31, i = i0 + rr;
            tile[rr][c] = (i < re && c < nbk && i > k0 + c) ? V[(size_t)(k0 + c) * n + i] : make_double2(0.0, 0.0);
        }
        __syncthreads();
#pragma unroll 4
        for (int rr = 0; rr < 32; ++rr) {
            double2 a[4], bb[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = tile[rr][4 * tq + q];
#pragma unroll
            for (int c = 0; c < 4; ++c) bb[c] = tile[rr][4 * tc + c];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    acc[q][c].x = fma(a[q].x, bb[c].x, fma(a[q].y, bb[c].y, acc[q][c].x));
                    acc[q][c].y = fma(a[q].x, bb[c].y, fma(-a[q].y, bb[c].x, acc[q][c].y));
                }
        }
        __syncthreads();
    }
    double2* g = gp + (size_t)(b * S + s) * NB * NB;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 4; ++c) g[(4 * tq + q) * NB + 4 * tc + c] = acc[q][c];
}

// T_b (NB x NB upper triangular) from the Gram partials, built by merging neighbouring
// groups of reflectors: for groups 1 (rows a..a+s) and 2 (rows c = a+s .. c+s) with
// factors T1 and T2, the merged factor has T12 = -T1 G12 T2, G12 = V1^H V2.  Six levels
// (s = 1 .. 32) from T = diag(tau), two barriers each.  Result in LDS (Ts, ld NB + 1).
template <int NB>
__device__ void wy_tfactor(int b, int S, int nbk, int k0, const double2* __restrict__ gp,
                           const double2* __restrict__ tau, double2 (*G)[NB + 1], double2 (*Ts)[NB + 1],
                           double2* X) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int e = tid; e < NB * NB; e += nt) {
        double gx = 0, gy = 0;
        for (int q = 0; q < S; ++q) {
            const double2 t = gp[(size_t)(b * S + q) * NB * NB + e];
            gx += t.x;
            gy += t.y;
        }
        const int i = e / NB, j = e % NB;
        G[i][j] = make_double2(gx, gy);
        Ts[i][j] = (i == j && i < nbk) ? tau[k0 + i] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int s = 1; s < NB; s *= 2) {
        const int cnt = (NB / (2 * s)) * s * s;
        // X_p = G12 T2: X(i, j) = sum_{q <= j} G(a + i, c + q) T(c + q, c + j)
        for (int e = tid; e < cnt; e += nt) {
            const int p = e / (s * s), i = (e / s) % s, j = e % s, a = 2 * s * p, c = a + s;
            double x = 0, y = 0;
#pragma unroll 4
            for (int q = 0; q <= j; ++q) {
                const double2 t = cmul(G[a + i][c + q], Ts[c + q][c + j]);
                x += t.x;
                y += t.y;
            }
            X[e] = make_double2(x, y);
        }
        __syncthreads();
        // T12 = -T1 X: T(a + i, c + j) = -sum_{q >= i} T(a + i, a + q) X(q, j)
        for (int e = tid; e < cnt; e += nt) {
            const int p = e / (s * s), i = (e / s) % s, j = e % s, a = 2 * s * p, c = a + s;
            double x = 0, y = 0;
#pragma unroll 4
            for (int q = i; q < s; ++q) {
                const double2 t = cmul(Ts[a + i][a + q], X[p * s * s + q * s + j]);
                x += t.x;
                y += t.y;
            }
            Ts[a + i][c + j] = make_double2(-x, -y);
        }
        __syncthreads();
    }
}

// T_b of every block (one workgroup each) into T (NB x NB per block, row-major)
template <int NB>
__global__ __launch_bounds__(256) void k_wy_tfac(int n, int S, const double2* __restrict__ gp,
                                                 const double2* __restrict__ tau, double2* __restrict__ T) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2(*G)[NB + 1] = reinterpret_cast<double2(*)[NB + 1]>(smem);
    double2(*Ts)[NB + 1] = G + NB;
    double2* X = reinterpret_cast<double2*>(Ts + NB);   // NB * NB / 4 entries (32 s at s = NB / 2)
    const int b = blockIdx.x, k0 = b * NB, nbk = min(NB, n - 1 - k0);
    wy_tfactor<NB>(b, S, nbk, k0, gp, tau, G, Ts, X);
    double2* Tb = T + (size_t)b * NB * NB;
    for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) Tb[e] = Ts[e / NB][e % NB];
}

// U_b = V_b T_b^H (rows k0 + 1 .. n - 1, stored where V_b is in V's layout), so that
// T_b V_b^H Z = U_b^H Z.  Grid (blocks, 64-row chunks).  Thread (row i, wave w) computes
// U(i, q) for the 16 q of wave w: T(q, c) is then wave-uniform in LDS, and a wave's loads
// of V(i, c) are 64 consecutive rows of one column.
template <int NB>
__global__ __launch_bounds__(256) void k_wy_u(int n, const double2* __restrict__ T, const double2* __restrict__ V,
                                              double2* __restrict__ U) {
    static_assert(NB == 64, "four waves of 16 q each");
    __shared__ double2 Ts[NB][NB + 1];
    const int b = blockIdx.x;
    const int k0 = b * NB, nbk = min(NB, n - 1 - k0), r0 = k0 + 1;
    const int i0 = r0 + blockIdx.y * 64;
    if (i0 >= n) return;   // whole workgroup: no barrier reached
    const double2* Tb = T + (size_t)b * NB * NB;
    for (int e = threadIdx.x; e < NB * NB; e += 256) Ts[e / NB][e % NB] = Tb[e];
    __syncthreads();
    const int r = threadIdx.x & 63, w = threadIdx.x >> 6, i = i0 + r;
    const int ii = i < n ? i : n - 1;   // rows past n load a valid row, their results are not stored
    double2 acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = make_double2(0.0, 0.0);
#pragma unroll 4
    for (int c = 16 * w; c < nbk; ++c) {
        const double2 v = V[(size_t)(k0 + c) * n + ii];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const double2 t = Ts[16 * w + q][c];   // zero for c < q: T is upper triangular
            acc[q].x = fma(v.x, t.x, fma(v.y, t.y, acc[q].x));
            acc[q].y = fma(v.y, t.x, fma(-v.x, t.y, acc[q].y));
        }
    }
    if (i < n) {
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (16 * w + q < nbk) U[(size_t)(k0 + 16 * w + q) * n + i] = acc[q];
    }
}

// ---- Cholesky G = R^H R and R^-1, register resident (single workgroup) ----
// 1024 threads as a 32 x 32 grid; thread (tr, tc) holds entries (tr + 32a, tc + 32b) of the
// upper triangle, a, b < P, so n <= 32 P.  Cholesky is right-looking with row j of the
// trailing matrix broadcast through LDS; R^-1 is formed in place by Gauss-Jordan
// elimination from the bottom row up (the eliminated slots of R hold R^-1).  One barrier
// per step, the broadcast buffers alternate between steps.
template <int P>
__global__ __launch_bounds__(1024) void k_chol_inv(int n, const double2* __restrict__ g, int ldg,
                                                   double2* __restrict__ rinv, int ldr, int* __restrict__ status,
                                                   double min_norm) {
    __shared__ double2 rowbuf[2][32 * P];
    __shared__ double2 colbuf[2][32 * P];
    __shared__ double diagr[32 * P];
    const int tr = threadIdx.x & 31, tc = threadIdx.x >> 5;
    double2 M[P][P];
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
        for (int b = 0; b < P; ++b) {
            if (a > b) continue;
            const int i = tr + 32 * a, c = tc + 32 * b;
            M[a][b] = (i <= c && c < n) ? g[i + (size_t)c * ldg] : make_double2(0.0, 0.0);
        }
    int bad = 0;
    for (int j = 0; j < n; ++j) {
        const int buf = j & 1;
        if (tr == (j & 31)) {
#pragma unroll
            for (int a = 0; a < P; ++a)
                if (a == (j >> 5)) {
#pragma unroll
                    for (int b = 0; b < P; ++b) {
                        const int c = tc + 32 * b;
                        if (a <= b && c >= j && c < n) rowbuf[buf][c] = M[a][b];
                    }
                }
        }
        __syncthreads();
        const double piv2 = rowbuf[buf][j].x;
        const double piv = piv2 > 0 ? sqrt(piv2) : 0.0;
        if (!(piv2 > 0) || piv < min_norm) {   // uniform: every thread read the same pivot
            bad = j + 1;
            break;
        }
        const double ip = 1.0 / piv, ip2 = 1.0 / piv2;
        if (threadIdx.x == 0) diagr[j] = piv;
#pragma unroll
        for (int a = 0; a < P; ++a) {
            const int i = tr + 32 * a;
            if (i < j || i >= n) continue;
            const double2 ri = rowbuf[buf][i];
#pragma unroll
            for (int b = 0; b < P; ++b) {
                const int c = tc + 32 * b;
                if (a > b || c < i || c >= n) continue;
                const double2 rc = rowbuf[buf][c];
                if (i == j) {
                    M[a][b] = (c == j) ? make_double2(piv, 0.0) : make_double2(rc.x * ip, rc.y * ip);
                } else {
                    const double2 q = cconjmul(ri, rc);
                    M[a][b].x -= q.x * ip2;
                    M[a][b].y -= q.y * ip2;
                }
            }
        }
    }
    if (bad) {   // zero R^-1 so that work queued behind a failed factorisation stays finite
        if (threadIdx.x == 0) status[0] = bad;
        for (size_t t = threadIdx.x; t < (size_t)n * n; t += blockDim.x)
            rinv[t % n + (t / n) * (size_t)ldr] = make_double2(0.0, 0.0);
        return;
    }
    __syncthreads();
    for (int r = n - 1; r >= 0; --r) {
        const int buf = r & 1;
        const double id = 1.0 / diagr[r];
        if (tr == (r & 31)) {   // row r of R^-1: (delta + accumulated) / R[r][r]
#pragma unroll
            for (int a = 0; a < P; ++a)
                if (a == (r >> 5)) {
#pragma unroll
                    for (int b = 0; b < P; ++b) {
                        const int c = tc + 32 * b;
                        if (a > b || c < r || c >= n) continue;
                        const double2 x = (c == r) ? make_double2(id, 0.0) : make_double2(M[a][b].x * id, M[a][b].y * id);
                        M[a][b] = x;
                        rowbuf[buf][c] = x;
                    }
                }
        }
        if (tc == (r & 31)) {   // multipliers R[k][r], k < r
#pragma unroll
            for (int b = 0; b < P; ++b)
                if (b == (r >> 5)) {
#pragma unroll
                    for (int a = 0; a < P; ++a) {
                        const int k = tr + 32 * a;
                        if (a <= b && k < r) colbuf[buf][k] = M[a][b];
                    }
                }
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < P; ++a) {
            const int k = tr + 32 * a;
            if (k >= r) continue;
            const double2 f = colbuf[buf][k];
#pragma unroll
            for (int b = 0; b < P; ++b) {
                const int c = tc + 32 * b;
                if (a > b || c < r || c >= n) continue;
                const double2 q = cmul(f, rowbuf[buf][c]);
                if (c == r) M[a][b] = make_double2(-q.x, -q.y);
                else {
                    M[a][b].x -= q.x;
                    M[a][b].y -= q.y;
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
        for (int b = 0; b < P; ++b) {
            const int i = tr + 32 * a, c = tc + 32 * b;
            if (i < n && c < n) rinv[i + (size_t)c * ldr] = (a <= b && i <= c) ? M[a][b] : make_double2(0.0, 0.0);
        }
}

}  // namespace

// Launches of the multi-workgroup reduction, whose workgroups wait on each other's ring
// values, may run at the same time on several streams only while all their workgroups fit
// on the GPU together (each workgroup fits one CU): then none can wait on a workgroup that
// cannot start, and ordinary kernels beside them always finish and free their CUs.  A
// device-wide ledger keeps the reductions not known to have finished (launch order, width);
// a new one whose width would take the running total past the CU count first makes its
// stream wait for the oldest of them, as many as needed.  Every set of reductions that can
// run together then has at most num_cu workgroups: a later one overlaps an earlier one only
// if it did not wait for it, and then the earlier one was counted against its width.
namespace {
struct SpinLedger {
    std::mutex m;
    std::deque<std::pair<hipEvent_t, int>> live;   // (completion event, workgroups)
    std::vector<hipEvent_t> pool;
};
SpinLedger g_ledger[64];

void spin_launch(const void* fn, dim3 grid, dim3 block, void** args, unsigned lds, hipStream_t s, int num_cu) {
    int dev = 0;
    APW_HIP(hipGetDevice(&dev));
    SpinLedger& L = g_ledger[dev & 63];
    std::lock_guard<std::mutex> lock(L.m);
    const int w = (int)grid.x;
    APW_CHECK(w <= num_cu, "tridiagonal reduction: %d workgroups on %d CUs", w, num_cu);
    while (!L.live.empty() && hipEventQuery(L.live.front().first) == hipSuccess) {
        L.pool.push_back(L.live.front().first);
        L.live.pop_front();
    }
    int used = 0;
    for (const auto& g : L.live) used += g.second;
    for (auto it = L.live.begin(); used + w > num_cu && it != L.live.end(); ++it) {
        APW_HIP(hipStreamWaitEvent(s, it->first, 0));
        used -= it->second;
    }
    APW_HIP(hipLaunchKernel(fn, grid, block, args, lds, s));
    hipEvent_t ev;
    if (!L.pool.empty()) {
        ev = L.pool.back();
        L.pool.pop_back();
    } else {
        APW_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    APW_HIP(hipEventRecord(ev, s));
    L.live.emplace_back(ev, w);
}
}  // namespace

// sizes of the blocked back-transform's work arrays (doubles) for n x mc
struct WySizes {
    int nblk, S;
    size_t gp, t, w;
    WySizes(int n, int mc) {
        nblk = (n - 1 + WY_NB - 1) / WY_NB;
        S = (n - 1 + WY_ROWS - 1) / WY_ROWS;
        gp = (size_t)2 * nblk * S * WY_NB * WY_NB;
        t = (size_t)2 * nblk * WY_NB * WY_NB;
        w = (size_t)2 * WY_NB * mc;
    }
    size_t total() const { return gp + t + w; }
};

void wy_reserve(EigWorkspace& ws, int n, int mc) {
    ws.a.reserve((size_t)2 * n * n);
    ws.zc.reserve((size_t)2 * n * mc);
    ws.wy.reserve(WySizes(n, mc).total());
    ws.cols.reserve(mc);
}

// vcc[:, j] = Q z_j for the mc columns j with need[j] != 0 (all when need is null), by WY
// blocks: gather, Gram partials, T_b, U_b = V_b T_b^H, two products per block, scatter.
void backtransform_wy(hipStream_t s, EigWorkspace& ws, GemmWorkspace& gws, int num_cu, int n, int nband,
                      const int* need, int mc, double* vcc, int ldv) {
    wy_reserve(ws, n, mc);
    const WySizes z(n, mc);
    double2* V = reinterpret_cast<double2*>(ws.v.get());
    const double2* tau = reinterpret_cast<const double2*>(ws.tau.get());
    double2* zc = reinterpret_cast<double2*>(ws.zc.get());
    double2* gp = reinterpret_cast<double2*>(ws.wy.get());
    double2* U = reinterpret_cast<double2*>(ws.a.get());   // the reduction's copy of hcc is spent
    double2* T = reinterpret_cast<double2*>(ws.wy.get() + z.gp);
    double* W = ws.wy.get() + z.gp + z.t;
    int* cols = ws.cols.get();
    hipLaunchKernelGGL(k_wy_cols, dim3(1), dim3(64), 0, s, need, nband, mc, cols);
    const dim3 gcol((n + NT - 1) / NT, mc);
    hipLaunchKernelGGL(k_wy_gather, gcol, dim3(NT), 0, s, n, cols, ws.z.get(), zc);
    hipLaunchKernelGGL(k_wy_zero, dim3(n - 1), dim3(WY_NB), 0, s, n, V);
    hipLaunchKernelGGL(k_wy_gram<WY_NB>, dim3(z.nblk, z.S), dim3(256), 0, s, n, z.S, V, gp);
    const size_t lds_t = (size_t)2 * WY_NB * (WY_NB + 1) * sizeof(double2) + (size_t)WY_NB * WY_NB / 4 * sizeof(double2);
    APW_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_wy_tfac<WY_NB>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_t));
    hipLaunchKernelGGL(k_wy_tfac<WY_NB>, dim3(z.nblk), dim3(256), lds_t, s, n, z.S, gp, tau, T);
    hipLaunchKernelGGL(k_wy_u<WY_NB>, dim3(z.nblk, (n - 1 + 63) / 64), dim3(256), 0, s, n, T, V, U);
    APW_LAUNCH_CHECK();
    const double* Vd = ws.v.get();
    const double* Ud = ws.a.get();
    double* zd = ws.zc.get();
    for (int b = z.nblk - 1; b >= 0; --b) {
        const int k0 = b * WY_NB, nbk = std::min(WY_NB, n - 1 - k0), r0 = k0 + 1, L = n - r0;
        const size_t off = 2 * ((size_t)k0 * n + r0);
        // W = U_b^H Z = T_b V_b^H Z, then Z -= V_b W over rows r0 .. n - 1
        gemm<double>(s, gws, num_cu, 'C', 'N', nbk, mc, L, 1.0, 0.0, Ud + off, n, zd + 2 * (size_t)r0, n, 0.0, 0.0,
                     W, WY_NB);
        gemm<double>(s, gws, num_cu, 'N', 'N', L, mc, nbk, -1.0, 0.0, Vd + off, n, W, WY_NB, 1.0, 0.0,
                     zd + 2 * (size_t)r0, n);
    }
    hipLaunchKernelGGL(k_wy_scatter, gcol, dim3(NT), 0, s, n, cols, zc, reinterpret_cast<double2*>(vcc), ldv);
    APW_LAUNCH_CHECK();
}

void heevx_reserve(EigWorkspace& ws, int nmax, int nband) {
    const size_t n = (size_t)std::max(nmax, 1);
    ws.a.reserve(2 * n * n);
    ws.v.reserve(2 * n * n);
    ws.tau.reserve(2 * n);
    ws.d.reserve(n);
    ws.e.reserve(n);
    ws.z.reserve(n * (size_t)std::max(nband, 1));
    ws.bar.reserve(64 + 256);
    ws.ring.reserve(2 * (2 * n * n));   // the larger of the two exchange rings below
    if ((int)n >= WY_MIN_N) wy_reserve(ws, (int)n, std::max(nband, 1));
}

void heevx_lowest(hipStream_t s, EigWorkspace& ws, GemmWorkspace& gws, int num_cu, int n, int ldh,
                  const double* hcc, int nband, double* eig, double* vcc, int ldv, Prof* prof) {
    APW_CHECK(ldv >= n, "heevx: ldv=%d below n=%d", ldv, n);
    heevx_values(s, ws, num_cu, n, ldh, hcc, nband, eig, prof);
    heevx_vectors(s, ws, gws, num_cu, n, nband, eig, vcc, ldv, nullptr, nband, prof);
}

void heevx_values(hipStream_t s, EigWorkspace& ws, int num_cu, int n, int ldh, const double* hcc, int nband,
                  double* eig, Prof* prof) {
    APW_CHECK(n >= 1 && nband >= 1 && nband <= n && ldh >= n, "heevx: bad sizes n=%d nband=%d", n, nband);
    APW_CHECK(n <= HEEVX_MAX, "heevx: n=%d above %d (LDS of the inverse-iteration kernel)", n, HEEVX_MAX);
    ws.a.reserve((size_t)2 * n * n);
    ws.v.reserve((size_t)2 * n * n);
    ws.tau.reserve((size_t)2 * n);
    ws.d.reserve(n);
    ws.e.reserve(n);
    ws.z.reserve((size_t)n * nband);
    ws.bar.reserve(64 + 256);
    double2* A = reinterpret_cast<double2*>(ws.a.get());
    hipLaunchKernelGGL(k_copy_herm, dim3(std::min(4096, (n * n + 255) / 256)), dim3(256), 0, s, n,
                       reinterpret_cast<const double2*>(hcc), ldh, A);
    APW_LAUNCH_CHECK();
    if (n > 1) {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, ws.bar.get(), 64 + 256);
        APW_LAUNCH_CHECK();
        // bar[1]: abort word of the exchange polls; V: n x n reflectors
        APW_CHECK(ws.bar.n >= 2 && ws.v.n >= (size_t)2 * n * n, "heevx: workspace too small for n=%d", n);
        double2* Vp = reinterpret_cast<double2*>(ws.v.get());
        double2* taup = reinterpret_cast<double2*>(ws.tau.get());
        double* dp = ws.d.get();
        double* ep = ws.e.get();
        unsigned* barp = ws.bar.get();
        // algorithmic HBM bytes: the matrix read once (16 n^2) and the reflectors written
        // (8 n^2); the per-column hemv and rank-2 update work on LDS-resident columns
        ProfScope ps(prof, s, "eig_tridiag", 24.0 * n * (double)n, 16.0 / 3.0 * n * (double)n * n);
        // LDS-resident form: the fewest workgroups whose column share fits the LDS budget
        // fused form: v_k, v_{k+1}, y and the gathered column (n each) + scratch, and the
        // own columns when they live in LDS
        const size_t fixed = (size_t)4 * n * sizeof(double2) + 64 * sizeof(double);
        int nwg = -1, ncol = 0;
        for (int w = std::max(1, std::min(n / 8, 16)); w <= std::min(num_cu, 256); ++w) {
            const int nc = (n + w - 1) / w;
            if (fixed + (size_t)nc * n * sizeof(double2) <= TRIDIAG_LDS_BYTES) {
                nwg = w;
                ncol = nc;
                break;
            }
        }
        const int R = (n + 63) / 64;
        // the row-split register form from n = 705 (measured faster than the LDS-column register
        // form there: 4.3 vs 4.5 ms at n = 768, 6.2 vs 8.0 ms at n = 1024) when n / 8
        // workgroups fit on the GPU; the LDS-column form below that, or to n = 1024 otherwise
        const bool rows_fit = n <= 2048 && (n + 7) / 8 <= std::min(num_cu, 256);
        if (n <= ws.single_max) {
            // one workgroup owns every column (kept in the copy A) and exchanges with itself:
            // an ordinary launch, not ordered against other streams' reductions, so many small
            // solves (batched k-points or cells) reduce at the same time
            const int ncol1 = n;
            const size_t lds = fixed;
            const size_t nring = (size_t)2 * n * (n - 1);
            ws.ring.reserve(2 * nring);
            double2* ring = reinterpret_cast<double2*>(ws.ring.get());
            hipLaunchKernelGGL(k_fill_ring, dim3((unsigned)std::min<size_t>(1024, (nring + 255) / 256)), dim3(256), 0, s,
                               ring, nring);
            APW_LAUNCH_CHECK();
            APW_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_tridiag_fused<true>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            hipLaunchKernelGGL(k_tridiag_fused<true>, dim3(1), dim3(NT), lds, s, A, n, ncol1, Vp, taup, dp, ep, ring,
                               barp + 1);
            APW_LAUNCH_CHECK();
        } else if (R <= 11 || (R <= 16 && !rows_fit)) {
            // register form: LDS = own columns + two gather buffers of 2n
            const size_t fixed_r = (size_t)4 * n * sizeof(double2);
            int w = std::max(1, std::min(n / 8, 16)), nc = 0;
            for (; w <= std::min(num_cu, 256); ++w) {
                nc = (n + w - 1) / w;
                if (fixed_r + (size_t)nc * n * sizeof(double2) <= TRIDIAG_LDS_BYTES) break;
            }
            APW_CHECK(w <= std::min(num_cu, 256), "heevx: no workgroup count fits n=%d", n);
            const int w_min = w;
            // about four columns a workgroup (at most one a wave in each step's update): with
            // the per-wave polling of the ring, spreading the columns beats fewer exchange
            // participants up to ~160 of them (measured n = 160..640: 5-17 % faster than the
            // LDS minimum; 192 and 256 workgroups lost again at n = 640)
            w = std::max(w, std::min(std::min(num_cu, 160), n / 4));
            const size_t nring = (size_t)2 * n * n;
            ws.ring.reserve(2 * nring);
            double2* ring = reinterpret_cast<double2*>(ws.ring.get());
            const void* fn = nullptr;
            // threads per workgroup: 512 for 6 <= R <= 12 (measured faster at n = 480, 640: each
            // wave updates fewer columns); 256 otherwise (small n: the extra waves cost more in
            // barriers than they save; R = 16: the register budget of v, y, w, 6 R doubles a lane)
            int ntr = 256;
            switch (R <= 6 ? R : (R <= 8 ? 8 : (R <= 10 ? 10 : (R <= 12 ? 12 : 16)))) {
                case 1: fn = reinterpret_cast<const void*>(k_tridiag_reg<1, 256>); ntr = 256; break;
                case 2: fn = reinterpret_cast<const void*>(k_tridiag_reg<2, 256>); ntr = 256; break;
                case 3: fn = reinterpret_cast<const void*>(k_tridiag_reg<3, 256>); ntr = 256; break;
                case 4: fn = reinterpret_cast<const void*>(k_tridiag_reg<4, 256>); ntr = 256; break;
                case 5: fn = reinterpret_cast<const void*>(k_tridiag_reg<5, 256>); ntr = 256; break;
                case 6: fn = reinterpret_cast<const void*>(k_tridiag_reg<6, 512>); ntr = 512; break;
                case 8: fn = reinterpret_cast<const void*>(k_tridiag_reg<8, 512>); ntr = 512; break;
                case 10: fn = reinterpret_cast<const void*>(k_tridiag_reg<10, 512>); ntr = 512; break;
                case 12: fn = reinterpret_cast<const void*>(k_tridiag_reg<12, 512>); ntr = 512; break;
                default: fn = reinterpret_cast<const void*>(k_tridiag_reg<16, 256>); ntr = 256; break;
            }
            // one reduction with wg workgroups (A is only read: it can run again)
            auto run_reg = [&](int wg) {
                int ncol_r = (n + wg - 1) / wg;
                const size_t lds = fixed_r + (size_t)ncol_r * n * sizeof(double2);
                hipLaunchKernelGGL(k_fill_ring, dim3((unsigned)std::min<size_t>(1024, (nring + 255) / 256)), dim3(256),
                                   0, s, ring, nring);
                APW_LAUNCH_CHECK();
                APW_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                const double2* Ac = A;
                unsigned* abortp = barp + 1;
                void* args[] = {&Ac, &n, &ncol_r, &Vp, &taup, &dp, &ep, &ring, &abortp};
                spin_launch(fn, dim3(wg), dim3(ntr), args, (unsigned)lds, s, num_cu);
            };
            // autotuned workgroup count
            const int tmode = tune_mode();
            if (tmode > 0) {
                char key[96];
                std::snprintf(key, sizeof(key), "tridiag reg n=%d cu=%d", (n + 31) / 32 * 32, num_cu);
                int a = 0, b = 0;
                const int wmax = std::min(num_cu, 256);
                if (tune_find(key, a, b) && a >= w_min && a <= wmax) {
                    w = a;
                } else if (tmode == 2) {
                    std::vector<int> cand;
                    for (int c : {w_min, n / 8, n / 6, n / 4, n / 3, n / 2, 128, 160, 192, 256}) {
                        c = std::max(w_min, std::min(c, wmax));
                        if (std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
                    }
                    double best = 1e300;
                    for (int c : cand) {
                        const double us = time_on_stream(s, 2, [&] { run_reg(c); });
                        if (us < best) {
                            best = us;
                            w = c;
                        }
                    }
                    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, ws.bar.get(), 64 + 256);
                    APW_LAUNCH_CHECK();
                    tune_store(key, w, 0, best);
                }
            }
            run_reg(w);
        } else if (rows_fit) {
            // row-split register form: up to 8 own columns a workgroup in registers, 1024
            // threads (16 waves x 2 row chunks; 8 waves x 4 chunks measured 1.7x slower)
            const int wmin = (n + 7) / 8, wmax = std::min(num_cu, 256);
            int w = wmax;
            const size_t nring = (size_t)2 * n * n;
            ws.ring.reserve(2 * nring);
            double2* ring = reinterpret_cast<double2*>(ws.ring.get());
            auto run_rows = [&](int wg) {
                int ncol_r = (n + wg - 1) / wg;
                constexpr int nw = 16;
                const size_t lds = (size_t)2 * n * sizeof(double2) + (size_t)(4 * 8 * nw + 4) * sizeof(double2) +
                                   (size_t)2 * nw * sizeof(double);
                const void* fn = reinterpret_cast<const void*>(k_tridiag_rows<2, 8, nw>);
                APW_CHECK(ncol_r <= 8 && n <= 64 * 32, "heevx: row-split form cannot take n=%d on %d workgroups", n, wg);
                hipLaunchKernelGGL(k_fill_ring, dim3((unsigned)std::min<size_t>(1024, (nring + 255) / 256)), dim3(256),
                                   0, s, ring, nring);
                APW_LAUNCH_CHECK();
                APW_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                const double2* Ac = A;
                unsigned* abortp = barp + 1;
                void* args[] = {&Ac, &n, &ncol_r, &Vp, &taup, &dp, &ep, &ring, &abortp};
                spin_launch(fn, dim3(wg), dim3(64 * nw), args, (unsigned)lds, s, num_cu);
            };
            const int tmode = tune_mode();
            if (tmode > 0) {
                char key[96];
                std::snprintf(key, sizeof(key), "tridiag rows n=%d cu=%d", (n + 31) / 32 * 32, num_cu);
                int a = 0, b = 0;
                if (tune_find(key, a, b) && a >= wmin && a <= wmax && b == 0) {
                    w = a;
                } else if (tmode == 2) {
                    std::vector<int> cand;
                    for (int c : {wmin, (n + 5) / 6, (n + 3) / 4, 192, 224, 256}) {
                        c = std::max(wmin, std::min(c, wmax));
                        if (std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
                    }
                    double best = 1e300;
                    for (int c : cand) {
                        const double us = time_on_stream(s, 2, [&] { run_rows(c); });
                        if (us < best) {
                            best = us;
                            w = c;
                        }
                    }
                    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, ws.bar.get(), 64 + 256);
                    APW_LAUNCH_CHECK();
                    tune_store(key, w, 0, best);
                }
            }
            run_rows(w);
        } else {
            // columns in LDS when the share fits, else in A itself (one column per
            // workgroup slot, L2/Infinity-Cache resident); the ring as above
            const bool gcol = nwg <= 0;
            if (gcol) {
                nwg = std::min(num_cu, 256);
                ncol = (n + nwg - 1) / nwg;
            }
            const size_t lds = gcol ? fixed : fixed + (size_t)ncol * n * sizeof(double2);
            const size_t nring = (size_t)2 * n * (n - 1);
            ws.ring.reserve(2 * nring);
            double2* ring = reinterpret_cast<double2*>(ws.ring.get());
            hipLaunchKernelGGL(k_fill_ring, dim3((unsigned)std::min<size_t>(1024, (nring + 255) / 256)), dim3(256), 0, s,
                               ring, nring);
            APW_LAUNCH_CHECK();
            const void* fn = gcol ? reinterpret_cast<const void*>(k_tridiag_fused<true>)
                                  : reinterpret_cast<const void*>(k_tridiag_fused<false>);
            APW_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            unsigned* abortp = barp + 1;
            void* args[] = {&A, &n, &ncol, &Vp, &taup, &dp, &ep, &ring, &abortp};
            spin_launch(fn, dim3(nwg), dim3(NT), args, (unsigned)lds, s, num_cu);
        }
    } else {
        // n == 1: T = [Re a00]; the real part is the first double of A
        APW_HIP(hipMemcpyAsync(ws.d.get(), A, sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    {
    ProfScope ps(prof, s, "eig_bisect", 0.0, 0.0);
    hipLaunchKernelGGL(k_bisect, dim3((nband + 3) / 4), dim3(256), (size_t)2 * n * sizeof(double), s, ws.d.get(),
                       ws.e.get(), n, nband, eig, n > 1 ? ws.bar.get() + 1 : nullptr);
    APW_LAUNCH_CHECK();
    }
}

void heevx_vectors(hipStream_t s, EigWorkspace& ws, GemmWorkspace& gws, int num_cu, int n, int nband,
                   const double* eig, double* vcc, int ldv, const int* need, int nneed, Prof* prof) {
    APW_CHECK(n >= 1 && nband >= 1 && nband <= n && ldv >= n, "heevx: bad sizes n=%d nband=%d", n, nband);
    const int mc = need ? nneed : nband;
    APW_CHECK(mc >= 0 && mc <= nband, "heevx: %d needed eigenvectors of %d", mc, nband);
    {
    ProfScope ps(prof, s, "eig_invit", 0.0, 0.0);
    const size_t lds_inv = (size_t)5 * n * sizeof(double) + 64 * sizeof(double) + (size_t)n;
    APW_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_invit), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_inv));
    hipLaunchKernelGGL(k_invit, dim3(nband), dim3(NT), lds_inv, s, ws.d.get(), ws.e.get(), n, nband, eig, ws.z.get(),
                       1e-7, need);
    APW_LAUNCH_CHECK();
    }
    ProfScope pb(prof, s, "eig_backtransform", 16.0 * n * (double)n / 2.0 + 16.0 * n * mc,
                 8.0 * n * (double)n * mc);
    if (n >= WY_MIN_N) {
        if (mc > 0) backtransform_wy(s, ws, gws, num_cu, n, nband, need, mc, vcc, ldv);
        return;
    }
    const double2* Vc = reinterpret_cast<const double2*>(ws.v.get());
    const double2* tc = reinterpret_cast<const double2*>(ws.tau.get());
    double2* out = reinterpret_cast<double2*>(vcc);
    const dim3 gw((nband + NT / 64 - 1) / (NT / 64));
    if (n < 2) {
        hipLaunchKernelGGL((k_backtransform_w<1, 4>), gw, dim3(NT), 0, s, n, nband, Vc, tc, ws.z.get(), out, ldv, need);
    } else if (n <= 128) {
        hipLaunchKernelGGL((k_backtransform_w<2, 4>), gw, dim3(NT), 0, s, n, nband, Vc, tc, ws.z.get(), out, ldv, need);
    } else if (n <= 256) {
        hipLaunchKernelGGL((k_backtransform_w<4, 4>), gw, dim3(NT), 0, s, n, nband, Vc, tc, ws.z.get(), out, ldv, need);
    } else if (n <= 384) {
        hipLaunchKernelGGL((k_backtransform_w<6, 4>), gw, dim3(NT), 0, s, n, nband, Vc, tc, ws.z.get(), out, ldv, need);
    } else {
        static_assert(WY_MIN_N <= 513, "one wave per column holds at most 512 rows");
        hipLaunchKernelGGL((k_backtransform_w<8, 4>), gw, dim3(NT), 0, s, n, nband, Vc, tc, ws.z.get(), out, ldv, need);
    }
    APW_LAUNCH_CHECK();
}

const char* eig_fault_text(unsigned word) {
    if (word & 1u) return "subspace eigensolver: an exchange poll of the tridiagonal reduction timed out";
    if (word & 2u) return "subspace eigensolver: the tridiagonal reduction produced a non-finite value";
    return "subspace eigensolver: no fault";
}

void chol_inv_upper(hipStream_t s, int n, const double* g, int ldg, double* rinv, int ldr, int* status,
                    double min_norm, Prof* prof) {
    APW_CHECK(n >= 1 && n <= CHOL_MAX, "chol_inv_upper: n=%d outside 1..%d", n, CHOL_MAX);
    // algorithmic: n^3/3 complex multiply-adds for the factor, n^3/3 for the inverse
    ProfScope ps(prof, s, "chol_inv", 32.0 * n * (double)n, 16.0 / 3.0 * n * (double)n * n);
    const double2* gp = reinterpret_cast<const double2*>(g);
    double2* rp = reinterpret_cast<double2*>(rinv);
    switch ((n + 31) / 32) {
        case 1: hipLaunchKernelGGL(k_chol_inv<1>, dim3(1), dim3(1024), 0, s, n, gp, ldg, rp, ldr, status, min_norm); break;
        case 2: hipLaunchKernelGGL(k_chol_inv<2>, dim3(1), dim3(1024), 0, s, n, gp, ldg, rp, ldr, status, min_norm); break;
        case 3: hipLaunchKernelGGL(k_chol_inv<3>, dim3(1), dim3(1024), 0, s, n, gp, ldg, rp, ldr, status, min_norm); break;
        default: hipLaunchKernelGGL(k_chol_inv<4>, dim3(1), dim3(1024), 0, s, n, gp, ldg, rp, ldr, status, min_norm); break;
    }
    APW_LAUNCH_CHECK();
}

}  // namespace apw

extern "C" int apw_dnevx_op(apw_ctx* ctx, int nstart, int ldh, const double* hcc, int nbands, double* eigenvalue,
                            double* vcc) {
    return apw::guarded([&] {
        APW_CHECK(ctx && hcc && eigenvalue && vcc, "apw_dnevx_op: null argument");
        apw::DeviceGuard g(ctx->device);
        apw::EigWorkspace& ws = ctx->ews;
        ws.single_max = ctx->eig_single_max;
        apw::heevx_lowest(ctx->stream, ws, ctx->gws, ctx->num_cu, nstart, ldh, hcc, nbands, eigenvalue, vcc, ldh,
                          &ctx->prof);
        APW_HIP(hipStreamSynchronize(ctx->stream));
        unsigned flag[2] = {0, 0};
        if (nstart > 1) APW_HIP(hipMemcpy(flag, ws.bar.get(), sizeof(flag), hipMemcpyDeviceToHost));
        APW_CHECK(flag[1] == 0, "apw_dnevx_op: %s", apw::eig_fault_text(flag[1]));
        return APW_OK;
    });
}
