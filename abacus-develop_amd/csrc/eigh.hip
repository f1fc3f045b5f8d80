// FP64 Hermitian subspace eigensolver on the GPU (lowest nband eigenpairs).
//
// Replaces dnevx_op<complex<double>, DEVICE_GPU> (dngvd_op.h:85-107), which on ROCm copies
// hcc to the host and calls LAPACK zheevx every Davidson iteration
// (rocm/dngvd_op.hip.cu:146-170).  Same algorithm family as zheevx, all on the device:
//   1. Householder reduction of the lower triangle to a real symmetric tridiagonal T
//      (zhetrd 'L' form, unblocked) in one cooperative kernel: the trailing matrix is
//      partitioned by columns over all waves of the grid, two grid barriers per column;
//   2. bisection with Sturm counts for eigenvalues 0..nband-1 of T (dstebz form);
//   3. inverse iteration with partial-pivoting tridiagonal LU and Gram-Schmidt inside
//      clusters of close eigenvalues (dstein form);
//   4. back-transformation z -> Q z with the stored reflectors (zunmtr form).
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "apw_internal.hpp"

namespace apw {
namespace {

constexpr int NT = 256;
constexpr unsigned SPIN_LIMIT = 1u << 24;

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cconjmul(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
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

__device__ void grid_barrier(unsigned* bar, unsigned& target, unsigned nwg, unsigned* abort_flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        target += nwg;
        __threadfence();
        atomicAdd(bar, 1u);
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
            if (++spins > SPIN_LIMIT) {
                atomicExch(abort_flag, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
    }
    __syncthreads();
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

// Householder tridiagonalisation (lower), cooperative, one grid barrier per column.
// Column c of the trailing matrix is owned by global wave c % nwaves for the whole
// reduction, so its reads (p = A v) and writes (rank-2 update) never cross waves.  What
// crosses workgroups is p (double-buffered by column parity, so a workgroup one column
// ahead cannot overwrite what a slower one still reads) and the column that yields the
// next reflector: its owner leaves it untouched and every workgroup recomputes its
// rank-2-updated values itself.  LDS: v, y, c (n c128 each) + reduction scratch.
// bar[0]: barrier counter, bar[1]: abort flag (set on a barrier timeout).
__global__ __launch_bounds__(NT) void k_tridiag(double2* __restrict__ A, int n, double2* __restrict__ V,
                                                double2* __restrict__ tau, double* __restrict__ d,
                                                double* __restrict__ e, double2* __restrict__ P, unsigned* bar) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* v = reinterpret_cast<double2*>(smem);
    double2* y = v + n;
    double2* cn = y + n;
    double* red = reinterpret_cast<double*>(cn + n);
    const unsigned nwg = gridDim.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gwave = blockIdx.x * (NT / 64) + wave;
    const int nwaves = nwg * (NT / 64);
    unsigned target = 0;
    // reflector of column 0
    double2 t;
    double beta;
    make_reflector(A + 1, n - 1, v, t, beta, red);
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < n - 1; i += NT) V[1 + i] = v[i];
        if (threadIdx.x == 0) {
            tau[0] = t;
            d[0] = A[0].x;
            e[0] = beta;
        }
    }
    for (int k = 0; k < n - 1; ++k) {
        const int m = n - k - 1;           // trailing matrix rows/cols k+1 .. n-1
        double2* Pk = P + (size_t)(k & 1) * n;
        // ---- p = A_trail v on owned columns c = k+1+jj ----
        {
            int c0 = k + 1 + ((gwave - (k + 1)) % nwaves + nwaves) % nwaves;
            for (int c = c0; c < n; c += nwaves) {
                const double2* col = A + (size_t)c * n + (k + 1);
                double pr = 0, pi = 0;
                for (int i = lane; i < m; i += 64) {
                    const double2 q = cconjmul(col[i], v[i]);   // conj(A[i][c]) v_i = A[c][i] v_i
                    pr += q.x;
                    pi += q.y;
                }
                pr = wave_sum(pr);
                pi = wave_sum(pi);
                if (lane == 0) Pk[c - (k + 1)] = make_double2(pr, pi);
            }
        }
        grid_barrier(bar, target, nwg, bar + 1);
        // ---- y = tau p - (|tau|^2 mu / 2) v,  mu = v^H p ----
        double mr = 0;
        for (int i = threadIdx.x; i < m; i += NT) {
            const double2 p = Pk[i];
            y[i] = p;
            mr += cconjmul(v[i], p).x;
        }
        const double mu = block_sum2(mr, 0.0, red).x;
        const double c2 = 0.5 * (t.x * t.x + t.y * t.y) * mu;
        for (int i = threadIdx.x; i < m; i += NT) {
            const double2 tp = cmul(t, y[i]);
            y[i] = make_double2(tp.x - c2 * v[i].x, tp.y - c2 * v[i].y);
        }
        __syncthreads();
        // ---- updated column k+1 (every workgroup): cn[i] = A[k+1+i][k+1] - y_i conj(v_0) - v_i conj(y_0)
        {
            const double2* col = A + (size_t)(k + 1) * n + (k + 1);
            const double2 v0 = v[0], y0 = y[0];
            for (int i = threadIdx.x; i < m; i += NT) {
                const double2 a1 = cconjmul(v0, y[i]);
                const double2 a2 = cconjmul(y0, v[i]);
                double2 a = col[i];
                a.x -= a1.x + a2.x;
                a.y -= a1.y + a2.y;
                cn[i] = a;
            }
        }
        __syncthreads();
        // ---- rank-2 update of owned columns k+2 .. n-1, rows k+2 .. n-1 ----
        {
            int c0 = k + 2 + ((gwave - (k + 2)) % nwaves + nwaves) % nwaves;
            for (int c = c0; c < n; c += nwaves) {
                double2* col = A + (size_t)c * n + (k + 1);
                const int jj = c - (k + 1);
                const double2 vj = v[jj], yj = y[jj];
                for (int i = 1 + lane; i < m; i += 64) {
                    const double2 a1 = cconjmul(vj, y[i]);
                    const double2 a2 = cconjmul(yj, v[i]);
                    double2 a = col[i];
                    a.x -= a1.x + a2.x;
                    a.y -= a1.y + a2.y;
                    col[i] = a;
                }
            }
        }
        __threadfence_block();
        __syncthreads();
        // ---- next reflector from the updated column (every workgroup, redundantly) ----
        if (m >= 2) {
            make_reflector(cn + 1, m - 1, v, t, beta, red);
            if (blockIdx.x == 0) {
                for (int i = threadIdx.x; i < m - 1; i += NT) V[(size_t)(k + 1) * n + (k + 2) + i] = v[i];
                if (threadIdx.x == 0) {
                    tau[k + 1] = t;
                    d[k + 1] = cn[0].x;
                    e[k + 1] = beta;
                }
            }
        } else if (blockIdx.x == 0 && threadIdx.x == 0) {
            d[k + 1] = cn[0].x;
        }
        __syncthreads();
    }
}

// Sturm count: number of eigenvalues of T strictly below x
__device__ int sturm(const double* d, const double* e, int n, double x, double pivmin) {
    int c = 0;
    double q = d[0] - x;
    if (fabs(q) <= pivmin) q = -pivmin;
    if (q < 0) ++c;
    for (int i = 1; i < n; ++i) {
        q = d[i] - x - e[i - 1] * e[i - 1] / q;
        if (fabs(q) <= pivmin) q = -pivmin;
        if (q < 0) ++c;
    }
    return c;
}

// Multisection: one wave per eigenvalue index j; each step the 64 lanes evaluate the Sturm
// count at 64 interior points of [lo, hi] and the wave keeps the sub-interval holding the
// j-th eigenvalue (a 65x shrink per step, ~9 steps to full precision).
__global__ __launch_bounds__(256) void k_bisect(const double* __restrict__ d, const double* __restrict__ e, int n,
                                                int nband, double* __restrict__ w) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= nband) return;   // whole wave
    double lo = 1e300, hi = -1e300, emax2 = 0;
    for (int i = lane; i < n; i += 64) {
        const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
        if (i < n - 1) emax2 = fmax(emax2, e[i] * e[i]);
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
        const int c = sturm(d, e, n, x, pivmin);
        const unsigned long long above = __ballot(c > j);
        const double x63 = __shfl(x, 63, 64);
        if (above == 0ull) {
            lo = x63;
        } else {
            const int l = __ffsll((long long)above) - 1;
            const double xl = __shfl(x, l, 64);
            const double xlm = __shfl(x, l > 0 ? l - 1 : 0, 64);
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
// cluster's first index does the whole cluster; the others exit).
// LDS: the iterate xv, the LU factors a, b, c, dd (n doubles each, touched by thread 0) and
// the pivots (n ints).
__global__ __launch_bounds__(NT) void k_invit(const double* __restrict__ d, const double* __restrict__ e, int n,
                                              int nband, const double* __restrict__ w, double* __restrict__ Z,
                                              double ctol_rel) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* xv = reinterpret_cast<double*>(smem);
    double* red = xv + n;
    double* a = red + 64;
    double* b = a + n;
    double* c = b + n;
    double* dd = c + n;
    int* piv = reinterpret_cast<int*>(dd + n);
    const int j0 = blockIdx.x;
    // ||T||_1 bound (Gershgorin)
    double tn = 0;
    for (int i = 0; i < n; ++i) {
        const double r = fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
        tn = fmax(tn, r);
    }
    tn = fmax(tn, DBL_MIN);
    const double ctol = ctol_rel * tn;
    if (j0 > 0 && w[j0] - w[j0 - 1] <= ctol) return;   // not a cluster head
    int j1 = j0;
    while (j1 + 1 < nband && w[j1 + 1] - w[j1] <= ctol) ++j1;
    const double eps = DBL_EPSILON;
    const double pertol = 10.0 * eps * tn;
    const double tiny = eps * tn;
    double lam_prev = -1e300;
    for (int j = j0; j <= j1; ++j) {
        double lam = w[j];
        if (j > j0 && lam - lam_prev < pertol) lam = lam_prev + pertol;
        lam_prev = lam;
        // LU of T - lam I with partial pivoting (thread 0, sequential)
        if (threadIdx.x == 0) {
            for (int i = 0; i < n; ++i) a[i] = d[i] - lam;
            for (int i = 0; i < n - 1; ++i) {
                b[i] = e[i];
                c[i] = e[i];
            }
            for (int k = 0; k < n - 1; ++k) {
                if (fabs(c[k]) <= fabs(a[k])) {
                    piv[k] = 0;
                    double ak = a[k];
                    if (fabs(ak) < tiny) ak = (ak >= 0 ? tiny : -tiny);
                    a[k] = ak;
                    const double mult = c[k] / ak;
                    c[k] = mult;
                    a[k + 1] -= mult * b[k];
                    if (k < n - 2) dd[k] = 0.0;
                } else {
                    piv[k] = 1;
                    const double mult = a[k] / c[k];
                    a[k] = c[k];
                    const double temp = a[k + 1];
                    a[k + 1] = b[k] - mult * temp;
                    if (k < n - 2) {
                        dd[k] = b[k + 1];
                        b[k + 1] = -mult * dd[k];
                    }
                    b[k] = temp;
                    c[k] = mult;
                }
            }
            if (fabs(a[n - 1]) < tiny) a[n - 1] = (a[n - 1] >= 0 ? tiny : -tiny);
        }
        // start vector
        for (int i = threadIdx.x; i < n; i += NT) {
            const unsigned h = hash_u32((unsigned)(i * 2654435761u) ^ (unsigned)(j * 40503u + 17u));
            xv[i] = ((double)h / 4294967296.0) - 0.5;
        }
        __syncthreads();
        for (int it = 0; it < 3; ++it) {
            if (threadIdx.x == 0) {
                // forward: L
                for (int k = 0; k < n - 1; ++k) {
                    if (piv[k] == 0) {
                        xv[k + 1] -= c[k] * xv[k];
                    } else {
                        const double t = xv[k];
                        xv[k] = xv[k + 1];
                        xv[k + 1] = t - c[k] * xv[k];
                    }
                }
                // backward: U (diag a, super b, second super dd)
                xv[n - 1] /= a[n - 1];
                if (n > 1) xv[n - 2] = (xv[n - 2] - b[n - 2] * xv[n - 1]) / a[n - 2];
                for (int k = n - 3; k >= 0; --k) xv[k] = (xv[k] - b[k] * xv[k + 1] - dd[k] * xv[k + 2]) / a[k];
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

// vcc[:, j] = Q z_j, Q = P_0 P_1 ... P_{n-2}; one workgroup per column.
__global__ __launch_bounds__(NT) void k_backtransform(int n, const double2* __restrict__ V,
                                                      const double2* __restrict__ tau, const double* __restrict__ Z,
                                                      double2* __restrict__ vcc, int ldv) {
    extern __shared__ __align__(16) unsigned char smem[];
    double2* z = reinterpret_cast<double2*>(smem);
    double* red = reinterpret_cast<double*>(z + n);
    const int j = blockIdx.x;
    for (int i = threadIdx.x; i < n; i += NT) z[i] = make_double2(Z[(size_t)j * n + i], 0.0);
    __syncthreads();
    for (int k = n - 2; k >= 0; --k) {
        const double2 t = tau[k];
        if (t.x == 0.0 && t.y == 0.0) continue;
        const double2* vk = V + (size_t)k * n;   // rows k+1..n-1
        double sr = 0, si = 0;
        for (int i = k + 1 + threadIdx.x; i < n; i += NT) {
            const double2 q = cconjmul(vk[i], z[i]);
            sr += q.x;
            si += q.y;
        }
        const double2 dot = block_sum2(sr, si, red);
        const double2 f = cmul(t, dot);
        for (int i = k + 1 + threadIdx.x; i < n; i += NT) {
            const double2 q = cmul(vk[i], f);
            z[i].x -= q.x;
            z[i].y -= q.y;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += NT) vcc[(size_t)j * ldv + i] = z[i];
}

// ---- Cholesky G = R^H R and R^-1 (single workgroup) ----
// W (n x n, ld n) work copy of G; on exit rinv holds R^-1 (upper, zeros below).
__global__ __launch_bounds__(1024) void k_chol_inv(int n, const double2* __restrict__ g, int ldg,
                                                   double2* __restrict__ W, double2* __restrict__ rinv,
                                                   int* __restrict__ status, double min_norm) {
    for (size_t t = threadIdx.x; t < (size_t)n * n; t += blockDim.x) {
        const int r = (int)(t % n), c = (int)(t / n);
        W[t] = g[r + (size_t)c * ldg];
    }
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        const double piv2 = W[j + (size_t)j * n].x;
        const double piv = piv2 > 0 ? sqrt(piv2) : 0.0;
        if (!(piv2 > 0) || piv < min_norm) {
            if (threadIdx.x == 0) bad = j + 1;
            __syncthreads();
            break;
        }
        const double ip = 1.0 / piv;
        // row j of R: R[j][c] = W[j][c] / piv for c > j
        for (int c = j + 1 + threadIdx.x; c < n; c += blockDim.x) {
            double2 x = W[j + (size_t)c * n];
            W[j + (size_t)c * n] = make_double2(x.x * ip, x.y * ip);
        }
        if (threadIdx.x == 0) W[j + (size_t)j * n] = make_double2(piv, 0.0);
        __syncthreads();
        // trailing update of the upper triangle: W[r][c] -= conj(R[j][r]) R[j][c], j < r <= c
        const int m = n - j - 1;
        for (size_t t = threadIdx.x; t < (size_t)m * m; t += blockDim.x) {
            const int r = j + 1 + (int)(t % m), c = j + 1 + (int)(t / m);
            if (r > c) continue;
            const double2 a = W[j + (size_t)r * n], b = W[j + (size_t)c * n];
            const double2 q = cconjmul(a, b);
            double2 x = W[r + (size_t)c * n];
            x.x -= q.x;
            x.y -= q.y;
            W[r + (size_t)c * n] = x;
        }
        __syncthreads();
    }
    if (bad) {
        if (threadIdx.x == 0) status[0] = bad;
        return;
    }
    // R^-1 column by column (thread per column): X[c][c] = 1/R[c][c];
    // X[r][c] = -(sum_{k=r+1..c} R[r][k] X[k][c]) / R[r][r]
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        double2* X = rinv + (size_t)c * n;
        for (int r = c + 1; r < n; ++r) X[r] = make_double2(0.0, 0.0);
        X[c] = make_double2(1.0 / W[c + (size_t)c * n].x, 0.0);
        for (int r = c - 1; r >= 0; --r) {
            double sr = 0, si = 0;
            for (int k = r + 1; k <= c; ++k) {
                const double2 q = cmul(W[r + (size_t)k * n], X[k]);
                sr += q.x;
                si += q.y;
            }
            const double id = 1.0 / W[r + (size_t)r * n].x;
            X[r] = make_double2(-sr * id, -si * id);
        }
    }
}

}  // namespace

void heevx_lowest(hipStream_t s, EigWorkspace& ws, int num_cu, int n, int ldh, const double* hcc, int nband,
                  double* eig, double* vcc, int ldv, Prof* prof) {
    APW_CHECK(n >= 1 && nband >= 1 && nband <= n && ldh >= n && ldv >= n, "heevx: bad sizes n=%d nband=%d", n, nband);
    APW_CHECK(n <= 1300, "heevx: n=%d too large (LDS of the reduction kernel)", n);
    ws.a.reserve((size_t)2 * n * n);
    ws.v.reserve((size_t)2 * n * n);
    ws.tau.reserve((size_t)2 * n);
    ws.d.reserve(n);
    ws.e.reserve(n);
    ws.p.reserve((size_t)4 * n);
    ws.z.reserve((size_t)n * nband);
    ws.bar.reserve(2);
    double2* A = reinterpret_cast<double2*>(ws.a.get());
    hipLaunchKernelGGL(k_copy_herm, dim3(std::min(4096, (n * n + 255) / 256)), dim3(256), 0, s, n,
                       reinterpret_cast<const double2*>(hcc), ldh, A);
    APW_LAUNCH_CHECK();
    if (n > 1) {
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, ws.bar.get(), 2);
        APW_LAUNCH_CHECK();
        const size_t lds = (size_t)3 * n * sizeof(double2) + 64 * sizeof(double);
        int occ = 0;
        APW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k_tridiag), NT, lds));
        int nwg = std::max(1, std::min(n / 16, std::min(num_cu, 64)));
        nwg = std::min(nwg, std::max(1, occ) * num_cu);
        double2* Vp = reinterpret_cast<double2*>(ws.v.get());
        double2* taup = reinterpret_cast<double2*>(ws.tau.get());
        double* dp = ws.d.get();
        double* ep = ws.e.get();
        double2* Pp = reinterpret_cast<double2*>(ws.p.get());
        unsigned* barp = ws.bar.get();
        void* args[] = {&A, &n, &Vp, &taup, &dp, &ep, &Pp, &barp};
        // algorithmic: per column a hemv read and a rank-2 read+write of the trailing matrix
        ProfScope ps(prof, s, "eig_tridiag", 16.0 * 3.0 * n * (double)n * n / 3.0, 16.0 / 3.0 * n * (double)n * n);
        APW_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_tridiag), dim3(nwg), dim3(NT), args,
                                           (unsigned)lds, s));
    } else {
        // n == 1: T = [Re a00]; the real part is the first double of A
        APW_HIP(hipMemcpyAsync(ws.d.get(), A, sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    {
    ProfScope ps(prof, s, "eig_bisect_invit", 0.0, 0.0);
    hipLaunchKernelGGL(k_bisect, dim3((nband + 3) / 4), dim3(256), 0, s, ws.d.get(), ws.e.get(), n, nband, eig);
    APW_LAUNCH_CHECK();
    const size_t lds_inv = (size_t)5 * n * sizeof(double) + 64 * sizeof(double) + (size_t)n * sizeof(int);
    hipLaunchKernelGGL(k_invit, dim3(nband), dim3(NT), lds_inv, s, ws.d.get(), ws.e.get(), n, nband, eig, ws.z.get(),
                       1e-7);
    APW_LAUNCH_CHECK();
    }
    const size_t lds_bt = (size_t)n * sizeof(double2) + 64 * sizeof(double);
    ProfScope pb(prof, s, "eig_backtransform", 16.0 * n * (double)n / 2.0 + 16.0 * n * nband,
                 8.0 * n * (double)n * nband);
    hipLaunchKernelGGL(k_backtransform, dim3(nband), dim3(NT), lds_bt, s, n,
                       reinterpret_cast<const double2*>(ws.v.get()), reinterpret_cast<const double2*>(ws.tau.get()),
                       ws.z.get(), reinterpret_cast<double2*>(vcc), ldv);
    APW_LAUNCH_CHECK();
}

void chol_inv_upper(hipStream_t s, int n, const double* g, int ldg, double* rinv, int* status, double min_norm) {
    // work copy lives in rinv's tail: caller passes rinv with room for 2 * n * n c128
    double2* W = reinterpret_cast<double2*>(rinv) + (size_t)n * n;
    hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(1024), 0, s, n, reinterpret_cast<const double2*>(g), ldg, W,
                       reinterpret_cast<double2*>(rinv), status, min_norm);
    APW_LAUNCH_CHECK();
}

}  // namespace apw

extern "C" int apw_dnevx_op(apw_ctx* ctx, int nstart, int ldh, const double* hcc, int nbands, double* eigenvalue,
                            double* vcc) {
    return apw::guarded([&] {
        APW_CHECK(ctx && hcc && eigenvalue && vcc, "apw_dnevx_op: null argument");
        apw::DeviceGuard g(ctx->device);
        apw::EigWorkspace ws;
        apw::heevx_lowest(ctx->stream, ws, ctx->num_cu, nstart, ldh, hcc, nbands, eigenvalue, vcc, ldh);
        APW_HIP(hipStreamSynchronize(ctx->stream));
        unsigned flag[2] = {0, 0};
        if (nstart > 1) APW_HIP(hipMemcpy(flag, ws.bar.get(), sizeof(flag), hipMemcpyDeviceToHost));
        APW_CHECK(flag[1] == 0, "apw_dnevx_op: grid barrier timed out");
        return APW_OK;
    });
}
