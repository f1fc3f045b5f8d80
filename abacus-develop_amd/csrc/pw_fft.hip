// Plane-wave 3D FFT passes for the H|psi> path (gfx950).
//
// Replaces, per band, the reference GPU sequence (pw_transform_k.cpp:336-476):
//   memset(box) -> set_3d_fft_box -> hipfftExecZ2Z(BACKWARD) -> copy -> veff_pw_op ->
//   copy -> hipfftExecZ2Z(FORWARD) -> set_real_to_recip_output
// with five batched passes that only touch the part of the box the sphere reaches:
//   1. z-backward on the ns sticks, gathering psi straight from the sphere;
//   2. y-backward on the x columns that hold sticks (nxr of nx), scattering the sticks;
//   3. x-backward, multiply by veff, x-forward, in one kernel (the only pass over the
//      full grid; the real-space box never goes to HBM);
//   4. y-forward, keeping the stick rows;
//   5. z-forward on the sticks, scaled by 1/nxyz and gathered into hpsi, with the kinetic
//      term gk2*tpiba2*psi fused in.
// Each pass is a batch of 1D transforms.  For the lengths compiled in (APW_FFT_SIZES) a
// length N = NA * NB transform runs as a four-step in registers: NB DFTs of length NA per
// line in the threads' registers, twiddles, one exchange through LDS, NA DFTs of length NB,
// every index and inner constant fixed at compile time; so one line crosses LDS once per
// direction, and the fused x pass twice in all.  Global accesses are runs of consecutive z
// (16 lines x 16 B) or of one stick, and the sphere's holes are read as zeros by raw
// buffer loads (no branches).  Other 2-3-5 lengths run a mixed-radix (4, 2, 3, 5) Stockham
// FFT in LDS with the plan read at run time.  Twiddles exp(-2 pi i k / n) are computed on
// the host in long double.
#include "pw_fft.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <type_traits>

namespace apw {

FftPlan make_plan(int n) {
    FftPlan p;
    p.n = n;
    int m = n;
    while (m % 4 == 0) { p.rad[p.nst++] = 4; m /= 4; }
    while (m % 2 == 0) { p.rad[p.nst++] = 2; m /= 2; }
    while (m % 3 == 0) { p.rad[p.nst++] = 3; m /= 3; }
    while (m % 5 == 0) { p.rad[p.nst++] = 5; m /= 5; }
    APW_CHECK(m == 1 && n > 0, "FFT length %d is not 2^a 3^b 5^c", n);
    APW_CHECK(p.nst <= 16, "FFT length %d: too many stages", n);
    return p;
}

namespace {

template <class R>
using cplx = typename std::conditional<std::is_same<R, double>::value, double2, float2>::type;

template <class C>
__device__ inline C cmul(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <class C>
__device__ inline C cadd(C a, C b) { C r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <class C>
__device__ inline C csub(C a, C b) { C r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// multiply by sgn * i
template <class C>
__device__ inline C cmuli(C a, int sgn) {
    C r;
    r.x = -sgn * a.y;
    r.y = sgn * a.x;
    return r;
}

template <class C, class R>
__device__ inline void dft_radix(C* v, int rdx, int sgn) {
    if (rdx == 2) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (rdx == 4) {
        C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        C t2 = cadd(v[1], v[3]), t3 = cmuli(csub(v[1], v[3]), sgn);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else if (rdx == 3) {
        const R h = (R)0.86602540378443864676;   // sqrt(3)/2
        C t = cadd(v[1], v[2]);
        C u = csub(v[1], v[2]);
        C a = v[0];
        C m;
        m.x = a.x - (R)0.5 * t.x;
        m.y = a.y - (R)0.5 * t.y;
        C w = cmuli(u, sgn);
        w.x *= h;
        w.y *= h;
        v[0] = cadd(a, t);
        v[1] = cadd(m, w);
        v[2] = csub(m, w);
    } else {  // 5
        const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
        const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
        C a = v[0];
        C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        C u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
        C m1, m2, q1, q2;
        m1.x = a.x + c1 * t1.x + c2 * t2.x;
        m1.y = a.y + c1 * t1.y + c2 * t2.y;
        m2.x = a.x + c2 * t1.x + c1 * t2.x;
        m2.y = a.y + c2 * t1.y + c1 * t2.y;
        q1.x = s1 * u1.x + s2 * u2.x;
        q1.y = s1 * u1.y + s2 * u2.y;
        q2.x = s2 * u1.x - s1 * u2.x;
        q2.y = s2 * u1.y - s1 * u2.y;
        q1 = cmuli(q1, sgn);
        q2 = cmuli(q2, sgn);
        v[0] = cadd(a, cadd(t1, t2));
        v[1] = cadd(m1, q1);
        v[4] = csub(m1, q1);
        v[2] = cadd(m2, q2);
        v[3] = csub(m2, q2);
    }
}

// In-LDS Stockham FFT of L interleaved lines of length p.n: element k of line l at
// x[k * ld + l].  sgn = +1 backward (exp +i), -1 forward.  tw: exp(-2 pi i j / n), j < n.
// Returns the buffer (x or y) that holds the result.  All threads of the block call it.
template <class R>
__device__ cplx<R>* lds_fft(cplx<R>* x, cplx<R>* y, const cplx<R>* tw, const FftPlan& p, int L, int ld, int sgn) {
    using C = cplx<R>;
    const int n = p.n;
    int Ns = 1;
    for (int st = 0; st < p.nst; ++st) {
        const int rdx = p.rad[st];
        const int nb = n / rdx;
        const int stride = n / (Ns * rdx);
        for (int b = threadIdx.x; b < nb * L; b += blockDim.x) {
            const int l = b % L;
            const int j = b / L;
            const int jm = j % Ns;
            C v[5];
            for (int r = 0; r < rdx; ++r) v[r] = x[(j + r * nb) * ld + l];
            if (Ns > 1) {
                for (int r = 1; r < rdx; ++r) {
                    C t = tw[(r * jm * stride) % n];
                    if (sgn > 0) t.y = -t.y;
                    v[r] = cmul(v[r], t);
                }
            }
            dft_radix<C, R>(v, rdx, sgn);
            const int base = (j / Ns) * Ns * rdx + jm;
            for (int r = 0; r < rdx; ++r) y[(base + r * Ns) * ld + l] = v[r];
        }
        __syncthreads();
        C* t = x;
        x = y;
        y = t;
        Ns *= rdx;
    }
    return x;
}

template <class R>
__device__ inline void load_tw(cplx<R>* tws, const cplx<R>* tw, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) tws[i] = tw[i];
}

template <class R>
__device__ inline cplx<R> to_c(double2 v) {
    cplx<R> r;
    r.x = (R)v.x;
    r.y = (R)v.y;
    return r;
}

constexpr int NT = 256;

// ---- compile-time lengths ---------------------------------------------------------------
// Lengths with register-resident kernels (every FFT length of the named workloads); other
// 2-3-5 lengths run the LDS Stockham kernels with the plan read at run time.
#define APW_FFT_SIZES(X) X(12) X(16) X(18) X(20) X(24) X(25) X(27) X(30) X(32) X(36) X(40) X(45) \
    X(48) X(50) X(54) X(60) X(64) X(72) X(80) X(90) X(96) X(100) X(120) X(128) X(144) X(160)

// lines per workgroup: two LDS buffers of n * (L+1) complex doubles within ~64 KB
constexpr int lines_for(int n) {
    int L = 16;
    while (L > 1 && 2 * n * (L + 1) * 16 > 66000) L /= 2;
    return L;
}


// ---- register-resident transforms ------------------------------------------------------
// A length-P DFT held in one thread's registers, with every index and constant fixed at
// compile time.  The x pass splits its length N = NA * NB (four-step: NB transforms of
// length NA, twiddles W_N^(n2 k1), one exchange through LDS, NA transforms of length NB),
// so each line crosses LDS once per direction instead of once per radix stage.

// cos and sin of 2 pi m / p, evaluated at compile time: exact rational reduction to the
// first octant, then the Taylor series in double-double arithmetic (Dekker's products), so
// the constants are the correctly rounded values to within one unit in the last place.
struct CtDD { double hi, lo; };
struct CtCis { double c, s; };
constexpr CtDD ct_two_sum(double a, double b) {
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
constexpr CtDD ct_split(double a) {
    const double c = 134217729.0 * a;
    const double hi = c - (c - a);
    return {hi, a - hi};
}
constexpr CtDD ct_two_prod(double a, double b) {
    const double p = a * b;
    const CtDD x = ct_split(a), y = ct_split(b);
    return {p, ((x.hi * y.hi - p) + x.hi * y.lo + x.lo * y.hi) + x.lo * y.lo};
}
constexpr CtDD ct_add(CtDD a, CtDD b) {
    const CtDD s = ct_two_sum(a.hi, b.hi);
    return ct_two_sum(s.hi, s.lo + a.lo + b.lo);
}
constexpr CtDD ct_mul(CtDD a, CtDD b) {
    const CtDD p = ct_two_prod(a.hi, b.hi);
    return ct_two_sum(p.hi, p.lo + a.hi * b.lo + a.lo * b.hi);
}
constexpr CtDD ct_div(CtDD a, double d) {
    const double q = a.hi / d;
    const CtDD p = ct_two_prod(q, d);
    return ct_two_sum(q, ((a.hi - p.hi) - p.lo + a.lo) / d);
}
constexpr CtDD ct_series(CtDD x, bool sine) {   // sin or cos of x in [0, pi/4]
    const CtDD x2 = ct_mul(x, x);
    CtDD t = sine ? x : CtDD{1.0, 0.0};
    CtDD s = t;
    for (int k = 1; k <= 14; ++k) {
        const double d = sine ? (2.0 * k) * (2.0 * k + 1.0) : (2.0 * k - 1.0) * (2.0 * k);
        t = ct_div(ct_mul(t, x2), -d);
        s = ct_add(s, t);
    }
    return s;
}
constexpr CtCis ct_cis(long m, long p) {
    m %= p;
    if (m < 0) m += p;
    const long oct = 8 * m / p, r = 8 * m - oct * p;   // 2 pi m / p = (oct + r / p) pi / 4
    const CtDD quarter_pi = {0.78539816339744830962, 3.0616169978683830179e-17};
    const CtDD x = ct_mul(quarter_pi, ct_div(CtDD{(double)((oct % 2 == 0) ? r : p - r), 0.0}, (double)p));
    const double c = ct_series(x, false).hi, s = ct_series(x, true).hi;
    if (oct % 2 == 0) {
        switch (oct / 2) {
            case 0: return {c, s};
            case 1: return {-s, c};
            case 2: return {-c, -s};
            default: return {s, -c};
        }
    }
    switch ((oct + 1) / 2) {   // angle = q pi / 2 - x
        case 1: return {s, c};
        case 2: return {-c, s};
        case 3: return {-s, -c};
        default: return {c, -s};
    }
}

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}
#define APW_CI(c) decltype(c)::value

// v * exp(2 pi i M / P); quarter turns are exact
template <class C, class R, int P, int M>
__device__ __forceinline__ C rot(C v) {
    constexpr int m = ((M % P) + P) % P;
    C r;
    if constexpr (m == 0) {
        r = v;
    } else if constexpr (2 * m == P) {
        r.x = -v.x;
        r.y = -v.y;
    } else if constexpr (4 * m == P) {
        r.x = -v.y;
        r.y = v.x;
    } else if constexpr (4 * m == 3 * P) {
        r.x = v.y;
        r.y = -v.x;
    } else {
        constexpr CtCis w = ct_cis(m, P);
        r.x = v.x * (R)w.c - v.y * (R)w.s;
        r.y = v.x * (R)w.s + v.y * (R)w.c;
    }
    return r;
}

constexpr int ct_first_radix(int p) { return p % 4 == 0 ? 4 : p % 5 == 0 ? 5 : p % 3 == 0 ? 3 : 2; }

// v[k] <- sum_n v[n] exp(SG 2 pi i n k / P), natural order in and out (P = 2^a 3^b 5^c)
template <class C, class R, int P, int SG>
__device__ __forceinline__ void rdft(C* v) {
    if constexpr (P == 2 || P == 3 || P == 4 || P == 5) {
        dft_radix<C, R>(v, P, SG);
    } else if constexpr (P > 1) {
        constexpr int PA = ct_first_radix(P), PB = P / PA;
        C a[P];
        static_for<0, PB>([&](auto n2c) {
            C t[PA];
            static_for<0, PA>([&](auto n1c) { t[APW_CI(n1c)] = v[PB * APW_CI(n1c) + APW_CI(n2c)]; });
            rdft<C, R, PA, SG>(t);
            static_for<0, PA>([&](auto k1c) {
                a[APW_CI(n2c) * PA + APW_CI(k1c)] = rot<C, R, P, SG * APW_CI(n2c) * APW_CI(k1c)>(t[APW_CI(k1c)]);
            });
        });
        static_for<0, PA>([&](auto k1c) {
            C t[PB];
            static_for<0, PB>([&](auto n2c) { t[APW_CI(n2c)] = a[APW_CI(n2c) * PA + APW_CI(k1c)]; });
            rdft<C, R, PB, SG>(t);
            static_for<0, PB>([&](auto k2c) { v[APW_CI(k1c) + PA * APW_CI(k2c)] = t[APW_CI(k2c)]; });
        });
    }
}

// x-pass split N = NA * NB: the largest NA with NA <= NB <= 16 (0: none)
constexpr int ct_split_a(int n) {
    for (int a = 16; a >= 2; --a)
        if (n % a == 0 && a * a <= n && n / a <= 16) return a;
    return 0;
}
constexpr int XL = 16;   // lines (consecutive z) per workgroup of the register x pass
constexpr int xreg_threads(int n) { return XL * (n / ct_split_a(n)); }
#define APW_XREG_CHECK(N) static_assert(ct_split_a(N) != 0, "x length without a register split");
APW_FFT_SIZES(APW_XREG_CHECK)
#undef APW_XREG_CHECK

// ---- pass 1: z-backward on sticks, gathered from the sphere --------------------------
// grid (ceil(ns/L), nb); LDS: 2 * nz * (L+1) + nz complex
template <class R>
__global__ __launch_bounds__(NT) void k_zbwd(const double2* __restrict__ psi, int ld, const int* __restrict__ stick_ig,
                                             int ns, cplx<R>* __restrict__ S, FftPlan pz,
                                             const cplx<R>* __restrict__ tw, int Lrt) {
    using C = cplx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int nz = pz.n;
    const int L = Lrt;
    const int ldl = L + 1;
    C* x = reinterpret_cast<C*>(smem);
    C* y = x + nz * ldl;
    C* tws = y + nz * ldl;
    const int band = blockIdx.y;
    const int s0 = blockIdx.x * L;
    const double2* pb = psi + (size_t)band * ld;
    load_tw<R>(tws, tw, nz);
    for (int t = threadIdx.x; t < L * nz; t += NT) {
        const int l = t / nz, iz = t - (t / nz) * nz;
        const int s = s0 + l;
        C v;
        v.x = 0;
        v.y = 0;
        if (s < ns) {
            const int ig = stick_ig[(size_t)s * nz + iz];
            if (ig >= 0) v = to_c<R>(pb[ig]);
        }
        x[iz * ldl + l] = v;
    }
    __syncthreads();
    C* res = lds_fft<R>(x, y, tws, pz, L, ldl, +1);
    C* Sb = S + (size_t)band * ns * nz;
    for (int t = threadIdx.x; t < L * nz; t += NT) {
        const int l = t / nz, iz = t - (t / nz) * nz;
        const int s = s0 + l;
        if (s < ns) Sb[(size_t)s * nz + iz] = res[iz * ldl + l];
    }
}

// ---- pass 2: y-backward on the x slots -----------------------------------------------
// grid (nxr, ceil(nz/L), nb); LDS 2 * ny * (L+1) + ny
template <class R>
__global__ __launch_bounds__(NT) void k_ybwd(const cplx<R>* __restrict__ S, int ns, int nz, int nxr,
                                             const int* __restrict__ slot_start, const int* __restrict__ slot_cnt,
                                             const int* __restrict__ stick_iy, cplx<R>* __restrict__ X, FftPlan py,
                                             const cplx<R>* __restrict__ tw, int Lrt) {
    using C = cplx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int ny = py.n;
    const int L = Lrt;
    const int ldl = L + 1;
    C* x = reinterpret_cast<C*>(smem);
    C* y = x + ny * ldl;
    C* tws = y + ny * ldl;
    const int slot = blockIdx.x;
    const int iz0 = blockIdx.y * L;
    const int band = blockIdx.z;
    load_tw<R>(tws, tw, ny);
    for (int t = threadIdx.x; t < ny * ldl; t += NT) {
        C z;
        z.x = 0;
        z.y = 0;
        x[t] = z;
    }
    __syncthreads();
    const int st0 = slot_start[slot], cnt = slot_cnt[slot];
    const C* Sb = S + (size_t)band * ns * nz;
    for (int t = threadIdx.x; t < cnt * L; t += NT) {
        const int q = t / L, dz = t - (t / L) * L;
        const int iz = iz0 + dz;
        const int s = st0 + q;
        if (iz < nz) x[stick_iy[s] * ldl + dz] = Sb[(size_t)s * nz + iz];
    }
    __syncthreads();
    C* res = lds_fft<R>(x, y, tws, py, L, ldl, +1);
    C* Xb = X + ((size_t)band * nxr + slot) * ny * nz;
    for (int t = threadIdx.x; t < ny * L; t += NT) {
        const int iy = t / L, dz = t - (t / L) * L;
        const int iz = iz0 + dz;
        if (iz < nz) Xb[(size_t)iy * nz + iz] = res[iy * ldl + dz];
    }
}

// ---- pass 3: x pass.  MODE 0: bwd, *veff, fwd (X -> X); 1: bwd, write full box; 2: read full box, fwd
// grid (ny, ceil(nz/L), nb); LDS 2 * nx * (L+1) + nx
template <class R, int MODE>
__global__ __launch_bounds__(NT) void k_xpass(cplx<R>* __restrict__ X, int nxr, int ny, int nz,
                                              const int* __restrict__ slot_ix, const R* __restrict__ veff,
                                              double2* __restrict__ box, int add, double factor, FftPlan px,
                                              const cplx<R>* __restrict__ tw, int Lrt) {
    using C = cplx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = px.n;
    const int L = Lrt;
    const int ldl = L + 1;
    C* x = reinterpret_cast<C*>(smem);
    C* y = x + nx * ldl;
    C* tws = y + nx * ldl;
    const int iy = blockIdx.x;
    const int iz0 = blockIdx.y * L;
    const int band = blockIdx.z;
    load_tw<R>(tws, tw, nx);
    C* Xb = X + (size_t)band * nxr * ny * nz;
    const size_t nxyz = (size_t)nx * ny * nz;
    double2* boxb = box ? box + (size_t)band * nxyz : nullptr;
    if (MODE == 2) {
        for (int t = threadIdx.x; t < nx * L; t += NT) {
            const int ix = t / L, dz = t - (t / L) * L;
            const int iz = iz0 + dz;
            C v;
            v.x = 0;
            v.y = 0;
            if (iz < nz) v = to_c<R>(boxb[((size_t)ix * ny + iy) * nz + iz]);
            x[ix * ldl + dz] = v;
        }
    } else {
        for (int t = threadIdx.x; t < nx * ldl; t += NT) {
            C z;
            z.x = 0;
            z.y = 0;
            x[t] = z;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < nxr * L; t += NT) {
            const int q = t / L, dz = t - (t / L) * L;
            const int iz = iz0 + dz;
            if (iz < nz) x[slot_ix[q] * ldl + dz] = Xb[((size_t)q * ny + iy) * nz + iz];
        }
    }
    __syncthreads();
    C* res = x;
    if (MODE != 2) res = lds_fft<R>(x, y, tws, px, L, ldl, +1);
    if (MODE == 1) {
        for (int t = threadIdx.x; t < nx * L; t += NT) {
            const int ix = t / L, dz = t - (t / L) * L;
            const int iz = iz0 + dz;
            if (iz < nz) {
                const C v = res[ix * ldl + dz];
                double2* bp = boxb + ((size_t)ix * ny + iy) * nz + iz;
                double2 o;
                if (add) {
                    o = *bp;
                    o.x += factor * (double)v.x;
                    o.y += factor * (double)v.y;
                } else {
                    o.x = (double)v.x;
                    o.y = (double)v.y;
                }
                *bp = o;
            }
        }
        return;
    }
    if (MODE == 0) {
        for (int t = threadIdx.x; t < nx * L; t += NT) {
            const int ix = t / L, dz = t - (t / L) * L;
            const int iz = iz0 + dz;
            if (iz < nz) {
                const R vr = veff[((size_t)ix * ny + iy) * nz + iz];
                C v = res[ix * ldl + dz];
                v.x *= vr;
                v.y *= vr;
                res[ix * ldl + dz] = v;
            }
        }
        __syncthreads();
    }
    C* other = (res == x) ? y : x;
    C* out = lds_fft<R>(res, other, tws, px, L, ldl, -1);
    for (int t = threadIdx.x; t < nxr * L; t += NT) {
        const int q = t / L, dz = t - (t / L) * L;
        const int iz = iz0 + dz;
        if (iz < nz) Xb[((size_t)q * ny + iy) * nz + iz] = out[slot_ix[q] * ldl + dz];
    }
}

// ---- density pass: x-backward of every band of the chunk, rho += w_b |psi_b(r)|^2 ------
// grid (ny, ceil(nz/L)); each workgroup owns lines (iy, iz0 .. iz0+L) for all nb bands and
// accumulates in LDS in band order (deterministic, no atomics), then adds to rho once.
// LDS 2 * nx * (L+1) + nx complex + nx * L doubles
__global__ __launch_bounds__(NT) void k_xrho(const cplx<double>* __restrict__ X, int nxr, int ny, int nz, int nb,
                                             const int* __restrict__ slot_ix, const double* __restrict__ w,
                                             double* __restrict__ rho, int add, FftPlan px,
                                             const cplx<double>* __restrict__ tw, int Lrt) {
    using C = cplx<double>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int nx = px.n;
    const int L = Lrt;
    const int ldl = L + 1;
    C* x = reinterpret_cast<C*>(smem);
    C* y = x + nx * ldl;
    C* tws = y + nx * ldl;
    double* acc = reinterpret_cast<double*>(tws + nx);
    const int iy = blockIdx.x;
    const int iz0 = blockIdx.y * L;
    load_tw<double>(tws, tw, nx);
    for (int t = threadIdx.x; t < nx * L; t += NT) acc[t] = 0.0;
    for (int b = 0; b < nb; ++b) {
        const C* Xb = X + (size_t)b * nxr * ny * nz;
        __syncthreads();
        for (int t = threadIdx.x; t < nx * ldl; t += NT) {
            C z;
            z.x = 0;
            z.y = 0;
            x[t] = z;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < nxr * L; t += NT) {
            const int q = t / L, dz = t - (t / L) * L;
            const int iz = iz0 + dz;
            if (iz < nz) x[slot_ix[q] * ldl + dz] = Xb[((size_t)q * ny + iy) * nz + iz];
        }
        __syncthreads();
        const C* res = lds_fft<double>(x, y, tws, px, L, ldl, +1);
        const double wb = w[b];
        for (int t = threadIdx.x; t < nx * L; t += NT) {
            const int ix = t / L, dz = t - (t / L) * L;
            const C v = res[ix * ldl + dz];
            acc[t] += wb * (v.x * v.x + v.y * v.y);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nx * L; t += NT) {
        const int ix = t / L, dz = t - (t / L) * L;
        const int iz = iz0 + dz;
        if (iz < nz) {
            double* rp = rho + ((size_t)ix * ny + iy) * nz + iz;
            *rp = add ? *rp + acc[t] : acc[t];
        }
    }
}

// ---- raw buffer access: a lane whose byte offset is outside [0, bytes) reads zeros and its
// store is dropped, so gathers and scatters with holes need no branches (and no waits per
// element).  Descriptor word 3 = 0x00020000 (32-bit data format), as for every gfx9 part.
constexpr unsigned BUF_OOB = 0x80000000u;
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ inline double2 buf_load(__amdgpu_buffer_rsrc_t r, unsigned off, double2*) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    double2 d;
    __builtin_memcpy(&d, &v, 16);
    return d;
}
__device__ inline float2 buf_load(__amdgpu_buffer_rsrc_t r, unsigned off, float2*) {
    const u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    float2 d;
    __builtin_memcpy(&d, &v, 8);
    return d;
}
__device__ inline void buf_store(double2 d, __amdgpu_buffer_rsrc_t r, unsigned off) {
    u32x4 v;
    __builtin_memcpy(&v, &d, 16);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 0);
}
__device__ inline void buf_store(float2 d, __amdgpu_buffer_rsrc_t r, unsigned off) {
    u32x2 v;
    __builtin_memcpy(&v, &d, 8);
    __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, 0);
}

// ---- pass 3, register form (every x length with a compile-time split) ------------------
// Workgroup: XL consecutive z lines at one iy; thread (l, g): l = line, g < NB.
//   step A (all threads, g = n2):  x = NB n1 + g, n1 < NA, DFT_NA, twiddle W_N^(+-g k1)
//   LDS exchange
//   step B (threads g < NA, g = k1): DFT_NB, giving x = g + NA k2
// MODE 0 runs A, B, * veff, then the forward transform with the roles swapped (B', A'),
// so the data crosses LDS twice in all.  MODE 1: A, B, store into the full box (add:
// box += factor * result); MODE 2: load the full box, B', A'.
// grid (ny, ceil(nz / XL), nb); static LDS N * XL complex
// The sphere's x slots are at most two runs of consecutive x (sorted order): slot of x =
// x - r.x for x in [r.x, r.x + r.y), r.y + x - r.z for x in [r.z, r.z + r.w), else none.
__device__ inline int xrun_slot(int4 r, int x) {
    return (unsigned)(x - r.x) < (unsigned)r.y ? x - r.x : ((unsigned)(x - r.z) < (unsigned)r.w ? r.y + x - r.z : -1);
}

template <class R, int N, int MODE>
__global__ __launch_bounds__(xreg_threads(N)) void k_xreg(cplx<R>* __restrict__ X, int nxr, int ny, int nz, int4 xrun,
                                                         const R* __restrict__ veff, double2* __restrict__ box,
                                                         int add, double factor, const cplx<R>* __restrict__ tw) {
    using C = cplx<R>;
    constexpr int NA = ct_split_a(N), NB = N / NA;
    __shared__ C buf[N * XL];
    __shared__ C tws[N];
    const int l = threadIdx.x % XL, g = threadIdx.x / XL;
    const int iy = blockIdx.x, iz = blockIdx.y * XL + l;
    const bool zin = iz < nz;
    const size_t plane = (size_t)ny * nz, off = (size_t)iy * nz + (zin ? iz : 0);
    const size_t band = blockIdx.z;
    const bool mid = g < NA;
    const __amdgpu_buffer_rsrc_t xr = make_rsrc(X + band * nxr * plane, (unsigned)(nxr * plane * sizeof(C)));
    double2* boxb = (MODE != 0) ? box + band * (size_t)N * plane : nullptr;
    // every global load is issued before the one wait at the barrier
    unsigned qo[NA];   // byte offsets of x = NB n1 + g in this band's slots (BUF_OOB: no slot)
    static_for<0, NA>([&](auto n1) {
        const int q = xrun_slot(xrun, NB * APW_CI(n1) + g);
        qo[APW_CI(n1)] = (zin && q >= 0) ? (unsigned)(((size_t)q * plane + off) * sizeof(C)) : BUF_OOB;
    });
    C v[NA], u[NB];
    R ve[NB];
    if constexpr (MODE != 2) {
        static_for<0, NA>([&](auto n1) { v[APW_CI(n1)] = buf_load(xr, qo[APW_CI(n1)], (C*)nullptr); });
    }
    if constexpr (MODE == 0) {
        if (mid) static_for<0, NB>([&](auto k2) { ve[APW_CI(k2)] = veff[(size_t)(g + NA * APW_CI(k2)) * plane + off]; });
    }
    if constexpr (MODE == 2) {
        if (mid) {
            static_for<0, NB>([&](auto k2) {
                C t = to_c<R>(boxb[(size_t)(g + NA * APW_CI(k2)) * plane + off]);
                if (!zin) t.x = t.y = 0;
                u[APW_CI(k2)] = t;
            });
        }
    }
    for (int t = threadIdx.x; t < N; t += xreg_threads(N)) tws[t] = tw[t];
    __syncthreads();
    if constexpr (MODE != 2) {
        rdft<C, R, NA, +1>(v);
        static_for<1, NA>([&](auto k1) {
            C w = tws[g * APW_CI(k1)];   // g k1 < N
            w.y = -w.y;
            v[APW_CI(k1)] = cmul(v[APW_CI(k1)], w);
        });
        static_for<0, NA>([&](auto k1) { buf[(APW_CI(k1) * NB + g) * XL + l] = v[APW_CI(k1)]; });
        __syncthreads();
        if (mid) {
            static_for<0, NB>([&](auto n2) { u[APW_CI(n2)] = buf[(g * NB + APW_CI(n2)) * XL + l]; });
            rdft<C, R, NB, +1>(u);
        }
        if constexpr (MODE == 1) {
            if (mid && zin) {
                static_for<0, NB>([&](auto k2) {
                    double2* bp = boxb + (size_t)(g + NA * APW_CI(k2)) * plane + off;
                    const C r = u[APW_CI(k2)];
                    double2 o;
                    if (add) {
                        o = *bp;
                        o.x += factor * (double)r.x;
                        o.y += factor * (double)r.y;
                    } else {
                        o.x = (double)r.x;
                        o.y = (double)r.y;
                    }
                    *bp = o;
                });
            }
            return;
        }
        if (mid) {
            static_for<0, NB>([&](auto k2) {
                u[APW_CI(k2)].x *= ve[APW_CI(k2)];
                u[APW_CI(k2)].y *= ve[APW_CI(k2)];
            });
        }
        __syncthreads();   // every step-B read of buf is done
    }
    if (mid) {
        rdft<C, R, NB, -1>(u);
        static_for<1, NB>([&](auto m1) { u[APW_CI(m1)] = cmul(u[APW_CI(m1)], tws[g * APW_CI(m1)]); });
        static_for<0, NB>([&](auto m1) { buf[(APW_CI(m1) * NA + g) * XL + l] = u[APW_CI(m1)]; });
    }
    __syncthreads();
    static_for<0, NA>([&](auto k1) { v[APW_CI(k1)] = buf[(g * NA + APW_CI(k1)) * XL + l]; });
    rdft<C, R, NA, -1>(v);
    static_for<0, NA>([&](auto m2) { buf_store(v[APW_CI(m2)], xr, qo[APW_CI(m2)]); });
}

// ---- density pass, register form: rho (+)= sum_b w_b |x-backward(psi_b)|^2 ------------
// grid (ny, ceil(nz / XL)); bands in order inside the workgroup, sums in registers.
template <int N>
__global__ __launch_bounds__(xreg_threads(N)) void k_xrho_reg(const cplx<double>* __restrict__ X, int nxr, int ny,
                                                             int nz, int nb, int4 xrun, const double* __restrict__ w,
                                                             double* __restrict__ rho, int add,
                                                             const cplx<double>* __restrict__ tw) {
    using C = cplx<double>;
    constexpr int NA = ct_split_a(N), NB = N / NA;
    __shared__ C buf[N * XL];
    const int l = threadIdx.x % XL, g = threadIdx.x / XL;
    const int iy = blockIdx.x, iz = blockIdx.y * XL + l;
    const bool zin = iz < nz;
    const size_t plane = (size_t)ny * nz, off = (size_t)iy * nz + (zin ? iz : 0);
    const bool mid = g < NA;
    unsigned qo[NA];
    C twa[NA];
    static_for<0, NA>([&](auto n1) {
        const int q = xrun_slot(xrun, NB * APW_CI(n1) + g);
        qo[APW_CI(n1)] = (zin && q >= 0) ? (unsigned)(((size_t)q * plane + off) * sizeof(C)) : BUF_OOB;
    });
    static_for<0, NA>([&](auto k1) {
        C t = tw[g * APW_CI(k1)];
        t.y = -t.y;
        twa[APW_CI(k1)] = t;
    });
    double acc[NB];
    static_for<0, NB>([&](auto k2) { acc[APW_CI(k2)] = 0.0; });
    C v[NA], vn[NA], u[NB];
    const unsigned band_bytes = (unsigned)(nxr * plane * sizeof(C));
    if (nb > 0) {
        const __amdgpu_buffer_rsrc_t xr = make_rsrc(X, band_bytes);
        static_for<0, NA>([&](auto n1) { v[APW_CI(n1)] = buf_load(xr, qo[APW_CI(n1)], (C*)nullptr); });
    }
    for (int b = 0; b < nb; ++b) {
        if (b + 1 < nb) {   // the next band's loads fly while this band is transformed
            const __amdgpu_buffer_rsrc_t xr = make_rsrc(X + (size_t)(b + 1) * nxr * plane, band_bytes);
            static_for<0, NA>([&](auto n1) { vn[APW_CI(n1)] = buf_load(xr, qo[APW_CI(n1)], (C*)nullptr); });
        }
        rdft<C, double, NA, +1>(v);
        static_for<1, NA>([&](auto k1) { v[APW_CI(k1)] = cmul(v[APW_CI(k1)], twa[APW_CI(k1)]); });
        __syncthreads();   // the previous band's reads of buf are done
        static_for<0, NA>([&](auto k1) { buf[(APW_CI(k1) * NB + g) * XL + l] = v[APW_CI(k1)]; });
        __syncthreads();
        if (mid) {
            static_for<0, NB>([&](auto n2) { u[APW_CI(n2)] = buf[(g * NB + APW_CI(n2)) * XL + l]; });
            rdft<C, double, NB, +1>(u);
            const double wb = w[b];
            static_for<0, NB>([&](auto k2) {
                const C r = u[APW_CI(k2)];
                acc[APW_CI(k2)] += wb * (r.x * r.x + r.y * r.y);
            });
        }
        static_for<0, NA>([&](auto n1) { v[APW_CI(n1)] = vn[APW_CI(n1)]; });
    }
    if (mid && zin) {
        static_for<0, NB>([&](auto k2) {
            double* rp = rho + (size_t)(g + NA * APW_CI(k2)) * plane + off;
            *rp = add ? *rp + acc[APW_CI(k2)] : acc[APW_CI(k2)];
        });
    }
}

// ---- passes 2 and 4, register form: y transforms of one x slot, XL consecutive z -------
// DIR +1: sticks -> X (every y); DIR -1: X -> the stick rows.  The slot's sticks are at most
// two runs of consecutive y (yrun[slot], same convention as xrun_slot), numbered from
// slot_start[slot] in y order.  Thread (l, g) as in k_xreg.  grid (nxr, ceil(nz / XL), nb)
template <class R, int N, int DIR>
__global__ __launch_bounds__(xreg_threads(N)) void k_yreg(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out,
                                                         int ns, int nz, int nxr, const int* __restrict__ slot_start,
                                                         const int4* __restrict__ yrun,
                                                         const cplx<R>* __restrict__ tw) {
    using C = cplx<R>;
    constexpr int NA = ct_split_a(N), NB = N / NA;
    __shared__ C buf[N * XL];
    __shared__ C tws[N];
    const int l = threadIdx.x % XL, g = threadIdx.x / XL;
    const int slot = blockIdx.x, iz = blockIdx.y * XL + l;
    const bool zin = iz < nz;
    const int izc = zin ? iz : 0;
    const size_t band = blockIdx.z;
    const bool mid = g < NA;
    const int st0 = slot_start[slot];
    const int4 r = yrun[slot];
    const size_t xoff = (band * nxr + slot) * (size_t)N * nz + izc;   // X at y = 0 of this line
    unsigned so[NA];   // byte offsets of the sticks at y = NB n1 + g (BUF_OOB: none)
    static_for<0, NA>([&](auto n1) {
        const int q = xrun_slot(r, NB * APW_CI(n1) + g);
        so[APW_CI(n1)] = (zin && q >= 0) ? (unsigned)(((size_t)(st0 + q) * nz + iz) * sizeof(C)) : BUF_OOB;
    });
    C v[NA], u[NB];
    if constexpr (DIR > 0) {
        const __amdgpu_buffer_rsrc_t sr = make_rsrc(in + band * ns * nz, (unsigned)((size_t)ns * nz * sizeof(C)));
        static_for<0, NA>([&](auto n1) { v[APW_CI(n1)] = buf_load(sr, so[APW_CI(n1)], (C*)nullptr); });
    } else {
        if (mid) static_for<0, NB>([&](auto k2) { u[APW_CI(k2)] = in[xoff + (size_t)(g + NA * APW_CI(k2)) * nz]; });
    }
    for (int t = threadIdx.x; t < N; t += xreg_threads(N)) tws[t] = tw[t];
    __syncthreads();
    if constexpr (DIR > 0) {
        rdft<C, R, NA, +1>(v);
        static_for<1, NA>([&](auto k1) {
            C w = tws[g * APW_CI(k1)];
            w.y = -w.y;
            v[APW_CI(k1)] = cmul(v[APW_CI(k1)], w);
        });
        static_for<0, NA>([&](auto k1) { buf[(APW_CI(k1) * NB + g) * XL + l] = v[APW_CI(k1)]; });
        __syncthreads();
        if (mid) {
            static_for<0, NB>([&](auto n2) { u[APW_CI(n2)] = buf[(g * NB + APW_CI(n2)) * XL + l]; });
            rdft<C, R, NB, +1>(u);
            if (zin) static_for<0, NB>([&](auto k2) { out[xoff + (size_t)(g + NA * APW_CI(k2)) * nz] = u[APW_CI(k2)]; });
        }
    } else {
        if (mid) {
            rdft<C, R, NB, -1>(u);
            static_for<1, NB>([&](auto m1) { u[APW_CI(m1)] = cmul(u[APW_CI(m1)], tws[g * APW_CI(m1)]); });
            static_for<0, NB>([&](auto m1) { buf[(APW_CI(m1) * NA + g) * XL + l] = u[APW_CI(m1)]; });
        }
        __syncthreads();
        static_for<0, NA>([&](auto k1) { v[APW_CI(k1)] = buf[(g * NA + APW_CI(k1)) * XL + l]; });
        rdft<C, R, NA, -1>(v);
        const __amdgpu_buffer_rsrc_t sr = make_rsrc(out + band * ns * nz, (unsigned)((size_t)ns * nz * sizeof(C)));
        static_for<0, NA>([&](auto m2) { buf_store(v[APW_CI(m2)], sr, so[APW_CI(m2)]); });
    }
}

// ---- passes 1 and 5, register form: z transforms of ZL sticks ---------------------------
// Thread (g, l): g = threadIdx % NB runs along z (consecutive lanes, consecutive z), l =
// stick.  DIR +1: psi (sphere, gathered through stick_ig) -> S; DIR -1: S -> hpsi with the
// gather mode of k_zfwd.  LDS rows padded to odd strides (conflict-free 16-byte accesses).
// grid (ceil(ns / ZL), nb)
constexpr int ZL = 16;
constexpr int zreg_threads(int n) { return ZL * (n / ct_split_a(n)); }
constexpr int odd_up(int n) { return n | 1; }

template <class R, int N, int DIR>
__global__ __launch_bounds__(zreg_threads(N)) void k_zreg(const double2* __restrict__ psi, int ld,
                                                         const int* __restrict__ stick_ig,
                                                         const int4* __restrict__ zrun, int ns,
                                                         cplx<R>* __restrict__ S, double2* __restrict__ hpsi, int ldh,
                                                         const double* __restrict__ gk2, double tpiba2, int mode,
                                                         double scale, const cplx<R>* __restrict__ tw) {
    using C = cplx<R>;
    constexpr int NA = ct_split_a(N), NB = N / NA;
    constexpr int NBP = odd_up(NB), NAP = odd_up(NA);
    constexpr int LDZ = odd_up(DIR > 0 ? NA * NBP : NB * NAP);
    __shared__ C buf[ZL * LDZ];
    __shared__ C tws[N];
    const int g = threadIdx.x % NB, l = threadIdx.x / NB;
    const int s = blockIdx.x * ZL + l;
    const bool sin = s < ns;
    const size_t band = blockIdx.y;
    const bool mid = g < NA;
    C* Sl = S + (band * ns + (sin ? s : 0)) * (size_t)N;   // this stick's z line
    C v[NA], u[NB];
    unsigned po[NA];   // byte offsets into the band's sphere of z = NB n1 + g (BUF_OOB: no G)
    if (zrun) {
        // the stick's plane waves as at most two runs of z, consecutive in the sphere:
        // iz in [a0, a0 + la) holds ig = igA + iz - a0, iz in [b0, b0 + lb) holds igB + iz - b0
        const int4 r = zrun[sin ? s : 0];
        const int a0 = r.x & 0xffff, la = r.x >> 16, b0 = r.z & 0xffff, lb = r.z >> 16;
        static_for<0, NA>([&](auto n1) {
            const int iz = NB * APW_CI(n1) + g;
            const int ig = (unsigned)(iz - a0) < (unsigned)la ? r.y + iz - a0
                         : ((unsigned)(iz - b0) < (unsigned)lb ? r.w + iz - b0 : -1);
            po[APW_CI(n1)] = (sin && ig >= 0) ? (unsigned)ig * 16u : BUF_OOB;
        });
    } else {
        const int* sig = stick_ig + (size_t)(sin ? s : 0) * N;
        static_for<0, NA>([&](auto n1) {
            const int ig = sig[NB * APW_CI(n1) + g];
            po[APW_CI(n1)] = (sin && ig >= 0) ? (unsigned)ig * 16u : BUF_OOB;
        });
    }
    double2 add[NA];   // DIR -1: what the result is added to (kinetic term or old hpsi)
    double kin[NA];
    const __amdgpu_buffer_rsrc_t hr = make_rsrc(hpsi + band * ldh, (unsigned)((size_t)ldh * 16));
    if constexpr (DIR > 0) {
        const __amdgpu_buffer_rsrc_t pr = make_rsrc(psi + band * ld, (unsigned)((size_t)ld * 16));
        static_for<0, NA>([&](auto n1) { v[APW_CI(n1)] = to_c<R>(buf_load(pr, po[APW_CI(n1)], (double2*)nullptr)); });
    } else {
        if (mid) static_for<0, NB>([&](auto k2) { u[APW_CI(k2)] = Sl[g + NA * APW_CI(k2)]; });
        if (mode == GATHER_ADD) {
            static_for<0, NA>([&](auto m2) { add[APW_CI(m2)] = buf_load(hr, po[APW_CI(m2)], (double2*)nullptr); });
        } else if (mode == GATHER_SET_KIN) {
            const __amdgpu_buffer_rsrc_t pr = make_rsrc(psi + band * ld, (unsigned)((size_t)ld * 16));
            const __amdgpu_buffer_rsrc_t kr = make_rsrc(gk2, (unsigned)((size_t)ld * 8));
            static_for<0, NA>([&](auto m2) {
                const unsigned o = po[APW_CI(m2)];
                add[APW_CI(m2)] = buf_load(pr, o, (double2*)nullptr);
                kin[APW_CI(m2)] = __builtin_bit_cast(
                    double, __builtin_amdgcn_raw_buffer_load_b64(kr, o == BUF_OOB ? o : o / 2u, 0, 0));
            });
        }
    }
    for (int t = threadIdx.x; t < N; t += zreg_threads(N)) tws[t] = tw[t];
    __syncthreads();
    C* row = buf + l * LDZ;
    if constexpr (DIR > 0) {
        rdft<C, R, NA, +1>(v);
        static_for<1, NA>([&](auto k1) {
            C w = tws[g * APW_CI(k1)];
            w.y = -w.y;
            v[APW_CI(k1)] = cmul(v[APW_CI(k1)], w);
        });
        static_for<0, NA>([&](auto k1) { row[APW_CI(k1) * NBP + g] = v[APW_CI(k1)]; });
        __syncthreads();
        if (mid) {
            static_for<0, NB>([&](auto n2) { u[APW_CI(n2)] = row[g * NBP + APW_CI(n2)]; });
            rdft<C, R, NB, +1>(u);
            if (sin) static_for<0, NB>([&](auto k2) { Sl[g + NA * APW_CI(k2)] = u[APW_CI(k2)]; });
        }
    } else {
        if (mid) {
            rdft<C, R, NB, -1>(u);
            static_for<1, NB>([&](auto m1) { u[APW_CI(m1)] = cmul(u[APW_CI(m1)], tws[g * APW_CI(m1)]); });
            static_for<0, NB>([&](auto m1) { row[APW_CI(m1) * NAP + g] = u[APW_CI(m1)]; });
        }
        __syncthreads();
        static_for<0, NA>([&](auto k1) { v[APW_CI(k1)] = row[g * NAP + APW_CI(k1)]; });
        rdft<C, R, NA, -1>(v);
        // z = g + NB m2 is the z of po[m2]
        if (mode == GATHER_ADD) {
            static_for<0, NA>([&](auto m2) {
                double2 h = add[APW_CI(m2)];
                h.x += scale * (double)v[APW_CI(m2)].x;
                h.y += scale * (double)v[APW_CI(m2)].y;
                buf_store(h, hr, po[APW_CI(m2)]);
            });
        } else if (mode == GATHER_SET_KIN) {
            static_for<0, NA>([&](auto m2) {
                const double k = kin[APW_CI(m2)] * tpiba2;
                const double2 p = add[APW_CI(m2)];
                double2 h;
                h.x = scale * (double)v[APW_CI(m2)].x + k * p.x;
                h.y = scale * (double)v[APW_CI(m2)].y + k * p.y;
                buf_store(h, hr, po[APW_CI(m2)]);
            });
        } else {
            static_for<0, NA>([&](auto m2) {
                double2 h;
                h.x = scale * (double)v[APW_CI(m2)].x;
                h.y = scale * (double)v[APW_CI(m2)].y;
                buf_store(h, hr, po[APW_CI(m2)]);
            });
        }
    }
}

// ---- pass 4: y-forward, keeping the stick rows -----------------------------------------
template <class R>
__global__ __launch_bounds__(NT) void k_yfwd(const cplx<R>* __restrict__ X, int ns, int nz, int nxr,
                                             const int* __restrict__ slot_start, const int* __restrict__ slot_cnt,
                                             const int* __restrict__ stick_iy, cplx<R>* __restrict__ S, FftPlan py,
                                             const cplx<R>* __restrict__ tw, int Lrt) {
    using C = cplx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int ny = py.n;
    const int L = Lrt;
    const int ldl = L + 1;
    C* x = reinterpret_cast<C*>(smem);
    C* y = x + ny * ldl;
    C* tws = y + ny * ldl;
    const int slot = blockIdx.x;
    const int iz0 = blockIdx.y * L;
    const int band = blockIdx.z;
    load_tw<R>(tws, tw, ny);
    const C* Xb = X + ((size_t)band * nxr + slot) * ny * nz;
    for (int t = threadIdx.x; t < ny * L; t += NT) {
        const int iy = t / L, dz = t - (t / L) * L;
        const int iz = iz0 + dz;
        C v;
        v.x = 0;
        v.y = 0;
        if (iz < nz) v = Xb[(size_t)iy * nz + iz];
        x[iy * ldl + dz] = v;
    }
    __syncthreads();
    C* res = lds_fft<R>(x, y, tws, py, L, ldl, -1);
    const int st0 = slot_start[slot], cnt = slot_cnt[slot];
    C* Sb = S + (size_t)band * ns * nz;
    for (int t = threadIdx.x; t < cnt * L; t += NT) {
        const int q = t / L, dz = t - (t / L) * L;
        const int iz = iz0 + dz;
        const int s = st0 + q;
        if (iz < nz) Sb[(size_t)s * nz + iz] = res[stick_iy[s] * ldl + dz];
    }
}

// ---- pass 5: z-forward on sticks, gathered into hpsi ----------------------------------
template <class R>
__global__ __launch_bounds__(NT) void k_zfwd(const cplx<R>* __restrict__ S, const int* __restrict__ stick_ig, int ns,
                                             double2* __restrict__ hpsi, int ldh, const double2* __restrict__ psi,
                                             int ld, const double* __restrict__ gk2, double tpiba2, int mode,
                                             double scale, FftPlan pz, const cplx<R>* __restrict__ tw, int Lrt) {
    using C = cplx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int nz = pz.n;
    const int L = Lrt;
    const int ldl = L + 1;
    C* x = reinterpret_cast<C*>(smem);
    C* y = x + nz * ldl;
    C* tws = y + nz * ldl;
    const int band = blockIdx.y;
    const int s0 = blockIdx.x * L;
    load_tw<R>(tws, tw, nz);
    const C* Sb = S + (size_t)band * ns * nz;
    for (int t = threadIdx.x; t < L * nz; t += NT) {
        const int l = t / nz, iz = t - (t / nz) * nz;
        const int s = s0 + l;
        C v;
        v.x = 0;
        v.y = 0;
        if (s < ns) v = Sb[(size_t)s * nz + iz];
        x[iz * ldl + l] = v;
    }
    __syncthreads();
    C* res = lds_fft<R>(x, y, tws, pz, L, ldl, -1);
    double2* hb = hpsi + (size_t)band * ldh;
    const double2* pb = psi ? psi + (size_t)band * ld : nullptr;
    for (int t = threadIdx.x; t < L * nz; t += NT) {
        const int l = t / nz, iz = t - (t / nz) * nz;
        const int s = s0 + l;
        if (s >= ns) continue;
        const int ig = stick_ig[(size_t)s * nz + iz];
        if (ig < 0) continue;
        const C v = res[iz * ldl + l];
        double2 o;
        o.x = scale * (double)v.x;
        o.y = scale * (double)v.y;
        if (mode == GATHER_ADD) {
            const double2 h = hb[ig];
            o.x += h.x;
            o.y += h.y;
        } else if (mode == GATHER_SET_KIN) {
            const double kin = gk2[ig] * tpiba2;
            const double2 p = pb[ig];
            o.x += kin * p.x;
            o.y += kin * p.y;
        }
        hb[ig] = o;
    }
}

double fft_flops(int n) { return 5.0 * n * std::log2((double)n); }

template <class R>
size_t lds_bytes(int n, int L) {
    return (2 * (size_t)n * (L + 1) + n) * sizeof(cplx<R>);
}

template <class R>
const cplx<R>* tw_ptr(const PwGrid& g, int off) {
    if (std::is_same<R, double>::value) return reinterpret_cast<const cplx<R>*>(g.tw64.get() + 2 * off);
    return reinterpret_cast<const cplx<R>*>(g.tw32.get() + 2 * off);
}

// x length for the register x pass; 0 (the run-time-plan kernel) when the x slots are more
// than two runs or a band's slots exceed the 31-bit byte offsets of its buffer loads
int xreg_len(const PwGrid& g, int sched) {
    return (!(sched & 1) && g.xrun_ok && g.bytes_X(sizeof(double2)) < ((size_t)1 << 31)) ? g.nx : 0;
}
// y length for the register y passes (0: run-time-plan kernels): sticks of every slot in
// at most two runs of y, a band's sticks within 31-bit byte offsets
int yreg_len(const PwGrid& g, int sched) {
    return (!(sched & 2) && g.yrun_ok && g.bytes_S(sizeof(double2)) < ((size_t)1 << 31)) ? g.ny : 0;
}
// z length for the register z passes: sphere offsets (16 B per plane wave) within 31 bits
int zreg_len(const PwGrid& g, int ld, int ldh, int sched) {
    const size_t m = (size_t)std::max(std::max(ld, ldh), g.npw) * 16;
    return (!(sched & 4) && m < ((size_t)1 << 31)) ? g.nz : 0;
}
// the schedule bits of the precision R (autotuner's choice; 0 = register kernels throughout)
template <class R>
int sched_of(const PwGrid& g) {
    return g.sched[std::is_same<R, float>::value ? 1 : 0];
}

template <class R>
void pass_forward_half(hipStream_t s, const PwGrid& g, R* S, R* X, int nb, double2* hpsi, int ldh,
                       const double2* psi, int ld, const double* gk2, double tpiba2, GatherMode mode, double scale,
                       Prof* prof) {
    using C = cplx<R>;
    const int Ly = lines_for(g.ny), Lz = lines_for(g.nz);
    const double cs = sizeof(C);
    const double nSz = (double)g.ns * g.nz, nX = (double)g.nxr * g.ny * g.nz;
    {
        ProfScope ps(prof, s, "fft_y_fwd", nb * cs * (nX + nSz), nb * fft_flops(g.ny) * g.nxr * g.nz, true);
#define APW_YREG_LAUNCH(N)                                                                                    \
    case N:                                                                                                  \
        APW_PLAUNCH(ps, (k_yreg<R, N, -1>), dim3(g.nxr, (g.nz + XL - 1) / XL, nb), dim3(xreg_threads(N)), 0, \
                           s, reinterpret_cast<const C*>(X), reinterpret_cast<C*>(S), g.ns, g.nz, g.nxr,         \
                           g.slot_start.get(), g.yrun.get(), tw_ptr<R>(g, g.twoff_y));                       \
        break;
        switch (yreg_len(g, sched_of<R>(g))) {
            APW_FFT_SIZES(APW_YREG_LAUNCH)
            default:
                APW_PLAUNCH(ps, (k_yfwd<R>), dim3(g.nxr, (g.nz + Ly - 1) / Ly, nb), dim3(NT),
                                   lds_bytes<R>(g.ny, Ly), s, reinterpret_cast<const C*>(X), g.ns, g.nz, g.nxr,
                                   g.slot_start.get(), g.slot_cnt.get(), g.stick_iy.get(), reinterpret_cast<C*>(S),
                                   g.py, tw_ptr<R>(g, g.twoff_y), Ly);
                break;
        }
#undef APW_YREG_LAUNCH
        APW_LAUNCH_CHECK();
    }
    const double pwio = 16.0 * g.npw * ((mode == GATHER_SET) ? 1 : 2);
    ProfScope ps(prof, s, "fft_z_fwd_gather", nb * (cs * nSz + pwio), nb * (fft_flops(g.nz) * g.ns + 4.0 * g.npw), true);
#define APW_ZREG_LAUNCH(N)                                                                                    \
    case N:                                                                                                  \
        APW_PLAUNCH(ps, (k_zreg<R, N, -1>), dim3((g.ns + ZL - 1) / ZL, nb), dim3(zreg_threads(N)), 0, s, psi, \
                           ld, g.stick_ig.get(), g.zrun_ok ? g.zrun.get() : nullptr, g.ns, reinterpret_cast<C*>(S), \
                           hpsi, ldh, gk2, tpiba2, (int)mode,                                               \
                           scale, tw_ptr<R>(g, g.twoff_z));                                                  \
        break;
    switch (zreg_len(g, ld, ldh, sched_of<R>(g))) {
        APW_FFT_SIZES(APW_ZREG_LAUNCH)
        default:
            APW_PLAUNCH(ps, (k_zfwd<R>), dim3((g.ns + Lz - 1) / Lz, nb), dim3(NT), lds_bytes<R>(g.nz, Lz), s,
                               reinterpret_cast<const C*>(S), g.stick_ig.get(), g.ns, hpsi, ldh, psi, ld, gk2, tpiba2,
                               (int)mode, scale, g.pz, tw_ptr<R>(g, g.twoff_z), Lz);
            break;
    }
#undef APW_ZREG_LAUNCH
    APW_LAUNCH_CHECK();
}

template <class R>
void pass_backward_half(hipStream_t s, const PwGrid& g, const double2* psi, int ld, R* S, R* X, int nb, Prof* prof) {
    using C = cplx<R>;
    const int Ly = lines_for(g.ny), Lz = lines_for(g.nz);
    const double cs = sizeof(C);
    const double nSz = (double)g.ns * g.nz, nX = (double)g.nxr * g.ny * g.nz;
    {
        ProfScope ps(prof, s, "fft_z_bwd_scatter", nb * (16.0 * g.npw + cs * nSz), nb * fft_flops(g.nz) * g.ns, true);
#define APW_ZREG_LAUNCH(N)                                                                                    \
    case N:                                                                                                  \
        APW_PLAUNCH(ps, (k_zreg<R, N, 1>), dim3((g.ns + ZL - 1) / ZL, nb), dim3(zreg_threads(N)), 0, s, psi,  \
                           ld, g.stick_ig.get(), g.zrun_ok ? g.zrun.get() : nullptr, g.ns, reinterpret_cast<C*>(S), \
                           (double2*)nullptr, 0,                                                            \
                           (const double*)nullptr, 0.0, 0, 0.0, tw_ptr<R>(g, g.twoff_z));                    \
        break;
        switch (zreg_len(g, ld, 0, sched_of<R>(g))) {
            APW_FFT_SIZES(APW_ZREG_LAUNCH)
            default:
                APW_PLAUNCH(ps, (k_zbwd<R>), dim3((g.ns + Lz - 1) / Lz, nb), dim3(NT), lds_bytes<R>(g.nz, Lz),
                                   s, psi, ld, g.stick_ig.get(), g.ns, reinterpret_cast<C*>(S), g.pz,
                                   tw_ptr<R>(g, g.twoff_z), Lz);
                break;
        }
#undef APW_ZREG_LAUNCH
        APW_LAUNCH_CHECK();
    }
    ProfScope ps(prof, s, "fft_y_bwd", nb * cs * (nSz + nX), nb * fft_flops(g.ny) * g.nxr * g.nz, true);
#define APW_YREG_LAUNCH(N)                                                                                    \
    case N:                                                                                                  \
        APW_PLAUNCH(ps, (k_yreg<R, N, 1>), dim3(g.nxr, (g.nz + XL - 1) / XL, nb), dim3(xreg_threads(N)), 0,  \
                           s, reinterpret_cast<const C*>(S), reinterpret_cast<C*>(X), g.ns, g.nz, g.nxr,         \
                           g.slot_start.get(), g.yrun.get(), tw_ptr<R>(g, g.twoff_y));                       \
        break;
    switch (yreg_len(g, sched_of<R>(g))) {
        APW_FFT_SIZES(APW_YREG_LAUNCH)
        default:
            APW_PLAUNCH(ps, (k_ybwd<R>), dim3(g.nxr, (g.nz + Ly - 1) / Ly, nb), dim3(NT), lds_bytes<R>(g.ny, Ly),
                               s, reinterpret_cast<const C*>(S), g.ns, g.nz, g.nxr, g.slot_start.get(),
                               g.slot_cnt.get(), g.stick_iy.get(), reinterpret_cast<C*>(X), g.py,
                               tw_ptr<R>(g, g.twoff_y), Ly);
            break;
    }
#undef APW_YREG_LAUNCH
    APW_LAUNCH_CHECK();
}

}  // namespace

bool xreg_available(const PwGrid& g) {
    switch (xreg_len(g, 0)) {
#define APW_LEN_CASE(N) case N: return true;
        APW_FFT_SIZES(APW_LEN_CASE)
        default: return false;
    }
}
bool yreg_available(const PwGrid& g) {
    switch (yreg_len(g, 0)) {
        APW_FFT_SIZES(APW_LEN_CASE)
        default: return false;
    }
}
bool zreg_available(const PwGrid& g, int ld) {
    switch (zreg_len(g, ld, ld, 0)) {
        APW_FFT_SIZES(APW_LEN_CASE)
#undef APW_LEN_CASE
        default: return false;
    }
}

void PwGrid::build(int nx_, int ny_, int nz_, int npw_, const int* ig2ixyz) {
    nx = nx_;
    ny = ny_;
    nz = nz_;
    npw = npw_;
    APW_CHECK(nx > 0 && ny > 0 && nz > 0 && npw > 0, "PwGrid: bad sizes");
    px = make_plan(nx);
    py = make_plan(ny);
    pz = make_plan(nz);
    const long nxyz = (long)nx * ny * nz;
    // sticks keyed by ix*ny + iy, ordered by (ix, iy)
    std::map<long, int> key2s;
    for (int ig = 0; ig < npw; ++ig) {
        const long idx = ig2ixyz[ig];
        APW_CHECK(idx >= 0 && idx < nxyz, "ig2ixyz[%d] = %ld outside the %dx%dx%d box", ig, idx, nx, ny, nz);
        key2s[idx / nz] = 0;
    }
    ns = 0;
    std::vector<int> h_iy, h_ix_of_s;
    for (auto& kv : key2s) {
        kv.second = ns++;
        h_ix_of_s.push_back((int)(kv.first / ny));
        h_iy.push_back((int)(kv.first % ny));
    }
    std::vector<int> h_ig((size_t)ns * nz, -1);
    for (int ig = 0; ig < npw; ++ig) {
        const long idx = ig2ixyz[ig];
        const int s = key2s[idx / nz];
        const int iz = (int)(idx % nz);
        APW_CHECK(h_ig[(size_t)s * nz + iz] < 0, "ig2ixyz maps two plane waves to box point %ld", idx);
        h_ig[(size_t)s * nz + iz] = ig;
    }
    // each stick's plane waves as runs of consecutive z that are consecutive in the sphere
    // (the plane-wave order of PW_Basis_K gives one run, or two when z wraps through 0)
    std::vector<int4> h_zrun(ns);
    zrun_ok = nz <= 0x7fff;
    for (int s = 0; s < ns && zrun_ok; ++s) {
        int nrun = 0;
        int st[2] = {0, 0}, ln[2] = {0, 0}, ig0[2] = {0, 0};
        for (int iz = 0; iz < nz; ++iz) {
            const int ig = h_ig[(size_t)s * nz + iz];
            if (ig < 0) continue;
            const bool cont = nrun > 0 && iz == st[nrun - 1] + ln[nrun - 1] && ig == ig0[nrun - 1] + ln[nrun - 1];
            if (!cont) {
                if (nrun == 2) {
                    zrun_ok = false;
                    break;
                }
                st[nrun] = iz;
                ig0[nrun] = ig;
                ++nrun;
            }
            ln[nrun - 1]++;
        }
        h_zrun[s] = make_int4(st[0] | (ln[0] << 16), ig0[0], st[1] | (ln[1] << 16), ig0[1]);
    }
    if (zrun_ok) {
        zrun.reserve(ns);
        APW_HIP(hipMemcpy(zrun.get(), h_zrun.data(), ns * sizeof(int4), hipMemcpyHostToDevice));
    }
    std::vector<int> h_slot_ix, h_slot_start, h_slot_cnt;
    for (int s = 0; s < ns; ++s) {
        if (h_slot_ix.empty() || h_slot_ix.back() != h_ix_of_s[s]) {
            h_slot_ix.push_back(h_ix_of_s[s]);
            h_slot_start.push_back(s);
            h_slot_cnt.push_back(0);
        }
        h_slot_cnt.back()++;
    }
    nxr = (int)h_slot_ix.size();
    // the slots as runs of consecutive x (a sphere gives one or two)
    std::vector<int> run_start, run_len;
    for (int q = 0; q < nxr; ++q) {
        if (q == 0 || h_slot_ix[q] != h_slot_ix[q - 1] + 1) {
            run_start.push_back(h_slot_ix[q]);
            run_len.push_back(0);
        }
        run_len.back()++;
    }
    xrun_ok = run_start.size() <= 2;
    xrun = make_int4(run_start[0], run_len[0], run_start.size() > 1 ? run_start[1] : 0,
                     run_start.size() > 1 ? run_len[1] : 0);
    // the sticks of each slot as runs of consecutive y
    std::vector<int4> h_yrun(nxr);
    yrun_ok = true;
    for (int q = 0; q < nxr; ++q) {
        int nrun = 0;
        int st[2] = {0, 0}, ln[2] = {0, 0};
        for (int k = 0; k < h_slot_cnt[q]; ++k) {
            const int iy = h_iy[h_slot_start[q] + k];
            if (k == 0 || iy != h_iy[h_slot_start[q] + k - 1] + 1) {
                if (nrun == 2) {
                    yrun_ok = false;
                    break;
                }
                st[nrun++] = iy;
            }
            ln[nrun - 1]++;
        }
        h_yrun[q] = make_int4(st[0], ln[0], st[1], ln[1]);
    }
    yrun.reserve(nxr);
    APW_HIP(hipMemcpy(yrun.get(), h_yrun.data(), nxr * sizeof(int4), hipMemcpyHostToDevice));
    auto up = [](DevBuf<int>& d, const std::vector<int>& h) {
        d.reserve(h.size());
        APW_HIP(hipMemcpy(d.get(), h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
    };
    up(stick_ig, h_ig);
    up(stick_iy, h_iy);
    up(slot_ix, h_slot_ix);
    up(slot_start, h_slot_start);
    up(slot_cnt, h_slot_cnt);
    // twiddles exp(-2 pi i k / n) in long double
    std::vector<double> t64;
    std::vector<float> t32;
    auto add_tw = [&](int n) {
        const int off = (int)(t64.size() / 2);
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int k = 0; k < n; ++k) {
            const long double a = -2.0L * pi * (long double)k / (long double)n;
            t64.push_back((double)cosl(a));
            t64.push_back((double)sinl(a));
            t32.push_back((float)cosl(a));
            t32.push_back((float)sinl(a));
        }
        return off;
    };
    twoff_x = add_tw(nx);
    twoff_y = add_tw(ny);
    twoff_z = add_tw(nz);
    tw64.reserve(t64.size());
    tw32.reserve(t32.size());
    APW_HIP(hipMemcpy(tw64.get(), t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice));
    APW_HIP(hipMemcpy(tw32.get(), t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice));
}

template <class R>
void local_hpsi(hipStream_t s, const PwGrid& g, const double* psi, int ld, double* hpsi, int ldh, int nb,
                const R* veff, const double* gk2, double tpiba2, GatherMode mode, double scale, R* S, R* X,
                Prof* prof) {
    using C = cplx<R>;
    if (nb <= 0) return;
    const double2* p2 = reinterpret_cast<const double2*>(psi);
    pass_backward_half<R>(s, g, p2, ld, S, X, nb, prof);
    const int Lx = lines_for(g.nx);
    {
        const double nX = (double)g.nxr * g.ny * g.nz, nxyz = (double)g.nx * g.ny * g.nz;
        ProfScope ps(prof, s, "fft_x_veff_fused", nb * (2.0 * sizeof(C) * nX + sizeof(R) * nxyz),
                     nb * (2.0 * fft_flops(g.nx) * g.ny * g.nz + 2.0 * nxyz), true);
#define APW_XREG_LAUNCH(N)                                                                                     \
    case N:                                                                                                   \
        APW_PLAUNCH(ps, (k_xreg<R, N, 0>), dim3(g.ny, (g.nz + XL - 1) / XL, nb), dim3(xreg_threads(N)), 0, s, \
                           reinterpret_cast<C*>(X), g.nxr, g.ny, g.nz, g.xrun, veff, (double2*)nullptr, 0,  \
                           0.0, tw_ptr<R>(g, g.twoff_x));                                                     \
        break;
        switch (xreg_len(g, sched_of<R>(g))) {
            APW_FFT_SIZES(APW_XREG_LAUNCH)
            default:
                APW_PLAUNCH(ps, (k_xpass<R, 0>), dim3(g.ny, (g.nz + Lx - 1) / Lx, nb), dim3(NT),
                                   lds_bytes<R>(g.nx, Lx), s, reinterpret_cast<C*>(X), g.nxr, g.ny, g.nz,
                                   g.slot_ix.get(), veff, (double2*)nullptr, 0, 0.0, g.px, tw_ptr<R>(g, g.twoff_x),
                                   Lx);
                break;
        }
#undef APW_XREG_LAUNCH
        APW_LAUNCH_CHECK();
    }
    pass_forward_half<R>(s, g, S, X, nb, reinterpret_cast<double2*>(hpsi), ldh, p2, ld, gk2, tpiba2, mode, scale, prof);
}

template void local_hpsi<double>(hipStream_t, const PwGrid&, const double*, int, double*, int, int, const double*,
                                 const double*, double, GatherMode, double, double*, double*, Prof*);
template void local_hpsi<float>(hipStream_t, const PwGrid&, const double*, int, double*, int, int, const float*,
                                const double*, double, GatherMode, double, float*, float*, Prof*);

void sphere_to_box(hipStream_t s, const PwGrid& g, const double* psi, int ld, double* box, int nb, bool add,
                   double factor, double* S, double* X) {
    using C = cplx<double>;
    pass_backward_half<double>(s, g, reinterpret_cast<const double2*>(psi), ld, S, X, nb, nullptr);
    const int Lx = lines_for(g.nx);
#define APW_XREG_LAUNCH(N)                                                                                         \
    case N:                                                                                                       \
        hipLaunchKernelGGL((k_xreg<double, N, 1>), dim3(g.ny, (g.nz + XL - 1) / XL, nb), dim3(xreg_threads(N)), 0, \
                           s, reinterpret_cast<C*>(X), g.nxr, g.ny, g.nz, g.xrun, (const double*)nullptr,   \
                           reinterpret_cast<double2*>(box), add ? 1 : 0, factor, tw_ptr<double>(g, g.twoff_x));    \
        break;
    switch (xreg_len(g, g.sched[0])) {
        APW_FFT_SIZES(APW_XREG_LAUNCH)
        default:
            hipLaunchKernelGGL((k_xpass<double, 1>), dim3(g.ny, (g.nz + Lx - 1) / Lx, nb), dim3(NT),
                               lds_bytes<double>(g.nx, Lx), s, reinterpret_cast<C*>(X), g.nxr, g.ny, g.nz,
                               g.slot_ix.get(), (const double*)nullptr, reinterpret_cast<double2*>(box), add ? 1 : 0,
                               factor, g.px, tw_ptr<double>(g, g.twoff_x), Lx);
            break;
    }
#undef APW_XREG_LAUNCH
    APW_LAUNCH_CHECK();
}

void box_to_sphere(hipStream_t s, const PwGrid& g, const double* box, double* hpsi, int ldh, int nb,
                   GatherMode mode, double scale, double* S, double* X) {
    using C = cplx<double>;
    const int Lx = lines_for(g.nx);
    double2* bx = const_cast<double2*>(reinterpret_cast<const double2*>(box));
#define APW_XREG_LAUNCH(N)                                                                                         \
    case N:                                                                                                       \
        hipLaunchKernelGGL((k_xreg<double, N, 2>), dim3(g.ny, (g.nz + XL - 1) / XL, nb), dim3(xreg_threads(N)), 0, \
                           s, reinterpret_cast<C*>(X), g.nxr, g.ny, g.nz, g.xrun, (const double*)nullptr,   \
                           bx, 0, 0.0, tw_ptr<double>(g, g.twoff_x));                                              \
        break;
    switch (xreg_len(g, g.sched[0])) {
        APW_FFT_SIZES(APW_XREG_LAUNCH)
        default:
            hipLaunchKernelGGL((k_xpass<double, 2>), dim3(g.ny, (g.nz + Lx - 1) / Lx, nb), dim3(NT),
                               lds_bytes<double>(g.nx, Lx), s, reinterpret_cast<C*>(X), g.nxr, g.ny, g.nz,
                               g.slot_ix.get(), (const double*)nullptr, bx, 0, 0.0, g.px, tw_ptr<double>(g, g.twoff_x),
                               Lx);
            break;
    }
#undef APW_XREG_LAUNCH
    APW_LAUNCH_CHECK();
    pass_forward_half<double>(s, g, S, X, nb, reinterpret_cast<double2*>(hpsi), ldh, nullptr, 0, nullptr, 0.0, mode,
                              scale, nullptr);
}

void psi_to_rho(hipStream_t s, const PwGrid& g, const double* psi, int ld, int nb, const double* w, double* rho,
                bool add, double* S, double* X, Prof* prof) {
    using C = cplx<double>;
    if (nb <= 0) return;
    pass_backward_half<double>(s, g, reinterpret_cast<const double2*>(psi), ld, S, X, nb, prof);
    const int Lx = lines_for(g.nx);
    const double nX = (double)g.nxr * g.ny * g.nz, nxyz = (double)g.nx * g.ny * g.nz;
    // algorithmic: the x slots of every band in, rho read (when adding) and written once
    ProfScope ps(prof, s, "fft_x_rho", nb * sizeof(C) * nX + (add ? 16.0 : 8.0) * nxyz,
                 nb * (fft_flops(g.nx) * g.ny * g.nz + 3.0 * nxyz), true);
    const size_t lds = lds_bytes<double>(g.nx, Lx) + (size_t)g.nx * Lx * sizeof(double);
#define APW_XREG_LAUNCH(N)                                                                                    \
    case N:                                                                                                  \
        APW_PLAUNCH(ps, (k_xrho_reg<N>), dim3(g.ny, (g.nz + XL - 1) / XL), dim3(xreg_threads(N)), 0, s,       \
                           reinterpret_cast<const C*>(X), g.nxr, g.ny, g.nz, nb, g.xrun, w, rho, add ? 1 : 0, \
                           tw_ptr<double>(g, g.twoff_x));                                                    \
        break;
    switch (xreg_len(g, g.sched[0])) {
        APW_FFT_SIZES(APW_XREG_LAUNCH)
        default:
            APW_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_xrho),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            APW_PLAUNCH(ps, k_xrho, dim3(g.ny, (g.nz + Lx - 1) / Lx), dim3(NT), lds, s,
                               reinterpret_cast<const C*>(X), g.nxr, g.ny, g.nz, nb, g.slot_ix.get(), w, rho,
                               add ? 1 : 0, g.px, tw_ptr<double>(g, g.twoff_x), Lx);
            break;
    }
#undef APW_XREG_LAUNCH
    APW_LAUNCH_CHECK();
}

}  // namespace apw
