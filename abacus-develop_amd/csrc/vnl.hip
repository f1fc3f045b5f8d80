// Projector generation for one k-point: real spherical harmonics of k+G, the structure
// factor, and the Kleinman-Bylander projectors beta(k+G) that the nonlocal operator
// consumes (pseudopot_cell_vnl::getvnl, VNL_in_pw.cpp:412-524).
//
// Three device functors with the reference's argument lists (cal_ylm_real_op,
// module_base/kernels/math_op.h; cal_sk_op, kernels/wf_op.h; cal_vnl_op, kernels/vnl_op.h)
// and one fused kernel, apw_getvnl, that produces vkb in a single pass: each thread owns
// one plane wave, forms Y_lm(k+G) in registers, the structure factor of a block of atoms
// with one sincos per atom, and writes every projector row of that plane wave.  Rows are
// nkb x npwx (projector-major, the layout the nonlocal GEMMs read), so a wave writes 64
// consecutive plane waves of one row: 1 KiB contiguous.  The pass is bound by the
// nkb * npw * 16 B it writes.
#include <algorithm>
#include <cmath>

#include "apw_ham.hpp"
#include "apw_internal.hpp"

namespace apw {
namespace {

constexpr int YLM_LMAX = 6;      // largest lmax of the Ylm functor
constexpr int VNL_LMAX = 3;      // largest angular momentum of a projector (f)
constexpr int VNL_NTYPE = 16;    // atom types of the fused kernel
constexpr int VNL_NH = 32;       // projectors per atom of one type (nhm)
constexpr int VNL_ATOMS = 8;     // atoms whose structure factor a thread holds at once

constexpr double PI = 3.14159265358979323846;
constexpr double SQRT2 = 1.41421356237309504880;

int grid1(size_t n, int nt = 256) { return (int)std::max<size_t>(1, (n + nt - 1) / nt); }

__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

__global__ void k_zero_i32(int* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// P_l^m(cos theta) and Y_lm for one direction, the reference's recursion
// (math_op.cpp:45-124): phi from atan with the quadrant fix-up (+-pi/2 when
// |g_x| <= 1e-9), cos theta = 0 when |g| < 1e-9; rows lm = l^2 + {0 | 2m-1 (cos) | 2m (sin)}.
// P[m * (L+1) + l] holds P_l^m for m <= l.
template <int L>
__device__ inline void ylm_point(double gx, double gy, double gz, double pi, double pi_half, double four_pi,
                                 double sqrt2, double (&y)[(L + 1) * (L + 1)], double (&P)[(L + 1) * (L + 1)]) {
    const double gmod = sqrt(gx * gx + gy * gy + gz * gz);
    const double cost = gmod < 1.0e-9 ? 0.0 : gz / gmod;
    double phi;
    if (gx > 1.0e-9) phi = atan(gy / gx);
    else if (gx < -1.0e-9) phi = atan(gy / gx) + pi;
    else phi = pi_half * ((gy >= 0.0) ? 1.0 : -1.0);
    const double x2 = fmax(0.0, 1.0 - cost * cost);
    int lm = -1;
#pragma unroll
    for (int l = 0; l <= L; ++l) {
        const double c = sqrt((2 * l + 1) / four_pi);
        if (l == 0) {
            P[0] = 1.0;
        } else if (l == 1) {
            P[0 * (L + 1) + 1] = cost;
            P[1 * (L + 1) + 1] = -sqrt(x2);
        } else {
            const int l1 = l - 1, l2 = l - 2, l3 = 2 * l - 1;
#pragma unroll
            for (int m = 0; m <= l2; ++m)
                P[m * (L + 1) + l] = (cost * l3 * P[m * (L + 1) + l1] - (l1 + m) * P[m * (L + 1) + l2]) / (l - m);
            P[l1 * (L + 1) + l] = cost * l3 * P[l1 * (L + 1) + l1];
            int semif = 1;
            for (int i = l3; i > 2; i -= 2) semif *= i;
            double pll = semif * pow(x2, (double)l / 2.0);
            if (l % 2 == 1) pll = -pll;
            P[l * (L + 1) + l] = pll;
        }
        y[++lm] = c * P[l];
#pragma unroll
        for (int m = 1; m <= l; ++m) {
            double fa = 1.0, fb = 1.0;
            for (int i = l - m; i > 1; --i) fa *= i;
            for (int i = l + m; i > 1; --i) fb *= i;
            const double same = c * sqrt(fa / fb) * sqrt2;
            double sp, cp;
            sincos(m * phi, &sp, &cp);
            y[++lm] = same * P[m * (L + 1) + l] * cp;
            y[++lm] = same * P[m * (L + 1) + l] * sp;
        }
    }
}

// cal_ylm_real_op: one thread per point; writes ylm and, as the reference does, the
// Legendre values into the caller's p[(m * (lmax+1) + l) * ng + ig]
template <int L>
__global__ void k_ylm_real(int ng, double sqrt2, double pi, double pi_half, double four_pi, double sqrt_inv_four_pi,
                           const double* __restrict__ g, double* __restrict__ p, double* __restrict__ ylm) {
    const int ig = blockIdx.x * blockDim.x + threadIdx.x;
    if (ig >= ng) return;
    if constexpr (L == 0) {
        ylm[ig] = sqrt_inv_four_pi;
        return;
    } else {
    double y[(L + 1) * (L + 1)], P[(L + 1) * (L + 1)];
    ylm_point<L>(g[3 * ig], g[3 * ig + 1], g[3 * ig + 2], pi, pi_half, four_pi, sqrt2, y, P);
#pragma unroll
    for (int q = 0; q < (L + 1) * (L + 1); ++q) ylm[(size_t)q * ng + ig] = y[q];
    if (p) {
#pragma unroll
        for (int l = 0; l <= L; ++l)
#pragma unroll
            for (int m = 0; m <= l; ++m) p[((size_t)m * (L + 1) + l) * ng + ig] = P[m * (L + 1) + l];
    }
    }
}

// cal_sk_op (wf_op.cpp:8-74): one thread per plane wave, all atoms
__global__ void k_cal_sk(int ik, int ntype, int nx, int ny, int nz, int rho_nx, int rho_ny, int rho_nz, int npw,
                         int npwx, int fftny, int e1, int e2, int e3, const int* __restrict__ atom_na,
                         const int* __restrict__ igl2isz, const int* __restrict__ is2fftixy, double two_pi,
                         const double* __restrict__ kvec_c, const double* __restrict__ tau,
                         const double2* __restrict__ eigts1, const double2* __restrict__ eigts2,
                         const double2* __restrict__ eigts3, double2* __restrict__ sk) {
    const int igl = blockIdx.x * blockDim.x + threadIdx.x;
    if (igl >= npw) return;
    const int isz = igl2isz[(size_t)ik * npwx + igl];
    int iz = isz % nz;
    const int ixy = is2fftixy[isz / nz];
    int ix = ixy / fftny, iy = ixy % fftny;
    if (ix >= nx / 2 + 1) ix -= nx;
    if (iy >= ny / 2 + 1) iy -= ny;
    if (iz >= nz / 2 + 1) iz -= nz;
    ix += rho_nx;
    iy += rho_ny;
    iz += rho_nz;
    int nat = 0;
    for (int it = 0; it < ntype; ++it) nat += atom_na[it];
    const double kx = kvec_c[3 * ik], ky = kvec_c[3 * ik + 1], kz = kvec_c[3 * ik + 2];
    for (int iat = 0; iat < nat; ++iat) {
        const double arg = two_pi * (kx * tau[3 * iat] + ky * tau[3 * iat + 1] + kz * tau[3 * iat + 2]);
        double sp, cp;
        sincos(arg, &sp, &cp);
        const double2 a = eigts1[(size_t)iat * e1 + ix], b = eigts2[(size_t)iat * e2 + iy],
                      c = eigts3[(size_t)iat * e3 + iz];
        const double2 kp = make_double2(cp, -sp);
        const double2 t1 = cmul(kp, a);
        const double2 t2 = cmul(t1, b);
        sk[(size_t)iat * npw + igl] = cmul(t2, c);
    }
}

// _polynomial_interpolation (vnl_tools.hpp:7-27): four-point Lagrange on tab[it][nb][:].
// A point past the table (iq + 3 >= tab_3; the reference reads beyond it) raises
// status[0] and gives 0.
__device__ inline double interp_tab(const double* __restrict__ tab, int it, int nb, int tab_2, int tab_3, double dq,
                                    double x, int* __restrict__ status) {
    const double pos = x / dq;
    if (!(pos >= 0.0 && pos < (double)(tab_3 - 3))) {
        status[0] = 1;
        return 0.0;
    }
    const int iq = (int)pos;
    const double x0 = pos - (double)iq, x1 = 1.0 - x0, x2 = 2.0 - x0, x3 = 3.0 - x0;
    const double* t = tab + ((size_t)it * tab_2 + nb) * tab_3 + iq;
    return t[0] * x1 * x2 * x3 / 6.0 + t[1] * x0 * x2 * x3 / 2.0 - t[2] * x1 * x0 * x3 / 2.0 + t[3] * x1 * x2 * x0 / 6.0;
}

// cal_vnl_op (vnl_op.cpp:8-96), one thread per plane wave; vkb1 (nhm x npw) is the
// reference's per-type scratch and is written the same way
__global__ void k_cal_vnl(int ntype, int npw, int npwx, int nhm, int tab_2, int tab_3, const int* __restrict__ atom_na,
                          const int* __restrict__ atom_nb, const int* __restrict__ atom_nh, double dq, double tpiba,
                          double2 neg_i, const double* __restrict__ gk, const double* __restrict__ ylm,
                          const double* __restrict__ indv, const double* __restrict__ nhtol,
                          const double* __restrict__ nhtolm, const double* __restrict__ tab, double* __restrict__ vkb1,
                          const double2* __restrict__ sk, double2* __restrict__ vkb, int* __restrict__ status) {
    const int ig = blockIdx.x * blockDim.x + threadIdx.x;
    if (ig >= npw) return;
    const double gnorm = sqrt(gk[3 * ig] * gk[3 * ig] + gk[3 * ig + 1] * gk[3 * ig + 1] + gk[3 * ig + 2] * gk[3 * ig + 2]) *
                         tpiba;
    int jkb = 0, iat = 0;
    for (int it = 0; it < ntype; ++it) {
        const int nh = atom_nh[it], nbeta = atom_nb[it];
        for (int nb = 0; nb < nbeta; ++nb) {
            const double vq = interp_tab(tab, it, nb, tab_2, tab_3, dq, gnorm, status);
            for (int ih = 0; ih < nh; ++ih)
                if (nb == (int)indv[it * nhm + ih])
                    vkb1[(size_t)ih * npw + ig] = ylm[(size_t)(int)nhtolm[it * nhm + ih] * npw + ig] * vq;
        }
        for (int ia = 0; ia < atom_na[it]; ++ia) {
            for (int ih = 0; ih < nh; ++ih) {
                double2 pref = make_double2(1.0, 0.0);
                for (int q = (int)nhtol[it * nhm + ih]; q > 0; --q) pref = cmul(pref, neg_i);
                const double v = vkb1[(size_t)ih * npw + ig];
                const double2 s = sk[(size_t)iat * npw + ig];
                vkb[(size_t)jkb * npwx + ig] = cmul(make_double2(v * s.x, v * s.y), pref);
                ++jkb;
            }
            ++iat;
        }
    }
}

// per-type projector descriptors of the fused kernel, passed by value
struct VnlDesc {
    int ntype;
    int na[VNL_NTYPE], nh[VNL_NTYPE];
    unsigned char nb[VNL_NTYPE][VNL_NH];   // radial channel (indv)
    unsigned char lm[VNL_NTYPE][VNL_NH];   // Ylm row (nhtolm)
    unsigned char l[VNL_NTYPE][VNL_NH];    // angular momentum (nhtol)
};

// fused getvnl: Y_lm, |k+G|-interpolated radial part, structure factor
// exp(-2 pi i (k+G).tau), (-i)^l, one pass; rows ordered type -> atom -> ih
template <int L>
__global__ __launch_bounds__(256) void k_getvnl(VnlDesc dsc, int npw, int npwx, int tab_2, int tab_3, double dq,
                                                double tpiba, const double* __restrict__ gk,
                                                const double* __restrict__ tau, const double* __restrict__ tab,
                                                double2* __restrict__ vkb, int* __restrict__ status) {
    const int ig = blockIdx.x * blockDim.x + threadIdx.x;
    if (ig >= npw) return;
    const double gx = gk[3 * ig], gy = gk[3 * ig + 1], gz = gk[3 * ig + 2];
    const double gnorm = sqrt(gx * gx + gy * gy + gz * gz) * tpiba;
    double y[(L + 1) * (L + 1)], P[(L + 1) * (L + 1)];
    if constexpr (L == 0) y[0] = 0.28209479177387814;   // 1 / sqrt(4 pi)
    else ylm_point<L>(gx, gy, gz, PI, 0.5 * PI, 4.0 * PI, SQRT2, y, P);
    const double two_pi = 2.0 * PI;
    size_t row = 0;
    int iat = 0;
    for (int it = 0; it < dsc.ntype; ++it) {
        const int nh = dsc.nh[it], na = dsc.na[it];
        for (int a0 = 0; a0 < na; a0 += VNL_ATOMS) {
            const int nab = min(VNL_ATOMS, na - a0);
            double2 s[VNL_ATOMS];
#pragma unroll
            for (int a = 0; a < VNL_ATOMS; ++a) {
                if (a < nab) {
                    const double* t = tau + 3 * (size_t)(iat + a0 + a);
                    double sp, cp;
                    sincos(two_pi * (gx * t[0] + gy * t[1] + gz * t[2]), &sp, &cp);
                    s[a] = make_double2(cp, -sp);
                }
            }
            for (int ih = 0; ih < nh; ++ih) {
                const int lm = dsc.lm[it][ih];
                double ylmv = 0.0;
#pragma unroll
                for (int q = 0; q < (L + 1) * (L + 1); ++q)
                    if (q == lm) ylmv = y[q];
                const double v = ylmv * interp_tab(tab, it, dsc.nb[it][ih], tab_2, tab_3, dq, gnorm, status);
                // v (-i)^l
                double2 vp;
                switch (dsc.l[it][ih] & 3) {
                    case 0: vp = make_double2(v, 0.0); break;
                    case 1: vp = make_double2(0.0, -v); break;
                    case 2: vp = make_double2(-v, 0.0); break;
                    default: vp = make_double2(0.0, v); break;
                }
#pragma unroll
                for (int a = 0; a < VNL_ATOMS; ++a)
                    if (a < nab) vkb[(row + (size_t)(a0 + a) * nh + ih) * npwx + ig] = cmul(vp, s[a]);
            }
        }
        row += (size_t)na * nh;
        iat += na;
    }
}

}  // namespace

// status word of the interpolation bound: cleared before a launch, read after it
int* vnl_status_clear(apw_ctx* ctx) {
    ctx->status.reserve(4);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, ctx->stream, ctx->status.get(), 4);
    APW_LAUNCH_CHECK();
    return ctx->status.get();
}

void vnl_status_check(apw_ctx* ctx, const char* what) {
    int h = 0;
    APW_HIP(hipMemcpyAsync(&h, ctx->status.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    APW_HIP(hipStreamSynchronize(ctx->stream));
    APW_CHECK(h == 0, "%s: |k+G| beyond the interpolation table (tab_3 too small for the cutoff)", what);
}

}  // namespace apw

using apw::guarded;

extern "C" {

int apw_cal_ylm_real_op(apw_ctx* ctx, int ng, int lmax, double sqrt2, double pi, double pi_half, double four_pi,
                        double sqrt_inverse_four_pi, const double* g, double* p, double* ylm) {
    return guarded([&] {
        APW_CHECK(ctx && g && ylm, "apw_cal_ylm_real_op: null argument");
        APW_CHECK(ng >= 0 && lmax >= 0 && lmax <= apw::YLM_LMAX, "apw_cal_ylm_real_op: lmax %d outside 0..%d", lmax,
                  apw::YLM_LMAX);
        if (ng == 0) return APW_OK;
        apw::DeviceGuard gd(ctx->device);
        const dim3 grid(apw::grid1(ng)), blk(256);
        hipStream_t s = ctx->stream;
#define APW_YLM_CASE(LL)                                                                                      \
    case LL:                                                                                                  \
        hipLaunchKernelGGL(apw::k_ylm_real<LL>, grid, blk, 0, s, ng, sqrt2, pi, pi_half, four_pi,             \
                           sqrt_inverse_four_pi, g, p, ylm);                                                  \
        break;
        switch (lmax) {
            APW_YLM_CASE(0)
            APW_YLM_CASE(1)
            APW_YLM_CASE(2)
            APW_YLM_CASE(3)
            APW_YLM_CASE(4)
            APW_YLM_CASE(5)
            default: APW_YLM_CASE(6)
        }
#undef APW_YLM_CASE
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_cal_sk_op(apw_ctx* ctx, int ik, int ntype, int nx, int ny, int nz, int rho_nx, int rho_ny, int rho_nz, int npw,
                  int npwx, int fftny, int eigts1_nc, int eigts2_nc, int eigts3_nc, const int* atom_na,
                  const int* igl2isz, const int* is2fftixy, double two_pi, const double* kvec_c, const double* atom_tau,
                  const double* eigts1, const double* eigts2, const double* eigts3, double* sk) {
    return guarded([&] {
        APW_CHECK(ctx && atom_na && igl2isz && is2fftixy && kvec_c && atom_tau && eigts1 && eigts2 && eigts3 && sk,
                  "apw_cal_sk_op: null argument");
        APW_CHECK(npw >= 0 && npwx >= npw && nx > 0 && ny > 0 && nz > 0 && fftny > 0, "apw_cal_sk_op: bad sizes");
        if (npw == 0) return APW_OK;
        apw::DeviceGuard gd(ctx->device);
        hipLaunchKernelGGL(apw::k_cal_sk, dim3(apw::grid1(npw)), dim3(256), 0, ctx->stream, ik, ntype, nx, ny, nz, rho_nx,
                           rho_ny, rho_nz, npw, npwx, fftny, eigts1_nc, eigts2_nc, eigts3_nc, atom_na, igl2isz,
                           is2fftixy, two_pi, kvec_c, atom_tau, reinterpret_cast<const double2*>(eigts1),
                           reinterpret_cast<const double2*>(eigts2), reinterpret_cast<const double2*>(eigts3),
                           reinterpret_cast<double2*>(sk));
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_cal_vnl_op(apw_ctx* ctx, int ntype, int npw, int npwx, int nhm, int tab_2, int tab_3, const int* atom_na,
                   const int* atom_nb, const int* atom_nh, double dq, double tpiba, double neg_imag_re,
                   double neg_imag_im, const double* gk, const double* ylm, const double* indv, const double* nhtol,
                   const double* nhtolm, const double* tab, double* vkb1, const double* sk, double* vkb_in) {
    return guarded([&] {
        APW_CHECK(ctx && atom_na && atom_nb && atom_nh && gk && ylm && indv && nhtol && nhtolm && tab && vkb1 && sk &&
                      vkb_in,
                  "apw_cal_vnl_op: null argument");
        APW_CHECK(npw >= 0 && npwx >= npw && nhm > 0 && tab_2 > 0 && tab_3 >= 4 && dq > 0,
                  "apw_cal_vnl_op: bad sizes");
        if (npw == 0) return APW_OK;
        apw::DeviceGuard gd(ctx->device);
        int* st = apw::vnl_status_clear(ctx);
        hipLaunchKernelGGL(apw::k_cal_vnl, dim3(apw::grid1(npw)), dim3(256), 0, ctx->stream, ntype, npw, npwx, nhm,
                           tab_2, tab_3, atom_na, atom_nb, atom_nh, dq, tpiba, make_double2(neg_imag_re, neg_imag_im),
                           gk, ylm, indv, nhtol, nhtolm, tab, vkb1, reinterpret_cast<const double2*>(sk),
                           reinterpret_cast<double2*>(vkb_in), st);
        APW_LAUNCH_CHECK();
        apw::vnl_status_check(ctx, "apw_cal_vnl_op");
        return APW_OK;
    });
}

int apw_getvnl(apw_ctx* ctx, int ntype, int npw, int npwx, int nhm, int tab_2, int tab_3, const int* atom_na,
               const int* atom_nb, const int* atom_nh, double dq, double tpiba, const double* gk, const double* tau,
               const double* indv, const double* nhtol, const double* nhtolm, const double* tab, double* vkb) {
    return guarded([&] {
        APW_CHECK(ctx && atom_na && atom_nb && atom_nh && gk && tau && indv && nhtol && nhtolm && tab && vkb,
                  "apw_getvnl: null argument");
        APW_CHECK(ntype >= 1 && ntype <= apw::VNL_NTYPE, "apw_getvnl: ntype %d outside 1..%d", ntype, apw::VNL_NTYPE);
        APW_CHECK(npw >= 0 && npwx >= npw && nhm >= 1 && nhm <= apw::VNL_NH && tab_2 >= 1 && tab_3 >= 4 && dq > 0,
                  "apw_getvnl: bad sizes (nhm %d, at most %d)", nhm, apw::VNL_NH);
        apw::VnlDesc dsc{};
        dsc.ntype = ntype;
        int lmax = 0, nat = 0;
        for (int it = 0; it < ntype; ++it) {
            APW_CHECK(atom_na[it] >= 0 && atom_nh[it] >= 0 && atom_nh[it] <= nhm && atom_nb[it] >= 0 &&
                          atom_nb[it] <= tab_2,
                      "apw_getvnl: type %d: na %d nh %d nbeta %d", it, atom_na[it], atom_nh[it], atom_nb[it]);
            dsc.na[it] = atom_na[it];
            dsc.nh[it] = atom_nh[it];
            nat += atom_na[it];
            for (int ih = 0; ih < atom_nh[it]; ++ih) {
                const int nb = (int)indv[it * nhm + ih], l = (int)nhtol[it * nhm + ih], lm = (int)nhtolm[it * nhm + ih];
                APW_CHECK(nb >= 0 && nb < atom_nb[it], "apw_getvnl: type %d projector %d: radial index %d", it, ih, nb);
                APW_CHECK(l >= 0 && l <= apw::VNL_LMAX && lm >= l * l && lm < (l + 1) * (l + 1),
                          "apw_getvnl: type %d projector %d: l %d lm %d (l at most %d)", it, ih, l, lm, apw::VNL_LMAX);
                dsc.nb[it][ih] = (unsigned char)nb;
                dsc.l[it][ih] = (unsigned char)l;
                dsc.lm[it][ih] = (unsigned char)lm;
                lmax = std::max(lmax, l);
            }
        }
        if (npw == 0 || nat == 0) return APW_OK;
        apw::DeviceGuard gd(ctx->device);
        // tau goes up as a small device array
        int* st = apw::vnl_status_clear(ctx);
        ctx->vnl_tau.reserve((size_t)3 * nat);
        APW_HIP(hipMemcpyAsync(ctx->vnl_tau.get(), tau, (size_t)3 * nat * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
        APW_HIP(hipStreamSynchronize(ctx->stream));
        const size_t nkb_rows = [&] {
            size_t r = 0;
            for (int it = 0; it < ntype; ++it) r += (size_t)atom_na[it] * atom_nh[it];
            return r;
        }();
        apw::ProfScope ps(&ctx->prof, ctx->stream, "getvnl", 16.0 * (double)nkb_rows * npw + 24.0 * npw, 0.0);
        const dim3 grid(apw::grid1(npw)), blk(256);
        double2* out = reinterpret_cast<double2*>(vkb);
        switch (lmax) {
            case 0: hipLaunchKernelGGL(apw::k_getvnl<0>, grid, blk, 0, ctx->stream, dsc, npw, npwx, tab_2, tab_3, dq, tpiba, gk, ctx->vnl_tau.get(), tab, out, st); break;
            case 1: hipLaunchKernelGGL(apw::k_getvnl<1>, grid, blk, 0, ctx->stream, dsc, npw, npwx, tab_2, tab_3, dq, tpiba, gk, ctx->vnl_tau.get(), tab, out, st); break;
            case 2: hipLaunchKernelGGL(apw::k_getvnl<2>, grid, blk, 0, ctx->stream, dsc, npw, npwx, tab_2, tab_3, dq, tpiba, gk, ctx->vnl_tau.get(), tab, out, st); break;
            default: hipLaunchKernelGGL(apw::k_getvnl<3>, grid, blk, 0, ctx->stream, dsc, npw, npwx, tab_2, tab_3, dq, tpiba, gk, ctx->vnl_tau.get(), tab, out, st); break;
        }
        APW_LAUNCH_CHECK();
        apw::vnl_status_check(ctx, "apw_getvnl");
        return APW_OK;
    });
}

}  // extern "C"
