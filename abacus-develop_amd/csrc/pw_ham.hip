// Hamiltonian of one k-point (HamiltPW operator chain) and the device functors of the
// H|psi> path, behind the C ABI of include/apw.h.
//
//   Operator::hPsi          operator.cpp:39-88       -> apw_hpsi
//   Ekinetic::act           ekinetic_pw.cpp:34-58    -> apw_act_ekinetic
//   Veff::act               veff_pw.cpp:38-105       -> apw_act_veff
//   Nonlocal::act           nonlocal_pw.cpp:211-304  -> apw_act_nonlocal
//
// apw_hpsi fuses the chain: one batched local pass writes hpsi = gk2*tpiba2*psi + V psi
// (kinetic fused into the last FFT pass), then the nonlocal part is accumulated with two
// MFMA GEMMs around a block-diagonal D scaling:
//   becp = vkb^H psi    (nkb x nb, split-K over npw)
//   ps   = D becp       (per atom, nh x nh block of deeq)
//   hpsi += vkb ps
#include <algorithm>
#include <cstring>
#include <vector>

#include "apw_ham.hpp"

namespace apw {
namespace {

// ---- functor mirrors ------------------------------------------------------------------
// ekinetic_pw_op: grid nband, one block strides over npw (cuda/ekinetic_op.cu:13-48)
__global__ void k_ekinetic(int npw, int max_npw, double tpiba2, const double* __restrict__ gk2,
                           double2* __restrict__ hpsi, const double2* __restrict__ psi) {
    const int ib = blockIdx.x;
    for (int ig = threadIdx.x; ig < npw; ig += blockDim.x) {
        const size_t i = (size_t)ib * max_npw + ig;
        const double k = gk2[ig] * tpiba2;
        double2 h = hpsi[i];
        const double2 p = psi[i];
        h.x += k * p.x;
        h.y += k * p.y;
        hpsi[i] = h;
    }
}

// elecstate_pw_op (nspin 1/2): rho[ir] += w1 |wfcr[ir]|^2
__global__ void k_elecstate(int nrxx, double w1, double* __restrict__ rho, const double2* __restrict__ wfcr) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nrxx; i += gridDim.x * blockDim.x) {
        const double2 a = wfcr[i];
        rho[i] += w1 * (a.x * a.x + a.y * a.y);
    }
}

// non-collinear form: charge from both spinor components; mode 2 = DOMAG (all of m),
// 1 = DOMAG_Z (m_z accumulated, m_x = m_y = 0), 0 = m set to 0
__global__ void k_elecstate_nc(int nrxx, int mode, double w1, double* __restrict__ r0, double* __restrict__ r1,
                               double* __restrict__ r2, double* __restrict__ r3, const double2* __restrict__ a,
                               const double2* __restrict__ b) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nrxx; i += gridDim.x * blockDim.x) {
        const double2 u = a[i], v = b[i];
        const double nu = u.x * u.x + u.y * u.y, nv = v.x * v.x + v.y * v.y;
        r0[i] += w1 * (nu + nv);
        if (mode == 2) {
            r1[i] += w1 * 2.0 * (u.x * v.x + u.y * v.y);
            r2[i] += w1 * 2.0 * (u.x * v.y - v.x * u.y);
            r3[i] += w1 * (nu - nv);
        } else if (mode == 1) {
            r1[i] = 0.0;
            r2[i] = 0.0;
            r3[i] += w1 * (nu - nv);
        } else {
            r1[i] = 0.0;
            r2[i] = 0.0;
            r3[i] = 0.0;
        }
    }
}

__global__ void k_veff(int size, double2* __restrict__ out, const double* __restrict__ in) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < size; i += gridDim.x * blockDim.x) {
        double2 o = out[i];
        o.x *= in[i];
        o.y *= in[i];
        out[i] = o;
    }
}

// nonlocal_pw_op: ps[(sum+ii*l3+kk)*l2 + jj] += sum_xx deeq[((spin*dx + iat+ii)*dy + xx)*dz + kk]
//                                               * becp[jj*nkb + sum + ii*l3 + xx]
__global__ void k_nonlocal_ref(int l1, int l2, int l3, int sum, int iat, int spin, int nkb, int dx, int dy, int dz,
                               const double* __restrict__ deeq, double2* __restrict__ ps,
                               const double2* __restrict__ becp) {
    const int total = l1 * l2 * l3;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int kk = t % l3;
        const int jj = (t / l3) % l2;
        const int ii = t / (l3 * l2);
        double sr = 0, si = 0;
        for (int xx = 0; xx < l3; ++xx) {
            const double d = deeq[(((size_t)spin * dx + iat + ii) * dy + xx) * dz + kk];
            const double2 b = becp[(size_t)jj * nkb + sum + ii * l3 + xx];
            sr += d * b.x;
            si += d * b.y;
        }
        double2* p = ps + (size_t)(sum + ii * l3 + kk) * l2 + jj;
        double2 o = *p;
        o.x += sr;
        o.y += si;
        *p = o;
    }
}

__global__ void k_set_box(int npwk, const int* __restrict__ idx, const double2* __restrict__ in,
                          double2* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npwk; i += gridDim.x * blockDim.x) out[idx[i]] = in[i];
}

__global__ void k_r2r_out(int n, int add, double factor, const double2* __restrict__ in, double2* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double2 v = in[i];
        if (add) {
            double2 o = out[i];
            o.x += factor * v.x;
            o.y += factor * v.y;
            out[i] = o;
        } else {
            out[i] = v;
        }
    }
}

__global__ void k_rr2_out(int npw, double scale, int add, const int* __restrict__ idx,
                          const double2* __restrict__ in, double2* __restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npw; i += gridDim.x * blockDim.x) {
        const double2 v = in[idx[i]];
        double2 o;
        if (add) {
            o = out[i];
            o.x += scale * v.x;
            o.y += scale * v.y;
        } else {
            o.x = v.x * scale;
            o.y = v.y * scale;
        }
        out[i] = o;
    }
}

// ---- path kernels ----------------------------------------------------------------------
// ps[ikb + ib*nkb] = sum_x deeq[(iat*nhm + x)*nhm + k] * becp[(off + x) + ib*nkb],  ikb = off + k
template <class C>
__global__ void k_dscale(int nkb, int nb, int nhm, const int* __restrict__ kb_iat, const int* __restrict__ kb_off,
                         const int* __restrict__ kb_nh, const double* __restrict__ deeq, const C* __restrict__ becp,
                         C* __restrict__ ps) {
    const int total = nkb * nb;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int ikb = t % nkb, ib = t / nkb;
        const int iat = kb_iat[ikb], off = kb_off[ikb], nh = kb_nh[ikb];
        const int k = ikb - off;
        double sr = 0, si = 0;
        for (int x = 0; x < nh; ++x) {
            const double d = deeq[((size_t)iat * nhm + x) * nhm + k];
            const C b = becp[(size_t)ib * nkb + off + x];
            sr += d * (double)b.x;
            si += d * (double)b.y;
        }
        C o;
        o.x = sr;
        o.y = si;
        ps[(size_t)ib * nkb + ikb] = o;
    }
}

__global__ void k_d2f(size_t n, const double* __restrict__ in, float* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

// copy nb columns of length rows, converting c128 -> c64 (ld_in -> ld_out)
__global__ void k_cols_d2f(int rows, int nb, const double2* __restrict__ in, int ldi, float2* __restrict__ out,
                           int ldo) {
    const size_t total = (size_t)rows * nb;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t % rows), c = (int)(t / rows);
        const double2 v = in[(size_t)c * ldi + r];
        out[(size_t)c * ldo + r] = make_float2((float)v.x, (float)v.y);
    }
}

// hpsi[:, c] += (c128) h32[:, c]
__global__ void k_cols_addf(int rows, int nb, const float2* __restrict__ in, int ldi, double2* __restrict__ out,
                            int ldo) {
    const size_t total = (size_t)rows * nb;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(t % rows), c = (int)(t / rows);
        const float2 v = in[(size_t)c * ldi + r];
        double2 o = out[(size_t)c * ldo + r];
        o.x += (double)v.x;
        o.y += (double)v.y;
        out[(size_t)c * ldo + r] = o;
    }
}

int grid_for(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 16384)); }

}  // namespace

void Ham::create(apw_ctx* c, const apw_ham_desc* d) {
    ctx = c;
    APW_CHECK(d, "apw_ham_create: null desc");
    APW_CHECK(d->npw > 0 && d->npwx >= d->npw, "apw_ham_create: need 0 < npw <= npwx");
    APW_CHECK(d->ig2ixyz && d->gk2 && d->veff, "apw_ham_create: null ig2ixyz/gk2/veff");
    APW_CHECK(d->nkb >= 0 && (d->nkb == 0 || (d->vkb && d->deeq && d->na && d->nh)), "apw_ham_create: nonlocal arrays");
    APW_CHECK(d->precision == APW_FP64 || d->precision == APW_FP32_MIXED, "apw_ham_create: bad precision");
    npw = d->npw;
    npwx = d->npwx;
    nkb = d->nkb;
    tpiba2 = d->tpiba2;
    precision = d->precision;
    grid.build(d->nx, d->ny, d->nz, d->npw, d->ig2ixyz);
    const size_t nxyz = (size_t)d->nx * d->ny * d->nz;
    gk2.reserve(npw);
    APW_HIP(hipMemcpy(gk2.get(), d->gk2, npw * sizeof(double), hipMemcpyHostToDevice));
    veff.reserve(nxyz);
    APW_HIP(hipMemcpy(veff.get(), d->veff, nxyz * sizeof(double), hipMemcpyHostToDevice));
    veff32.reserve(nxyz);
    hipLaunchKernelGGL(k_d2f, dim3(grid_for(nxyz)), dim3(256), 0, ctx->stream, nxyz, veff.get(), veff32.get());
    APW_LAUNCH_CHECK();
    if (nkb > 0) {
        // vkb: row-major nkb x npwx == column-major npwx x nkb
        vkb.reserve((size_t)2 * npwx * nkb);
        APW_HIP(hipMemcpy(vkb.get(), d->vkb, (size_t)2 * npwx * nkb * sizeof(double), hipMemcpyHostToDevice));
        vkb32.reserve((size_t)2 * npwx * nkb);
        hipLaunchKernelGGL(k_d2f, dim3(grid_for((size_t)2 * npwx * nkb)), dim3(256), 0, ctx->stream,
                           (size_t)2 * npwx * nkb, vkb.get(), vkb32.get());
        APW_LAUNCH_CHECK();
        std::vector<int> iat_of(nkb), off_of(nkb), nh_of(nkb);
        int ikb = 0, iat = 0;
        nhm = d->nhm;
        for (int it = 0; it < d->ntype; ++it) {
            APW_CHECK(d->nh[it] <= d->nhm, "apw_ham_create: nh[%d] > nhm", it);
            for (int ia = 0; ia < d->na[it]; ++ia) {
                for (int ih = 0; ih < d->nh[it]; ++ih) {
                    APW_CHECK(ikb < nkb, "apw_ham_create: sum(na*nh) exceeds nkb");
                    iat_of[ikb] = iat;
                    off_of[ikb] = ikb - ih;
                    nh_of[ikb] = d->nh[it];
                    ++ikb;
                }
                ++iat;
            }
        }
        APW_CHECK(ikb == nkb, "apw_ham_create: sum(na*nh) = %d != nkb = %d", ikb, nkb);
        nat = iat;
        deeq.reserve((size_t)nat * nhm * nhm);
        APW_HIP(hipMemcpy(deeq.get(), d->deeq, (size_t)nat * nhm * nhm * sizeof(double), hipMemcpyHostToDevice));
        auto up = [](DevBuf<int>& dst, const std::vector<int>& h) {
            dst.reserve(h.size());
            APW_HIP(hipMemcpy(dst.get(), h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice));
        };
        up(kb_iat, iat_of);
        up(kb_off, off_of);
        up(kb_nh, nh_of);
    }
    // band chunk of the local pass (fixed rule): keep the S + X intermediates of one chunk
    // near the size of the Infinity Cache so the five passes mostly hit it; FP32
    // intermediates take half the bytes: twice the bands per chunk
    const size_t per_band = grid.bytes_S(16) + grid.bytes_X(16);
    chunk = (int)std::max<size_t>(1, std::min<size_t>(256, (size_t)(192u << 20) / per_band));
    chunk32 = std::min(256, 2 * chunk);
    if (tune_mode() > 0) tune_local();
    const size_t bands16 = (size_t)std::max(chunk, (chunk32 + 1) / 2);   // c128-equivalent bands
    S.reserve(2 * grid.bytes_S(1) * bands16);   // sized here: no allocation inside H|psi>
    X.reserve(2 * grid.bytes_X(1) * bands16);
    APW_HIP(hipStreamSynchronize(ctx->stream));
}

// Autotuner of the local pass, per precision: the bands per launch and the pass schedule
// (register four-step kernels, or the LDS Stockham kernels with the run-time radix plan on
// the x, y or z passes), each candidate timed over 64 bands of the sphere.
void Ham::tune_local() {
    const int nbt = 64;
    const size_t per_band16 = grid.bytes_S(16) + grid.bytes_X(16);
    for (int pi = 0; pi < 2; ++pi) {
        int& ch = pi ? chunk32 : chunk;
        char key[160];
        std::snprintf(key, sizeof(key), "fft %s %dx%dx%d npw=%d cu=%d", pi ? "f32" : "f64", grid.nx, grid.ny, grid.nz,
                      npw, ctx->num_cu);
        int a = 0, b = 0;
        if (tune_find(key, a, b) && a >= 1 && a <= 256 && b >= 0 && b <= 7) {
            ch = a;
            grid.sched[pi] = b;
            continue;
        }
        if (tune_mode() != 2) continue;
        DevBuf<double> tp, th;
        tp.reserve((size_t)2 * npwx * nbt);
        th.reserve((size_t)2 * npwx * nbt);
        APW_HIP(hipMemsetAsync(tp.get(), 0, (size_t)16 * npwx * nbt, ctx->stream));
        std::vector<int> chunks;
        const int rule = ch;
        const size_t per_band = pi ? per_band16 / 2 : per_band16;
        for (int c : {rule / 2, rule, 2 * rule, 4 * rule, 8, 16, 32, 64}) {
            c = std::max(1, std::min(c, nbt));
            if ((size_t)c * per_band > ((size_t)2 << 30) && c > rule) continue;   // at most 2 GiB of intermediates
            if (std::find(chunks.begin(), chunks.end(), c) == chunks.end()) chunks.push_back(c);
        }
        std::vector<int> scheds = {0};
        if (xreg_available(grid)) scheds.push_back(1);
        if (yreg_available(grid)) scheds.push_back(2);
        if (zreg_available(grid, npwx)) scheds.push_back(4);
        double best = 1e300;
        int best_c = rule, best_s = 0;
        const int prec = pi ? APW_FP32_MIXED : APW_FP64;
        for (int sc : scheds)
            for (int c : chunks) {
                grid.sched[pi] = sc;
                ch = c;
                const double us = time_on_stream(ctx->stream, 2, [&] {
                    local(tp.get(), npwx, th.get(), npwx, nbt, GATHER_SET, prec);
                });
                if (us < best) {
                    best = us;
                    best_c = c;
                    best_s = sc;
                }
            }
        ch = best_c;
        grid.sched[pi] = best_s;
        tune_store(key, best_c, best_s, best);
    }
}

void Ham::local(const double* psi, int ld, double* hpsi, int ldh, int nb, GatherMode mode, int prec) {
    const double scale = 1.0 / (double)((size_t)grid.nx * grid.ny * grid.nz);
    const int cb = prec == APW_FP64 ? chunk : chunk32;
    for (int b0 = 0; b0 < nb; b0 += cb) {
        const int n = std::min(cb, nb - b0);
        if (prec == APW_FP64) {
            S.reserve(2 * grid.bytes_S(1) * n);
            X.reserve(2 * grid.bytes_X(1) * n);
            local_hpsi<double>(ctx->stream, grid, psi + (size_t)2 * b0 * ld, ld, hpsi + (size_t)2 * b0 * ldh, ldh, n,
                               veff.get(), gk2.get(), tpiba2, mode, scale, S.get(), X.get(), &ctx->prof);
        } else {
            S.reserve(grid.bytes_S(1) * n);
            X.reserve(grid.bytes_X(1) * n);
            local_hpsi<float>(ctx->stream, grid, psi + (size_t)2 * b0 * ld, ld, hpsi + (size_t)2 * b0 * ldh, ldh, n,
                              veff32.get(), gk2.get(), tpiba2, mode, scale, reinterpret_cast<float*>(S.get()),
                              reinterpret_cast<float*>(X.get()), &ctx->prof);
        }
    }
}

void Ham::nonlocal(const double* psi, int ld, double* hpsi, int ldh, int nb, int prec) {
    if (nkb == 0 || nb == 0) return;
    hipStream_t s = ctx->stream;
    const int ncu = ctx->num_cu;
    if (prec == APW_FP64) {
        becp.reserve((size_t)2 * nkb * nb);
        ps.reserve((size_t)2 * nkb * nb);
        gemm<double>(s, gws, ncu, 'C', 'N', nkb, nb, npw, 1.0, 0.0, vkb.get(), npwx, psi, ld, 0.0, 0.0, becp.get(),
                     nkb, &ctx->prof);
        hipLaunchKernelGGL(k_dscale<double2>, dim3(grid_for((size_t)nkb * nb)), dim3(256), 0, s, nkb, nb, nhm,
                           kb_iat.get(), kb_off.get(), kb_nh.get(), deeq.get(),
                           reinterpret_cast<const double2*>(becp.get()), reinterpret_cast<double2*>(ps.get()));
        APW_LAUNCH_CHECK();
        gemm<double>(s, gws, ncu, 'N', 'N', npw, nb, nkb, 1.0, 0.0, vkb.get(), npwx, ps.get(), nkb, 1.0, 0.0, hpsi,
                     ldh, &ctx->prof);
    } else {
        // FP32 arithmetic, FP32 accumulation; psi converted on the way in, the product
        // added to the FP64 hpsi on the way out
        psi32.reserve((size_t)npw * nb);                       // c64 = 1 double each
        h32.reserve((size_t)npw * nb);
        becp.reserve((size_t)nkb * nb);
        ps.reserve((size_t)nkb * nb);
        float* p32 = reinterpret_cast<float*>(psi32.get());
        float* hh = reinterpret_cast<float*>(h32.get());
        float* b32 = reinterpret_cast<float*>(becp.get());
        float* s32 = reinterpret_cast<float*>(ps.get());
        hipLaunchKernelGGL(k_cols_d2f, dim3(grid_for((size_t)npw * nb)), dim3(256), 0, s, npw, nb,
                           reinterpret_cast<const double2*>(psi), ld, reinterpret_cast<float2*>(p32), npw);
        APW_LAUNCH_CHECK();
        gemm<float>(s, gws, ncu, 'C', 'N', nkb, nb, npw, 1.0, 0.0, vkb32.get(), npwx, p32, npw, 0.0, 0.0, b32, nkb, &ctx->prof);
        hipLaunchKernelGGL(k_dscale<float2>, dim3(grid_for((size_t)nkb * nb)), dim3(256), 0, s, nkb, nb, nhm,
                           kb_iat.get(), kb_off.get(), kb_nh.get(), deeq.get(), reinterpret_cast<const float2*>(b32),
                           reinterpret_cast<float2*>(s32));
        APW_LAUNCH_CHECK();
        gemm<float>(s, gws, ncu, 'N', 'N', npw, nb, nkb, 1.0, 0.0, vkb32.get(), npwx, s32, nkb, 0.0, 0.0, hh, npw, &ctx->prof);
        hipLaunchKernelGGL(k_cols_addf, dim3(grid_for((size_t)npw * nb)), dim3(256), 0, s, npw, nb,
                           reinterpret_cast<const float2*>(hh), npw, reinterpret_cast<double2*>(hpsi), ldh);
        APW_LAUNCH_CHECK();
    }
}

void Ham::hpsi(const double* psi, int ld, double* out, int ldh, int nb, int prec) {
    APW_CHECK(ld >= npw && ldh >= npw, "hpsi: leading dimension < npw");
    local(psi, ld, out, ldh, nb, GATHER_SET_KIN, prec);
    nonlocal(psi, ld, out, ldh, nb, prec);
}

void Ham::act_ekinetic(int nb, int nbasis, const double* psi, double* out) {
    APW_CHECK(nbasis >= npw, "act_ekinetic: nbasis < npw");
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_ekinetic, dim3(nb), dim3(256), 0, ctx->stream, npw, nbasis, tpiba2, gk2.get(),
                       reinterpret_cast<double2*>(out), reinterpret_cast<const double2*>(psi));
    APW_LAUNCH_CHECK();
}

}  // namespace apw

using apw::guarded;

extern "C" {

int apw_ham_create(apw_ctx* ctx, const apw_ham_desc* desc, apw_ham** out) {
    return guarded([&] {
        APW_CHECK(ctx && out, "apw_ham_create: null argument");
        apw::DeviceGuard g(ctx->device);
        auto* h = new apw_ham();
        try {
            h->h.create(ctx, desc);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return APW_OK;
    });
}

int apw_ham_destroy(apw_ham* ham) {
    return guarded([&] {
        if (!ham) return APW_OK;
        apw::DeviceGuard g(ham->h.ctx->device);
        (void)hipStreamSynchronize(ham->h.ctx->stream);
        ham->h.ctx->prof.scrub(ham->h.ctx->stream);   // profiled launches must not pin the freed buffers
        delete ham;
        return APW_OK;
    });
}

int apw_ham_set_precision(apw_ham* ham, int precision) {
    return guarded([&] {
        APW_CHECK(ham, "null ham");
        APW_CHECK(precision == APW_FP64 || precision == APW_FP32_MIXED, "bad precision %d", precision);
        ham->h.precision = precision;
        return APW_OK;
    });
}

int apw_ham_set_vkb(apw_ham* ham, const double* vkb, int ldvkb) {
    return guarded([&] {
        APW_CHECK(ham && vkb, "apw_ham_set_vkb: null argument");
        apw::Ham& h = ham->h;
        APW_CHECK(h.nkb > 0, "apw_ham_set_vkb: the Hamiltonian has no projectors");
        APW_CHECK(ldvkb >= h.npw, "apw_ham_set_vkb: ldvkb %d < npw %d", ldvkb, h.npw);
        apw::DeviceGuard g(h.ctx->device);
        // rows of npw plane waves into the npwx-strided buffer, then the FP32 shadow
        APW_HIP(hipMemcpy2DAsync(h.vkb.get(), (size_t)2 * h.npwx * sizeof(double), vkb, (size_t)2 * ldvkb * sizeof(double),
                                 (size_t)2 * h.npw * sizeof(double), h.nkb, hipMemcpyDeviceToDevice, h.ctx->stream));
        const size_t n = (size_t)2 * h.npwx * h.nkb;
        hipLaunchKernelGGL(apw::k_d2f, dim3(apw::grid_for(n)), dim3(256), 0, h.ctx->stream, n, h.vkb.get(), h.vkb32.get());
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_hpsi(apw_ham* ham, const double* psi, int ldpsi, double* hpsi, int ldhpsi, int nbands) {
    return guarded([&] {
        APW_CHECK(ham && psi && hpsi, "apw_hpsi: null argument");
        APW_CHECK(nbands >= 0, "apw_hpsi: nbands < 0");
        apw::DeviceGuard g(ham->h.ctx->device);
        ham->h.hpsi(psi, ldpsi, hpsi, ldhpsi, nbands, ham->h.precision);
        return APW_OK;
    });
}

int apw_act_ekinetic(apw_ham* ham, int nbands, int nbasis, const double* psi, double* hpsi) {
    return guarded([&] {
        APW_CHECK(ham && psi && hpsi, "apw_act_ekinetic: null argument");
        apw::DeviceGuard g(ham->h.ctx->device);
        ham->h.act_ekinetic(nbands, nbasis, psi, hpsi);
        return APW_OK;
    });
}

int apw_act_veff(apw_ham* ham, int nbands, int nbasis, const double* psi, double* hpsi) {
    return guarded([&] {
        APW_CHECK(ham && psi && hpsi, "apw_act_veff: null argument");
        APW_CHECK(nbasis >= ham->h.npw, "apw_act_veff: nbasis < npw");
        apw::DeviceGuard g(ham->h.ctx->device);
        ham->h.local(psi, nbasis, hpsi, nbasis, nbands, apw::GATHER_ADD, ham->h.precision);
        return APW_OK;
    });
}

int apw_act_nonlocal(apw_ham* ham, int nbands, int nbasis, const double* psi, double* hpsi) {
    return guarded([&] {
        APW_CHECK(ham && psi && hpsi, "apw_act_nonlocal: null argument");
        APW_CHECK(nbasis >= ham->h.npw, "apw_act_nonlocal: nbasis < npw");
        apw::DeviceGuard g(ham->h.ctx->device);
        ham->h.nonlocal(psi, nbasis, hpsi, nbasis, nbands, ham->h.precision);
        return APW_OK;
    });
}

int apw_recip_to_real(apw_ham* ham, const double* in, double* out, int add, double factor) {
    return guarded([&] {
        APW_CHECK(ham && in && out, "apw_recip_to_real: null argument");
        apw::DeviceGuard g(ham->h.ctx->device);
        apw::Ham& h = ham->h;
        h.S.reserve(2 * h.grid.bytes_S(1));
        h.X.reserve(2 * h.grid.bytes_X(1));
        apw::sphere_to_box(h.ctx->stream, h.grid, in, h.npw, out, 1, add != 0, factor, h.S.get(), h.X.get());
        return APW_OK;
    });
}

int apw_real_to_recip(apw_ham* ham, const double* in, double* out, int add, double factor) {
    return guarded([&] {
        APW_CHECK(ham && in && out, "apw_real_to_recip: null argument");
        apw::DeviceGuard g(ham->h.ctx->device);
        apw::Ham& h = ham->h;
        h.S.reserve(2 * h.grid.bytes_S(1));
        h.X.reserve(2 * h.grid.bytes_X(1));
        const double nxyz = (double)((size_t)h.grid.nx * h.grid.ny * h.grid.nz);
        // reference: out = box/nxyz, or out += factor/nxyz * box  (pw_op.cpp:45-67)
        const double scale = add ? factor / nxyz : 1.0 / nxyz;
        apw::box_to_sphere(h.ctx->stream, h.grid, in, out, h.npw, 1, add ? apw::GATHER_ADD : apw::GATHER_SET, scale,
                           h.S.get(), h.X.get());
        return APW_OK;
    });
}

int apw_ekinetic_pw_op(apw_ctx* ctx, int nband, int npw, int max_npw, double tpiba2, const double* gk2,
                       double* hpsi, const double* psi) {
    return guarded([&] {
        APW_CHECK(ctx && gk2 && hpsi && psi, "apw_ekinetic_pw_op: null argument");
        APW_CHECK(nband >= 0 && npw >= 0 && max_npw >= npw, "apw_ekinetic_pw_op: bad sizes");
        if (nband == 0) return APW_OK;
        apw::DeviceGuard g(ctx->device);
        hipLaunchKernelGGL(apw::k_ekinetic, dim3(nband), dim3(256), 0, ctx->stream, npw, max_npw, tpiba2, gk2,
                           reinterpret_cast<double2*>(hpsi), reinterpret_cast<const double2*>(psi));
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_veff_pw_op(apw_ctx* ctx, int size, double* out, const double* in) {
    return guarded([&] {
        APW_CHECK(ctx && out && in && size >= 0, "apw_veff_pw_op: bad argument");
        if (size == 0) return APW_OK;
        apw::DeviceGuard g(ctx->device);
        hipLaunchKernelGGL(apw::k_veff, dim3(apw::grid_for(size)), dim3(256), 0, ctx->stream, size,
                           reinterpret_cast<double2*>(out), in);
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_elecstate_pw_op(apw_ctx* ctx, int spin, int nrxx, double w1, double* const* rho, const double* wfcr) {
    return guarded([&] {
        APW_CHECK(ctx && rho && wfcr && nrxx >= 0 && spin >= 0, "apw_elecstate_pw_op: bad argument");
        APW_CHECK(rho[spin], "apw_elecstate_pw_op: null rho[spin]");
        if (nrxx == 0) return APW_OK;
        apw::DeviceGuard g(ctx->device);
        hipLaunchKernelGGL(apw::k_elecstate, dim3(apw::grid_for(nrxx)), dim3(256), 0, ctx->stream, nrxx, w1,
                           rho[spin], reinterpret_cast<const double2*>(wfcr));
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_elecstate_pw_op_nc(apw_ctx* ctx, int domag, int domag_z, int nrxx, double w1, double* const* rho,
                           const double* wfcr, const double* wfcr_another_spin) {
    return guarded([&] {
        APW_CHECK(ctx && rho && wfcr && wfcr_another_spin && nrxx >= 0, "apw_elecstate_pw_op_nc: bad argument");
        APW_CHECK(rho[0] && rho[1] && rho[2] && rho[3], "apw_elecstate_pw_op_nc: null rho component");
        if (nrxx == 0) return APW_OK;
        apw::DeviceGuard g(ctx->device);
        const int mode = domag ? 2 : (domag_z ? 1 : 0);
        hipLaunchKernelGGL(apw::k_elecstate_nc, dim3(apw::grid_for(nrxx)), dim3(256), 0, ctx->stream, nrxx, mode, w1,
                           rho[0], rho[1], rho[2], rho[3], reinterpret_cast<const double2*>(wfcr),
                           reinterpret_cast<const double2*>(wfcr_another_spin));
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_psi_to_rho(apw_ham* ham, const double* psi, int ldpsi, int nbands, const double* wg, double omega,
                   double* rho, int add) {
    return guarded([&] {
        APW_CHECK(ham && rho && (nbands == 0 || (psi && wg)), "apw_psi_to_rho: null argument");
        APW_CHECK(nbands >= 0 && omega > 0.0, "apw_psi_to_rho: nbands=%d omega=%g", nbands, omega);
        apw::Ham& h = ham->h;
        APW_CHECK(ldpsi >= h.npw, "apw_psi_to_rho: ldpsi=%d < npw=%d", ldpsi, h.npw);
        apw::DeviceGuard g(h.ctx->device);
        hipStream_t s = h.ctx->stream;
        const size_t nxyz = (size_t)h.grid.nx * h.grid.ny * h.grid.nz;
        if (nbands == 0) {
            if (!add) APW_HIP(hipMemsetAsync(rho, 0, nxyz * sizeof(double), s));
            APW_HIP(hipStreamSynchronize(s));
            return APW_OK;
        }
        // weights wg_b / omega (elecstate_pw.cpp:200), uploaded once
        std::vector<double> w(nbands);
        for (int b = 0; b < nbands; ++b) w[b] = wg[b] / omega;
        h.wts.reserve(nbands);
        APW_HIP(hipMemcpyAsync(h.wts.get(), w.data(), nbands * sizeof(double), hipMemcpyHostToDevice, s));
        for (int b0 = 0; b0 < nbands; b0 += h.chunk) {
            const int n = std::min(h.chunk, nbands - b0);
            h.S.reserve(2 * h.grid.bytes_S(1) * n);
            h.X.reserve(2 * h.grid.bytes_X(1) * n);
            apw::psi_to_rho(s, h.grid, psi + (size_t)2 * b0 * ldpsi, ldpsi, n, h.wts.get() + b0, rho,
                            add != 0 || b0 > 0, h.S.get(), h.X.get(), &h.ctx->prof);
        }
        APW_HIP(hipStreamSynchronize(s));
        return APW_OK;
    });
}

int apw_nonlocal_pw_op(apw_ctx* ctx, int l1, int l2, int l3, int* sum, int* iat, int spin, int nkb, int deeq_x,
                       int deeq_y, int deeq_z, const double* deeq, double* ps, const double* becp) {
    return guarded([&] {
        APW_CHECK(ctx && sum && iat && deeq && ps && becp, "apw_nonlocal_pw_op: null argument");
        APW_CHECK(l1 >= 0 && l2 >= 0 && l3 >= 0 && l3 <= deeq_y && l3 <= deeq_z, "apw_nonlocal_pw_op: bad sizes");
        apw::DeviceGuard g(ctx->device);
        const size_t total = (size_t)l1 * l2 * l3;
        if (total > 0) {
            hipLaunchKernelGGL(apw::k_nonlocal_ref, dim3(apw::grid_for(total)), dim3(256), 0, ctx->stream, l1, l2, l3,
                               *sum, *iat, spin, nkb, deeq_x, deeq_y, deeq_z, deeq, reinterpret_cast<double2*>(ps),
                               reinterpret_cast<const double2*>(becp));
            APW_LAUNCH_CHECK();
        }
        *sum += l1 * l3;
        *iat += l1;
        return APW_OK;
    });
}

int apw_set_3d_fft_box_op(apw_ctx* ctx, int npwk, const int* box_index, const double* in, double* out) {
    return guarded([&] {
        APW_CHECK(ctx && box_index && in && out && npwk >= 0, "apw_set_3d_fft_box_op: bad argument");
        if (npwk == 0) return APW_OK;
        apw::DeviceGuard g(ctx->device);
        hipLaunchKernelGGL(apw::k_set_box, dim3(apw::grid_for(npwk)), dim3(256), 0, ctx->stream, npwk, box_index,
                           reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out));
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_set_recip_to_real_output_op(apw_ctx* ctx, int nrxx, int add, double factor, const double* in, double* out) {
    return guarded([&] {
        APW_CHECK(ctx && in && out && nrxx >= 0, "apw_set_recip_to_real_output_op: bad argument");
        if (nrxx == 0) return APW_OK;
        apw::DeviceGuard g(ctx->device);
        hipLaunchKernelGGL(apw::k_r2r_out, dim3(apw::grid_for(nrxx)), dim3(256), 0, ctx->stream, nrxx, add, factor,
                           reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out));
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_set_real_to_recip_output_op(apw_ctx* ctx, int npw_k, int nxyz, int add, double factor, const int* box_index,
                                    const double* in, double* out) {
    return guarded([&] {
        APW_CHECK(ctx && box_index && in && out && npw_k >= 0 && nxyz > 0, "apw_set_real_to_recip_output_op: bad argument");
        if (npw_k == 0) return APW_OK;
        apw::DeviceGuard g(ctx->device);
        const double scale = add ? factor / (double)nxyz : 1.0 / (double)nxyz;
        hipLaunchKernelGGL(apw::k_rr2_out, dim3(apw::grid_for(npw_k)), dim3(256), 0, ctx->stream, npw_k, scale, add,
                           box_index, reinterpret_cast<const double2*>(in), reinterpret_cast<double2*>(out));
        APW_LAUNCH_CHECK();
        return APW_OK;
    });
}

int apw_fft3d(apw_ctx* ctx, int nx, int ny, int nz, double* box, int dir) {
    return guarded([&] {
        APW_CHECK(ctx && box && (dir == 1 || dir == -1), "apw_fft3d: bad argument");
        apw::DeviceGuard g(ctx->device);
        // the full box as a "sphere" in z-fastest order: every stick and x slot present
        const size_t n = (size_t)nx * ny * nz;
        APW_CHECK(n < (size_t)1 << 31, "apw_fft3d: box too large");
        std::vector<int> ident(n);
        for (size_t i = 0; i < n; ++i) ident[i] = (int)i;
        apw::PwGrid grid;
        grid.build(nx, ny, nz, (int)n, ident.data());
        apw::DevBuf<double> S, X;
        S.reserve(2 * grid.bytes_S(1));
        X.reserve(2 * grid.bytes_X(1));
        if (dir == 1) {
            apw::sphere_to_box(ctx->stream, grid, box, (int)n, box, 1, false, 1.0, S.get(), X.get());
        } else {
            apw::box_to_sphere(ctx->stream, grid, box, box, (int)n, 1, apw::GATHER_SET, 1.0, S.get(), X.get());
        }
        APW_HIP(hipStreamSynchronize(ctx->stream));
        return APW_OK;
    });
}

}  // extern "C"
