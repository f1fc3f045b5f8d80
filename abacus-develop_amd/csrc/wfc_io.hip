// Plane-wave wavefunction files of one k-point (ModuleIO::write_wfc_pw / read_wfc_pw,
// module_io/write_wfc_pw.cpp:13-283, read_wfc_pw.cpp:12-341), npol 1.
//
// The file layouts are the reference's, so files move between the two programs:
// * ".dat" (out_wfc_pw 2): Fortran-style records, each framed by its byte count as a
//   4-byte int: header {ik+1, nkstot, kvec_c[3], weight, npwtot, nbands, ecutwfc, lat0,
//   tpiba} (72 B), the reciprocal vectors G (9 doubles), the FFT-box indices (ix, iy, iz)
//   of every plane wave in basis order, then one record of npwtot complex doubles per band;
// * ".txt" (out_wfc_pw 1): the same content as formatted text (written only; the reference
//   reads .dat files alone).
// psi stays in HBM: the writer stages it through pinned host memory band block by band
// block; the reader uploads the file's bands and a gather kernel puts each plane wave at
// its place in this basis (the file's Miller order may differ from the basis order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <memory>
#include <string>
#include <vector>

#include "apw_internal.hpp"

namespace apw {
namespace {

// wfc[b * ld + i] = in[b * ntot + perm[i]]
__global__ void k_wfc_gather(int nb, int npw, int ntot, const double2* __restrict__ in, const int* __restrict__ perm,
                             double2* __restrict__ out, int ld) {
    const size_t total = (size_t)nb * npw;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int b = (int)(t / npw), i = (int)(t % npw);
        out[(size_t)b * ld + i] = in[(size_t)b * ntot + perm[i]];
    }
}

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

struct Pinned {
    void* p = nullptr;
    explicit Pinned(size_t bytes) {
        APW_HIP(hipHostMalloc(&p, std::max<size_t>(bytes, 16), hipHostMallocDefault));
    }
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

template <class T>
void put(FILE* f, const T& v) {
    APW_CHECK(std::fwrite(&v, sizeof(T), 1, f) == 1, "write_wfc_pw: write failed");
}

template <class T>
void get(FILE* f, T& v, const char* what) {
    APW_CHECK(std::fread(&v, sizeof(T), 1, f) == 1, "read_wfc_pw: file ends inside the %s record", what);
}

// record frame: the byte count before and after
void frame(FILE* f, int expect, const char* what) {
    int n = 0;
    get(f, n, what);
    APW_CHECK(n == expect, "read_wfc_pw: %s record of %d bytes, expected %d", what, n, expect);
}

std::string suffix(const char* fn) {
    const std::string s(fn);
    return s.size() >= 3 ? s.substr(s.size() - 3) : s;
}

}  // namespace
}  // namespace apw

using apw::guarded;

extern "C" {

int apw_write_wfc_pw(apw_ctx* ctx, const char* filename, int ik_global, int nkstot, const double* kvec_c,
                     double weight, int nbands, double ecutwfc, double lat0, double tpiba, const double* G, int nx,
                     int ny, int nz, int npw, const int* ig2ixyz, const double* psi, int ldpsi) {
    return guarded([&] {
        APW_CHECK(ctx && filename && kvec_c && G && ig2ixyz && (psi || nbands == 0), "apw_write_wfc_pw: null argument");
        APW_CHECK(npw >= 0 && ldpsi >= npw && nbands >= 0 && nx > 0 && ny > 0 && nz > 0,
                  "apw_write_wfc_pw: bad sizes");
        const std::string ft = apw::suffix(filename);
        APW_CHECK(ft == "dat" || ft == "txt", "apw_write_wfc_pw: unknown file type %s", ft.c_str());
        apw::DeviceGuard gd(ctx->device);
        const size_t nxyz = (size_t)nx * ny * nz;
        std::vector<int> mil((size_t)3 * npw);
        for (int i = 0; i < npw; ++i) {
            const int idx = ig2ixyz[i];
            APW_CHECK(idx >= 0 && (size_t)idx < nxyz, "apw_write_wfc_pw: ig2ixyz[%d] = %d outside the box", i, idx);
            mil[3 * i] = idx / (ny * nz);
            mil[3 * i + 1] = (idx / nz) % ny;
            mil[3 * i + 2] = idx % nz;
        }
        // bands come down in blocks through pinned memory
        const int bb = std::max(1, std::min(nbands, (int)std::max<size_t>(1, (64u << 20) / (16 * (size_t)std::max(npw, 1)))));
        apw::Pinned stage((size_t)16 * npw * bb);
        double* hb = static_cast<double*>(stage.p);
        auto fetch = [&](int b0, int nb) {
            APW_HIP(hipMemcpy2DAsync(hb, (size_t)16 * npw, psi + (size_t)2 * b0 * ldpsi, (size_t)16 * ldpsi,
                                     (size_t)16 * npw, nb, hipMemcpyDeviceToHost, ctx->stream));
            APW_HIP(hipStreamSynchronize(ctx->stream));
        };
        if (ft == "dat") {
            apw::File f;
            f.f = std::fopen(filename, "wb");
            APW_CHECK(f.f, "apw_write_wfc_pw: can't open file %s", filename);
            using apw::put;
            put(f.f, 72);
            put(f.f, ik_global + 1);
            put(f.f, nkstot);
            for (int d = 0; d < 3; ++d) put(f.f, kvec_c[d]);
            put(f.f, weight);
            put(f.f, npw);
            put(f.f, nbands);
            put(f.f, ecutwfc);
            put(f.f, lat0);
            put(f.f, tpiba);
            put(f.f, 72);
            put(f.f, 72);
            for (int d = 0; d < 9; ++d) put(f.f, G[d]);
            put(f.f, 72);
            put(f.f, npw * 12);
            if (npw) APW_CHECK(std::fwrite(mil.data(), sizeof(int), mil.size(), f.f) == mil.size(), "apw_write_wfc_pw: write failed");
            put(f.f, npw * 12);
            for (int b0 = 0; b0 < nbands; b0 += bb) {
                const int nb = std::min(bb, nbands - b0);
                fetch(b0, nb);
                for (int b = 0; b < nb; ++b) {
                    put(f.f, npw * 16);
                    if (npw) APW_CHECK(std::fwrite(hb + (size_t)2 * b * npw, 16, npw, f.f) == (size_t)npw,
                                       "apw_write_wfc_pw: write failed");
                    put(f.f, npw * 16);
                }
            }
            APW_CHECK(std::fflush(f.f) == 0, "apw_write_wfc_pw: write failed");
        } else {
            std::ofstream o(filename);
            APW_CHECK(o.good(), "apw_write_wfc_pw: can't open file %s", filename);
            o << std::setprecision(6);
            o << std::setw(10) << "Kpoint" << std::setw(10) << "nKpoint" << std::setw(10) << "kv.x" << std::setw(10)
              << "kv.y" << std::setw(10) << "kv.z" << std::setw(10) << "weight" << std::setw(10) << "ngtot"
              << std::setw(10) << "nband" << std::setw(10) << "ecut" << std::setw(10) << "lat0" << std::setw(10)
              << "2pi/lat0" << std::endl;
            o << std::setw(10) << ik_global + 1 << std::setw(10) << nkstot << std::setw(10) << kvec_c[0]
              << std::setw(10) << kvec_c[1] << std::setw(10) << kvec_c[2] << std::setw(10) << weight << std::setw(10)
              << npw << std::setw(10) << nbands << std::setw(10) << ecutwfc << std::setw(10) << lat0 << std::setw(10)
              << tpiba << std::endl;
            o << "\n<Reciprocal Lattice Vector>" << std::endl;
            for (int r = 0; r < 3; ++r)
                o << std::setw(10) << G[3 * r] << std::setw(10) << G[3 * r + 1] << std::setw(10) << G[3 * r + 2]
                  << std::endl;
            o << "<Reciprocal Lattice Vector>\n" << std::endl;
            o << "<G vectors>" << std::endl;
            for (int i = 0; i < npw; ++i)
                o << std::setw(10) << mil[3 * i] << std::setw(10) << mil[3 * i + 1] << std::setw(10) << mil[3 * i + 2]
                  << std::endl;
            o << "<G vectors>\n" << std::endl;
            for (int b0 = 0; b0 < nbands; b0 += bb) {
                const int nb = std::min(bb, nbands - b0);
                fetch(b0, nb);
                for (int b = 0; b < nb; ++b) {
                    const double* v = hb + (size_t)2 * b * npw;
                    o << "\n< Band " << b0 + b + 1 << " >" << std::endl;
                    o << std::scientific;
                    for (int ig = 0; ig < npw; ++ig) {
                        if (ig % 4 == 0 && ig != 0) o << "\n";
                        o << std::setw(15) << v[2 * ig] << std::setw(15) << v[2 * ig + 1];
                    }
                    o << "\n< Band " << b0 + b + 1 << " >" << std::endl;
                    o << std::defaultfloat;
                }
            }
            APW_CHECK(o.good(), "apw_write_wfc_pw: write failed");
        }
        return APW_OK;
    });
}

int apw_read_wfc_pw(apw_ctx* ctx, const char* filename, int ik, int nkstot, int nbands, const double* kvec_c,
                    double lat0, double tpiba, const double* G, int nx, int ny, int nz, int npw, const int* ig2ixyz,
                    double* wfc, int ldwfc) {
    return guarded([&] {
        APW_CHECK(ctx && filename && kvec_c && G && ig2ixyz && (wfc || nbands == 0), "apw_read_wfc_pw: null argument");
        APW_CHECK(npw >= 0 && ldwfc >= npw && nbands >= 0 && nx > 0 && ny > 0 && nz > 0, "apw_read_wfc_pw: bad sizes");
        const std::string ft = apw::suffix(filename);
        APW_CHECK(ft == "dat" || ft == "txt", "apw_read_wfc_pw: unknown file type %s", ft.c_str());
        apw::File f;
        f.f = std::fopen(filename, "rb");
        APW_CHECK(f.f, "apw_read_wfc_pw: can't open file %s", filename);
        APW_CHECK(ft == "dat", "apw_read_wfc_pw: %s: only binary (.dat) wavefunction files are read", filename);
        using apw::get;
        int ik_in = 0, nks_in = 0, npw_in = 0, nb_in = 0;
        double kv[3], w = 0, ecut = 0, lat0_in = 0, tpiba_in = 0, g_in[9];
        apw::frame(f.f, 72, "header");
        get(f.f, ik_in, "header");
        get(f.f, nks_in, "header");
        for (double& x : kv) get(f.f, x, "header");
        get(f.f, w, "header");
        get(f.f, npw_in, "header");
        get(f.f, nb_in, "header");
        get(f.f, ecut, "header");
        get(f.f, lat0_in, "header");
        get(f.f, tpiba_in, "header");
        apw::frame(f.f, 72, "header");
        // the reference's consistency checks, compared exactly as it compares them
        APW_CHECK(ik_in == ik + 1 && nks_in == nkstot && npw_in == npw && nb_in == nbands,
                  "apw_read_wfc_pw: ikstot_in != ikstot || nkstot_in != nkstot || npwtot_in != npwtot || "
                  "nbands_in != nbands (file %d %d %d %d, expected %d %d %d %d)",
                  ik_in, nks_in, npw_in, nb_in, ik + 1, nkstot, npw, nbands);
        APW_CHECK(kv[0] == kvec_c[0] && kv[1] == kvec_c[1] && kv[2] == kvec_c[2],
                  "apw_read_wfc_pw: k vector in file is not the same as the one in memory");
        APW_CHECK(lat0_in == lat0 && tpiba_in == tpiba, "apw_read_wfc_pw: lat0_in != lat0 || tpiba_in != tpiba");
        apw::frame(f.f, 72, "reciprocal vectors");
        for (double& x : g_in) get(f.f, x, "reciprocal vectors");
        apw::frame(f.f, 72, "reciprocal vectors");
        for (int d = 0; d < 9; ++d) APW_CHECK(g_in[d] == G[d], "apw_read_wfc_pw: G_in != G");
        // Miller indices -> file order of every box point; then this basis' order
        const size_t nxyz = (size_t)nx * ny * nz;
        std::vector<int> mil((size_t)3 * npw);
        apw::frame(f.f, npw * 12, "Miller index");
        if (npw)
            APW_CHECK(std::fread(mil.data(), sizeof(int), mil.size(), f.f) == mil.size(),
                      "apw_read_wfc_pw: file ends inside the Miller index record");
        apw::frame(f.f, npw * 12, "Miller index");
        std::vector<int> order(nxyz, -1), perm(std::max(npw, 1));
        for (int i = 0; i < npw; ++i) {
            const int x = mil[3 * i], y = mil[3 * i + 1], z = mil[3 * i + 2];
            APW_CHECK(x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz,
                      "apw_read_wfc_pw: Miller index (%d, %d, %d) outside the %dx%dx%d box", x, y, z, nx, ny, nz);
            order[((size_t)x * ny + y) * nz + z] = i;
        }
        for (int i = 0; i < npw; ++i) {
            const int idx = ig2ixyz[i];
            APW_CHECK(idx >= 0 && (size_t)idx < nxyz && order[idx] >= 0,
                      "apw_read_wfc_pw: plane wave %d of the basis is not in the file", i);
            perm[i] = order[idx];
        }
        if (nbands == 0 || npw == 0) return APW_OK;
        apw::DeviceGuard gd(ctx->device);
        // bands: host (pinned) -> device staging -> gather into the basis order
        const size_t band_bytes = (size_t)16 * npw;
        apw::Pinned stage(band_bytes * nbands);
        char* hb = static_cast<char*>(stage.p);
        for (int b = 0; b < nbands; ++b) {
            apw::frame(f.f, npw * 16, "band");
            APW_CHECK(std::fread(hb + b * band_bytes, 16, npw, f.f) == (size_t)npw,
                      "apw_read_wfc_pw: file ends inside band %d", b + 1);
            apw::frame(f.f, npw * 16, "band");
        }
        ctx->io_stage.reserve((size_t)2 * npw * nbands);
        ctx->io_perm.reserve(npw);
        APW_HIP(hipMemcpyAsync(ctx->io_stage.get(), hb, band_bytes * nbands, hipMemcpyHostToDevice, ctx->stream));
        APW_HIP(hipMemcpyAsync(ctx->io_perm.get(), perm.data(), sizeof(int) * npw, hipMemcpyHostToDevice, ctx->stream));
        const size_t total = (size_t)nbands * npw;
        hipLaunchKernelGGL(apw::k_wfc_gather, dim3((unsigned)std::min<size_t>(8192, (total + 255) / 256)), dim3(256), 0,
                           ctx->stream, nbands, npw, npw, reinterpret_cast<const double2*>(ctx->io_stage.get()),
                           ctx->io_perm.get(), reinterpret_cast<double2*>(wfc), ldwfc);
        APW_LAUNCH_CHECK();
        APW_HIP(hipStreamSynchronize(ctx->stream));
        return APW_OK;
    });
}

}  // extern "C"
