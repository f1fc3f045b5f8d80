// Stick-pruned mixed-radix (2, 3, 4, 5) 3D FFT of the plane-wave sphere, staged in LDS.
#pragma once

#include "apw_internal.hpp"

namespace apw {

struct FftPlan {
    int n = 0;
    int nst = 0;
    int rad[16] = {0};
};

FftPlan make_plan(int n);

// Sphere <-> box bookkeeping derived from the caller's ig2ixyz map.
//
// Data layout of one band through the transform (complex, precision R):
//   sphere  psi[ig]                                   (caller layout, ld per band)
//   S       [band][stick s][iz]       ns * nz         (z lines of the sticks that hold G)
//   X       [band][x slot][iy][iz]    nxr * ny * nz   (y/z planes of the x columns that
//                                                      intersect the sphere)
//   box     [ix][iy][iz]              nx * ny * nz    (full real-space grid; only the
//                                                      fused x pass touches it, never stored
//                                                      on the H|psi> path)
struct PwGrid {
    int nx = 0, ny = 0, nz = 0, npw = 0;
    int ns = 0;    // sticks (distinct (ix, iy) of the sphere)
    int nxr = 0;   // x slots (distinct ix of the sphere)
    FftPlan px, py, pz;
    DevBuf<int> stick_ig;     // ns * nz, -1 where no G
    DevBuf<int4> zrun;        // ns: each stick's G as at most two runs of z consecutive in
                              // the sphere (iz start | length << 16, ig of the start) x 2
    bool zrun_ok = false;     // every stick is at most two such runs
    DevBuf<int> stick_iy;     // ns
    DevBuf<int> slot_ix;      // nxr
    DevBuf<int> slot_start;   // nxr (first stick of the slot; sticks sorted by ix, iy)
    DevBuf<int> slot_cnt;     // nxr
    int4 xrun{};              // slots as runs of x: slot = x - xrun.x on [xrun.x, +xrun.y),
                              // xrun.y + x - xrun.z on [xrun.z, +xrun.w)
    bool xrun_ok = false;     // the slots are at most two runs
    DevBuf<int4> yrun;        // nxr: each slot's sticks as runs of y (same convention)
    bool yrun_ok = false;     // every slot's sticks are at most two runs
    // pass schedule per precision ([0] FP64, [1] FP32), chosen by the autotuner: bit 0 runs
    // the x pass, bit 1 the y passes, bit 2 the z passes on the LDS Stockham kernels with the
    // run-time radix plan instead of the register four-step kernels (0: register throughout)
    int sched[2] = {0, 0};
    DevBuf<double> tw64;      // twiddles exp(-2 pi i k / n) for x, y, z (c128), offsets below
    DevBuf<float> tw32;
    int twoff_x = 0, twoff_y = 0, twoff_z = 0;

    void build(int nx_, int ny_, int nz_, int npw_, const int* ig2ixyz_host);
    size_t bytes_S(size_t elem) const { return (size_t)ns * nz * elem; }
    size_t bytes_X(size_t elem) const { return (size_t)nxr * ny * nz * elem; }
};

// whether the register four-step kernels can run the x, y and z passes of this grid (the
// alternative schedules the autotuner may choose)
bool xreg_available(const PwGrid& g);
bool yreg_available(const PwGrid& g);
bool zreg_available(const PwGrid& g, int ld);

enum GatherMode { GATHER_SET = 0, GATHER_ADD = 1, GATHER_SET_KIN = 2 };

// Local part of H|psi> for nb bands: G->r, * veff, r->G, gathered into hpsi with the mode
// above (GATHER_SET_KIN also adds gk2*tpiba2*psi).  S, X: workspace for nb bands.
template <class R>
void local_hpsi(hipStream_t s, const PwGrid& g, const double* psi, int ld, double* hpsi, int ldh, int nb,
                const R* veff, const double* gk2, double tpiba2, GatherMode mode, double scale, R* S, R* X,
                Prof* prof = nullptr);

// Sphere -> full box (backward, unnormalised): box = FFT^+(scatter(psi)), nb bands of
// nx*ny*nz each.  add: box += factor * result.
void sphere_to_box(hipStream_t s, const PwGrid& g, const double* psi, int ld, double* box, int nb, bool add,
                   double factor, double* S, double* X);
// Full box -> sphere (forward): hpsi = scale * gather(FFT^-(box)) (mode SET or ADD).
void box_to_sphere(hipStream_t s, const PwGrid& g, const double* box, double* hpsi, int ldh, int nb,
                   GatherMode mode, double scale, double* S, double* X);

// rho(r) (+)= sum_b w[b] |FFT^+(psi_b)(r)|^2 over nb bands (w: device doubles), bands in
// order, no atomics.  S, X hold nb bands of intermediates.
void psi_to_rho(hipStream_t s, const PwGrid& g, const double* psi, int ld, int nb, const double* w, double* rho,
                bool add, double* S, double* X, Prof* prof = nullptr);

}  // namespace apw
