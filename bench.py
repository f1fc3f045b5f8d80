"""Benchmark: Davidson time-to-1e-9 Ry and H|psi> GFLOP/s on one MI355X, FP32-mixed vs FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload si216] [--no-cpu-baseline]

A step is one full block-Davidson solve (DiagoDavid::diag semantics, eigenvalue-change
threshold 1e-9 Ry) of the 216-atom synthetic cell (228,487 plane waves, 160^3 FFT grid,
480 bands, 1728 projectors, subspace up to 1920) from the same seeded start vectors, in
the FP32-mixed form (FP32 H|psi> until the corrected bands' residuals fall below 1e-5,
then FP64).  The plain FP64 form is timed once beside it, and each form's eigenvalues are
compared with the compiled reference's own (tests/golden/ref_eigs.json).  Inputs are
resident in HBM before the timed region.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import re
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "abacus-develop_amd")]

METRIC = "Davidson time-to-1e-9Ry (s) + hpsi GFLOP/s, FP32-mixed vs FP64, 1 MI355X"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}


def flops_per_band(s):
    """Algorithmic FLOPs of one band through H|psi>: two 3D FFTs (5 N log2 N each), the
    real-space multiply (2N), kinetic (4 npw), becp + back-projection (16 npw nkb), D (4 nh nkb)."""
    n = s.nxyz
    nh = int(s.nh[0])
    return 10.0 * n * math.log2(n) + 2.0 * n + 4.0 * s.npw + 16.0 * s.npw * s.nkb + 4.0 * nh * s.nkb


def run_solve(pw, ctx, ham, s, psi0_dev, precision):
    import numpy as np
    import torch
    dav = pw.DiagoDavid(ctx, s.precondition, s.nband, s.npw, s.david_ndim)
    psi = psi0_dev.clone()
    eig = np.zeros(s.nband)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it = dav.diag(ham, None, s.npwx, psi, eig, 1e-9, 500, 5, 0, precision=precision)
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = dict(dav.stats)
    dav.close()
    return t1 - t0, it, eig, st


# library profiler name -> kernel-name prefixes in the rocprofv3 traffic summary
_PMC_FAMILY = {   # regular expressions over the kernel names of the committed traffic summary
    "eig_tridiag": r"apw::k_tridiag_(reg|rows|fused|lds)",
    # GEMMs: the path's updates are op(A) = N (tiled), its inner products op(A) = C (split-K)
    # (template arguments: MFMA type, storage type, op(A), ...)
    "zgemm_f64_tiled": r"apw::zgemm_kernel<double, double, 0, ",
    "zgemm_f64_splitk": r"apw::zgemm_kernel<double, double, [12], ",
    "cgemm_f32_tiled": r"apw::zgemm_kernel<float, float, 0, ", "cgemm_f32_splitk": r"apw::zgemm_kernel<float, float, [12], ",
    "zgemm_x32_tiled": r"apw::zgemm_kernel<float, double, 0, ",
    "zgemm_x32_splitk": r"apw::zgemm_kernel<float, double, [12], ",
    "fft_x_veff_fused": r"apw::k_xreg<\w+, \d+, 0>|apw::k_xpass<\w+, 0>",
    "fft_y_bwd": r"apw::k_yreg<\w+, \d+, 1>|apw::k_ybwd", "fft_y_fwd": r"apw::k_yreg<\w+, \d+, -1>|apw::k_yfwd",
    "fft_z_bwd_scatter": r"apw::k_zreg<\w+, \d+, 1>|apw::k_zbwd",
    "fft_z_fwd_gather": r"apw::k_zreg<\w+, \d+, -1>|apw::k_zfwd",
    "eig_bisect": r"apw::k_bisect", "eig_invit": r"apw::k_invit",
    "eig_backtransform": r"apw::k_(backtransform|wy_)|back-transform apw::(zgemm_kernel<double|splitk_reduce<double)",
    "chol_inv": r"apw::k_chol_inv",
}
# entries made of several dispatches per launch: the dispatches that occur once per launch
_PMC_CALLS = {"eig_backtransform": r"apw::k_(backtransform_w|wy_cols)"}
# kernels whose time is set by dependent latency chains, not by a byte or flop roof
_LATENCY_BOUND = {
    "eig_tridiag": "latency-bound: one cross-workgroup exchange per column of the reduction",
    "eig_bisect": "latency-bound: sequential Sturm recurrences", "eig_invit": "latency-bound: sequential solves",
    "chol_inv": "latency-bound",
}
_NOTES = dict(_LATENCY_BOUND, eig_backtransform="n > 512: WY blocks of 64 reflectors, two FP64 matrix-core "
              "products each (8 n^2 flops per vector); n <= 512: one wave per vector")


WORKLOAD = "si216"


def _pmc_traffic(name):
    """HBM bytes per dispatch (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of the kernels behind a
    profiler entry, dispatch-weighted, from the committed rocprofv3 summary of the benched
    workload; None if absent."""
    path = os.path.join(ROOT, "profiles", f"{WORKLOAD}_davidson_hbm_traffic.json")
    if name not in _PMC_FAMILY or not os.path.exists(path):
        return None
    with open(path) as f:
        ks = json.load(f)["kernels"]
    tot = cnt = 0.0
    for k, v in ks.items():
        if re.match(_PMC_FAMILY[name], k):
            tot += (v["fetch_bytes"] + v["write_bytes"]) * v["dispatches"]
            cnt += v["dispatches"]
    if name in _PMC_CALLS:
        cnt = sum(v["dispatches"] for k, v in ks.items() if re.match(_PMC_CALLS[name], k))
    return round(tot / cnt) if cnt else None


def _roofline(name, e):
    avg_s = e["ms"] * 1e-3 / max(1, e["launches"])
    if name.startswith(("zgemm", "cgemm")) or name == "eig_backtransform":
        bound, unit = "mfma", "TFLOP/s"
        achieved = e["flops"] / max(1, e["launches"]) / avg_s / 1e12
        # zgemm_x32: c128 data on the FP32 matrix cores, priced against the FP32 peak
        peak = PEAK_TFLOPS["f64" if name.startswith(("zgemm_f64", "eig_")) else "f32"]
    else:
        bound, unit = "hbm", "GB/s"
        achieved = e["bytes"] / max(1, e["launches"]) / avg_s / 1e9
        peak = HBM_PEAK_GBS
    out = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
           "frac": round(achieved / peak, 4), "traffic": _pmc_traffic(name), "kernel": name,
           "avg_launch_us": round(avg_s * 1e6, 2), "launches": e["launches"]}
    if name in _NOTES:
        out["note"] = _NOTES[name]
    return out


def roofline_from_profile(prof):
    """Dominant kernel (largest total device time) and its achieved rate vs the MI355X peak;
    achieved = algorithmic bytes (or flops) per launch / mean launch time (HIP events)."""
    if not prof:
        return None, {}
    name, e = max(prof.items(), key=lambda kv: kv[1]["ms"])
    return _roofline(name, e), e


_CPU_SAMPLE = r"""
import json, os, sys, time
sys.path[:0] = [sys.argv[1], os.path.join(sys.argv[1], "abacus-develop_amd")]
from apw import configs
from oracle import ref_david
s = configs.make(sys.argv[2])
nb = int(sys.argv[3])
ref_david.hpsi(s, s.psi0[:1], 1, s.npwx)   # library load, FFT plans
t0 = time.perf_counter()
ref_david.hpsi(s, s.psi0[:nb], nb, s.npwx)
print(json.dumps({"dt": time.perf_counter() - t0, "bands": nb}))
"""


def _ref_fixture(workload):
    path = os.path.join(ROOT, "tests", "golden", "ref_eigs.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(f"{workload}:0")


def cpu_baseline(workload, s, fp64_hpsi_bands, fp64_iters, nb_sample=64):
    """The reference CPU path (oracle/_ref: ABACUS's Operator::hPsi with its CPU ekinetic,
    veff and nonlocal functors, MKL BLAS and FFTW3 API) on a bounded sample of the same
    workload: H|psi> of nb_sample of its start vectors, timed on this host, scaled by H|psi>
    band count to the FP64 solve (whose trajectory is the reference's, so its band count is
    the reference's).  The subspace work (Gram-Schmidt, Rayleigh-Ritz, zheevx) is not in the
    sample, so the figure is a lower bound; the full reference solve measured in the build
    container is attached.  Runs in a child process.  Falls back to the numpy oracle's
    H|psi> when the reference build is absent."""
    import subprocess
    threads = min(16, os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), MKL_NUM_THREADS=str(threads))
    fix = _ref_fixture(workload)
    full = None
    if fix and "plain" in fix:
        full = {"seconds": fix["plain"]["seconds"], "iterations": fix["plain"]["iterations"],
                "threads": fix.get("threads", 8), "where": "8-core build container (not this host)"}
    try:
        from oracle import ref_david
        if ref_david.available(False):
            r = subprocess.run([sys.executable, "-c", _CPU_SAMPLE, ROOT, workload, str(nb_sample)], env=env,
                               capture_output=True, text=True, timeout=900)
            if r.returncode == 0:
                res = json.loads(r.stdout.strip().splitlines()[-1])
                est = res["dt"] * fp64_hpsi_bands / max(1, res["bands"])
                out = {"value": round(est, 3), "unit": "s", "cores": threads, "kind": "reference",
                       "estimate": "extrapolated (H|psi> only: a lower bound)",
                       "sample": f"reference Operator::hPsi (ABACUS CPU ekinetic/veff/nonlocal functors, MKL "
                                 f"BLAS and FFTW3 API, {threads} threads) on {res['bands']} start vectors of the "
                                 f"same system, {res['dt']:.1f} s measured, scaled by H|psi> band count to the FP64 "
                                 f"solve's {fp64_hpsi_bands} bands ({fp64_iters} iterations); Gram-Schmidt, "
                                 f"Rayleigh-Ritz and zheevx not counted",
                       "sample_s": round(res["dt"], 2)}
                if full:
                    out["full_solve_measured"] = full
                return out
            print(f"cpu_baseline: reference build exited {r.returncode}: {r.stderr.strip().splitlines()[-1:]}",
                  file=sys.stderr)
    except Exception as exc:
        print(f"cpu_baseline: reference build unusable ({exc}); using the oracle", file=sys.stderr)
    from oracle import pw_oracle
    ham = pw_oracle.HamiltOracle(s)
    nb = 4
    t0 = time.perf_counter()
    ham.hpsi(s.psi0[:nb, : s.npw].reshape(-1), nb, s.npw)
    dt = time.perf_counter() - t0
    est = dt / nb * fp64_hpsi_bands
    return {"value": round(est, 3), "unit": "s", "cores": 1, "kind": "port", "estimate": "extrapolated",
            "sample": f"numpy oracle H|psi> of {nb} bands ({dt:.2f} s) scaled to the FP64 GPU solve's "
                      f"{fp64_hpsi_bands} bands (H|psi> only; subspace work not counted)"}


def measure_next(pw, ctx, ham, s, psi0_dev, eig_david, reps=5):
    """The components on either side of the Davidson path, on the same workload: projector
    generation (getvnl, HBM-write bound: nkb x npw c128), the charge density of the nband
    start vectors (psiToRho, five FFT passes per band), the Diago_DavSubspace solve from the
    same start vectors, and one write + read of the bands as a .dat wavefunction file."""
    import tempfile
    import numpy as np
    import torch
    from apw import synth
    out = {}
    # ---- getvnl
    t = synth.projector_tables(s)
    gk = torch.from_numpy(np.ascontiguousarray(t["gk"])).cuda()
    tab = torch.from_numpy(t["tab"]).cuda()
    vkb = torch.empty((s.nkb, s.npwx), dtype=torch.complex128, device="cuda")
    args = (ctx, t["ntype"], t["npw"], t["npwx"], t["nhm"], t["tab_2"], t["tab_3"], t["atom_na"], t["atom_nb"],
            t["atom_nh"], t["dq"], t["tpiba"], gk, t["tau"], t["indv"], t["nhtol"], t["nhtolm"], tab, vkb)
    pw.getvnl(*args)
    ctx.synchronize()
    ctx.set_profiling(True)
    for _ in range(reps):
        pw.getvnl(*args)
    ctx.synchronize()
    e = ctx.profile()["getvnl"]
    ctx.set_profiling(False)
    avg = e["ms"] * 1e-3 / e["launches"]
    gbs = e["bytes"] / e["launches"] / avg / 1e9
    sign = np.tile(t["sign"], s.nat)[:, None]
    err = float(np.abs(vkb[:8].cpu().numpy() * sign[:8] - s.vkb[:8]).max() / np.abs(s.vkb).max())
    out["getvnl"] = {"avg_launch_us": round(avg * 1e6, 2), "bytes_per_launch": e["bytes"] / e["launches"],
                     "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs / HBM_PEAK_GBS, 4)},
                     "rel_err_vs_analytic_projectors": err}
    # ---- psiToRho of the nband start vectors (weights 2 / nband each)
    wg = np.full(s.nband, 2.0 / s.nband)
    rho = torch.zeros(s.nxyz, dtype=torch.float64, device="cuda")
    ham.psi_to_rho(psi0_dev, s.nband, wg, s.lat.omega, rho=rho, ldpsi=s.npwx)
    ctx.synchronize()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        ham.psi_to_rho(psi0_dev, s.nband, wg, s.lat.omega, rho=rho, ldpsi=s.npwx)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / reps
    prof = ctx.profile()
    ctx.set_profiling(False)
    out["psi_to_rho"] = {"bands": s.nband, "time_ms": round(dt * 1e3, 3),
                         "bands_per_s": round(s.nband / dt, 1),
                         "kernels": {k: {"avg_us": round(1e3 * v["ms"] / max(1, v["launches"]), 2),
                                         "GB_s": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1)}
                                     for k, v in prof.items()}}
    # ---- Diago_DavSubspace from the same start vectors
    # HSolverPW::update_precondition for dav_subspace: |k+G|^2 tpiba2 (hsolver_pw.cpp:570-577)
    pre = np.ones(s.npwx)
    pre[: s.npw] = np.asarray(s.gk2[: s.npw]) * s.tpiba2
    ds = pw.DiagoDavSubspace(ctx, pre, s.nband, s.npw, s.david_ndim, 1e-9, 500)
    psi = psi0_dev.clone()
    eig = np.zeros(s.nband)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it = ds.diag(ham, psi, s.npwx, eig)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    st = dict(ds.stats)
    ds.close()
    # residual norms of the returned vectors from an independent H|psi> (the solver reports none)
    hp = ham.hPsi(psi, s.nband)
    ctx.synchronize()
    res = torch.linalg.vector_norm(hp[:, : s.npw] - torch.from_numpy(eig).cuda()[:, None] * psi[:, : s.npw], dim=1)
    out["dav_subspace"] = {"time_s": round(dt, 4), "iterations": it, "hpsi_bands": st["hpsi_bands"],
                           "notconv": st["notconv"], "max_residual": float(res.max().item()),
                           "max_abs_eig_diff_vs_david_Ry": float(np.abs(eig - eig_david).max())}
    del hp, res
    # ---- wavefunction file round trip (at most 32 bands: the file goes through the host filesystem)
    kc = s.lat.gcar(s.basis.kvec_d[s.ik])
    nb_io = min(s.nband, 32)
    with tempfile.TemporaryDirectory() as td:
        fn = os.path.join(td, "WAVEFUNC1.dat")
        t0 = time.perf_counter()
        pw.write_wfc_pw(ctx, fn, s.ik, 1, kc, 1.0, nb_io, s.meta["ecutwfc"], s.lat.lat0, s.lat.tpiba, s.lat.G,
                        s.nx, s.ny, s.nz, s.npw, s.ig2ixyz, psi0_dev, s.npwx)
        tw = time.perf_counter() - t0
        back = torch.empty_like(psi0_dev[:nb_io])
        t0 = time.perf_counter()
        pw.read_wfc_pw(ctx, fn, s.ik, 1, nb_io, kc, s.lat.lat0, s.lat.tpiba, s.lat.G, s.nx, s.ny, s.nz, s.npw,
                       s.ig2ixyz, back, s.npwx)
        tr = time.perf_counter() - t0
        same = bool(torch.equal(back[:, : s.npw], psi0_dev[:nb_io, : s.npw]))
    mb = 16.0 * s.npw * nb_io / 1e6
    out["wfc_io"] = {"bands": nb_io, "MB": round(mb, 1), "write_MB_s": round(mb / tw, 1),
                     "read_MB_s": round(mb / tr, 1), "round_trip_identical": same,
                     "filesystem": "tempfile default (host)"}
    out["batched_kpoints"] = measure_batched(pw)
    out["batched_cells"] = measure_cells(pw)
    return out


def _batched_run(pw, systems, ncontexts, precision, single_wg, thr=1e-9):
    """apw_david_diag_batched over the systems spread round-robin on ncontexts contexts;
    single_wg: subspace reductions up to that size in one workgroup (no device-wide ordering
    of the eigensolves, so the units' solves overlap).  Returns (seconds, iterations, eigs)."""
    import numpy as np
    import torch
    ctxs = [pw.Context(0) for _ in range(ncontexts)]
    for c in ctxs:
        c.set_eig_single_wg(single_wg)
    units = []
    for i, s in enumerate(systems):
        c = ctxs[i % ncontexts]
        units.append((pw.DiagoDavid(c, s.precondition, s.nband, s.npw, s.david_ndim), pw.HamiltPW(c, s),
                      torch.from_numpy(np.ascontiguousarray(s.psi0)).cuda(), s.npwx, np.zeros(s.nband)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = pw.david_diag_batched(units, thr, 500, precision=precision)
    for c in ctxs:
        c.synchronize()
    dt = time.perf_counter() - t0
    eigs = [u[4].copy() for u in units]
    for u in units:
        u[0].close()
        u[1].close()
    for c in ctxs:
        c.close()
    return dt, int(sum(its)), eigs


def _err_vs(eigs, ref):
    """(largest per-unit max |difference|, units within 1e-9 Ry)"""
    import numpy as np
    d = [float(np.abs(a - b).max()) for a, b in zip(eigs, ref)]
    return max(d), int(sum(x < 1e-9 for x in d))


def measure_cells(pw, ncells=128, ncontexts=4):
    """The uneven-cell batch (all 128 cells: 2-32 atoms, 18^3-72^3 grids, nbands = 3 atoms):
    the plain FP64 per-cell loop on one context vs the batched solve on ncontexts contexts (one
    stream each; the GPU has 4 hardware queues), FP64 and FP32-mixed.  Each form's per-cell
    eigenvalue error is taken against a converged FP64 solve (thr 1e-12 Ry), since the
    reference's stopping rule (eigenvalue change below 1e-9 Ry) bounds the change between
    iterations, not the distance from the converged values."""
    from apw import configs
    systems = [configs.make_cell(i) for i in range(ncells)]
    res = {}
    _, _, eig_conv = _batched_run(pw, systems, ncontexts, "fp64", 0, thr=1e-12)
    dt, it, eig64 = _batched_run(pw, systems, 1, "fp64", 0)
    e, n = _err_vs(eig64, eig_conv)
    res["fp64_loop_1_context"] = {"time_s": round(dt, 4), "iterations": it, "max_abs_eig_err_vs_converged_Ry": e,
                                  "cells_within_1e-9_Ry_of_converged": n}
    for label, prec in (("fp64_batched", "fp64"), ("fp32_mixed_batched", "fp32_mixed")):
        dt, it, eigs = _batched_run(pw, systems, ncontexts, prec, 0)
        e, n = _err_vs(eigs, eig_conv)
        e64, n64 = _err_vs(eigs, eig64)
        res[f"{label}_{ncontexts}_contexts"] = {"time_s": round(dt, 4), "iterations": it,
                                               "max_abs_eig_err_vs_converged_Ry": e,
                                               "cells_within_1e-9_Ry_of_converged": n,
                                               "max_abs_eig_diff_vs_fp64_loop_Ry": e64,
                                               "cells_within_1e-9_Ry_of_fp64_loop": n64}
    res["cells"] = ncells
    res["npw_range"] = [int(min(s.npw for s in systems)), int(max(s.npw for s in systems))]
    res["grids"] = sorted({f"{s.nx}^3" for s in systems}, key=lambda g: int(g.split("^")[0]))
    return res


def measure_batched(pw, nk=64, ncontexts=4):
    """The batched form on all 64 k-points of the 16-atom 4x4x4 mesh (uneven plane-wave
    counts, shared 48x48x96 grid): the FP64 per-k loop on one context vs the batched solve on
    ncontexts contexts (one stream each; the GPU has 4 hardware queues), FP64 and FP32-mixed,
    and the FP64 batch with one-workgroup subspace reductions (no device-wide ordering of the
    eigensolves) beside the cooperative ones."""
    from apw import configs
    systems = [configs.make("kpts16", ik=k) for k in range(0, 64, 64 // nk)]
    nbx = max(s.nband * s.david_ndim for s in systems)
    res = {}
    dt, it, eig_loop = _batched_run(pw, systems, 1, "fp64", 0)
    res["fp64_loop_1_context"] = {"time_s": round(dt, 4), "iterations": it}
    for label, prec, single in (("fp64_batched", "fp64", 0), ("fp32_mixed_batched", "fp32_mixed", 0),
                                ("fp64_batched_one_workgroup_eigensolves", "fp64", nbx)):
        dt, it, eigs = _batched_run(pw, systems, ncontexts, prec, single)
        res[f"{label}_{ncontexts}_contexts"] = {"time_s": round(dt, 4), "iterations": it,
                                               "max_abs_eig_diff_vs_fp64_loop_Ry": _err_vs(eigs, eig_loop)[0]}
    res["kpoints"] = len(systems)
    res["npw_range"] = [int(min(s.npw for s in systems)), int(max(s.npw for s in systems))]
    return res


def run_gemm_shapes():
    """The tall update (basis x vcc: 228487 x 256 x 1024, NN) and the inner product
    (basis^H hpsi: 1024 x 256 x 228487, CN) of the 216-atom solve at their largest, with the
    tuned plans; prints one JSON line with each shape's mean time and algorithmic TFLOP/s."""
    import torch
    from apw import pw
    if os.path.exists(pw.TUNE_CACHE_FILE):
        pw.tune_load()
        pw.tune_set_mode(pw.TUNE_CACHE)
    ctx = pw.Context(0)
    npw = 228487
    out = {}
    for ta, m, n, k in (("N", npw, 256, 1024), ("C", 1024, 256, npw)):
        rows = m if ta == "N" else k
        gen = torch.Generator(device="cuda").manual_seed(7)
        A = torch.randn(k if ta == "N" else m, rows, dtype=torch.complex128, device="cuda", generator=gen)
        B = torch.randn(n, k, dtype=torch.complex128, device="cuda", generator=gen)
        C = torch.zeros(n, m, dtype=torch.complex128, device="cuda")
        pw.gemm_op(ctx, ta, "N", m, n, k, 1.0, A, rows, B, k, 0.0, C, m)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(4):
            pw.gemm_op(ctx, ta, "N", m, n, k, 1.0, A, rows, B, k, 0.0, C, m)
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / 4
        out[f"{ta}N {m}x{n}x{k}"] = {"ms": round(1e3 * dt, 3), "tflops_algorithmic": round(8.0 * m * n * k / dt / 1e12, 2)}
        del A, B, C
    ctx.close()
    print(json.dumps({"gemm_shapes": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="si216")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="fixed rules only (do not load the autotuner's cache abacus-develop_amd/tune_cache.json)")
    ap.add_argument("--no-next", action="store_true", help="skip the getvnl / psiToRho / DavSubspace / wfc-file lines")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (the reference for HBM-byte counters) and exit")
    ap.add_argument("--gemm-shapes", action="store_true",
                    help="only run the two largest FP64 GEMM shapes of the 216-atom solve with the tuned plans "
                         "(4 launches each; for matrix-core counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 27                                    # 2^27 doubles = 1 GiB each way
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"bytes_read_per_copy": 8 * n, "bytes_written_per_copy": 8 * n,
                                               "copies": 3}}), flush=True)
        return
    if args.gemm_shapes:
        run_gemm_shapes()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(k-points and cells would shard across GPUs with no exchange inside Davidson)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from apw import configs, pw

    global WORKLOAD
    WORKLOAD = args.workload
    # autotuned plans (GEMM tiles/mapping/splits, eigensolver workgroups, FFT schedule and
    # chunk) from the committed cache, measured beforehand by abacus-develop_amd/tune_workloads.py
    tune = None
    if not args.no_tune and os.path.exists(pw.TUNE_CACHE_FILE):
        tune = {"cache": os.path.relpath(pw.TUNE_CACHE_FILE, ROOT), "entries": pw.tune_load()}
        pw.tune_set_mode(pw.TUNE_CACHE)
    t_build = time.perf_counter()
    s = configs.make(args.workload)
    t_build = time.perf_counter() - t_build
    print(f"[bench] built {args.workload}: npw={s.npw} nkb={s.nkb} grid={s.nx}x{s.ny}x{s.nz} nband={s.nband} "
          f"({t_build:.1f} s)", file=sys.stderr, flush=True)
    ctx = pw.Context(0)
    ham = pw.HamiltPW(ctx, s)
    psi0 = torch.from_numpy(np.ascontiguousarray(s.psi0)).cuda()
    torch.cuda.synchronize()

    warm = []
    for w in range(args.warmup):
        dt, it, _, st = run_solve(pw, ctx, ham, s, psi0, "fp32_mixed")
        warm.append(dt)
        print(f"[bench] warmup {w}: {dt:.3f} s, {it} iterations", file=sys.stderr, flush=True)

    # timed region: K FP32-mixed solves, kernels bracketed by HIP events on the library stream
    ctx.set_profiling(True)
    torch.cuda.synchronize()
    times, mixed = [], None
    t0 = time.perf_counter()
    for k in range(args.steps):
        dt, it, eig_m, st = run_solve(pw, ctx, ham, s, psi0, "fp32_mixed")
        times.append(dt)
        mixed = (it, eig_m, st)
        print(f"[bench] step {k}: {dt:.3f} s, {it} iterations, switch at {st['fp64_switch_iter']}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = ctx.profile()
    ctx.set_profiling(False)

    # FP64 form beside it (not part of the timed steps), and the mixed form with the
    # pre-phase's subspace products on the FP32 matrix cores too
    dt64, it64, eig64, st64 = run_solve(pw, ctx, ham, s, psi0, "fp64")
    print(f"[bench] fp64: {dt64:.3f} s, {it64} iterations", file=sys.stderr, flush=True)
    dtx, itx, eigx, stx = run_solve(pw, ctx, ham, s, psi0, "fp32_mixed_x32")
    print(f"[bench] fp32_mixed_x32: {dtx:.3f} s, {itx} iterations", file=sys.stderr, flush=True)

    if tune is not None:
        st = pw.tune_stats()
        tune.update(hits=st["hits"], misses=st["misses"])
        # the same FP32-mixed solve with the fixed rules (a Hamiltonian built without the table)
        pw.tune_set_mode(pw.TUNE_FIXED)
        ham0 = pw.HamiltPW(ctx, s)
        dt0, it0, _, _ = run_solve(pw, ctx, ham0, s, psi0, "fp32_mixed")
        ham0.close()
        pw.tune_set_mode(pw.TUNE_CACHE)
        tune["fixed_rules_fp32_mixed_s"] = round(dt0, 4)
        print(f"[bench] fixed rules: {dt0:.3f} s, {it0} iterations", file=sys.stderr, flush=True)

    fb = flops_per_band(s)
    it_m, eig_m, st_m = mixed
    roof, _ = roofline_from_profile(prof)
    top = sorted(prof.items(), key=lambda kv: -kv[1]["ms"])
    # each form's eigenvalues against the compiled reference DiagoDavid's own (same start vectors)
    fix = _ref_fixture(args.workload)
    err = {}
    if fix:
        for label in ("plain", "detsign"):
            if label in fix:
                r = np.array(fix[label]["eig"])
                err[label] = (float(np.abs(eig64 - r).max()), float(np.abs(eig_m - r).max()),
                              float(np.abs(eigx - r).max()))
    out = {
        "metric": METRIC,
        "value": round(sum(times) / len(times), 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / args.steps, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 H|psi> + c128 subspace (FP32-mixed); c128 (FP64 form)",
        "data": "synthetic: seeded Gaussian-well + noise local potential, 8 KB projectors/atom, random start vectors",
        "config": {"workload": f"{args.workload}: {s.nat}-atom cubic cell a={s.lat.lat0} bohr, "
                               f"ecutwfc {s.meta['ecutwfc']} Ry, {s.npw} plane waves, FFT {s.nx}x{s.ny}x{s.nz}, "
                               f"{s.nband} bands, {s.nkb} projectors, david_ndim {s.david_ndim}, thr 1e-9 Ry",
                   "npw": s.npw, "nband": s.nband, "nkb": s.nkb, "fft": [s.nx, s.ny, s.nz]},
        "fp32_mixed": {"time_s": [round(t, 4) for t in times], "iterations": it_m,
                       "fp64_switch_iter": st_m["fp64_switch_iter"], "hpsi_bands": st_m["hpsi_bands"],
                       "hpsi_gflops": round(st_m["hpsi_bands"] * fb / max(st_m["t_hpsi_s"], 1e-12) / 1e9, 1),
                       "max_residual": st_m["max_residual"]},
        "fp32_mixed_x32": {"time_s": round(dtx, 4), "iterations": itx, "fp64_switch_iter": stx["fp64_switch_iter"],
                           "speedup_vs_fp64": round(dt64 / dtx, 3),
                           "max_abs_eig_diff_vs_fp64_Ry": float(np.abs(eigx - eig64).max()),
                           "max_residual": stx["max_residual"]},
        "fp64": {"time_s": round(dt64, 4), "iterations": it64, "hpsi_bands": st64["hpsi_bands"],
                 "hpsi_gflops": round(st64["hpsi_bands"] * fb / max(st64["t_hpsi_s"], 1e-12) / 1e9, 1),
                 "max_residual": st64["max_residual"]},
        "hpsi_flop_count": "conventional: 10 N log2 N (two full-box 3D FFTs) + 2N + 4 npw + 16 npw nkb + 4 nh nkb "
                           "per band; the stick-pruned FFT passes perform fewer",
        "speedup_mixed_vs_fp64": round(dt64 / (sum(times) / len(times)), 3),
        "max_abs_eig_diff_mixed_vs_fp64_Ry": float(np.abs(eig_m - eig64).max()),
        "eig_err_vs_reference_Ry": ({label: {"fp64": e[0], "fp32_mixed": e[1], "fp32_mixed_x32": e[2]}
                                     for label, e in err.items()}
                                    if err else None),
        "warmup_s": [round(t, 4) for t in warm],
        "autotune": tune,
        "roofline": roof,
        "rooflines": [_roofline(k, v) for k, v in top[:8]],
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3),
                        "avg_us": round(1e3 * v["ms"] / max(1, v["launches"]), 2),
                        "GB_s": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1),
                        "TFLOP_s": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 2)} for k, v in top},
    }
    if not args.no_next:
        out["next_components"] = measure_next(pw, ctx, ham, s, psi0, eig64)
        print(f"[bench] next components: {json.dumps(out['next_components'])}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, s, st64["hpsi_bands"], it64)
    print(json.dumps(out), flush=True)
    ham.close()
    ctx.close()


if __name__ == "__main__":
    main()
