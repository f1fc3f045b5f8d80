"""Autotune the library on the named workloads and write the tuning cache.

    python abacus-develop_amd/tune_workloads.py [--out abacus-develop_amd/tune_cache.json]
                                                [--workloads si216,si64,kpts16,cells]

Runs with the autotuner measuring (mode 2): every GEMM shape class, subspace-eigensolver size
and FFT grid met while solving the workloads (FP32-mixed, FP32-mixed x32 on the single cells,
and FP64 Davidson, as benchmarked) is
timed over its candidate plans on the GPU and the fastest is recorded; the table is then
saved.  Needs a GPU.  Workloads: si216 (216-atom cell), si64 (64-atom cell), kpts16 (every
fourth k-point of the 16-atom 4x4x4 mesh), cells (the first 32 cells of the uneven batch).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [HERE]


def _solve(pw, ctx, s, ham, precision):
    import numpy as np
    import torch
    dav = pw.DiagoDavid(ctx, s.precondition, s.nband, s.npw, s.david_ndim)
    psi = torch.from_numpy(np.ascontiguousarray(s.psi0)).cuda()
    eig = np.zeros(s.nband)
    it = dav.diag(ham, None, s.npwx, psi, eig, 1e-9, 500, 5, 0, precision=precision)
    dav.close()
    ctx.synchronize()
    return it


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tune_cache.json"))
    ap.add_argument("--workloads", default="si216,si64,kpts16,cells")
    args = ap.parse_args()
    from apw import configs, pw
    pw.tune_set_mode(pw.TUNE_MEASURE)
    if os.path.exists(args.out):
        print(f"[tune] starting from {args.out}: {pw.tune_load(args.out)} entries", flush=True)
    ctx = pw.Context(0)
    for name in args.workloads.split(","):
        if name == "kpts16":
            systems = [configs.make("kpts16", ik=k) for k in range(0, 64, 4)]
        elif name == "cells":
            systems = [configs.make_cell(i) for i in range(32)]
        else:
            systems = [configs.make(name)]
        for i, s in enumerate(systems):
            t0 = time.perf_counter()
            ham = pw.HamiltPW(ctx, s)
            precs = ("fp32_mixed", "fp32_mixed_x32", "fp64") if name in ("si216", "si64") else ("fp32_mixed", "fp64")
            its = [_solve(pw, ctx, s, ham, prec) for prec in precs]
            ham.close()
            st = pw.tune_stats()
            print(f"[tune] {name}[{i}] npw={s.npw} grid={s.nx}x{s.ny}x{s.nz}: iterations {its}, "
                  f"{time.perf_counter() - t0:.1f} s, table {st['entries']} entries", flush=True)
        pw.tune_save(args.out)
    ctx.close()
    print(f"[tune] wrote {pw.tune_save(args.out)} entries to {args.out}", flush=True)


if __name__ == "__main__":
    main()
