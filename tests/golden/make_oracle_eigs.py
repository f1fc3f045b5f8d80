"""Write the CPU oracle's converged Davidson eigenvalues for the multi-k and uneven-cell
workloads as a JSON fixture (tests/golden/oracle_eigs.json).

The oracle (oracle/pw_oracle.py) is pinned separately against the reference's golden
vectors, dense diagonalisation and the compiled reference DiagoDavid; this fixture only
saves the GPU parity tests from re-running it (15 s per k-point) on every call.  The
cases: the 16-atom tetragonal cell at k-points 0 and 37 of its 4x4x4 mesh (uneven sphere
sizes), and cells 1, 3, 6 of the 128-cell uneven batch.  Threshold 1e-9 Ry, the
workloads' own start vectors, ntry_max 5, notconv_max 0.

    python tests/golden/make_oracle_eigs.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "abacus-develop_amd")]

from apw import configs  # noqa: E402
from oracle import pw_oracle as O  # noqa: E402

CASES = [("kpts16", 0), ("kpts16", 37), ("cell1", 0), ("cell3", 0), ("cell6", 0)]


def case_key(name, ik):
    return f"{name}:{ik}"


def solve(name, ik):
    s = configs.make(name, ik=ik) if name == "kpts16" else configs.make(name)
    eig, _, iters, _ = O.david_solve(s, thr=1e-9)
    return {"npw": s.npw, "npwx": s.npwx, "grid": [s.nx, s.ny, s.nz], "nband": s.nband, "nkb": s.nkb,
            "iterations": int(iters), "eig": [float(x) for x in eig]}


def main():
    out = {case_key(n, k): solve(n, k) for n, k in CASES}
    with open(os.path.join(HERE, "oracle_eigs.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
