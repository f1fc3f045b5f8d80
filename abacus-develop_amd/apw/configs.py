"""Named workloads of the plane-wave Davidson path.

Sizes (atoms, lattice, cutoff, FFT grid, bands, Davidson block) of each workload;
the arrays themselves come from :func:`apw.synth.build_system`.

======  ==========================================================================
name    workload
======  ==========================================================================
si8     8-atom diamond-like cubic cell, a = 10.26 bohr, ecutwfc 30 Ry, FFT 36^3,
        k = 0 (full complex sphere), nbands 24, Davidson block 2*nbands
si64    64-atom cubic cell, a = 20.52, ecutwfc 50 Ry, FFT 96^3, nbands 160,
        512 projectors, Davidson block 4*nbands
kpts16  16-atom tetragonal cell a=b=10.26, c=20.52, ecutwfc 40 Ry, 4x4x4
        Monkhorst-Pack k mesh (64 k-points, no symmetry), FFT 48x48x96, nbands 48
si216   216-atom cubic cell, a = 30.78, ecutwfc 60 Ry, FFT 160^3, nbands 480,
        1728 projectors
cells   128 independent random cells, 2-32 atoms, a in [5.13, 15.4], ecutwfc in
        [30, 50], grid from 4*ecutwfc, nbands = 3*atoms
tiny    2-atom cubic cell, a = 6.0, ecutwfc 12 Ry (grid from 4*ecutwfc), nbands 8:
        small enough for dense-matrix checks in unit tests
======  ==========================================================================
"""
from __future__ import annotations

import numpy as np

from .basis import Lattice, PWBasisK
from .synth import build_system, diamond_positions

CONFIG_SEED = {"si8": 1001, "si64": 1002, "kpts16": 1003, "si216": 1004, "tiny": 1099}


def mp_mesh(n1, n2, n3) -> np.ndarray:
    """Monkhorst-Pack mesh in direct coordinates, no symmetry reduction."""
    def axis(n):
        return (2.0 * np.arange(1, n + 1) - n - 1) / (2.0 * n)
    a, b, c = np.meshgrid(axis(n1), axis(n2), axis(n3), indexing="ij")
    return np.stack([a, b, c], axis=-1).reshape(-1, 3)


def make(name: str, ik: int = 0, nband: int | None = None):
    """Build the synthetic Hamiltonian of workload ``name`` (k-point ``ik``)."""
    if name == "si8":
        lat = Lattice(10.26, np.eye(3))
        return build_system(name, lat, diamond_positions(), 30.0, nband or 24, 2,
                            CONFIG_SEED[name], grid=(36, 36, 36))
    if name == "si64":
        lat = Lattice(20.52, np.eye(3))
        return build_system(name, lat, diamond_positions((2, 2, 2)), 50.0, nband or 160, 4,
                            CONFIG_SEED[name], grid=(96, 96, 96))
    if name == "si216":
        lat = Lattice(30.78, np.eye(3))
        return build_system(name, lat, diamond_positions((3, 3, 3)), 60.0, nband or 480, 4,
                            CONFIG_SEED[name], grid=(160, 160, 160))
    if name == "kpts16":
        lat = Lattice(10.26, np.diag([1.0, 1.0, 2.0]))
        kd = mp_mesh(4, 4, 4)
        basis = PWBasisK(lat, 40.0, 48, 48, 96, kvec_d=kd)
        pos = diamond_positions((1, 1, 2))
        return build_system(name, lat, pos, 40.0, nband or 48, 4, CONFIG_SEED[name],
                            basis=basis, ik=ik)
    if name == "tiny":
        lat = Lattice(6.0, np.eye(3))
        pos = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])
        return build_system(name, lat, pos, 12.0, nband or 8, 4, CONFIG_SEED[name])
    if name.startswith("cell"):
        return make_cell(int(name[4:]), nband)
    raise KeyError(name)


def make_cell(i: int, nband: int | None = None):
    """Random independent cell ``i`` of the uneven 128-cell batch (seed 5000 + i)."""
    rng = np.random.default_rng(5000 + i)
    nat = int(rng.integers(2, 33))
    a = float(rng.uniform(5.13, 15.4))
    ecut = float(rng.uniform(30.0, 50.0))
    # random positions with a minimum separation of 1.6 bohr (periodic images included); a
    # cell too dense for random placement (cell 15: 31 atoms at a = 5.46) is enlarged by 5 %
    # and placed again, so every cell terminates
    while True:
        pos, tries = [], 0
        while len(pos) < nat and tries < 200000:
            tries += 1
            p = rng.uniform(0.0, 1.0, 3)
            ok = True
            for q in pos:
                d = p - q
                d -= np.round(d)
                if np.linalg.norm(d * a) < 1.6:
                    ok = False
                    break
            if ok:
                pos.append(p)
        if len(pos) == nat:
            break
        a *= 1.05
    lat = Lattice(a, np.eye(3))
    return build_system(f"cell{i}", lat, np.array(pos), ecut, nband or 3 * nat, 4, 5000 + i)
