"""Plane-wave basis for one or more k-points: G-sphere, FFT grid and index maps.

Mirrors the data that ``ModulePW::PW_Basis_K`` hands to the H|psi> path:

* FFT grid choice (``initgrids``: smallest 2-3-5 factorisable box that holds the
  ``gridecut`` sphere) -- source/module_basis/module_pw/pw_init.cpp:24-149, or an
  explicit ``nx, ny, nz`` -- pw_init.cpp:156-210.
* Stick (x, y) ordering of the plane waves: sticks ordered by ``ixy = x*ny + y``
  (x, y shifted into [0, n)), plane waves of one stick ordered by increasing signed
  ``iz`` -- source/module_basis/module_pw/pw_distributeg.cpp:59-135 (count_pw_st)
  and :149-195 (get_ig2isz_is2fftixy).
* Per-k selection ``|k+G|^2 <= gk_ecut`` keeping the big-sphere order
  -- source/module_basis/module_pw/pw_basis_k.cpp:118-168 (setupIndGk); the big
  sphere radius is ``sqrt(gk_ecut) + |k|max`` -- pw_basis_k.cpp:74-80.
* ``gk2`` = |k+G|^2 in units of tpiba^2 -- pw_basis_k.cpp:189-265.
* ``ig2ixyz_k = iz + iy*nz + ix*ny*nz`` -- pw_basis_k.cpp:319-344.

Everything here is set-up work done once on the host; the arrays it produces are
uploaded to the device by the C ABI (``apw_ham_create``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

TWO_PI = 2.0 * math.pi


def _factorisable_235(n: int) -> bool:
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def next_235(n: int) -> int:
    """Smallest m >= n with only prime factors 2, 3, 5 (pw_init.cpp:93-140)."""
    m = max(1, int(n))
    while not _factorisable_235(m):
        m += 1
    return m


@dataclass
class Lattice:
    lat0: float                 # bohr
    latvec: np.ndarray          # 3x3, rows are lattice vectors in units of lat0

    def __post_init__(self):
        self.latvec = np.asarray(self.latvec, dtype=np.float64).reshape(3, 3)
        self.tpiba = TWO_PI / self.lat0
        self.tpiba2 = self.tpiba * self.tpiba
        self.omega = abs(np.linalg.det(self.latvec)) * self.lat0 ** 3
        self.GT = np.linalg.inv(self.latvec)      # pw_init.cpp:35
        self.G = self.GT.T.copy()                 # rows: reciprocal vectors, units tpiba
        self.GGT = self.G @ self.GT               # metric: |f|^2 = f GGT f

    def gcar(self, f: np.ndarray) -> np.ndarray:
        """Cartesian vectors (units tpiba) of direct reciprocal coords ``f`` (..., 3)."""
        return f @ self.G


def grid_from_ecut(lat: Lattice, gridecut: float) -> tuple[int, int, int]:
    """FFT box for a cutoff ``gridecut`` (Ry) -- pw_init.cpp:40-140."""
    gc = gridecut / lat.tpiba2
    ibox = [int(math.sqrt(gc) * math.sqrt(float(lat.latvec[i] @ lat.latvec[i]))) + 1 for i in range(3)]
    rng = [np.arange(-b, b + 1) for b in ibox]
    fx, fy, fz = np.meshgrid(rng[0], rng[1], rng[2], indexing="ij")
    f = np.stack([fx, fy, fz], axis=-1).reshape(-1, 3).astype(np.float64)
    mod = np.einsum("ni,ij,nj->n", f, lat.GGT, f)
    inside = f[mod <= gc]
    n = np.abs(inside).max(axis=0).astype(int)
    return tuple(next_235(2 * int(v) + 1) for v in n)


def gridecut_of_box(lat: Lattice, nx: int, ny: int, nz: int) -> float:
    """Largest |G|^2 (tpiba^2 units) fully inside an explicit box -- pw_init.cpp:156-210."""
    ib = [(nx - 1) // 2 + 1, (ny - 1) // 2 + 1, (nz - 1) // 2 + 1]
    rng = [np.arange(-b, b + 1) for b in ib]
    fx, fy, fz = np.meshgrid(rng[0], rng[1], rng[2], indexing="ij")
    f = np.stack([fx, fy, fz], axis=-1).reshape(-1, 3)
    inner = (np.abs(f[:, 0]) <= ib[0] - 1) & (np.abs(f[:, 1]) <= ib[1] - 1) & (np.abs(f[:, 2]) <= ib[2] - 1)
    f = f[~inner].astype(np.float64)
    mod = np.einsum("ni,ij,nj->n", f, lat.GGT, f)
    return float(mod.min()) - 1e-6


@dataclass
class PWBasisK:
    """G-sphere and FFT maps for a set of k-points (direct coordinates ``kvec_d``)."""
    lat: Lattice
    ecutwfc: float                       # Ry
    nx: int = 0
    ny: int = 0
    nz: int = 0
    kvec_d: np.ndarray = field(default_factory=lambda: np.zeros((1, 3)))
    gridecut: float | None = None        # Ry; None: 4 * ecutwfc (esolver_fp.cpp:70)

    def __post_init__(self):
        lat = self.lat
        self.kvec_d = np.asarray(self.kvec_d, dtype=np.float64).reshape(-1, 3)
        self.nks = self.kvec_d.shape[0]
        if self.nx <= 0:
            # the wavefunction grid is the density grid: gridecut = 4 * ecutwfc (esolver_fp.cpp:70)
            gridecut = 4.0 * self.ecutwfc if self.gridecut is None else float(self.gridecut)
            self.nx, self.ny, self.nz = grid_from_ecut(lat, gridecut)
            self.gridecut_lat = gridecut / lat.tpiba2
        else:
            self.gridecut_lat = gridecut_of_box(lat, self.nx, self.ny, self.nz)
        nx, ny, nz = self.nx, self.ny, self.nz
        self.nxyz = nx * ny * nz
        self.tpiba2 = lat.tpiba2

        kcar = lat.gcar(self.kvec_d)
        kmaxmod = float(np.sqrt((kcar ** 2).sum(axis=1)).max()) if self.nks else 0.0
        self.gk_ecut = self.ecutwfc / lat.tpiba2
        self.ggecut = (math.sqrt(self.gk_ecut) + kmaxmod) ** 2
        if self.ggecut > self.gridecut_lat:
            self.ggecut = self.gridecut_lat
            self.gk_ecut = (math.sqrt(self.ggecut) - kmaxmod) ** 2

        # --- big sphere in stick order (pw_distributeg.cpp) ---
        ixe, iye, ize = nx // 2 + 1, ny // 2 + 1, nz // 2 + 1
        fx, fy, fz = np.meshgrid(np.arange(-ixe, ixe + 1), np.arange(-iye, iye + 1),
                                 np.arange(-ize, ize + 1), indexing="ij")
        f = np.stack([fx, fy, fz], axis=-1).reshape(-1, 3)
        ff = f.astype(np.float64)
        mod = np.einsum("ni,ij,nj->n", ff, lat.GGT, ff)
        f = f[mod <= self.ggecut]
        x = np.where(f[:, 0] < 0, f[:, 0] + nx, f[:, 0])
        y = np.where(f[:, 1] < 0, f[:, 1] + ny, f[:, 1])
        ixy = x * ny + y
        order = np.lexsort((f[:, 2], ixy))          # primary ixy, then signed iz
        f = f[order]
        ixy = ixy[order]
        z = np.where(f[:, 2] < 0, f[:, 2] + nz, f[:, 2])
        key = ixy.astype(np.int64) * nz + z
        if np.unique(key).size != key.size:
            raise ValueError("G sphere does not fit in the FFT box")
        self.npw_big = f.shape[0]
        self.f_big = f                                # signed direct coords, stick order
        self.ixyz_big = (z + (ixy % ny) * nz + (ixy // ny) * ny * nz).astype(np.int32)

        # --- per-k selection (pw_basis_k.cpp:118-168) ---
        self.npwk = np.zeros(self.nks, dtype=np.int32)
        sel = []
        for ik in range(self.nks):
            fk = self.f_big + self.kvec_d[ik]
            gk2 = np.einsum("ni,ij,nj->n", fk, lat.GGT, fk)
            m = gk2 <= self.gk_ecut
            sel.append((m, gk2[m]))
            self.npwk[ik] = int(m.sum())
        self.npwk_max = int(self.npwk.max())
        self.ig2ixyz_k = np.zeros((self.nks, self.npwk_max), dtype=np.int32)
        self.gk2 = np.zeros((self.nks, self.npwk_max), dtype=np.float64)
        self.gcar_k = np.zeros((self.nks, self.npwk_max, 3), dtype=np.float64)   # k+G, units tpiba
        for ik, (m, g2) in enumerate(sel):
            n = self.npwk[ik]
            self.ig2ixyz_k[ik, :n] = self.ixyz_big[m]
            self.gk2[ik, :n] = g2
            self.gcar_k[ik, :n] = lat.gcar(self.f_big[m] + self.kvec_d[ik])

    def box_gvectors(self) -> np.ndarray:
        """Signed direct coordinates of every FFT box point, shape (nx, ny, nz, 3)."""
        def signed(n):
            i = np.arange(n)
            return np.where(i >= n // 2 + 1, i - n, i)
        gx, gy, gz = np.meshgrid(signed(self.nx), signed(self.ny), signed(self.nz), indexing="ij")
        return np.stack([gx, gy, gz], axis=-1)
