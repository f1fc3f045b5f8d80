"""Seeded synthetic plane-wave Hamiltonians: the inputs of the Davidson / H|psi> path.

The path consumes exactly what ABACUS's ``HamiltPW`` + ``PW_Basis_K`` +
``pseudopot_cell_vnl`` hand to ``Operator::hPsi`` for one k-point:

* ``gk2`` (|k+G|^2, tpiba^2 units), ``tpiba2`` -> kinetic term
  (source/module_hamilt_pw/hamilt_pwdft/kernels/ekinetic_op.cpp:22);
* ``veff(r)`` on the FFT box, z fastest (operator_pw/veff_pw.cpp:63);
* ``vkb`` (nkb x npwx, projector-major, ordered type -> atom -> ih, with the
  (-i)^l and structure-factor phases; VNL_in_pw.cpp:386-401) and
  ``deeq[spin, iat, ih, jh]`` (nonlocal_op.cpp:33-37);
* the Davidson preconditioner ``1 + g + sqrt(1 + (g-1)^2)`` with g = gk2*tpiba2
  (source/module_hsolver/hsolver_pw.cpp:581) and 1.0 past npw;
* start vectors ``(N(0,1) + i N(0,1)) / (1 + |k+G|^2_Ry)`` (the reference's random
  start divides by gk2+1, wf_atomic.cpp:660-678).

Synthetic pseudopotential (all energies in Ry, lengths in bohr):

* local potential: a Gaussian well ``-2 exp(-|r-tau|^2 / (2 sigma^2))``, sigma = 1.2,
  on every atom, built in G space on the FFT box so it is exactly periodic, plus
  smooth random noise (Hermitian coefficients for |G| < 4 bohr^-1) scaled to an RMS
  of 0.05 Ry in real space;
* 8 Kleinman-Bylander projectors per atom (2 s + 2 x 3 p):
  ``beta(q) = c_l(w) Omega^-1/2 |q|^l exp(-(w |q|)^2 / 2) Y_lm(q^) (-i)^l exp(-i q.tau)``,
  q = k+G, one width w ~ U[0.8, 1.6] per radial channel, ``D`` diagonal with one
  value ~ U[-1, 1] per radial channel (shared by its m components).  ``c_l(w)``
  normalises the projector, sum_G |beta(G)|^2 -> 1 in the continuum limit:
  c_0 = sqrt((2 pi)^3 / (sqrt(pi) / (4 w^3))), c_1 = sqrt((2 pi)^3 / (3 sqrt(pi) / (8 w^5))).

Random draws use numpy's PCG64 with a fixed draw order, so a (config, seed) pair
gives bit-identical inputs on every machine.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .basis import Lattice, PWBasisK

SIGMA_LOCAL = 1.2
DEPTH_LOCAL = 2.0
NOISE_RMS = 0.05
NOISE_GMAX = 4.0
NH_PER_ATOM = 8
# radial channel of each of the 8 projectors and its angular momentum
_PROJ_CHANNEL = np.array([0, 1, 2, 2, 2, 3, 3, 3])
_PROJ_L = np.array([0, 0, 1, 1, 1, 1, 1, 1])
_PROJ_M = np.array([0, 0, 0, 1, 2, 0, 1, 2])      # p: x, y, z components


def diamond_positions(ncell=(1, 1, 1)) -> np.ndarray:
    """Direct coordinates of a diamond-like cubic supercell (8 atoms per cube)."""
    fcc = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]], dtype=np.float64)
    basis = np.concatenate([fcc, fcc + 0.25])
    out = []
    for i in range(ncell[0]):
        for j in range(ncell[1]):
            for k in range(ncell[2]):
                out.append((basis + np.array([i, j, k])) / np.array(ncell, dtype=np.float64))
    return np.concatenate(out)


@dataclass
class PWSystem:
    """One k-point of a synthetic Hamiltonian, in the reference's array layouts."""
    name: str
    lat: Lattice
    basis: PWBasisK
    ik: int
    tau: np.ndarray            # (nat, 3) cartesian bohr
    nband: int
    david_ndim: int
    veff: np.ndarray           # (nx*ny*nz,) float64, index iz + iy*nz + ix*ny*nz
    vkb: np.ndarray            # (nkb, npwx) complex128 (row = projector), zero past npw
    deeq: np.ndarray           # (nspin=1, nat, nh, nh) float64
    na: np.ndarray             # atoms per type
    nh: np.ndarray             # projectors per atom, per type
    precondition: np.ndarray   # (npwx,) float64
    psi0: np.ndarray           # (nband, npwx) complex128, zero past npw
    seed: int
    meta: dict = field(default_factory=dict)

    @property
    def nx(self): return self.basis.nx

    @property
    def ny(self): return self.basis.ny

    @property
    def nz(self): return self.basis.nz

    @property
    def nxyz(self): return self.basis.nxyz

    @property
    def npw(self): return int(self.basis.npwk[self.ik])

    @property
    def npwx(self): return int(self.basis.npwk_max)

    @property
    def tpiba2(self): return float(self.basis.tpiba2)

    @property
    def gk2(self): return self.basis.gk2[self.ik]

    @property
    def ig2ixyz(self): return self.basis.ig2ixyz_k[self.ik]

    @property
    def nkb(self): return int(self.vkb.shape[0])

    @property
    def nat(self): return int(self.tau.shape[0])


def _local_potential(lat: Lattice, basis: PWBasisK, tau: np.ndarray, rng) -> np.ndarray:
    f = basis.box_gvectors().astype(np.float64)
    gc = lat.gcar(f) * lat.tpiba                        # bohr^-1
    g2 = (gc ** 2).sum(-1)
    pref = -DEPTH_LOCAL * (2.0 * math.pi * SIGMA_LOCAL ** 2) ** 1.5 / lat.omega
    vg = np.zeros(g2.shape, dtype=np.complex128)
    gauss = pref * np.exp(-0.5 * SIGMA_LOCAL ** 2 * g2)
    for t in tau:
        vg += gauss * np.exp(-1j * (gc @ t))
    # smooth Hermitian noise for |G| < NOISE_GMAX
    mask = np.sqrt(g2) < NOISE_GMAX
    nmask = int(mask.sum())
    c = np.zeros(g2.shape, dtype=np.complex128)
    c[mask] = rng.standard_normal(nmask) + 1j * rng.standard_normal(nmask)
    # c(-G) index: negate signed coords -> (-i) mod n
    nx, ny, nz = basis.nx, basis.ny, basis.nz
    cm = c[(-np.arange(nx)) % nx][:, (-np.arange(ny)) % ny][:, :, (-np.arange(nz)) % nz]
    c = 0.5 * (c + np.conj(cm))
    noise_r = np.real(np.fft.ifftn(c) * basis.nxyz)
    rms = math.sqrt(float(np.mean(noise_r ** 2)))
    if rms > 0:
        c *= NOISE_RMS / rms
    vr = np.real(np.fft.ifftn(vg + c) * basis.nxyz)
    return np.ascontiguousarray(vr.reshape(-1))


def _projectors(lat: Lattice, basis: PWBasisK, ik: int, tau: np.ndarray, widths: np.ndarray) -> np.ndarray:
    npw = int(basis.npwk[ik])
    npwx = int(basis.npwk_max)
    q = basis.gcar_k[ik, :npw] * lat.tpiba                 # k+G, bohr^-1
    qn = np.sqrt((q ** 2).sum(-1))
    y00 = 1.0 / math.sqrt(4.0 * math.pi)
    y1 = math.sqrt(3.0 / (4.0 * math.pi))
    inv_sqrt_omega = 1.0 / math.sqrt(lat.omega)
    radial = [np.exp(-0.5 * (w * qn) ** 2) for w in widths]
    ch_l = [0, 0, 1, 1]
    norm = []
    for w, l in zip(widths, ch_l):
        integral = math.sqrt(math.pi) / (4.0 * w ** 3) if l == 0 else 3.0 * math.sqrt(math.pi) / (8.0 * w ** 5)
        norm.append(math.sqrt((2.0 * math.pi) ** 3 / integral))
    vkb1 = np.zeros((NH_PER_ATOM, npw), dtype=np.complex128)
    for ih in range(NH_PER_ATOM):
        ch, l, m = _PROJ_CHANNEL[ih], _PROJ_L[ih], _PROJ_M[ih]
        ang = np.full(npw, y00) if l == 0 else y1 * q[:, m]     # |q|^l Y_lm
        vkb1[ih] = norm[ch] * inv_sqrt_omega * radial[ch] * ang * ((-1j) ** l)
    nat = tau.shape[0]
    vkb = np.zeros((nat * NH_PER_ATOM, npwx), dtype=np.complex128)
    for ia in range(nat):
        sk = np.exp(-1j * (q @ tau[ia]))
        vkb[ia * NH_PER_ATOM:(ia + 1) * NH_PER_ATOM, :npw] = vkb1 * sk
    return vkb


# reference Ylm row (nhtolm, Y_lm order of cal_ylm_real_op) of each of the 8 projectors and
# the sign that maps it onto the x, y, z components used above: Y_1 rows are z (lm 1),
# -x (lm 2) and -y (lm 3)
_PROJ_LM = np.array([0, 0, 2, 3, 1, 2, 3, 1])
_PROJ_SIGN = np.array([1.0, 1.0, -1.0, -1.0, 1.0, -1.0, -1.0, 1.0])


def projector_tables(sysm: "PWSystem", dq: float = 0.01) -> dict:
    """The same projectors in the form pseudopot_cell_vnl::getvnl consumes
    (VNL_in_pw.cpp:412-524): a radial table tab[it][nb][iq] = beta_nb(iq dq) sampled every
    dq bohr^-1 (the reference's default, input_parameter dq = 0.01) up to past the cutoff,
    indv / nhtol / nhtolm per projector, k+G in 2 pi / lat0 units and tau in lat0 units.
    getvnl of these reproduces ``sysm.vkb`` row by row up to ``sign`` and the
    interpolation error; the radial part of an l = 1 channel carries the |q| of
    |q| Y_1m(q^) = y1 q_m."""
    lat, basis, ik = sysm.lat, sysm.basis, sysm.ik
    npw = sysm.npw
    gk = np.ascontiguousarray(basis.gcar_k[ik, :npw])                       # 2 pi / lat0
    qmax = float(np.sqrt((gk ** 2).sum(-1)).max()) * lat.tpiba
    nqx = int(qmax / dq) + 8
    q = np.arange(nqx) * dq
    widths = np.asarray(sysm.meta["widths"])
    ch_l = [0, 0, 1, 1]
    inv_sqrt_omega = 1.0 / math.sqrt(lat.omega)
    tab = np.zeros((1, 4, nqx))
    for ch, (w, l) in enumerate(zip(widths, ch_l)):
        integral = math.sqrt(math.pi) / (4.0 * w ** 3) if l == 0 else 3.0 * math.sqrt(math.pi) / (8.0 * w ** 5)
        norm = math.sqrt((2.0 * math.pi) ** 3 / integral)
        tab[0, ch] = norm * inv_sqrt_omega * np.exp(-0.5 * (w * q) ** 2) * (q if l == 1 else 1.0)
    return {"ntype": 1, "npw": npw, "npwx": sysm.npwx, "nhm": NH_PER_ATOM, "tab_2": 4, "tab_3": nqx,
            "atom_na": np.array([sysm.nat], dtype=np.int32), "atom_nb": np.array([4], dtype=np.int32),
            "atom_nh": np.array([NH_PER_ATOM], dtype=np.int32), "dq": dq, "tpiba": lat.tpiba, "gk": gk,
            "tau": sysm.tau / lat.lat0, "indv": _PROJ_CHANNEL.astype(np.float64),
            "nhtol": _PROJ_L.astype(np.float64), "nhtolm": _PROJ_LM.astype(np.float64),
            "tab": tab.reshape(-1), "sign": _PROJ_SIGN}


def build_system(name: str, lat: Lattice, pos_direct: np.ndarray, ecutwfc: float, nband: int,
                 david_ndim: int, seed: int, grid=None, kvec_d=None, ik: int = 0,
                 basis: PWBasisK | None = None) -> PWSystem:
    """Build one k-point of a synthetic Hamiltonian (see the module docstring)."""
    rng = np.random.default_rng(seed)
    if basis is None:
        nx, ny, nz = grid if grid is not None else (0, 0, 0)
        basis = PWBasisK(lat, ecutwfc, nx, ny, nz,
                         kvec_d=np.zeros((1, 3)) if kvec_d is None else kvec_d)
    tau = np.asarray(pos_direct, dtype=np.float64) @ lat.latvec * lat.lat0
    # draw order: widths, D, local-potential noise, start vectors
    widths = rng.uniform(0.8, 1.6, size=4)
    dch = rng.uniform(-1.0, 1.0, size=4)
    veff = _local_potential(lat, basis, tau, rng)
    vkb = _projectors(lat, basis, ik, tau, widths)
    nat = tau.shape[0]
    d = dch[_PROJ_CHANNEL]
    deeq = np.zeros((1, nat, NH_PER_ATOM, NH_PER_ATOM), dtype=np.float64)
    for ia in range(nat):
        deeq[0, ia] = np.diag(d)
    npw = int(basis.npwk[ik])
    npwx = int(basis.npwk_max)
    g = basis.gk2[ik, :npw] * basis.tpiba2
    precond = np.ones(npwx, dtype=np.float64)
    precond[:npw] = 1.0 + g + np.sqrt(1.0 + (g - 1.0) ** 2)
    # the per-k random stream continues from the shared draws
    krng = np.random.default_rng([seed, ik])
    psi0 = np.zeros((nband, npwx), dtype=np.complex128)
    z = krng.standard_normal((nband, npw)) + 1j * krng.standard_normal((nband, npw))
    psi0[:, :npw] = z / (1.0 + g)[None, :]
    return PWSystem(name=name, lat=lat, basis=basis, ik=ik, tau=tau, nband=nband,
                    david_ndim=david_ndim, veff=veff, vkb=vkb, deeq=deeq,
                    na=np.array([nat], dtype=np.int32), nh=np.array([NH_PER_ATOM], dtype=np.int32),
                    precondition=precond, psi0=psi0, seed=seed,
                    meta={"widths": widths.tolist(), "D": dch.tolist(), "ecutwfc": ecutwfc})
