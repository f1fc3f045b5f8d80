"""ctypes access to oracle/_ref/libref_david.so -- the reference's own CPU DiagoDavid
(TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg).

Build it with ``make -C oracle`` in a container that holds the reference tree.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libref_david.so")
LIB_DETSIGN = os.path.join(HERE, "_ref", "libref_david_detsign.so")
_libs = {}


def available(detsign: bool = False) -> bool:
    return os.path.exists(LIB_DETSIGN if detsign else LIB)


def load(detsign: bool = False):
    """The reference build (``detsign``: with zheevx's all-eigenpair path re-phased to the
    inverse-iteration convention, see ref_lapack_signs.cpp)."""
    path = LIB_DETSIGN if detsign else LIB
    _lib = _libs.get(path)
    if _lib is None:
        _lib = C.CDLL(path)
        _libs[path] = _lib
        P = C.c_void_p
        _lib.ref_david_diag.restype = C.c_int
        _lib.ref_david_diag.argtypes = [C.c_int] * 5 + [P, P, C.c_double, P, C.c_int, P, C.c_int, P, P, C.c_int, P,
                                                        P, C.c_int, C.c_int, C.c_int, P, P, C.c_double, C.c_int,
                                                        C.c_int, C.c_int, P, P, C.c_int, P, C.c_int, P]
        _lib.ref_dav_subspace_diag.restype = C.c_int
        _lib.ref_dav_subspace_diag.argtypes = ([C.c_int] * 5 + [P, P, C.c_double, P, C.c_int, P, C.c_int, P, P,
                                                                C.c_int, P, P, C.c_int, C.c_int, C.c_int, P, P,
                                                                C.c_double, C.c_int, C.c_int, C.c_int, P])
        _lib.ref_hpsi.restype = None
        _lib.ref_hpsi.argtypes = [C.c_int] * 5 + [P, P, C.c_double, P, C.c_int, P, C.c_int, P, P, C.c_int, P, P, P,
                                                  C.c_int, C.c_int]
    return _lib


def _arrays(s):
    a = dict(
        ig2ixyz=np.ascontiguousarray(s.ig2ixyz[: s.npw], dtype=np.int32),
        gk2=np.ascontiguousarray(s.gk2[: s.npw], dtype=np.float64),
        veff=np.ascontiguousarray(s.veff, dtype=np.float64),
        vkb=np.ascontiguousarray(s.vkb, dtype=np.complex128),
        na=np.ascontiguousarray(s.na, dtype=np.int32),
        nh=np.ascontiguousarray(s.nh, dtype=np.int32),
        deeq=np.ascontiguousarray(s.deeq[0], dtype=np.float64),
        pre=np.ascontiguousarray(s.precondition, dtype=np.float64),
    )
    return a


def _p(x):
    return x.ctypes.data_as(C.c_void_p)


def hpsi(s, psi, nb, ld):
    """One reference Operator::hPsi on nb bands (row = band, leading dimension ld)."""
    lib = load()
    a = _arrays(s)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros_like(psi)
    lib.ref_hpsi(s.nx, s.ny, s.nz, s.npw, s.npwx, _p(a["ig2ixyz"]), _p(a["gk2"]), s.tpiba2, _p(a["veff"]), s.nkb,
                 _p(a["vkb"]), len(a["na"]), _p(a["na"]), _p(a["nh"]), s.deeq.shape[-1], _p(a["deeq"]), _p(psi),
                 _p(out), nb, ld)
    return out


def david(s, thr=1e-9, maxiter=500, ntry_max=5, notconv_max=0, psi0=None, nband=None, dump_call=-1,
          detsign=False):
    """Reference DiagoDavid::diag from the system's start vectors.  Returns
    (eigenvalues, psi, iterations, hpsi_calls, hpsi_bands)."""
    lib = load(detsign)
    a = _arrays(s)
    nband = nband or s.nband
    psi = np.array(s.psi0[:nband] if psi0 is None else psi0, dtype=np.complex128, copy=True, order="C")
    eig = np.zeros(nband)
    counts = np.zeros(2, dtype=np.int64)
    cap = 4096
    trace = np.zeros(cap, dtype=np.int32)
    etrace = np.zeros(cap * nband)
    dump = np.zeros((s.david_ndim * nband, s.npw), dtype=np.complex128)
    it = lib.ref_david_diag(s.nx, s.ny, s.nz, s.npw, s.npwx, _p(a["ig2ixyz"]), _p(a["gk2"]), s.tpiba2, _p(a["veff"]),
                            s.nkb, _p(a["vkb"]), len(a["na"]), _p(a["na"]), _p(a["nh"]), s.deeq.shape[-1],
                            _p(a["deeq"]), _p(a["pre"]), nband, s.david_ndim, s.npwx, _p(psi), _p(eig), thr, maxiter,
                            ntry_max, notconv_max, _p(counts), _p(trace), cap, _p(etrace), dump_call,
                            _p(dump))
    ncall = int(counts[0])
    david.last_trace = trace[:ncall].tolist()
    david.last_etrace = etrace[: ncall * nband].reshape(ncall, nband)
    david.last_dump = dump
    return eig, psi, it, int(counts[0]), int(counts[1])


def dav_subspace(s, thr=1e-9, maxiter=500, need_subspace=False, scf=True, psi0=None, nband=None, precondition=None):
    """Reference Diago_DavSubspace::diag from the system's start vectors, with the
    dav_subspace preconditioner of HSolverPW (|k+G|^2 tpiba2 unless given).  Returns
    (eigenvalues, psi, iterations, hpsi_calls, hpsi_bands)."""
    lib = load(False)
    a = _arrays(s)
    nband = nband or s.nband
    if precondition is None:
        pre = np.ones(s.npwx)
        pre[: s.npw] = np.asarray(s.gk2[: s.npw]) * s.tpiba2
    else:
        pre = np.ascontiguousarray(precondition, dtype=np.float64)
    psi = np.array(s.psi0[:nband] if psi0 is None else psi0, dtype=np.complex128, copy=True, order="C")
    eig = np.zeros(nband)
    counts = np.zeros(2, dtype=np.int64)
    it = lib.ref_dav_subspace_diag(s.nx, s.ny, s.nz, s.npw, s.npwx, _p(a["ig2ixyz"]), _p(a["gk2"]), s.tpiba2,
                                   _p(a["veff"]), s.nkb, _p(a["vkb"]), len(a["na"]), _p(a["na"]), _p(a["nh"]),
                                   s.deeq.shape[-1], _p(a["deeq"]), _p(pre), nband, s.david_ndim, s.npwx, _p(psi),
                                   _p(eig), thr, maxiter, int(need_subspace), int(scf), _p(counts))
    return eig, psi, it, int(counts[0]), int(counts[1])
