"""Host-side mirror of the reference interface of the plane-wave Davidson path.

Names and argument meanings follow ABACUS:

* :class:`HamiltPW` -- ``hamilt::HamiltPW`` with ``ops->hPsi`` (operator.cpp:39-88) and the
  ``act`` of Ekinetic / Veff / Nonlocal (accumulate semantics);
* :class:`DiagoDavid` -- ``hsolver::DiagoDavid`` (diago_david.h:21-84): constructor
  ``(precondition, nband, dim, david_ndim, use_paw, ...)`` and
  ``diag(hpsi_func, spsi_func, ldPsi, psi, eigenvalue, david_diag_thr, david_maxiter,
  ntry_max=5, notconv_max=0)`` returning the total iteration count;
* the device functors ``ekinetic_pw_op``, ``veff_pw_op``, ``nonlocal_pw_op``,
  ``set_3d_fft_box_op``, ``set_recip_to_real_output_op``, ``set_real_to_recip_output_op``,
  ``gemm_op``, ``dnevx_op`` and ``FFT::fft3D_forward/backward``.

All compute goes through libapw.so (HIP, gfx950); device arrays are torch tensors used
only as device memory.  Errors are raised as :class:`apw._lib.ApwError` (the reference
aborts instead).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import APW_FP32_MIXED, APW_FP32_MIXED_X32, APW_FP64, check

_PREC = {"fp64": APW_FP64, "double": APW_FP64, "fp32_mixed": APW_FP32_MIXED, "mixed": APW_FP32_MIXED,
         "fp32_mixed_x32": APW_FP32_MIXED_X32,
         APW_FP64: APW_FP64, APW_FP32_MIXED: APW_FP32_MIXED, APW_FP32_MIXED_X32: APW_FP32_MIXED_X32}


def _torch():
    import torch
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _dev_c128(t):
    torch = _torch()
    if t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous():
        raise TypeError("expected a contiguous complex128 CUDA tensor")
    return t


class Context:
    """One device and one stream.  ``stream``: a HIP stream handle to share (e.g. a non-default
    torch stream's ``cuda_stream``); None lets the library create its own.  Calls made
    through this module wait for pending torch work on the device first, so torch-written
    inputs are complete; results are complete after :meth:`synchronize`."""

    def __init__(self, device: int = 0, stream=None):
        _torch().cuda.init()
        self.lib = _lib.load()
        self.device = device
        self.handle = C.c_void_p()
        check(self.lib.apw_ctx_create(C.byref(self.handle), device, C.c_void_p(stream or 0)))

    def pre(self):
        _torch().cuda.synchronize(self.device)

    def set_eig_single_wg(self, nmax: int):
        """Subspace matrices up to nmax are reduced by one workgroup (no device-wide ordering
        of the reductions: concurrent small solves on several contexts)."""
        check(self.lib.apw_ctx_set_eig_single_wg(self.handle, int(nmax)))

    def set_profiling(self, on: bool):
        check(self.lib.apw_ctx_set_profiling(self.handle, int(on)))

    def profile(self):
        """Per-kernel totals {name: dict(launches, ms, bytes, flops)} of the profiler."""
        n = check(self.lib.apw_ctx_profile_count(self.handle))
        out = {}
        buf = C.create_string_buffer(128)
        for i in range(n):
            la, ms, by, fl = C.c_longlong(), C.c_double(), C.c_double(), C.c_double()
            check(self.lib.apw_ctx_profile_entry(self.handle, i, buf, 128, C.byref(la), C.byref(ms), C.byref(by),
                                                 C.byref(fl)))
            out[buf.value.decode()] = dict(launches=la.value, ms=ms.value, bytes=by.value, flops=fl.value)
        return out

    def synchronize(self):
        check(self.lib.apw_ctx_synchronize(self.handle))

    def close(self):
        if self.handle:
            self.lib.apw_ctx_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HamiltPW:
    """H of one k-point built from host arrays (see apw_ham_desc in include/apw.h)."""

    def __init__(self, ctx: Context, system, precision="fp64"):
        self.ctx = ctx
        self.lib = ctx.lib
        s = system
        self.npw, self.npwx = s.npw, s.npwx
        self.nx, self.ny, self.nz = s.nx, s.ny, s.nz
        self._keep = dict(
            ig2ixyz=np.ascontiguousarray(s.ig2ixyz[: s.npw], dtype=np.int32),
            gk2=np.ascontiguousarray(s.gk2[: s.npw], dtype=np.float64),
            veff=np.ascontiguousarray(s.veff, dtype=np.float64),
            vkb=np.ascontiguousarray(s.vkb, dtype=np.complex128),
            na=np.ascontiguousarray(s.na, dtype=np.int32),
            nh=np.ascontiguousarray(s.nh, dtype=np.int32),
            deeq=np.ascontiguousarray(s.deeq[0], dtype=np.float64),
        )
        k = self._keep
        d = _lib.HamDesc()
        d.nx, d.ny, d.nz, d.npw, d.npwx = s.nx, s.ny, s.nz, s.npw, s.npwx
        d.ig2ixyz = k["ig2ixyz"].ctypes.data_as(C.POINTER(C.c_int))
        d.gk2 = k["gk2"].ctypes.data_as(C.POINTER(C.c_double))
        d.tpiba2 = s.tpiba2
        d.veff = k["veff"].ctypes.data_as(C.POINTER(C.c_double))
        d.nkb = s.vkb.shape[0]
        d.vkb = k["vkb"].ctypes.data_as(C.POINTER(C.c_double))
        d.ntype = len(k["na"])
        d.na = k["na"].ctypes.data_as(C.POINTER(C.c_int))
        d.nh = k["nh"].ctypes.data_as(C.POINTER(C.c_int))
        d.nhm = s.deeq.shape[-1]
        d.deeq = k["deeq"].ctypes.data_as(C.POINTER(C.c_double))
        d.precision = _PREC[precision]
        self.handle = C.c_void_p()
        check(self.lib.apw_ham_create(ctx.handle, C.byref(d), C.byref(self.handle)))

    def set_precision(self, precision):
        check(self.lib.apw_ham_set_precision(self.handle, _PREC[precision]))

    def hPsi(self, psi, nbands, ldpsi=None, hpsi=None):
        """Operator::hPsi: returns hpsi (nbands x ld) = H psi."""
        torch = _torch()
        ld = ldpsi or psi.shape[-1]
        if hpsi is None:
            hpsi = torch.empty((nbands, ld), dtype=torch.complex128, device=psi.device)
        self.ctx.pre()
        check(self.lib.apw_hpsi(self.handle, _ptr(_dev_c128(psi)), ld, _ptr(_dev_c128(hpsi)), hpsi.shape[-1], nbands))
        return hpsi

    def act_ekinetic(self, nbands, nbasis, psi, hpsi):
        self.ctx.pre()
        check(self.lib.apw_act_ekinetic(self.handle, nbands, nbasis, _ptr(psi), _ptr(hpsi)))

    def act_veff(self, nbands, nbasis, psi, hpsi):
        self.ctx.pre()
        check(self.lib.apw_act_veff(self.handle, nbands, nbasis, _ptr(psi), _ptr(hpsi)))

    def act_nonlocal(self, nbands, nbasis, psi, hpsi):
        self.ctx.pre()
        check(self.lib.apw_act_nonlocal(self.handle, nbands, nbasis, _ptr(psi), _ptr(hpsi)))

    def recip_to_real(self, inp, out, add=False, factor=1.0):
        self.ctx.pre()
        check(self.lib.apw_recip_to_real(self.handle, _ptr(inp), _ptr(out), int(add), float(factor)))

    def real_to_recip(self, inp, out, add=False, factor=1.0):
        self.ctx.pre()
        check(self.lib.apw_real_to_recip(self.handle, _ptr(inp), _ptr(out), int(add), float(factor)))

    def set_vkb(self, vkb, ldvkb: int):
        """Use device projectors (nkb rows of ``ldvkb``, c128) from now on."""
        self.ctx.pre()
        check(self.lib.apw_ham_set_vkb(self.handle, _ptr(vkb), int(ldvkb)))

    def psi_to_rho(self, psi, nbands, wg, omega, rho=None, add=False, ldpsi=None):
        """ElecStatePW::psiToRho for this k-point (nspin 1): rho(r) (+)= sum_b wg[b]/omega
        |psi_b(r)|^2; ``psi`` device c128 (nbands, ld), ``rho`` device float64 (nxyz)."""
        torch = _torch()
        ld = ldpsi or psi.shape[-1]
        if rho is None:
            rho = torch.zeros(self.nx * self.ny * self.nz, dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        w = (C.c_double * max(1, nbands))(*[float(x) for x in wg[:nbands]])
        self.ctx.pre()
        check(self.lib.apw_psi_to_rho(self.handle, _ptr(_dev_c128(psi)), ld, nbands, w, float(omega), _ptr(rho),
                                      int(add)))
        return rho

    def close(self):
        if self.handle:
            self.lib.apw_ham_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DiagoDavid:
    """hsolver::DiagoDavid on the GPU; ``precondition`` is a host array (len >= dim)."""

    def __init__(self, ctx: Context, precondition, nband: int, dim: int, david_ndim: int, use_paw: bool = False):
        if use_paw:
            raise NotImplementedError("PAW is outside the plane-wave norm-conserving path")
        self.ctx = ctx
        self.lib = ctx.lib
        self.nband, self.dim = nband, dim
        pre = np.ascontiguousarray(precondition[:dim], dtype=np.float64)
        self.handle = C.c_void_p()
        check(self.lib.apw_david_create(ctx.handle, pre.ctypes.data_as(C.POINTER(C.c_double)), nband, dim,
                                        david_ndim, C.byref(self.handle)))
        self.stats = None

    def diag(self, hpsi_func, spsi_func, ldPsi, psi, eigenvalue, david_diag_thr, david_maxiter, ntry_max=5,
             notconv_max=0, precision="fp64"):
        """DiagoDavid::diag.  ``hpsi_func`` is a :class:`HamiltPW` (H applied inside the
        library) or a callable ``f(X, HX, nvec, dim, id_start, id_end)`` on torch views
        (X: nvec x dim head of the block, HX: the destination rows).  ``spsi_func``: None
        for S = I or ``f(X, SX, nrow, npw, nbands)``.  ``psi``: CUDA complex128
        (nband x ldPsi); ``eigenvalue``: host float64 array, filled in place."""
        torch = _torch()
        _dev_c128(psi)
        eig = np.zeros(self.nband, dtype=np.float64)
        stats = _lib.DiagStats()
        ham = None
        hcb = _lib.HPSI_CB()
        scb = _lib.SPSI_CB()
        errors = []
        if isinstance(hpsi_func, HamiltPW):
            ham = hpsi_func.handle
        else:
            dim = self.dim

            def _h(user, X, HX, nvec, d, i0, i1):
                try:
                    xs = _wrap(X, nvec * d, psi.device).view(nvec, d)
                    hs = _wrap(HX, (i1 - i0 + 1) * d, psi.device).view(i1 - i0 + 1, d)
                    hpsi_func(xs, hs, nvec, d, i0, i1)
                    torch.cuda.synchronize(psi.device)
                    return 0
                except Exception as exc:  # reported after the call returns
                    errors.append(exc)
                    return 1

            hcb = _lib.HPSI_CB(_h)
            if spsi_func is not None:
                def _s(user, X, SX, nrow, npw, nb):
                    try:
                        xs = _wrap(X, nb * nrow, psi.device).view(nb, nrow)
                        ss = _wrap(SX, nb * nrow, psi.device).view(nb, nrow)
                        spsi_func(xs, ss, nrow, npw, nb)
                        torch.cuda.synchronize(psi.device)
                        return 0
                    except Exception as exc:
                        errors.append(exc)
                        return 1

                scb = _lib.SPSI_CB(_s)
            del dim
        torch.cuda.synchronize(psi.device)
        rc = self.lib.apw_david_diag(self.handle, ham, hcb, scb, None, ldPsi, _ptr(psi),
                                     eig.ctypes.data_as(C.POINTER(C.c_double)), float(david_diag_thr),
                                     int(david_maxiter), int(ntry_max), int(notconv_max), _PREC[precision],
                                     C.byref(stats))
        if errors:
            raise errors[0]
        check(rc)
        eigenvalue[: self.nband] = eig
        self.stats = stats.as_dict()
        return rc

    def close(self):
        if self.handle:
            self.lib.apw_david_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _wrap(ptr, count, device):
    """A complex128 torch view of device memory owned by the library (no copy)."""
    torch = _torch()

    class _Arr:
        pass

    a = _Arr()
    a.__cuda_array_interface__ = {"shape": (count,), "typestr": "<c16", "data": (int(ptr), False), "version": 2,
                                  "strides": None}
    return torch.as_tensor(a, device=device)


# ---------------------------------------------------------------------------- functors
class DiagoDavSubspace:
    """hsolver::Diago_DavSubspace on the GPU (the dav_subspace method): ``precondition``
    is a host array (len >= dim; HSolverPW passes |k+G|^2 tpiba2 for this method)."""

    def __init__(self, ctx: Context, precondition, nband: int, dim: int, david_ndim: int, diag_thr: float,
                 diag_nmax: int, need_subspace: bool = False):
        self.ctx = ctx
        self.lib = ctx.lib
        self.nband, self.dim = nband, dim
        pre = np.ascontiguousarray(precondition[:dim], dtype=np.float64)
        self.handle = C.c_void_p()
        check(self.lib.apw_dav_subspace_create(ctx.handle, pre.ctypes.data_as(C.POINTER(C.c_double)), nband, dim,
                                               david_ndim, float(diag_thr), int(diag_nmax), int(need_subspace),
                                               C.byref(self.handle)))
        self.stats = None

    def diag(self, hpsi_func, psi, psi_in_dmax, eigenvalue_in, is_occupied=None, scf_type=True):
        """Diago_DavSubspace::diag(hpsi_func, psi_in, psi_in_dmax, eigenvalue_in,
        is_occupied, scf_type).  ``hpsi_func``: a :class:`HamiltPW` or a callable
        ``f(X, HX, nvec, dim, id_start, id_end)`` on torch views; ``psi``: CUDA complex128
        (nband x psi_in_dmax); ``eigenvalue_in``: host float64 array, filled in place."""
        torch = _torch()
        _dev_c128(psi)
        eig = np.zeros(self.nband, dtype=np.float64)
        stats = _lib.DiagStats()
        ham = None
        hcb = _lib.HPSI_CB()
        errors = []
        if isinstance(hpsi_func, HamiltPW):
            ham = hpsi_func.handle
        else:
            def _h(user, X, HX, nvec, d, i0, i1):
                try:
                    xs = _wrap(X, nvec * d, psi.device).view(nvec, d)
                    hs = _wrap(HX, (i1 - i0 + 1) * d, psi.device).view(i1 - i0 + 1, d)
                    hpsi_func(xs, hs, nvec, d, i0, i1)
                    torch.cuda.synchronize(psi.device)
                    return 0
                except Exception as exc:
                    errors.append(exc)
                    return 1

            hcb = _lib.HPSI_CB(_h)
        occ = None
        if is_occupied is not None:
            occ = (C.c_int * self.nband)(*[int(bool(x)) for x in is_occupied[: self.nband]])
        torch.cuda.synchronize(psi.device)
        rc = self.lib.apw_dav_subspace_diag(self.handle, ham, hcb, None, int(psi_in_dmax), _ptr(psi),
                                            eig.ctypes.data_as(C.POINTER(C.c_double)), occ, int(bool(scf_type)),
                                            C.byref(stats))
        if errors:
            raise errors[0]
        check(rc)
        eigenvalue_in[: self.nband] = eig
        self.stats = stats.as_dict()
        return rc

    def close(self):
        if self.handle:
            self.lib.apw_dav_subspace_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ekinetic_pw_op(ctx, nband, npw, max_npw, tpiba2, gk2, hpsi, psi):
    ctx.pre()
    check(ctx.lib.apw_ekinetic_pw_op(ctx.handle, nband, npw, max_npw, float(tpiba2), _ptr(gk2), _ptr(hpsi), _ptr(psi)))


def veff_pw_op(ctx, size, out, inp):
    ctx.pre()
    check(ctx.lib.apw_veff_pw_op(ctx.handle, size, _ptr(out), _ptr(inp)))


def elecstate_pw_op(ctx, spin, nrxx, w1, rho, wfcr):
    """elecstate_pw_op: rho[spin] += w1 |wfcr|^2; ``rho`` a list of device float64 arrays."""
    ptrs = (C.c_void_p * len(rho))(*[_ptr(r) for r in rho])
    ctx.pre()
    check(ctx.lib.apw_elecstate_pw_op(ctx.handle, spin, nrxx, float(w1), ptrs, _ptr(wfcr)))


def elecstate_pw_op_nc(ctx, domag, domag_z, nrxx, w1, rho, wfcr, wfcr_another_spin):
    ptrs = (C.c_void_p * 4)(*[_ptr(r) for r in rho])
    ctx.pre()
    check(ctx.lib.apw_elecstate_pw_op_nc(ctx.handle, int(domag), int(domag_z), nrxx, float(w1), ptrs, _ptr(wfcr),
                                         _ptr(wfcr_another_spin)))


def nonlocal_pw_op(ctx, l1, l2, l3, sum_, iat, spin, nkb, deeq_x, deeq_y, deeq_z, deeq, ps, becp):
    s = C.c_int(sum_)
    ia = C.c_int(iat)
    ctx.pre()
    check(ctx.lib.apw_nonlocal_pw_op(ctx.handle, l1, l2, l3, C.byref(s), C.byref(ia), spin, nkb, deeq_x, deeq_y,
                                     deeq_z, _ptr(deeq), _ptr(ps), _ptr(becp)))
    return s.value, ia.value


def set_3d_fft_box_op(ctx, npwk, box_index, inp, out):
    ctx.pre()
    check(ctx.lib.apw_set_3d_fft_box_op(ctx.handle, npwk, _ptr(box_index), _ptr(inp), _ptr(out)))


def set_recip_to_real_output_op(ctx, nrxx, add, factor, inp, out):
    ctx.pre()
    check(ctx.lib.apw_set_recip_to_real_output_op(ctx.handle, nrxx, int(add), float(factor), _ptr(inp), _ptr(out)))


def set_real_to_recip_output_op(ctx, npw_k, nxyz, add, factor, box_index, inp, out):
    ctx.pre()
    check(ctx.lib.apw_set_real_to_recip_output_op(ctx.handle, npw_k, nxyz, int(add), float(factor), _ptr(box_index),
                                                  _ptr(inp), _ptr(out)))


def fft3d(ctx, nx, ny, nz, box, direction):
    """FFT::fft3D_backward (direction=+1) / fft3D_forward (-1), in place, unnormalised."""
    ctx.pre()
    check(ctx.lib.apw_fft3d(ctx.handle, nx, ny, nz, _ptr(box), int(direction)))


def gemm_op(ctx, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, x32=False):
    """gemm_op<complex<double>|complex<float>, DEVICE_GPU>: column-major, dtype from ``c``.
    ``x32``: a complex128 product on the FP32 matrix cores (apw_zgemm_x32)."""
    single = str(getattr(c, "dtype", "")) == "torch.complex64"
    if x32 and single:
        raise ValueError("gemm_op: x32 takes complex128 data")
    ct, fn = (C.c_float, ctx.lib.apw_cgemm) if single else (
        C.c_double, ctx.lib.apw_zgemm_x32 if x32 else ctx.lib.apw_zgemm)
    al = (ct * 2)(complex(alpha).real, complex(alpha).imag)
    be = (ct * 2)(complex(beta).real, complex(beta).imag)
    ctx.pre()
    check(fn(ctx.handle, transa.encode(), transb.encode(), m, n, k, al, _ptr(a), lda, _ptr(b), ldb, be, _ptr(c), ldc))


def david_diag_batched(units, david_diag_thr, david_maxiter, ntry_max=5, notconv_max=0, precision="fp64"):
    """Batched DiagoDavid::diag over independent units.  ``units``: list of
    (DiagoDavid, HamiltPW, psi, ldPsi, eigenvalue) with psi CUDA complex128 (nband x ldPsi)
    and eigenvalue a host float64 array filled in place.  Units whose solver and
    Hamiltonian were built on different contexts run concurrently.  Returns the per-unit
    iteration counts; each DiagoDavid's ``stats`` is set."""
    import torch
    n = len(units)
    if n == 0:
        return []
    lib = units[0][0].lib
    dav = (C.c_void_p * n)(*[u[0].handle for u in units])
    ham = (C.c_void_p * n)(*[u[1].handle for u in units])
    ld = (C.c_int * n)(*[int(u[3]) for u in units])
    psi = (C.c_void_p * n)(*[_ptr(u[2]) for u in units])
    eigs = [np.zeros(u[0].nband, dtype=np.float64) for u in units]
    eig = (C.c_void_p * n)(*[e.ctypes.data for e in eigs])
    stats = (_lib.DiagStats * n)()
    status = (C.c_int * n)()
    for u in units:
        _dev_c128(u[2])
    torch.cuda.synchronize()
    check(lib.apw_david_diag_batched(n, dav, ham, ld, psi, eig, float(david_diag_thr), int(david_maxiter),
                                     int(ntry_max), int(notconv_max), _PREC[precision], stats, status))
    for k, u in enumerate(units):
        u[4][: u[0].nband] = eigs[k]
        u[0].stats = stats[k].as_dict()
    return [int(status[k]) for k in range(n)]


def dnevx_op(ctx, nstart, ldh, hcc, nbands, eigenvalue, vcc):
    ctx.pre()
    check(ctx.lib.apw_dnevx_op(ctx.handle, nstart, ldh, _ptr(hcc), nbands, _ptr(eigenvalue), _ptr(vcc)))


def _hptr(a, dtype):
    """Host pointer of a numpy copy (kept alive by the returned pair)."""
    arr = np.ascontiguousarray(a, dtype=dtype)
    return arr, C.c_void_p(arr.ctypes.data)


def cal_ylm_real_op(ctx, ng, lmax, g, p, ylm):
    """ModuleBase::cal_ylm_real_op (math_op.h) with the reference's constants; device
    tensors (``p`` may be None)."""
    import math
    ctx.pre()
    check(ctx.lib.apw_cal_ylm_real_op(ctx.handle, ng, lmax, math.sqrt(2.0), math.pi, 0.5 * math.pi, 4.0 * math.pi,
                                      1.0 / math.sqrt(4.0 * math.pi), _ptr(g), None if p is None else _ptr(p),
                                      _ptr(ylm)))


def cal_sk_op(ctx, ik, ntype, nx, ny, nz, rho_nx, rho_ny, rho_nz, npw, npwx, fftny, eigts1_nc, eigts2_nc, eigts3_nc,
              atom_na, igl2isz, is2fftixy, two_pi, kvec_c, atom_tau, eigts1, eigts2, eigts3, sk):
    """hamilt::cal_sk_op (wf_op.h); all arrays device tensors."""
    ctx.pre()
    check(ctx.lib.apw_cal_sk_op(ctx.handle, ik, ntype, nx, ny, nz, rho_nx, rho_ny, rho_nz, npw, npwx, fftny,
                                eigts1_nc, eigts2_nc, eigts3_nc, _ptr(atom_na), _ptr(igl2isz), _ptr(is2fftixy),
                                float(two_pi), _ptr(kvec_c), _ptr(atom_tau), _ptr(eigts1), _ptr(eigts2),
                                _ptr(eigts3), _ptr(sk)))


def cal_vnl_op(ctx, ntype, npw, npwx, nhm, tab_2, tab_3, atom_na, atom_nb, atom_nh, dq, tpiba, neg_imag_unit, gk,
               ylm, indv, nhtol, nhtolm, tab, vkb1, sk, vkb_in):
    """hamilt::cal_vnl_op (vnl_op.h); all arrays device tensors."""
    ctx.pre()
    ni = complex(neg_imag_unit)
    check(ctx.lib.apw_cal_vnl_op(ctx.handle, ntype, npw, npwx, nhm, tab_2, tab_3, _ptr(atom_na), _ptr(atom_nb),
                                 _ptr(atom_nh), float(dq), float(tpiba), ni.real, ni.imag, _ptr(gk), _ptr(ylm),
                                 _ptr(indv), _ptr(nhtol), _ptr(nhtolm), _ptr(tab), _ptr(vkb1), _ptr(sk),
                                 _ptr(vkb_in)))


def getvnl(ctx, ntype, npw, npwx, nhm, tab_2, tab_3, atom_na, atom_nb, atom_nh, dq, tpiba, gk, tau, indv, nhtol,
           nhtolm, tab, vkb):
    """pseudopot_cell_vnl::getvnl for one k-point in one pass.  ``atom_na/nb/nh``,
    ``indv/nhtol/nhtolm`` (ntype x nhm) and ``tau`` (nat x 3, lat0 units) are host arrays;
    ``gk`` (npw x 3, 2 pi / lat0), ``tab`` and ``vkb`` (nkb x npwx c128) device tensors."""
    ctx.pre()
    keep = [_hptr(atom_na, np.int32), _hptr(atom_nb, np.int32), _hptr(atom_nh, np.int32), _hptr(tau, np.float64),
            _hptr(indv, np.float64), _hptr(nhtol, np.float64), _hptr(nhtolm, np.float64)]
    na, nb, nh, ta, iv, hl, hlm = (k[1] for k in keep)
    check(ctx.lib.apw_getvnl(ctx.handle, ntype, npw, npwx, nhm, tab_2, tab_3, na, nb, nh, float(dq), float(tpiba),
                             _ptr(gk), ta, iv, hl, hlm, _ptr(tab), _ptr(vkb)))


def write_wfc_pw(ctx, filename, ik_global, nkstot, kvec_c, weight, nbands, ecutwfc, lat0, tpiba, G, nx, ny, nz, npw,
                 ig2ixyz, psi, ldpsi):
    """ModuleIO::write_wfc_pw for one k-point: ``filename`` ends in .dat (binary) or .txt;
    ``kvec_c`` (3), ``G`` (3 x 3) and ``ig2ixyz`` (npw) host arrays, ``psi`` a CUDA c128
    tensor (band b at offset b * ldpsi)."""
    ctx.pre()
    kv, pk = _hptr(kvec_c, np.float64)
    g, pg = _hptr(np.asarray(G, dtype=np.float64).reshape(9), np.float64)
    ix, pi = _hptr(ig2ixyz[:npw], np.int32)
    check(ctx.lib.apw_write_wfc_pw(ctx.handle, os.fsencode(filename), ik_global, nkstot, pk, float(weight), nbands,
                                   float(ecutwfc), float(lat0), float(tpiba), pg, nx, ny, nz, npw, pi,
                                   _ptr(_dev_c128(psi)), ldpsi))


def read_wfc_pw(ctx, filename, ik, nkstot, nbands, kvec_c, lat0, tpiba, G, nx, ny, nz, npw, ig2ixyz, wfc, ldwfc):
    """ModuleIO::read_wfc_pw for one k-point into the CUDA c128 tensor ``wfc``."""
    ctx.pre()
    kv, pk = _hptr(kvec_c, np.float64)
    g, pg = _hptr(np.asarray(G, dtype=np.float64).reshape(9), np.float64)
    ix, pi = _hptr(ig2ixyz[:npw], np.int32)
    check(ctx.lib.apw_read_wfc_pw(ctx.handle, os.fsencode(filename), ik, nkstot, nbands, pk, float(lat0),
                                  float(tpiba), pg, nx, ny, nz, npw, pi, _ptr(_dev_c128(wfc)), ldwfc))


def dngvx_op(ctx, nstart, ldh, hcc, scc, m, eigenvalue, vcc):
    ctx.pre()
    check(ctx.lib.apw_dngvx_op(ctx.handle, nstart, ldh, _ptr(hcc), _ptr(scc), m, _ptr(eigenvalue), _ptr(vcc)))


def dngvd_op(ctx, nstart, ldh, hcc, scc, eigenvalue, vcc):
    ctx.pre()
    check(ctx.lib.apw_dngvd_op(ctx.handle, nstart, ldh, _ptr(hcc), _ptr(scc), _ptr(eigenvalue), _ptr(vcc)))


def probe_mfma_layout(ctx):
    torch = _torch()
    out = torch.zeros(4 * 64 * 4, dtype=torch.float64, device=f"cuda:{ctx.device}")
    check(ctx.lib.apw_probe_mfma_layout(ctx.handle, _ptr(out)))
    ctx.synchronize()
    return out.view(4, 64, 4).cpu().numpy()


# ---------------------------------------------------------------- autotuner
TUNE_FIXED, TUNE_CACHE, TUNE_MEASURE = 0, 1, 2
TUNE_CACHE_FILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tune_cache.json")


def _tlib():
    # torch before the library: libapw.so then binds to torch's HIP runtime (a process that
    # loads the system runtime first holds two, and the second finds no device)
    _torch()
    return _lib.load()


def tune_set_mode(mode: int):
    """0: fixed rules; 1: the measured table, fixed rules on a miss; 2: as 1, and a miss is
    measured where it occurs (every candidate plan timed on the calling stream)."""
    check(_tlib().apw_tune_set_mode(int(mode)))


def tune_load(path: str = TUNE_CACHE_FILE) -> int:
    return check(_tlib().apw_tune_load(path.encode()))


def tune_save(path: str = TUNE_CACHE_FILE) -> int:
    return check(_tlib().apw_tune_save(path.encode()))


def tune_clear():
    check(_tlib().apw_tune_clear())


def tune_stats():
    """{entries, hits, misses, tuned} of the process-wide table."""
    v = [C.c_longlong() for _ in range(4)]
    check(_tlib().apw_tune_stats(*[C.byref(x) for x in v]))
    return dict(zip(("entries", "hits", "misses", "tuned"), (x.value for x in v)))


def tune_entries():
    """[(key, (a, b), us)] of the table (a, b: the recorded choice)."""
    lib = _tlib()
    n = tune_stats()["entries"]
    buf = C.create_string_buffer(256)
    out = []
    for i in range(n):
        a, b, us = C.c_int(), C.c_int(), C.c_double()
        check(lib.apw_tune_entry(i, buf, 256, C.byref(a), C.byref(b), C.byref(us)))
        out.append((buf.value.decode(), (a.value, b.value), us.value))
    return out
