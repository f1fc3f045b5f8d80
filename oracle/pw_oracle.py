"""CPU oracle of the plane-wave Davidson path (TEST INFRASTRUCTURE ONLY).

This module restates, in numpy, the reference algorithm the HIP library must
reproduce.  It is imported only by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``, always as the checker -- never as the thing
measured or shipped.  Each function cites the reference file:line it follows
(paths relative to the ABACUS source tree).

How each part is pinned (checks live in tests/test_oracle.py):

* ekinetic_pw_op   -- reference golden vector ekinetic_op_test.cpp (27 x 8) + closed form
* veff_pw_op       -- reference golden vector veff_op_test.cpp (1728 points)
* nonlocal_pw_op   -- reference golden vector nonlocal_op_test.cpp (ps, sum, iat)
* set_3d_fft_box / set_recip_to_real_output / set_real_to_recip_output
                   -- reference golden vectors pw_op_test.cpp (npwk 59, nxyz 1728)
* fft3d            -- brute-force O(N^2) DFT on small boxes, and the pw_op_test
                      chain (box -> backward FFT reproduces the veff_op_test input)
* hpsi             -- dense H = kinetic + V(G-G') + sum beta D beta^H built
                      independently, applied to random vectors
* DiagoDavid       -- converged eigenvalues vs dense LAPACK eigh; iteration-level
                      trace vs the compiled reference DiagoDavid (oracle/_ref) when
                      that library is present
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla


# ---------------------------------------------------------------------------
# kernel functors (CPU forms of the reference device ops)
# ---------------------------------------------------------------------------

def ekinetic_pw_op(nband, npw, max_npw, tpiba2, gk2, hpsi, psi):
    """hpsi[ib*max_npw+ig] += gk2[ig]*tpiba2*psi[...]  -- kernels/ekinetic_op.cpp:6-26."""
    for ib in range(nband):
        s = slice(ib * max_npw, ib * max_npw + npw)
        hpsi[s] += gk2[:npw] * tpiba2 * psi[s]
    return hpsi


def veff_pw_op(size, out, veff):
    """out[ir] *= veff[ir]  -- kernels/veff_op.cpp:10-20."""
    out[:size] *= veff[:size]
    return out


def nonlocal_pw_op(l1, l2, l3, sum_, iat, spin, nkb, deeq_x, deeq_y, deeq_z, deeq, ps, becp):
    """ps[(sum+ii*l3+kk)*l2+jj] += sum_xx deeq[spin,iat+ii,xx,kk] * becp[jj*nkb+sum+ii*l3+xx]
    -- kernels/nonlocal_op.cpp:6-38.  Returns (sum, iat) advanced as the reference does."""
    d = np.asarray(deeq).reshape(-1, deeq_x, deeq_y, deeq_z)
    b = np.asarray(becp).reshape(l2, nkb)
    for ii in range(l1):
        dm = d[spin, iat + ii, :l3, :l3]                 # [xx, kk]
        blk = b[:, sum_ + ii * l3: sum_ + (ii + 1) * l3]  # [jj, xx]
        contrib = blk @ dm                                # [jj, kk]
        for kk in range(l3):
            base = (sum_ + ii * l3 + kk) * l2
            ps[base: base + l2] += contrib[:, kk]
    return sum_ + l1 * l3, iat + l1


def set_3d_fft_box_op(npwk, box_index, inp, out):
    """out[box_index[ig]] = in[ig]  -- module_pw/kernels/pw_op.cpp:6-19."""
    out[np.asarray(box_index[:npwk])] = inp[:npwk]
    return out


def set_recip_to_real_output_op(nrxx, add, factor, inp, out):
    """out = in, or out += factor*in  -- pw_op.cpp:21-43."""
    if add:
        out[:nrxx] += factor * inp[:nrxx]
    else:
        out[:nrxx] = inp[:nrxx]
    return out


def set_real_to_recip_output_op(npw_k, nxyz, add, factor, box_index, inp, out):
    """out[ig] (+)= (factor/nxyz) * in[box_index[ig]]  -- pw_op.cpp:45-67."""
    idx = np.asarray(box_index[:npw_k])
    if add:
        out[:npw_k] += factor / float(nxyz) * inp[idx]
    else:
        out[:npw_k] = inp[idx] / float(nxyz)
    return out


def fft3d_backward(box, nx, ny, nz):
    """Unnormalised G->r transform, exponent +i, z fastest -- fft.cpp:786-821 (BACKWARD)."""
    return (np.fft.ifftn(box.reshape(nx, ny, nz)) * (nx * ny * nz)).reshape(-1)


def fft3d_forward(box, nx, ny, nz):
    """Unnormalised r->G transform, exponent -i -- fft.cpp:773-785 (FORWARD)."""
    return np.fft.fftn(box.reshape(nx, ny, nz)).reshape(-1)


def dft_bruteforce(x, sign):
    """O(N^2) 1D DFT, used to pin the FFT on tiny sizes."""
    n = x.shape[-1]
    k = np.arange(n)
    w = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return x @ w.T


# ---------------------------------------------------------------------------
# the operator chain (Operator::hPsi -> Ekinetic, Veff, Nonlocal)
# ---------------------------------------------------------------------------

class HamiltOracle:
    """H|psi> of one k-point, following operator.cpp:39-88 and the three act()s."""

    def __init__(self, sysm):
        self.s = sysm
        self.nx, self.ny, self.nz = sysm.nx, sysm.ny, sysm.nz
        self.nxyz = sysm.nxyz
        self.npw = sysm.npw
        self.npwx = sysm.npwx
        self.box_index = np.asarray(sysm.ig2ixyz[: sysm.npw], dtype=np.int64)
        self.gk2 = np.asarray(sysm.gk2, dtype=np.float64)
        self.tpiba2 = sysm.tpiba2
        self.veff = sysm.veff
        self.vkb = sysm.vkb                      # (nkb, npwx) row = projector
        self.deeq = sysm.deeq
        self.na = list(sysm.na)
        self.nh = list(sysm.nh)
        self.nkb = sysm.vkb.shape[0]

    def act_ekinetic(self, nbands, nbasis, psi, hpsi):
        """ekinetic_pw.cpp:34-58."""
        return ekinetic_pw_op(nbands, self.npw, nbasis, self.tpiba2, self.gk2, hpsi, psi)

    def act_veff(self, nbands, nbasis, psi, hpsi):
        """veff_pw.cpp:38-105 with the transforms of pw_transform_k.cpp:336-476."""
        for ib in range(nbands):
            p = psi[ib * nbasis: ib * nbasis + self.npw]
            box = np.zeros(self.nxyz, dtype=np.complex128)
            set_3d_fft_box_op(self.npw, self.box_index, p, box)
            box = fft3d_backward(box, self.nx, self.ny, self.nz)
            veff_pw_op(self.nxyz, box, self.veff)
            box = fft3d_forward(box, self.nx, self.ny, self.nz)
            out = hpsi[ib * nbasis: ib * nbasis + self.npw]
            set_real_to_recip_output_op(self.npw, self.nxyz, True, 1.0, self.box_index, box, out)
        return hpsi

    def act_nonlocal(self, nbands, nbasis, psi, hpsi):
        """nonlocal_pw.cpp:211-304: becp = vkb^H psi; ps = D becp; hpsi += vkb ps^T."""
        npw = self.npw
        P = psi.reshape(nbands, nbasis)[:, :npw]                  # bands x npw
        V = self.vkb[:, :npw]                                      # nkb x npw
        becp = np.conj(V) @ P.T                                    # nkb x nbands (col-major becp[ib*nkb+ikb])
        becp_flat = np.ascontiguousarray(becp.T).reshape(-1)       # becp[ib*nkb + ikb]
        ps = np.zeros(self.nkb * nbands, dtype=np.complex128)      # ps[ikb*nbands + ib]
        sum_, iat = 0, 0
        for it in range(len(self.na)):
            sum_, iat = nonlocal_pw_op(self.na[it], nbands, self.nh[it], sum_, iat, 0, self.nkb,
                                       self.deeq.shape[1], self.deeq.shape[2], self.deeq.shape[3],
                                       self.deeq, ps, becp_flat)
        psm = ps.reshape(self.nkb, nbands)                          # [ikb, ib]
        add = (V.T @ psm).T                                         # bands x npw
        H = hpsi.reshape(nbands, nbasis)
        H[:, :npw] += add
        return hpsi

    def hpsi(self, psi, nbands, nbasis):
        """Operator::hPsi: zero the output, then every act() accumulates (operator.cpp:134-167)."""
        psi = np.asarray(psi, dtype=np.complex128).reshape(-1)
        hpsi = np.zeros(nbands * nbasis, dtype=np.complex128)
        self.act_ekinetic(nbands, nbasis, psi, hpsi)
        self.act_veff(nbands, nbasis, psi, hpsi)
        self.act_nonlocal(nbands, nbasis, psi, hpsi)
        return hpsi

    def dense(self):
        """Independent dense H_{GG'} = |k+G|^2 d + V(G-G') + sum_ij beta_i(G) D_ji beta_j(G')^*."""
        npw = self.npw
        ix = self.box_index
        nx, ny, nz = self.nx, self.ny, self.nz
        vg = np.fft.fftn(self.veff.reshape(nx, ny, nz)) / self.nxyz
        gx, rem = np.divmod(ix, ny * nz)
        gy, gz = np.divmod(rem, nz)
        dx = (gx[:, None] - gx[None, :]) % nx
        dy = (gy[:, None] - gy[None, :]) % ny
        dz = (gz[:, None] - gz[None, :]) % nz
        H = vg[dx, dy, dz].astype(np.complex128)
        H[np.arange(npw), np.arange(npw)] += self.gk2[:npw] * self.tpiba2
        V = self.vkb[:, :npw]
        D = np.zeros((self.nkb, self.nkb))
        off, iat = 0, 0
        for it in range(len(self.na)):
            nh = self.nh[it]
            for ia in range(self.na[it]):
                D[off:off + nh, off:off + nh] = self.deeq[0, iat][:nh, :nh].T   # D[k, x] = deeq[x, k]
                off += nh
                iat += 1
        H += V.T @ D @ np.conj(V)
        return H


# ---------------------------------------------------------------------------
# DiagoDavid (block Davidson) -- source/module_hsolver/diago_david.cpp
# ---------------------------------------------------------------------------

class DavidAbort(RuntimeError):
    pass


def householder_tridiag_lower(A):
    """Reflectors of the reduction of Hermitian A to real tridiagonal form, lower
    triangle, LAPACK zhetrd('L')/zlarfg convention: P_k = I - tau_k v_k v_k^H, v_k[k+1] = 1,
    P_k^H x = beta e_1 with beta = -sign(Re alpha) * ||x||.  Returns (V, tau)."""
    A = np.array(A, dtype=np.complex128, copy=True)
    n = A.shape[0]
    V = np.zeros((n, n), dtype=np.complex128)
    tau = np.zeros(n, dtype=np.complex128)
    for k in range(n - 1):
        x = A[k + 1:, k].copy()
        alpha = x[0]
        xn2 = float(np.sum(np.abs(x[1:]) ** 2))
        v = np.zeros(n - k - 1, dtype=np.complex128)
        v[0] = 1.0
        if xn2 == 0.0 and alpha.imag == 0.0:
            t = 0.0
        else:
            beta = -np.copysign(np.sqrt(abs(alpha) ** 2 + xn2), alpha.real)
            t = (beta - alpha) / beta
            v[1:] = x[1:] / (alpha - beta)
        V[k + 1:, k] = v
        tau[k] = t
        if t != 0.0:
            At = A[k + 1:, k + 1:]
            p = At @ v
            mu = np.vdot(v, p).real
            y = t * p - (abs(t) ** 2 * mu / 2.0) * v
            A[k + 1:, k + 1:] = At - np.outer(y, v.conj()) - np.outer(v, y.conj())
    return V, tau


def _apply_qh(V, tau, Z):
    Z = np.array(Z, dtype=np.complex128, copy=True)
    for k in range(V.shape[0] - 1):
        v = V[k + 1:, k]
        Z[k + 1:, :] -= np.conj(tau[k]) * np.outer(v, v.conj() @ Z[k + 1:, :])
    return Z


def dnevx(nbase, hcc, nband):
    """Lowest nband eigenpairs of the Hermitian nbase x nbase leading block, lower
    triangle referenced (dnevx_op -> zheevx, dngvd_op.cpp:142-235).

    Eigenvector phases follow LAPACK's reduction + inverse-iteration path of zheevx
    (zhetrd 'L', dstein): with A = Q T Q^H, each vector is Q z with z real and its
    largest-magnitude component positive.  zheevx uses that convention whenever fewer
    than all eigenpairs are requested; for the all-eigenpairs case (nbase == nband) its QR
    path leaves implementation-dependent signs, which the rule above replaces."""
    a = hcc[:nbase, :nbase]
    A = np.tril(a) + np.tril(a, -1).conj().T
    A[np.diag_indices(nbase)] = A.diagonal().real
    w, v = sla.eigh(A, subset_by_index=[0, nband - 1])
    V, tau = householder_tridiag_lower(A)
    z = _apply_qh(V, tau, v)
    for j in range(nband):
        i = int(np.argmax(np.abs(z[:, j])))
        ph = z[i, j] / abs(z[i, j])
        v[:, j] = v[:, j] / ph
    return w, v


class DiagoDavidOracle:
    """Restatement of DiagoDavid<complex<double>, CPU> (diago_david.cpp:34-1255).

    Matrices use the reference's column-major layouts as numpy arrays indexed
    [row, col]: basis/hpsi/spsi are (dim, nbase_x), hcc/scc/vcc (nbase_x, nbase_x).
    ``hpsi_func(X, HX_view, nvec, dim, i0, i1)`` writes H X[:, i0:i1+1] into HX_view;
    ``spsi_func`` is S = I unless given.
    """

    def __init__(self, precondition, nband, dim, david_ndim):
        assert david_ndim > 1 and david_ndim * nband < dim     # diago_david.cpp:60-61
        self.nband, self.dim, self.ndim = nband, dim, david_ndim
        self.nbase_x = david_ndim * nband
        self.pre = np.asarray(precondition[:dim], dtype=np.float64)
        self.trace = []            # per iteration: (nbase, notconv, eigenvalues)
        self.hpsi_calls = 0
        self.hpsi_bands = 0

    # -- diago_david.cpp:950-1075 (the block plan of planSchmidtOrth only regroups the
    # same inner products; classical Gram-Schmidt with a Pythagorean norm) --
    def _schmidt(self, basis, spsi, m):
        psi_m = basis[:, m]
        lag = np.conj(basis[:, : m + 1]).T @ spsi[:, m]          # <basis_j | S psi_m>, j <= m
        psi_norm = lag[m].real
        if m > 0:
            basis[:, m] = psi_m - basis[:, :m] @ lag[:m]
            psi_norm -= float(np.real(np.vdot(lag[:m], lag[:m])))
        if not psi_norm > 0.0:
            raise DavidAbort("SchmidtOrth: non-positive norm")
        psi_norm = np.sqrt(psi_norm)
        if psi_norm < 1.0e-12:
            raise DavidAbort("SchmidtOrth: psi_norm < 1e-12")
        basis[:, m] /= psi_norm

    def diag(self, hpsi_func, psi, eig, thr, maxiter, ntry_max=5, notconv_max=0, spsi_func=None):
        """diago_david.cpp:1173-1200; psi is (nband, ldpsi) row = band (updated in place)."""
        ntry, total = 0, 0
        self.notconv = 0
        while True:
            total += self._diag_once(hpsi_func, spsi_func, psi, eig, thr, maxiter)
            ntry += 1
            if ntry > ntry_max or self.notconv <= notconv_max:    # check_block_conv :1223-1240
                break
        return total

    def _spsi(self, spsi_func, x):
        return x.copy() if spsi_func is None else spsi_func(x)

    def _diag_once(self, hpsi_func, spsi_func, psi, eig, thr, maxiter):
        nband, dim, nbx = self.nband, self.dim, self.nbase_x
        basis = np.zeros((dim, nbx), dtype=np.complex128)
        hp = np.zeros((dim, nbx), dtype=np.complex128)
        sp = np.zeros((dim, nbx), dtype=np.complex128)
        hcc = np.zeros((nbx, nbx), dtype=np.complex128)
        scc = np.zeros((nbx, nbx), dtype=np.complex128)
        vcc = np.zeros((nbx, nbx), dtype=np.complex128)
        unconv = list(range(nband))
        notconv = nband
        nbase = 0
        # initial S-orthonormalisation (:186-228)
        for m in range(nband):
            sp[:, m] = self._spsi(spsi_func, psi[m, :dim])
        for m in range(nband):
            basis[:, m] = psi[m, :dim]
            self._schmidt(basis, sp, m)
            sp[:, m] = self._spsi(spsi_func, basis[:, m])
        hpsi_func(basis, hp[:, 0:nband], nbx, dim, 0, nband - 1)
        self.hpsi_calls += 1
        self.hpsi_bands += nband
        nbase = self._cal_elem(basis, hp, sp, hcc, scc, nbase, notconv)
        w, v = dnevx(nbase, hcc, nband)
        ev = np.zeros(nbx)
        ev[:nband] = w
        vcc[:nbase, :nband] = v
        eig[:nband] = w
        it = 0
        while True:
            it += 1
            self._cal_grad(hpsi_func, spsi_func, basis, hp, sp, vcc, unconv, notconv, nbase, ev)
            nbase = self._cal_elem(basis, hp, sp, hcc, scc, nbase, notconv)
            w, v = dnevx(nbase, hcc, nband)
            ev[:nband] = w
            # zheevx writes only Z(1:nbase, 1:nband); the rest of vcc keeps old values
            vcc[:nbase, :nband] = v
            notconv = 0
            unconv = []
            for m in range(nband):
                if not abs(w[m] - eig[m]) < thr:                 # :270-280
                    unconv.append(m)
                    notconv += 1
                eig[m] = w[m]
            self.notconv = notconv
            self.trace.append((nbase, notconv, w.copy()))
            if notconv == 0 or nbase + notconv > nbx or it == maxiter:
                psi[:, :] = 0
                psi[:, :dim] = (basis[:, :nbase] @ vcc[:nbase, :nband]).T   # :292-306
                if notconv == 0 or it == maxiter:
                    break
                nbase = self._refresh(basis, hp, sp, hcc, scc, vcc, psi, eig, nbase)
        return it

    def _cal_grad(self, hpsi_func, spsi_func, basis, hp, sp, vcc, unconv, notconv, nbase, ev):
        """diago_david.cpp:359-611."""
        if notconv == 0:
            return
        dim = self.dim
        vc = vcc[:nbase, unconv].copy()                            # (nbase, notconv)
        new = hp[:, :nbase] @ vc
        vc = vc * (-ev[unconv])[None, :]
        new = new + sp[:, :nbase] @ vc
        new = new / self.pre[:, None]
        basis[:, nbase:nbase + notconv] = new
        for m in range(notconv):
            sp[:, nbase + m] = self._spsi(spsi_func, basis[:, nbase + m])
        for m in range(notconv):
            self._schmidt(basis, sp, nbase + m)
            sp[:, nbase + m] = self._spsi(spsi_func, basis[:, nbase + m])
        hpsi_func(basis, hp[:, nbase:nbase + notconv], self.nbase_x, dim, nbase, nbase + notconv - 1)
        self.hpsi_calls += 1
        self.hpsi_bands += notconv

    def _cal_elem(self, basis, hp, sp, hcc, scc, nbase, notconv):
        """diago_david.cpp:626-717: hcc[nbase:nbase+nc, :nbase+nc] = basis_new^H hpsi."""
        if notconv == 0:
            return nbase
        bn = basis[:, nbase:nbase + notconv]
        hcc[nbase:nbase + notconv, :nbase + notconv] = np.conj(bn).T @ hp[:, :nbase + notconv]
        scc[nbase:nbase + notconv, :nbase + notconv] = np.conj(bn).T @ sp[:, :nbase + notconv]
        return nbase + notconv

    def _refresh(self, basis, hp, sp, hcc, scc, vcc, psi, eig, nbase):
        """diago_david.cpp:797-934."""
        nband, dim = self.nband, self.dim
        h_new = hp[:, :nbase] @ vcc[:nbase, :nband]
        s_new = sp[:, :nbase] @ vcc[:nbase, :nband]
        hp[:, :nband] = h_new
        sp[:, :nband] = s_new
        basis[:, :] = 0
        basis[:, :nband] = psi[:, :dim].T
        hcc[:, :] = 0
        scc[:, :] = 0
        # vcc is NOT cleared (diago_david.cpp:866-870 zero hcc and scc only): only its
        # diagonal is set to one, the off-diagonal entries of the previous solve remain
        # and enter the next cal_grad through vcc[:nband, unconv].
        for i in range(nband):
            hcc[i, i] = eig[i]
            scc[i, i] = 1.0
            vcc[i, i] = 1.0
        return nband


def david_solve(sysm, thr=1e-9, maxiter=500, ntry_max=5, notconv_max=0, nband=None, psi0=None):
    """Run the oracle Davidson on a synthetic system; returns (eig, psi, iters, oracle)."""
    ham = HamiltOracle(sysm)
    nband = nband or sysm.nband
    dim = sysm.npw

    def hpsi_func(X, HXv, nvec, d, i0, i1):
        nb = i1 - i0 + 1
        cols = X[:, i0:i1 + 1]
        out = ham.hpsi(np.ascontiguousarray(cols.T).reshape(-1), nb, d)
        HXv[:, :] = out.reshape(nb, d).T

    psi = np.array(sysm.psi0[:nband] if psi0 is None else psi0, dtype=np.complex128, copy=True)
    eig = np.zeros(nband)
    dav = DiagoDavidOracle(sysm.precondition, nband, dim, sysm.david_ndim)
    iters = dav.diag(hpsi_func, psi, eig, thr, maxiter, ntry_max, notconv_max)
    return eig, psi, iters, dav
