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
* dnevx            -- Householder reduction vs LAPACK zhetrd; eigenpairs vs LAPACK
* elecstate_pw_op(_nc) -- reference golden vectors elecstate_op_test.cpp
* psi_to_rho       -- sum rule (integral of rho = sum of weights for normalised
                      bands), one-band closed form (plane wave -> constant density)
* dngvx / dngvd    -- scipy.linalg.eigh (LAPACK zhegvd) on the same pencil
* DiagoDavSubspace -- converged eigenvalues vs dense eigh; iteration count and
                      eigenvalues vs the compiled reference Diago_DavSubspace
* cal_ylm_real     -- reference golden vector math_op_test.cpp (lmax 1) and
                      scipy.special.sph_harm_y for l <= 4 (sqrt2 Re / Im of Y_l^m)
* cal_sk           -- reference golden vector wf_op_test.cpp and the closed form
                      exp(-2 pi i (k+G).tau)
* cal_vnl / getvnl -- reference golden vector vnl_op_test.cpp; the cubic
                      interpolation is exact on cubic polynomials
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


def elecstate_pw_op(spin, nrxx, w1, rho, wfcr):
    """rho[spin][ir] += w1 * |wfcr[ir]|^2 -- elecstate_op.cpp:8-25 (nspin 1/2 form)."""
    rho[spin][:nrxx] += w1 * (wfcr[:nrxx].real ** 2 + wfcr[:nrxx].imag ** 2)


def elecstate_pw_op_nc(domag, domag_z, nrxx, w1, rho, wfcr, wfcr_b):
    """Non-collinear form, elecstate_op.cpp:27-81: charge from both spinor components;
    magnetisation (rho[1..3]) accumulated with DOMAG, only m_z with DOMAG_Z (m_x, m_y set
    to 0), else m set to 0."""
    a, b = wfcr[:nrxx], wfcr_b[:nrxx]
    na = a.real ** 2 + a.imag ** 2
    nb = b.real ** 2 + b.imag ** 2
    rho[0][:nrxx] += w1 * (na + nb)
    if domag:
        rho[1][:nrxx] += w1 * 2.0 * (a.real * b.real + a.imag * b.imag)
        rho[2][:nrxx] += w1 * 2.0 * (a.real * b.imag - b.real * a.imag)
        rho[3][:nrxx] += w1 * (na - nb)
    elif domag_z:
        rho[1][:nrxx] = 0.0
        rho[2][:nrxx] = 0.0
        rho[3][:nrxx] += w1 * (na - nb)
    else:
        for i in (1, 2, 3):
            rho[i][:nrxx] = 0.0


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


# ---------------------------------------------------------------------------
# Diago_DavSubspace -- source/module_hsolver/diago_dav_subspace.cpp
# ---------------------------------------------------------------------------

def dngvx(nbase, hcc, scc, nband):
    """Lowest nband eigenpairs of the Hermitian-definite pencil (hcc, scc) on the nbase x
    nbase leading blocks (dngvx_op -> zhegvx ITYPE 1, dngvd_op.cpp:238-330); both
    matrices are held fully.  Eigenvectors are S-orthonormal; their phases are those of
    scipy (the DavSubspace trajectory does not depend on them: refresh() resets vcc)."""
    h = hcc[:nbase, :nbase]
    sm = scc[:nbase, :nbase]
    h = 0.5 * (h + h.conj().T)
    sm = 0.5 * (sm + sm.conj().T)
    return sla.eigh(h, sm, subset_by_index=[0, nband - 1])


def dngvd(n, hcc, scc):
    """All eigenpairs of the pencil (dngvd_op -> zhegvd, dngvd_op.cpp:82-140)."""
    return dngvx(n, hcc, scc, n)


class DiagoDavSubspaceOracle:
    """Restatement of Diago_DavSubspace<complex<double>, CPU> (diago_dav_subspace.cpp).

    Differences from DiagoDavid that the algorithm itself makes: the correction vectors
    are preconditioned with the shifted QE form 0.5 (1 + x + sqrt(1 + (x-1)^2)),
    x = |precondition - e| (:372-393), normalised but not orthogonalised (:397-414); the
    projected problem is the pencil (hcc, scc) (:577-705), except the first solve of
    each diag_once, which uses hcc alone (dnevx, :612-624); refresh() resets vcc to the
    identity (:709-798)."""

    def __init__(self, precondition, nband, dim, david_ndim, diag_thr, diag_nmax, need_subspace=False):
        assert david_ndim > 1 and david_ndim * nband < dim        # :37-38
        self.n_band, self.dim, self.nbase_x = nband, dim, nband * david_ndim
        self.diag_thr, self.iter_nmax, self.is_subspace = diag_thr, diag_nmax, need_subspace
        self.pre = np.asarray(precondition[:dim], dtype=np.float64)
        nbx = self.nbase_x
        # persistent work arrays (zero-initialised in the constructor, :41-56)
        self.psi_iter = np.zeros((dim, nbx), dtype=np.complex128)
        self.hphi = np.zeros((dim, nbx), dtype=np.complex128)
        self.hcc = np.zeros((nbx, nbx), dtype=np.complex128)
        self.scc = np.zeros((nbx, nbx), dtype=np.complex128)
        self.vcc = np.zeros((nbx, nbx), dtype=np.complex128)
        self.notconv = 0
        self.trace = []

    @staticmethod
    def test_exit_cond(ntry, notconv, scf):
        """:841-858"""
        return ntry <= 5 and ((not scf and notconv > 0) or (scf and notconv > 5))

    def diag(self, hpsi_func, psi_in, eigenvalue, is_occupied=None, scf=True):
        """:804-839; psi_in (nband, ld) row = band, updated in place; returns the iterations."""
        is_occupied = [True] * self.n_band if is_occupied is None else is_occupied
        self.notconv = 0
        total, ntry = 0, 0
        while True:
            if self.is_subspace or ntry > 0:
                self.diagH_subspace(psi_in, eigenvalue, hpsi_func)
            total += self.diag_once(hpsi_func, psi_in, eigenvalue, is_occupied)
            ntry += 1
            if not self.test_exit_cond(ntry, self.notconv, scf):
                break
        return total

    def diagH_subspace(self, psi_in, en, hpsi_func):
        """:860-958: Rayleigh-Ritz of the nband input vectors, pencil solved in full."""
        nb, dim = self.n_band, self.dim
        X = np.ascontiguousarray(psi_in[:nb, :dim].T)
        hphi = np.zeros((dim, nb), dtype=np.complex128)
        hpsi_func(X, hphi, nb, dim, 0, nb - 1)
        hcc = X.conj().T @ hphi
        scc = X.conj().T @ X
        w, v = dngvd(nb, hcc, scc)
        en[:nb] = w
        psi_in[:nb, :dim] = (X @ v).T

    def diag_once(self, hpsi_func, psi_in, eig_in, is_occupied):
        """:86-291"""
        nb, dim, nbx = self.n_band, self.dim, self.nbase_x
        psi_iter, hphi, hcc, scc, vcc = self.psi_iter, self.hphi, self.hcc, self.scc, self.vcc
        eig_iter = np.zeros(nbx)
        unconv = list(range(nb))
        notconv = nb
        nbase = 0
        psi_iter[:, :nb] = psi_in[:nb, :dim].T
        # H on all nbase_x columns (:122-125); only the first nb are used before being
        # overwritten
        hpsi_func(psi_iter, hphi, nbx, dim, 0, nbx - 1)
        nbase = self._cal_elem(nbase, notconv)
        self._diag_zhegvx(nbase, eig_iter, init=True, is_subspace=self.is_subspace)
        eig_in[:nb] = eig_iter[:nb]
        it = 0
        while True:
            it += 1
            self._cal_grad(hpsi_func, nbase, notconv, unconv, eig_iter)
            nbase = self._cal_elem(nbase, notconv)
            self._diag_zhegvx(nbase, eig_iter, init=False, is_subspace=False)
            notconv = 0
            unconv = [0] * nb
            for m in range(nb):
                thr = self.diag_thr if is_occupied[m] else max(self.diag_thr * 5.0, 1e-5)
                if not abs(eig_iter[m] - eig_in[m]) < thr:
                    unconv[notconv] = m
                    notconv += 1
                eig_in[m] = eig_iter[m]
            self.notconv = notconv
            self.trace.append((nbase, notconv, eig_iter[:nb].copy()))
            if notconv == 0 or nbase + notconv + 1 > nbx or it == self.iter_nmax:
                psi_in[:nb, :] = 0
                psi_in[:nb, :dim] = (psi_iter[:, :nbase] @ vcc[:nbase, :nb]).T
                psi_iter[:, :nb] = psi_in[:nb, :dim].T
                nbase = self._refresh(nbase, eig_in)
                if notconv == 0 or it == self.iter_nmax:
                    break
        return it

    def _cal_grad(self, hpsi_func, nbase, notconv, unconv, eig_iter):
        """:294-429"""
        dim, nbx = self.dim, self.nbase_x
        psi_iter, hphi, vcc = self.psi_iter, self.hphi, self.vcc
        for i in range(notconv):
            if unconv[i] != i:
                vcc[:nbase, i] = vcc[:nbase, unconv[i]]
                eig_iter[i] = eig_iter[unconv[i]]
        new = psi_iter[:, :nbase] @ vcc[:nbase, :notconv]
        new = new * (-eig_iter[:notconv])[None, :]
        new = new + hphi[:, :nbase] @ vcc[:nbase, :notconv]
        for m in range(notconv):
            x = np.abs(self.pre - eig_iter[m])
            new[:, m] /= 0.5 * (1.0 + x + np.sqrt(1.0 + (x - 1.0) ** 2))
        for m in range(notconv):
            nrm = float(np.real(np.vdot(new[:, m], new[:, m])))
            assert nrm > 0.0
            new[:, m] /= np.sqrt(nrm)
        psi_iter[:, nbase:nbase + notconv] = new
        hpsi_func(psi_iter, hphi[:, nbase:nbase + notconv], nbx, dim, nbase, nbase + notconv - 1)

    def _cal_elem(self, nbase, notconv):
        """:431-575: new columns of hcc, scc (rows 0 .. nbase+notconv), Hermitian fill."""
        psi_iter, hphi, hcc, scc = self.psi_iter, self.hphi, self.hcc, self.scc
        n1 = nbase + notconv
        hcc[:n1, nbase:n1] = psi_iter[:, :n1].conj().T @ hphi[:, nbase:n1]
        scc[:n1, nbase:n1] = psi_iter[:, :n1].conj().T @ psi_iter[:, nbase:n1]
        last = nbase
        for i in range(n1):
            if i >= last:
                hcc[i, i] = hcc[i, i].real
                scc[i, i] = scc[i, i].real
            for j in range(max(i + 1, last), n1):
                hcc[j, i] = np.conj(hcc[i, j])
                scc[j, i] = np.conj(scc[i, j])
        hcc[n1:, n1:] = 0
        scc[n1:, n1:] = 0
        return n1

    def _diag_zhegvx(self, nbase, eig_iter, init, is_subspace):
        """:577-705"""
        nb = self.n_band
        if is_subspace:
            for m in range(nb):
                eig_iter[m] = self.hcc[m, m].real
                self.vcc[m, m] = 1.0
            return
        if init:
            w, v = dnevx(nbase, self.hcc, nb)
        else:
            w, v = dngvx(nbase, self.hcc, self.scc, nb)
        eig_iter[:nb] = w
        self.vcc[:nbase, :nb] = v
        self.hcc[:nbase, nbase:] = 0
        self.hcc[nbase:, :nbase] = 0
        self.scc[:nbase, nbase:] = 0
        self.scc[nbase:, :nbase] = 0

    def _refresh(self, nbase, eig_in):
        """:709-798"""
        nb = self.n_band
        self.hphi[:, :nb] = self.hphi[:, :nbase] @ self.vcc[:nbase, :nb]
        self.hcc[:nb, :nb] = 0
        self.scc[:nb, :nb] = 0
        self.vcc[:nb, :nb] = 0
        for i in range(nb):
            self.hcc[i, i] = eig_in[i]
            self.scc[i, i] = 1.0
            self.vcc[i, i] = 1.0
        return nb


def dav_subspace_precondition(sysm):
    """HSolverPW::update_precondition for dav_subspace (hsolver_pw.cpp:562-579): the
    kinetic energy |k+G|^2 tpiba2 on the sphere, 1.0 beyond npw."""
    pre = np.ones(sysm.npwx)
    pre[: sysm.npw] = np.asarray(sysm.gk2[: sysm.npw]) * sysm.tpiba2
    return pre


def dav_subspace_solve(sysm, thr=1e-9, maxiter=500, need_subspace=False, scf=True, nband=None, psi0=None):
    """Run the oracle DavSubspace on a synthetic system; returns (eig, psi, iters, oracle)."""
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
    dav = DiagoDavSubspaceOracle(dav_subspace_precondition(sysm), nband, dim, sysm.david_ndim, thr, maxiter,
                                 need_subspace)
    iters = dav.diag(hpsi_func, psi, eig, scf=scf)
    return eig, psi, iters, dav


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


def psi_to_rho(sysm, psi, wg, omega, rho=None):
    """ElecStatePW::rhoBandK for nspin 1 (elecstate_pw.cpp:142-214): for every band with
    nonzero weight, psi(G) -> psi(r) by recip_to_real (pw_transform_k.cpp:404-476: box
    scatter, unnormalised backward FFT) and rho(r) += wg_b / omega * |psi_b(r)|^2.
    ``psi`` is (nband, >= npw) in the sphere order; returns rho (nxyz real)."""
    npw, nxyz = sysm.npw, sysm.nxyz
    rho = np.zeros(nxyz) if rho is None else rho
    bi = np.asarray(sysm.ig2ixyz, dtype=np.int64)[:npw]
    for ib in range(len(wg)):
        w1 = wg[ib] / omega
        if w1 == 0.0:
            continue
        box = np.zeros(nxyz, dtype=np.complex128)
        box[bi] = psi[ib, :npw]
        wr = fft3d_backward(box, sysm.nx, sysm.ny, sysm.nz)
        elecstate_pw_op(0, nxyz, w1, [rho], wr)
    return rho


# ---------------------------------------------------------------------------
# projector generation: Ylm_Real, structure factor, beta(k+G) (getvnl)
# ---------------------------------------------------------------------------

def cal_ylm_real(ng, lmax, g):
    """Real spherical harmonics Y_lm(g^) for l <= lmax at ng points g (ng x 3, flat)
    -- cal_ylm_real_op, module_base/kernels/math_op.cpp:27-126.

    Associated Legendre P_l^m(cos theta) by the recursion on l (Condon-Shortley phase,
    P_l^l = (-1)^l (2l-1)!! (1-x^2)^(l/2)); phi from atan with the quadrant fix-up and
    phi = +-pi/2 when |g_x| <= 1e-9; cos theta = 0 when |g| < 1e-9.  Output rows
    lm = l^2 + {0: m=0, 2m-1: cos(m phi), 2m: sin(m phi)}, each row ng long:
    Y_l0 = c P_l0, Y_lm^{cos,sin} = c sqrt((l-m)!/(l+m)!) sqrt2 P_lm {cos, sin}(m phi),
    c = sqrt((2l+1)/(4 pi)).  Returns ((lmax+1)^2, ng)."""
    g = np.asarray(g, dtype=np.float64).reshape(ng, 3)
    ylm = np.zeros(((lmax + 1) ** 2, ng))
    if lmax == 0:
        ylm[0] = 1.0 / np.sqrt(4.0 * np.pi)
        return ylm
    gmod = np.sqrt((g ** 2).sum(1))
    safe = np.where(gmod < 1e-9, 1.0, gmod)
    cost = np.where(gmod < 1e-9, 0.0, g[:, 2] / safe)
    gx, gy = g[:, 0], g[:, 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        at = np.arctan(gy / np.where(np.abs(gx) > 1e-9, gx, 1.0))
    phi = np.where(gx > 1e-9, at, np.where(gx < -1e-9, at + np.pi, 0.5 * np.pi * np.where(gy >= 0.0, 1.0, -1.0)))
    P = np.zeros((lmax + 1, lmax + 1, ng))          # P[m, l]
    x2 = np.maximum(0.0, 1.0 - cost * cost)
    lm = -1
    for l in range(lmax + 1):
        c = np.sqrt((2 * l + 1) / (4.0 * np.pi))
        if l == 0:
            P[0, 0] = 1.0
        elif l == 1:
            P[0, 1] = cost
            P[1, 1] = -np.sqrt(x2)
        else:
            l1, l2, l3 = l - 1, l - 2, 2 * l - 1
            for m in range(l2 + 1):
                P[m, l] = (cost * l3 * P[m, l1] - (l1 + m) * P[m, l2]) / (l - m)
            P[l1, l] = cost * l3 * P[l1, l1]
            semif = 1
            for i in range(l3, 2, -2):
                semif *= i
            P[l, l] = semif * x2 ** (l / 2.0) * (-1.0 if l % 2 == 1 else 1.0)
        lm += 1
        ylm[lm] = c * P[0, l]
        for m in range(1, l + 1):
            same = c * np.sqrt(float(np.prod(np.arange(1, l - m + 1, dtype=np.float64))) /
                               float(np.prod(np.arange(1, l + m + 1, dtype=np.float64)))) * np.sqrt(2.0)
            lm += 1
            ylm[lm] = same * P[m, l] * np.cos(m * phi)
            lm += 1
            ylm[lm] = same * P[m, l] * np.sin(m * phi)
    return ylm


def cal_sk(ik, ntype, nx, ny, nz, rho_nx, rho_ny, rho_nz, npw, npwx, fftny, eigts1_nc, eigts2_nc, eigts3_nc,
           atom_na, igl2isz, is2fftixy, two_pi, kvec_c, atom_tau, eigts1, eigts2, eigts3):
    """Structure factor sk[iat*npw + ig] = e^{-2 pi i k.tau} eigts1[ix] eigts2[iy] eigts3[iz]
    with (ix, iy, iz) the signed Miller indices of G shifted by rho_n* -- cal_sk_op,
    kernels/wf_op.cpp:8-74 (= exp(-2 pi i (k+G).tau) when the eigts are e^{-2 pi i G_j.tau})."""
    igl2isz = np.asarray(igl2isz, dtype=np.int64)
    is2fftixy = np.asarray(is2fftixy, dtype=np.int64)
    kvec_c = np.asarray(kvec_c, dtype=np.float64)
    atom_tau = np.asarray(atom_tau, dtype=np.float64)
    nat = int(np.sum(atom_na[:ntype]))
    sk = np.zeros(nat * npw, dtype=np.complex128)
    isz = igl2isz[ik * npwx: ik * npwx + npw]
    iz = isz % nz
    ixy = is2fftixy[isz // nz]
    ix, iy = ixy // fftny, ixy % fftny
    ix = np.where(ix >= nx // 2 + 1, ix - nx, ix) + rho_nx
    iy = np.where(iy >= ny // 2 + 1, iy - ny, iy) + rho_ny
    iz = np.where(iz >= nz // 2 + 1, iz - nz, iz) + rho_nz
    for iat in range(nat):
        arg = two_pi * float(np.dot(kvec_c[ik * 3: ik * 3 + 3], atom_tau[iat * 3: iat * 3 + 3]))
        kphase = np.cos(arg) - 1j * np.sin(arg)
        sk[iat * npw:(iat + 1) * npw] = (kphase * eigts1[iat * eigts1_nc + ix] * eigts2[iat * eigts2_nc + iy]
                                         * eigts3[iat * eigts3_nc + iz])
    return sk


def polynomial_interpolation(tab, it, nb, tab_2, tab_3, dq, x):
    """Four-point Lagrange interpolation of tab[it, nb, :] (spacing dq) at x
    -- _polynomial_interpolation, kernels/vnl_tools.hpp:7-27."""
    pos = np.asarray(x, dtype=np.float64) / dq
    iq = pos.astype(np.int64)
    x0 = pos - iq
    x1, x2, x3 = 1.0 - x0, 2.0 - x0, 3.0 - x0
    base = (it * tab_2 + nb) * tab_3 + iq
    t = np.asarray(tab, dtype=np.float64)
    return (t[base] * x1 * x2 * x3 / 6.0 + t[base + 1] * x0 * x2 * x3 / 2.0
            - t[base + 2] * x1 * x0 * x3 / 2.0 + t[base + 3] * x1 * x2 * x0 / 6.0)


def cal_vnl(ntype, npw, npwx, nhm, tab_2, tab_3, atom_na, atom_nb, atom_nh, dq, tpiba, gk, ylm, indv, nhtol,
            nhtolm, tab, sk, vkb=None):
    """beta projectors vkb[jkb*npwx + ig] = ylm[lm(ih), ig] vq_nb(|k+G| tpiba) sk[iat, ig] (-i)^l(ih),
    jkb running over types, atoms, ih -- cal_vnl_op, kernels/vnl_op.cpp:8-96 (getvnl,
    VNL_in_pw.cpp:412-524).  Columns past npw are left as given."""
    nat = int(np.sum(atom_na[:ntype]))
    gk = np.asarray(gk, dtype=np.float64).reshape(-1, 3)[:npw]
    gnorm = np.sqrt((gk ** 2).sum(1)) * tpiba
    if vkb is None:
        vkb = np.zeros(int(sum(atom_na[t] * atom_nh[t] for t in range(ntype))) * npwx, dtype=np.complex128)
    pref_tab = np.array([1.0, -1j, -1.0, 1j])
    ylm = np.asarray(ylm, dtype=np.float64).reshape(-1, npw)
    sk = np.asarray(sk).reshape(nat, npw)
    jkb = iat = 0
    for it in range(ntype):
        nh = atom_nh[it]
        vkb1 = np.zeros((nh, npw))
        for nb in range(atom_nb[it]):
            vq = polynomial_interpolation(tab, it, nb, tab_2, tab_3, dq, gnorm)
            for ih in range(nh):
                if nb == int(indv[it * nhm + ih]):
                    vkb1[ih] = ylm[int(nhtolm[it * nhm + ih])] * vq
        for _ in range(atom_na[it]):
            for ih in range(nh):
                pref = pref_tab[int(nhtol[it * nhm + ih]) % 4]
                vkb[jkb * npwx: jkb * npwx + npw] = vkb1[ih] * sk[iat] * pref
                jkb += 1
            iat += 1
    return vkb


def getvnl(ntype, npw, npwx, nhm, tab_2, tab_3, atom_na, atom_nb, atom_nh, dq, tpiba, gk, tau, indv, nhtol,
           nhtolm, tab, vkb=None):
    """getvnl for one k-point (VNL_in_pw.cpp:412-524): Ylm_Real of k+G up to the largest l
    in nhtol, sk = exp(-2 pi i (k+G).tau) (gk: k+G cartesian in 2 pi / a, tau: cartesian
    in a), then cal_vnl."""
    lmaxkb = int(max(int(nhtol[it * nhm + ih]) for it in range(ntype) for ih in range(atom_nh[it])))
    gk = np.asarray(gk, dtype=np.float64).reshape(-1, 3)[:npw]
    ylm = cal_ylm_real(npw, lmaxkb, gk)
    tau = np.asarray(tau, dtype=np.float64).reshape(-1, 3)
    sk = np.exp(-2j * np.pi * (tau @ gk.T)).reshape(-1)
    return cal_vnl(ntype, npw, npwx, nhm, tab_2, tab_3, atom_na, atom_nb, atom_nh, dq, tpiba, gk, ylm, indv,
                   nhtol, nhtolm, tab, sk, vkb)


# ---------------------------------------------------------------------------
# wavefunction files (module_io/write_wfc_pw.cpp, read_wfc_pw.cpp), npol 1
# ---------------------------------------------------------------------------

def _miller(ig2ixyz, ny, nz):
    idx = np.asarray(ig2ixyz, dtype=np.int64)
    return np.stack([idx // (ny * nz), (idx // nz) % ny, idx % nz], axis=1).astype(np.int32)


def write_wfc_pw(path, ik_global, nkstot, kvec_c, weight, nbands, ecutwfc, lat0, tpiba, G, nx, ny, nz, ig2ixyz,
                 psi):
    """ModuleIO::write_wfc_pw for one k-point (write_wfc_pw.cpp:86-218): ``.dat`` records
    [int bytes][payload][int bytes] -- header (4 int + 7 double = 72 B), G (9 double),
    (ix, iy, iz) per plane wave, one complex record per band; ``.txt`` the same as text
    (setw(10), precision 6 header; setw(15) scientific values, four per line).
    ``psi``: (nbands, npw) complex."""
    import struct
    npw = len(ig2ixyz)
    mil = _miller(ig2ixyz, ny, nz)
    G = np.asarray(G, dtype=np.float64).reshape(9)
    psi = np.asarray(psi, dtype=np.complex128)[:nbands, :npw]
    if path.endswith(".dat"):
        with open(path, "wb") as f:
            f.write(struct.pack("<3i4d2i3di", 72, ik_global + 1, nkstot, *kvec_c, weight, npw, nbands, ecutwfc, lat0,
                                tpiba, 72))
            f.write(struct.pack("<i9di", 72, *G, 72))
            f.write(struct.pack("<i", npw * 12) + mil.astype("<i4").tobytes() + struct.pack("<i", npw * 12))
            for b in range(nbands):
                f.write(struct.pack("<i", npw * 16) + psi[b].astype("<c16").tobytes() + struct.pack("<i", npw * 16))
        return
    w = lambda x: "%10s" % _cfmt(x)   # noqa: E731
    lines = ["".join("%10s" % h for h in ("Kpoint", "nKpoint", "kv.x", "kv.y", "kv.z", "weight", "ngtot", "nband",
                                            "ecut", "lat0", "2pi/lat0")),
             "".join(w(v) for v in (ik_global + 1, nkstot, *kvec_c, weight, npw, nbands, ecutwfc, lat0, tpiba)),
             "", "<Reciprocal Lattice Vector>"]
    lines += ["".join(w(G[3 * r + c]) for c in range(3)) for r in range(3)]
    lines += ["<Reciprocal Lattice Vector>", "", "<G vectors>"]
    lines += ["".join("%10d" % v for v in m) for m in mil]
    lines += ["<G vectors>", ""]
    out = "\n".join(lines) + "\n"
    for b in range(nbands):
        out += "\n< Band %d >\n" % (b + 1)
        vals = []
        for ig in range(npw):
            if ig % 4 == 0 and ig != 0:
                vals.append("\n")
            vals.append("%15.6e%15.6e" % (psi[b, ig].real, psi[b, ig].imag))
        out += "".join(vals) + "\n< Band %d >\n" % (b + 1)
    with open(path, "w") as f:
        f.write(out)


def _cfmt(x):
    """C++ iostream default float format at precision 6 (printf %g), ints as is."""
    if isinstance(x, (int, np.integer)):
        return "%d" % x
    return "%g" % x


def read_wfc_pw(path, ik, nkstot, nbands, kvec_c, lat0, tpiba, G, nx, ny, nz, ig2ixyz):
    """ModuleIO::read_wfc_pw for one k-point (read_wfc_pw.cpp:96-323, .dat only): header
    checks (exact comparisons), then wfc[b, i] = file[b, order[ig2ixyz[i]]] with order the
    file position of each box index (ix ny + iy) nz + iz.  Raises ValueError with the
    reference's message on a mismatch."""
    import struct
    npw = len(ig2ixyz)
    with open(path, "rb") as f:
        data = f.read()
    h = struct.unpack_from("<3i4d2i3di", data, 0)
    ik_in, nks_in, kv, npw_in, nb_in = h[1], h[2], h[3:6], h[7], h[8]
    lat0_in, tpiba_in = h[10], h[11]
    if ik_in != ik + 1 or nks_in != nkstot or npw_in != npw or nb_in != nbands:
        raise ValueError("ikstot_in != ikstot || nkstot_in != nkstot || npwtot_in != npwtot || "
                         "nbands_in != GlobalV::NBANDS")
    if tuple(kv) != tuple(float(x) for x in kvec_c):
        raise ValueError("k vector in file is not the same as the one in memory")
    if lat0_in != lat0 or tpiba_in != tpiba:
        raise ValueError("lat0_in != pw_wfc->lat0 || tpiba_in != pw_wfc->tpiba")
    off = 80
    g_in = struct.unpack_from("<i9di", data, off)[1:10]
    if tuple(g_in) != tuple(np.asarray(G, dtype=np.float64).reshape(9)):
        raise ValueError("G_in != G")
    off += 80 + 4
    mil = np.frombuffer(data, dtype="<i4", count=3 * npw, offset=off).reshape(npw, 3)
    off += 12 * npw + 4
    order = np.full(nx * ny * nz, -1, dtype=np.int64)
    order[(mil[:, 0].astype(np.int64) * ny + mil[:, 1]) * nz + mil[:, 2]] = np.arange(npw)
    perm = order[np.asarray(ig2ixyz, dtype=np.int64)]
    wfc = np.zeros((nbands, npw), dtype=np.complex128)
    for b in range(nbands):
        off += 4
        wfc[b] = np.frombuffer(data, dtype="<c16", count=npw, offset=off)[perm]
        off += 16 * npw + 4
    return wfc
