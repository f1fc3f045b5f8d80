"""Pins the CPU oracle against the reference's own test vectors and against mathematics."""
import json
import os

import numpy as np
import pytest

from apw import configs
from oracle import pw_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _c(d):
    return np.array(d["re"]) + 1j * np.array(d["im"])


def test_ekinetic_golden():
    d = _load("ref_ekinetic_op.json")
    psi = _c(d["psi"])
    hpsi = np.zeros_like(psi)
    O.ekinetic_pw_op(d["band"], d["dim"], d["dim"], d["tpiba2"], np.array(d["gk2"]), hpsi, psi)
    assert np.abs(hpsi - _c(d["expected_hpsi"])).max() < 1e-6      # ekinetic_op_test.cpp:73


def test_veff_golden():
    d = _load("ref_veff_op.json")
    out = _c(d["out"])
    O.veff_pw_op(d["size"], out, np.array(d["in"]))
    assert np.abs(out - _c(d["expected_out"])).max() < 6e-5         # veff_op_test.cpp:63


def test_nonlocal_golden():
    d = _load("ref_nonlocal_op.json")
    ps = np.zeros(len(d["expected_ps"]["re"]), dtype=np.complex128)
    s, iat = O.nonlocal_pw_op(d["l1"], d["l2"], d["l3"], 0, 0, d["spin"], d["nkb"], d["deeq_x"],
                              d["deeq_y"], d["deeq_z"], np.array(d["deeq"]), ps, _c(d["becp"]))
    assert np.abs(ps - _c(d["expected_ps"])).max() < 5e-6             # nonlocal_op_test.cpp:116
    assert (s, iat) == (d["expected_sum"], d["expected_iat"])


def test_nonlocal_planted_transpose_fails():
    """A transposed D block (deeq[k][x] instead of deeq[x][k]) must not pass the golden check."""
    d = _load("ref_nonlocal_op.json")
    deeq = np.array(d["deeq"]).reshape(d["deeq_x"], d["deeq_y"], d["deeq_z"])
    # make D asymmetric so a transpose is visible
    deeq = deeq + np.triu(np.ones((4, 4)), 1)[None]
    becp = _c(d["becp"])
    a = np.zeros(64, complex)
    b = np.zeros(64, complex)
    O.nonlocal_pw_op(2, 8, 4, 0, 0, 0, 8, 2, 4, 4, deeq.reshape(-1), a, becp)
    O.nonlocal_pw_op(2, 8, 4, 0, 0, 0, 8, 2, 4, 4, np.transpose(deeq, (0, 2, 1)).reshape(-1), b, becp)
    assert np.abs(a - b).max() > 1e-3


def test_pw_op_golden():
    d = _load("ref_pw_op.json")
    box = np.zeros(d["nxyz"], dtype=np.complex128)
    O.set_3d_fft_box_op(d["npwk"], d["box_index"], _c(d["in_1"]), box)
    assert np.abs(box - _c(d["out_1"])).max() < 1e-12
    out = np.zeros(d["nxyz"], dtype=np.complex128)
    O.set_recip_to_real_output_op(d["nxyz"], False, 1.0, _c(d["in_2"]), out)
    assert np.abs(out - _c(d["out_2"])).max() < 1e-12
    o3 = _c(d["out_3_init"])
    O.set_real_to_recip_output_op(d["npwk"], d["nxyz"], True, 1.0, d["box_index"], _c(d["in_3"]), o3)
    assert np.abs(o3 - _c(d["out_3"])).max() < 5e-6


def test_fft_chain_against_reference_fixtures():
    """The pw_op / veff_op fixtures are one band's local H|psi> on a 12^3 box:
    box(in_1) --backward--> in_2 (= veff_op 'out'); veff*psi --forward--> in_3."""
    p = _load("ref_pw_op.json")
    v = _load("ref_veff_op.json")
    r = O.fft3d_backward(_c(p["out_1"]), 12, 12, 12)
    assert np.abs(r - _c(p["in_2"])).max() < 2e-5          # fixtures carry 6 significant digits
    g = O.fft3d_forward(_c(v["expected_out"]), 12, 12, 12)
    assert np.abs(g - _c(p["in_3"])).max() / np.abs(_c(p["in_3"])).max() < 2e-6
    # a planted normalisation error (numpy's 1/N inverse) must fail; the fixture box is
    # inversion-symmetric, so a sign error of the exponent would not be visible here
    # (test_fft_matches_bruteforce pins the sign)
    bad = np.fft.ifftn(_c(p["out_1"]).reshape(12, 12, 12)).reshape(-1)
    assert np.abs(bad - _c(p["in_2"])).max() > 1e-2


@pytest.mark.parametrize("n", [2, 3, 5, 12, 18, 45])
def test_fft_matches_bruteforce(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n, n, 3)) + 1j * rng.standard_normal((2, n, n, 3))
    x = x[0]
    ref = x.copy()
    for ax in range(3):
        ref = np.moveaxis(O.dft_bruteforce(np.moveaxis(ref, ax, -1), +1), -1, ax)
    got = O.fft3d_backward(x.reshape(-1), n, n, 3).reshape(n, n, 3)
    assert np.abs(got - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())


def test_hpsi_matches_dense_tiny():
    s = configs.make("tiny")
    h = O.HamiltOracle(s)
    H = h.dense()
    assert np.abs(H - H.conj().T).max() < 1e-14
    rng = np.random.default_rng(1)
    X = rng.standard_normal((4, s.npw)) + 1j * rng.standard_normal((4, s.npw))
    hp = h.hpsi(X.reshape(-1), 4, s.npw).reshape(4, s.npw)
    assert np.abs(hp - (H @ X.T).T).max() < 1e-12 * np.abs(hp).max()


def test_david_oracle_tiny_matches_dense():
    s = configs.make("tiny")
    w = np.linalg.eigvalsh(O.HamiltOracle(s).dense())
    eig, psi, it, dav = O.david_solve(s, thr=1e-9)
    assert np.abs(eig - w[: s.nband]).max() < 1e-9
    # eigenvectors orthonormal and residuals small
    P = psi[:, : s.npw]
    assert np.abs(P.conj() @ P.T - np.eye(s.nband)).max() < 1e-10


@pytest.mark.slow
def test_david_oracle_si8_matches_dense():
    s = configs.make("si8")
    w = np.linalg.eigvalsh(O.HamiltOracle(s).dense())
    eig, psi, it, dav = O.david_solve(s, thr=1e-9)
    assert np.abs(eig - w[: s.nband]).max() < 1e-9


@pytest.mark.parametrize("n", [2, 7, 48])
def test_householder_matches_lapack_zhetrd(n):
    """The oracle's reduction reproduces LAPACK zhetrd('L') exactly in convention: the same
    reflectors v_k (v_k[k+1] = 1, rest in the strict lower part), the same tau_k, and
    Q^H A Q = tridiag(d, e) with zhetrd's real e (beta = -sign(Re alpha) ||x||)."""
    import scipy.linalg.lapack as L
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    V, tau = O.householder_tridiag_lower(A)
    c, d, e, tau_l, info = L.zhetrd(A, lower=1)
    assert info == 0
    assert np.abs(tau[: n - 1] - tau_l[: n - 1]).max() < 1e-13
    for k in range(n - 2):
        assert np.abs(V[k + 2:, k] - c[k + 2:, k]).max() < 1e-12
    Q = np.eye(n, dtype=complex)
    for k in range(n - 1):
        v = V[:, k]
        Q = Q @ (np.eye(n) - tau[k] * np.outer(v, v.conj()))
    T = Q.conj().T @ A @ Q
    assert np.abs(np.diag(T).real - d).max() < 1e-12 * n
    assert np.abs(np.diag(T, -1) - e).max() < 1e-12 * n


def test_dnevx_phase_convention():
    """dnevx's eigenvectors are Q z with z = Q^H v real and its first largest-magnitude
    component positive (LAPACK zheevx's inverse-iteration path)."""
    rng = np.random.default_rng(3)
    n, nb = 30, 12
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A + A.conj().T
    w, v = O.dnevx(n, np.tril(A), nb)
    assert np.abs(w - np.linalg.eigvalsh(A)[:nb]).max() < 1e-12
    V, tau = O.householder_tridiag_lower(A)
    z = O._apply_qh(V, tau, v)
    assert np.abs(z.imag).max() < 1e-12
    for j in range(nb):
        i = int(np.argmax(np.abs(z[:, j])))
        assert z[i, j].real > 0


def test_oracle_eig_fixture_is_current():
    """The committed multi-k / uneven-cell eigenvalue fixture matches the oracle as it is
    now (cheapest case recomputed; the GPU parity tests compare against the fixture)."""
    ref = _load("oracle_eigs.json")["cell3:0"]
    s = configs.make("cell3")
    assert (s.npw, s.nband, [s.nx, s.ny, s.nz]) == (ref["npw"], ref["nband"], ref["grid"])
    eig, _, iters, _ = O.david_solve(s, thr=1e-9)
    assert iters == ref["iterations"]
    assert np.abs(eig - np.array(ref["eig"])).max() < 1e-12
