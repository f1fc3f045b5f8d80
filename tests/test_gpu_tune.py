"""The autotuner on the GPU, through the C ABI: a measured GEMM plan and FFT schedule give
the same products as the fixed rules (FP64 rounding bounds), the table is then used (hits),
and H|psi> of a Hamiltonian built while measuring matches the oracle."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


@pytest.fixture()
def tuner():
    from apw import pw
    pw.tune_clear()
    yield pw
    pw.tune_set_mode(pw.TUNE_FIXED)
    pw.tune_clear()


def _zgemm(pw, ctx, ta, m, n, k, A, lda, B, Cm):
    pw.gemm_op(ctx, ta, "N", m, n, k, 1.0, A, lda, B, k, 0.0, Cm, m)
    ctx.synchronize()
    return Cm.cpu().numpy().copy()


@pytest.mark.parametrize("ta,m,n,k", [("C", 96, 160, 20011), ("N", 20011, 48, 96)])
def test_tuned_gemm_matches_fixed(torch, tuner, ta, m, n, k):
    pw = tuner
    ctx = pw.Context(0)
    rng = np.random.default_rng(3)
    rows, cols = (k, m) if ta == "C" else (m, k)
    A = rng.standard_normal((cols, rows)) + 1j * rng.standard_normal((cols, rows))
    B = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    out = torch.zeros((n, m), dtype=torch.complex128, device="cuda")
    pw.tune_set_mode(pw.TUNE_FIXED)
    fixed = _zgemm(pw, ctx, ta, m, n, k, dA, rows, dB, out)
    pw.tune_set_mode(pw.TUNE_MEASURE)
    measured = _zgemm(pw, ctx, ta, m, n, k, dA, rows, dB, out)
    st = pw.tune_stats()
    assert st["tuned"] == 1 and st["entries"] == 1
    pw.tune_set_mode(pw.TUNE_CACHE)
    cached = _zgemm(pw, ctx, ta, m, n, k, dA, rows, dB, out)
    assert pw.tune_stats()["hits"] >= 1
    assert np.array_equal(cached, measured)      # same plan from the table
    mat = A.T
    opa = mat.conj().T if ta == "C" else mat
    want = (opa @ B.T).T
    bound = 8 * k * np.finfo(float).eps * (np.abs(opa) @ np.abs(B.T)).max()
    for got in (fixed, measured):
        assert np.abs(got - want).max() < bound
    (key, choice, us), = pw.tune_entries()
    print(key, "->", choice, f"{us:.1f} us")
    ctx.close()


def test_tuned_fft_hpsi_matches_oracle(torch, tuner):
    from apw import configs
    from oracle import pw_oracle
    pw = tuner
    s = configs.make("si8")
    ctx = pw.Context(0)
    pw.tune_set_mode(pw.TUNE_MEASURE)
    ham = pw.HamiltPW(ctx, s)
    keys = [k for k, _, _ in pw.tune_entries()]
    assert any(k.startswith("fft f64") for k in keys) and any(k.startswith("fft f32") for k in keys)
    rng = np.random.default_rng(5)
    nb = 5
    X = np.zeros((nb, s.npwx), complex)
    X[:, : s.npw] = rng.standard_normal((nb, s.npw)) + 1j * rng.standard_normal((nb, s.npw))
    out = ham.hPsi(torch.from_numpy(X).cuda(), nb)
    ctx.synchronize()
    got = out.cpu().numpy()[:, : s.npw]
    ref = pw_oracle.HamiltOracle(s).hpsi(X[:, : s.npw].reshape(-1), nb, s.npw).reshape(nb, s.npw)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
    # a second Hamiltonian of the same grid takes the recorded schedule (no new measurement)
    pw.tune_set_mode(pw.TUNE_CACHE)
    before = pw.tune_stats()
    ham2 = pw.HamiltPW(ctx, s)
    after = pw.tune_stats()
    assert after["tuned"] == before["tuned"] and after["hits"] >= before["hits"] + 2
    out2 = ham2.hPsi(torch.from_numpy(X).cuda(), nb)
    ctx.synchronize()
    assert np.array_equal(out2.cpu().numpy(), out.cpu().numpy())
    ham.close()
    ham2.close()
    ctx.close()


# every GEMM plan the autotuner can choose (tile/map variant 0..11, with and without a
# k split), forced through the table, against numpy: FP64 and FP32 three-product forms and
# the x32 form (which never takes the 128 x 64 tiles, variants 2, 3 and 8)
NVAR = 12


@pytest.mark.parametrize("prec", ["f64", "f32", "x32"])
@pytest.mark.parametrize("ta,m,n,k", [("N", 3001, 70, 100), ("C", 70, 90, 5003)])
def test_every_gemm_plan(torch, tuner, tmp_path, prec, ta, m, n, k):
    import json
    pw = tuner
    ctx = pw.Context(0)
    rng = np.random.default_rng(11)
    rows, cols = (k, m) if ta == "C" else (m, k)
    A = rng.standard_normal((cols, rows)) + 1j * rng.standard_normal((cols, rows))
    B = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    Cm = rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))
    dt = torch.complex64 if prec == "f32" else torch.complex128
    if prec == "f32":
        A, B, Cm = (x.astype(np.complex64).astype(np.complex128) for x in (A, B, Cm))
    dA, dB = torch.from_numpy(A).to(dt).cuda(), torch.from_numpy(B).to(dt).cuda()
    alpha, beta = 0.7 - 0.2j, 0.3 + 0.1j
    opa = A.T.conj().T if ta == "C" else A.T
    want = alpha * (opa @ B.T) + beta * Cm.T
    mag = np.abs(alpha) * (np.abs(opa) @ np.abs(B.T)) + np.abs(beta) * np.abs(Cm.T)
    eps = 6e-8 if prec in ("f32", "x32") else np.finfo(float).eps
    bound = 8 * np.sqrt(k) * eps * mag + 1e-13

    def run():
        out = torch.from_numpy(Cm).to(dt).cuda()
        pw.gemm_op(ctx, ta, "N", m, n, k, alpha, dA, rows, dB, k, beta, out, m, x32=(prec == "x32"))
        ctx.synchronize()
        return out.cpu().numpy().astype(np.complex128).T

    pw.tune_set_mode(pw.TUNE_MEASURE)
    run()
    (key, _, _), = pw.tune_entries()        # the table key of this shape
    path = str(tmp_path / "plan.json")
    tried = 0
    for v in range(NVAR):
        if prec == "x32" and v in (2, 3, 8):
            continue
        for occ in (0, 2):
            with open(path, "w") as f:
                f.write('{"version": 1, "entries": [\n' + json.dumps({"key": key, "choice": [v, occ], "us": 1.0}) + "\n]}\n")
            pw.tune_clear()
            assert pw.tune_load(path) == 1
            pw.tune_set_mode(pw.TUNE_CACHE)
            got = run()
            assert pw.tune_stats()["hits"] >= 1
            err = np.abs(got - want)
            assert (err <= bound).all(), (prec, ta, v, occ, (err / bound).max())
            tried += 1
    print(prec, ta, key, tried, "plans")
    ctx.close()
