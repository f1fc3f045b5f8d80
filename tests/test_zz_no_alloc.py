"""apw_zgemm_noalloc (split-K partials in the library's fixed device array, no allocation)
against apw_zgemm (partials in a workspace the library allocates), through the C ABI.

Both run the same plan, so the products agree bit for bit; two calls in a row, the second
smaller, show that the shared array's leftover contents from the first call do not leak into
the second.  A numpy product checks both against zgemm semantics (FP64 error bound).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


def _call(ctx, fn, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc):
    from apw import _lib
    al = (C.c_double * 2)(alpha.real, alpha.imag)
    be = (C.c_double * 2)(beta.real, beta.imag)
    ctx.pre()
    _lib.check(fn(ctx.handle, ta.encode(), tb.encode(), m, n, k, al, C.c_void_p(A.data_ptr()), lda,
                  C.c_void_p(B.data_ptr()), ldb, be, C.c_void_p(Cm.data_ptr()), ldc))
    ctx.synchronize()


def _rand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("ta,tb", [("C", "N"), ("N", "N")])
def test_zgemm_noalloc_matches_zgemm(torch, ta, tb):
    from apw import pw
    ctx = pw.Context(0)
    rng = np.random.default_rng(11)
    # first a large product (inner product: split over k into many partials; update: tall C),
    # then a smaller one on the same fixed array
    shapes = [(160, 96, 51627), (40, 24, 3001)] if ta == "C" else [(51627, 64, 160), (3001, 24, 40)]
    for m, n, k in shapes:
        print(f"zgemm_noalloc {ta}{tb} m={m} n={n} k={k}", flush=True)
        # column-major operands stored as row-major transposes: a (lda x cols) array is X.T
        a_rows, a_cols = (k, m) if ta == "C" else (m, k)
        A = _rand(rng, a_cols, a_rows)             # column-major a_rows x a_cols, lda = a_rows
        B = _rand(rng, n, k)                       # column-major k x n, ldb = k
        C0 = _rand(rng, n, m)                      # column-major m x n, ldc = m
        alpha, beta = 0.75 - 0.5j, 0.25 + 1.0j
        dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        c_ref = torch.from_numpy(C0).cuda()
        c_new = c_ref.clone()
        _call(ctx, ctx.lib.apw_zgemm, ta, tb, m, n, k, alpha, dA, a_rows, dB, k, beta, c_ref, m)
        _call(ctx, ctx.lib.apw_zgemm_noalloc, ta, tb, m, n, k, alpha, dA, a_rows, dB, k, beta, c_new, m)
        got_ref, got_new = c_ref.cpu().numpy(), c_new.cpu().numpy()
        assert np.array_equal(got_ref, got_new), np.abs(got_ref - got_new).max()
        mat_a = A.T                                # the a_rows x a_cols matrix
        op_a = mat_a.conj().T if ta == "C" else mat_a
        want = (alpha * (op_a @ B.T) + beta * C0.T).T
        bound = 8 * k * np.finfo(float).eps * (np.abs(op_a) @ np.abs(B.T)).max() * abs(alpha) + 1e-12
        assert np.abs(got_new - want).max() < bound
    ctx.close()
