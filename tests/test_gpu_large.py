"""GPU parity at the two large single-cell workloads, through the C ABI:

* si64  -- 64-atom cell, 51,627 plane waves, 96^3 grid, 160 bands, 512 projectors;
* si216 -- 216-atom cell, 228,487 plane waves, 160^3 grid, 480 bands, 1728 projectors
  (subspace up to 1920).

Checks, per workload:
* H|psi> of a few bands against the oracle's operator chain, 1e-12 relative;
* the FP64 Davidson eigenvalues against the compiled reference DiagoDavid's own
  (tests/golden/ref_eigs.json, written by tests/golden/make_ref_eigs.py) -- the ``detsign``
  build follows the same trajectory as the GPU solver (1e-10 Ry), the plain build differs
  only in zheevx's eigenvector phases on its all-eigenpair calls;
* the FP32-mixed eigenvalues within 1e-9 Ry of the FP64 reference values;
* invariants of the returned vectors, computed independently of the solver: psi^H psi = I,
  diag(psi^H H psi) = eigenvalues, and the residual norms ||H psi - e psi|| match the solver's.
"""
import json
import os

import numpy as np
import pytest

from apw import configs
from oracle import pw_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _ref(key):
    path = os.path.join(GOLD, "ref_eigs.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


@pytest.fixture(scope="module")
def ctx(torch):
    from apw import pw
    c = pw.Context(0)
    yield c
    c.close()


def _solve(torch, ctx, s, ham, precision):
    from apw import pw
    dav = pw.DiagoDavid(ctx, s.precondition, s.nband, s.npw, s.david_ndim)
    psi = torch.from_numpy(np.ascontiguousarray(s.psi0)).cuda()
    eig = np.zeros(s.nband)
    it = dav.diag(ham, None, s.npwx, psi, eig, 1e-9, 500, 5, 0, precision=precision)
    st = dict(dav.stats)
    dav.close()
    ctx.synchronize()
    return psi, eig, it, st


def _invariants(torch, ctx, s, ham, psi, eig, st):
    """psi^H psi = I, Rayleigh quotients = eig, residual norms from an independent H|psi>."""
    hp = ham.hPsi(psi, s.nband)
    ctx.synchronize()
    P = psi[:, : s.npw]
    HP = hp[:, : s.npw]
    ovl = (P.conj() @ P.T).cpu().numpy()
    rq = torch.sum(P.conj() * HP, dim=1).real.cpu().numpy()
    res = torch.linalg.vector_norm(HP - torch.from_numpy(eig).cuda()[:, None] * P, dim=1).cpu().numpy()
    return np.abs(ovl - np.eye(s.nband)).max(), np.abs(rq - eig).max(), res.max()


@pytest.mark.parametrize("name", ["si64", "si216"])
def test_large_hpsi_and_david(torch, ctx, name):
    from apw import pw
    print(f"building {name}", flush=True)
    s = configs.make(name)
    ham = pw.HamiltPW(ctx, s)
    # ---- H|psi> of 4 random bands against the oracle's operator chain
    nb = 4
    rng = np.random.default_rng(7)
    X = np.zeros((nb, s.npwx), complex)
    X[:, : s.npw] = rng.standard_normal((nb, s.npw)) + 1j * rng.standard_normal((nb, s.npw))
    out = ham.hPsi(torch.from_numpy(X).cuda(), nb)
    ctx.synchronize()
    got = out.cpu().numpy()[:, : s.npw]
    ref = O.HamiltOracle(s).hpsi(X[:, : s.npw].reshape(-1), nb, s.npw).reshape(nb, s.npw)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(name, "H|psi> rel err vs oracle", err, flush=True)
    assert err < 1e-12
    del out, got, ref
    # ---- FP64 and FP32-mixed Davidson
    psi, e64, it64, st64 = _solve(torch, ctx, s, ham, "fp64")
    ovl, rq, res = _invariants(torch, ctx, s, ham, psi, e64, st64)
    print(name, "fp64", it64, "iterations; |psi^H psi - I|", ovl, "|rq - e|", rq, "max residual", res,
          "(solver:", st64["max_residual"], ")", flush=True)
    assert ovl < 1e-10
    assert rq < 1e-10
    assert abs(res - st64["max_residual"]) < 1e-8
    del psi
    psim, emx, itm, stm = _solve(torch, ctx, s, ham, "fp32_mixed")
    ovl_m, rq_m, res_m = _invariants(torch, ctx, s, ham, psim, emx, stm)
    dm = np.abs(emx - e64).max()
    print(name, "mixed", itm, "iterations (FP64 from", stm["fp64_switch_iter"], "); |mixed - fp64|", dm,
          "|psi^H psi - I|", ovl_m, "|rq - e|", rq_m, flush=True)
    assert dm < 1e-9
    assert ovl_m < 1e-10 and rq_m < 1e-10
    del psim
    # the mixed form with the pre-phase's subspace products on the FP32 matrix cores too
    psix, ex32, itx, stx = _solve(torch, ctx, s, ham, "fp32_mixed_x32")
    ovl_x, rq_x, res_x = _invariants(torch, ctx, s, ham, psix, ex32, stx)
    dx = np.abs(ex32 - e64).max()
    print(name, "mixed x32", itx, "iterations (FP64 from", stx["fp64_switch_iter"], "); |x32 - fp64|", dx,
          "|psi^H psi - I|", ovl_x, "|rq - e|", rq_x, flush=True)
    assert dx < 1e-9
    assert ovl_x < 1e-10 and rq_x < 1e-10
    del psix
    ham.close()
    ref = _ref(f"{name}:0")
    if not ref or not ("plain" in ref or "detsign" in ref):
        pytest.skip(f"no reference eigenvalues for {name} in tests/golden/ref_eigs.json")
    assert (ref["npw"], ref["nband"]) == (s.npw, s.nband)
    if "detsign" in ref:
        d_det = np.abs(e64 - np.array(ref["detsign"]["eig"])).max()
        print(name, "fp64 vs reference (same phases)", d_det, "iterations", it64, "vs",
              ref["detsign"]["iterations"])
        assert d_det < 1e-10
    if "plain" in ref:
        # the build as shipped takes implementation-dependent eigenvector phases on its
        # all-eigenpair zheevx calls, so its trajectory differs; both stop within one threshold
        d_plain = np.abs(e64 - np.array(ref["plain"]["eig"])).max()
        dm_plain = np.abs(emx - np.array(ref["plain"]["eig"])).max()
        dx_plain = np.abs(ex32 - np.array(ref["plain"]["eig"])).max()
        print(name, "fp64 vs reference as built", d_plain, "| mixed vs reference as built", dm_plain,
              "| mixed x32", dx_plain)
        assert d_plain < 1e-9
        assert dm_plain < 1e-9
        assert dx_plain < 1e-9
