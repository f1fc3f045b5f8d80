"""GPU parity tests of projector generation (getvnl): the three device functors
(cal_ylm_real_op, cal_sk_op, cal_vnl_op) and the fused apw_getvnl against the
reference's golden vectors and the oracle, and the generated projectors driving H|psi>."""
import json
import os

import numpy as np
import pytest

from apw import configs
from oracle import pw_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


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


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _c(d):
    return np.array(d["re"]) + 1j * np.array(d["im"])


def _dev(torch, a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype=dtype)))
    return t.cuda()


def _host(ctx, t):
    ctx.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("lmax", [0, 1, 2, 3, 4, 6])
def test_ylm_vs_oracle(torch, ctx, lmax):
    from apw import pw
    rng = np.random.default_rng(lmax)
    g = rng.standard_normal((300, 3))
    g[0] = 0.0                      # the reference's g = 0 convention
    g[1] = [0.0, 0.5, 0.2]          # x = 0 plane
    g[2] = [1e-12, -0.5, 0.2]
    ng = len(g)
    y = torch.zeros(((lmax + 1) ** 2, ng), dtype=torch.float64, device="cuda")
    p = torch.zeros(((lmax + 1) ** 2, ng), dtype=torch.float64, device="cuda")
    pw.cal_ylm_real_op(ctx, ng, lmax, _dev(torch, g), p, y)
    assert np.abs(_host(ctx, y) - O.cal_ylm_real(ng, lmax, g)).max() < 1e-12


def test_ylm_golden(torch, ctx):
    from apw import pw
    d = _load("ref_math_op.json")
    y = torch.zeros(len(d["expected_ylm"]), dtype=torch.float64, device="cuda")
    pw.cal_ylm_real_op(ctx, d["ng"], d["lmax"], _dev(torch, d["g"], np.float64), None, y)
    assert np.abs(_host(ctx, y) - np.array(d["expected_ylm"])).max() < 1e-6     # math_op_test.cpp:332 (6e-5)


def test_sk_golden(torch, ctx):
    from apw import pw
    w = _load("ref_wf_op.json")
    sk = torch.zeros(len(w["expected_sk"]["re"]), dtype=torch.complex128, device="cuda")
    pw.cal_sk_op(ctx, w["ik"], w["ntype"], w["nx"], w["ny"], w["nz"], w["rho_nx"], w["rho_ny"], w["rho_nz"], w["npw"],
                 w["npwx"], w["fftny"], w["eigts1_nc"], w["eigts2_nc"], w["eigts3_nc"],
                 _dev(torch, w["atom_na"], np.int32), _dev(torch, w["igl2isz"], np.int32),
                 _dev(torch, w["is2fftixy"], np.int32), w["TWO_PI"], _dev(torch, w["kvec_c"], np.float64),
                 _dev(torch, w["atom_tau"], np.float64), _dev(torch, _c(w["eigts1"])), _dev(torch, _c(w["eigts2"])),
                 _dev(torch, _c(w["eigts3"])), sk)
    assert np.abs(_host(ctx, sk) - _c(w["expected_sk"])).max() < 1e-12        # wf_op_test.cpp:465 (6e-5)


def test_cal_vnl_golden(torch, ctx):
    from apw import pw
    v = _load("ref_vnl_op.json")
    vkb = torch.zeros(len(v["expected_vkb"]["re"]), dtype=torch.complex128, device="cuda")
    vkb1 = torch.zeros(v["nhm"] * v["npw"], dtype=torch.float64, device="cuda")
    pw.cal_vnl_op(ctx, v["ntype"], v["npw"], v["npwx"], v["nhm"], v["tab_2"], v["tab_3"],
                  _dev(torch, v["atom_na"], np.int32), _dev(torch, v["atom_nb"], np.int32),
                  _dev(torch, v["atom_nh"], np.int32), v["DQ"], v["tpiba"], -1j, _dev(torch, v["gk"], np.float64),
                  _dev(torch, v["ylm"], np.float64), _dev(torch, v["indv"], np.float64),
                  _dev(torch, v["nhtol"], np.float64), _dev(torch, v["nhtolm"], np.float64),
                  _dev(torch, v["tab"], np.float64), vkb1, _dev(torch, _c(v["sk"])), vkb)
    assert np.abs(_host(ctx, vkb) - _c(v["expected_vkb"])).max() < 1e-10      # vnl_op_test.cpp:4108 (6e-5)


def _getvnl(torch, ctx, t, vkb):
    from apw import pw
    pw.getvnl(ctx, t["ntype"], t["npw"], t["npwx"], t["nhm"], t["tab_2"], t["tab_3"], t["atom_na"], t["atom_nb"],
              t["atom_nh"], t["dq"], t["tpiba"], _dev(torch, t["gk"], np.float64), t["tau"], t["indv"], t["nhtol"],
              t["nhtolm"], _dev(torch, t["tab"], np.float64), vkb)


def test_getvnl_golden(torch, ctx):
    """Fused form on the reference's 2-atom inputs: Ylm and sk formed from k+G and tau."""
    v = _load("ref_vnl_op.json")
    w = _load("ref_wf_op.json")
    t = {"ntype": v["ntype"], "npw": v["npw"], "npwx": v["npwx"], "nhm": v["nhm"], "tab_2": v["tab_2"],
         "tab_3": v["tab_3"], "atom_na": v["atom_na"], "atom_nb": v["atom_nb"], "atom_nh": v["atom_nh"],
         "dq": v["DQ"], "tpiba": v["tpiba"], "gk": v["gk"], "tau": w["atom_tau"], "indv": v["indv"],
         "nhtol": v["nhtol"], "nhtolm": v["nhtolm"], "tab": v["tab"]}
    vkb = torch.zeros(len(v["expected_vkb"]["re"]), dtype=torch.complex128, device="cuda")
    _getvnl(torch, ctx, t, vkb)
    assert np.abs(_host(ctx, vkb) - _c(v["expected_vkb"])).max() < 1e-10


def test_getvnl_table_bound_is_reported(torch, ctx):
    """|k+G| beyond the radial table is an error, not an out-of-bounds read."""
    from apw import pw
    from apw._lib import ApwError
    v = _load("ref_vnl_op.json")
    w = _load("ref_wf_op.json")
    t = {"ntype": 1, "npw": v["npw"], "npwx": v["npwx"], "nhm": v["nhm"], "tab_2": v["tab_2"], "tab_3": 20,
         "atom_na": v["atom_na"], "atom_nb": v["atom_nb"], "atom_nh": v["atom_nh"], "dq": v["DQ"],
         "tpiba": v["tpiba"], "gk": v["gk"], "tau": w["atom_tau"], "indv": v["indv"], "nhtol": v["nhtol"],
         "nhtolm": v["nhtolm"], "tab": np.array(v["tab"]).reshape(v["tab_2"], v["tab_3"])[:, :20].reshape(-1)}
    vkb = torch.zeros(len(v["expected_vkb"]["re"]), dtype=torch.complex128, device="cuda")
    with pytest.raises(ApwError, match="interpolation table"):
        _getvnl(torch, ctx, t, vkb)


@pytest.mark.parametrize("name", ["si8", "kpts16:37"])
def test_getvnl_synthetic_vs_oracle(torch, ctx, name):
    """Synthetic cells (general k: complex structure factors): GPU getvnl vs the oracle's,
    and vs the analytically built projectors to the interpolation error."""
    from apw import synth
    nm, _, ik = name.partition(":")
    s = configs.make(nm, ik=int(ik)) if ik else configs.make(nm)
    t = synth.projector_tables(s)
    nkb = s.nkb
    vkb = torch.zeros((nkb, t["npwx"]), dtype=torch.complex128, device="cuda")
    _getvnl(torch, ctx, t, vkb)
    got = _host(ctx, vkb)
    sign = np.tile(t["sign"], s.nat)[:, None]
    scale = np.abs(s.vkb).max()
    assert np.abs(got * sign - s.vkb).max() < 1e-7 * scale
    rows = np.arange(nkb) if nkb <= 64 else np.r_[0:8, nkb - 8:nkb]
    ref = O.getvnl(t["ntype"], t["npw"], t["npwx"], t["nhm"], t["tab_2"], t["tab_3"], t["atom_na"], t["atom_nb"],
                   t["atom_nh"], t["dq"], t["tpiba"], t["gk"], t["tau"], t["indv"], t["nhtol"], t["nhtolm"],
                   t["tab"]).reshape(nkb, t["npwx"])
    assert np.abs(got[rows, : t["npw"]] - ref[rows, : t["npw"]]).max() < 1e-13 * scale


def test_getvnl_feeds_hpsi(torch, ctx):
    """H|psi> with device-generated projectors (apw_ham_set_vkb) equals H|psi> with the
    uploaded ones to the interpolation error (signs absorbed: |beta><beta| is even)."""
    from apw import pw, synth
    s = configs.make("si8")
    t = synth.projector_tables(s)
    vkb = torch.zeros((s.nkb, t["npwx"]), dtype=torch.complex128, device="cuda")
    _getvnl(torch, ctx, t, vkb)
    rng = np.random.default_rng(4)
    X = np.zeros((6, s.npwx), complex)
    X[:, : s.npw] = rng.standard_normal((6, s.npw)) + 1j * rng.standard_normal((6, s.npw))
    ham = pw.HamiltPW(ctx, s)
    a = _host(ctx, ham.hPsi(_dev(torch, X), 6))
    ham.set_vkb(vkb, t["npwx"])
    b = _host(ctx, ham.hPsi(_dev(torch, X), 6))
    ham.close()
    assert np.abs(a - b).max() < 1e-7 * np.abs(a).max()
    assert np.abs(a - b).max() > 0.0        # the new projectors are in use
