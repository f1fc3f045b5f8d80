"""GPU tests of the wavefunction files (write_wfc_pw / read_wfc_pw through the C ABI):
the reference's own test file read into HBM (its checked values, bit for bit), files
written from HBM byte-identical to the reference's and to the oracle's, the reference's
consistency errors, and a round trip on a general k-point."""
import json
import os
import struct

import numpy as np
import pytest

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


@pytest.fixture(scope="module")
def case():
    from apw.basis import Lattice, PWBasisK
    with open(os.path.join(GOLD, "ref_read_wfc_pw.json")) as f:
        d = json.load(f)
    lat = Lattice(d["lat0"], np.array(d["latvec"]).reshape(3, 3))
    b = PWBasisK(lat, d["ecutwfc"], *d["grid"], kvec_d=np.zeros((d["nkstot"], 3)))
    return d, lat, b


def _read(ctx, torch, d, lat, b, path, nbands=None, kvec=(0.0, 0.0, 0.0), lat0=None, G=None, ld=None):
    from apw import pw
    npw = int(b.npwk[0])
    nb = d["nbands"] if nbands is None else nbands
    ld = ld or npw
    wfc = torch.zeros((nb, ld), dtype=torch.complex128, device="cuda")
    pw.read_wfc_pw(ctx, path, 0, d["nkstot"], nb, list(kvec), d["lat0"] if lat0 is None else lat0, lat.tpiba,
                   lat.G if G is None else G, *d["grid"], npw, b.ig2ixyz_k[0, :npw], wfc, ld)
    ctx.synchronize()
    return wfc


def test_read_golden_file(torch, ctx, case):
    d, lat, b = case
    npw = int(b.npwk[0])
    path = os.path.join(GOLD, d["file"])
    w = _read(ctx, torch, d, lat, b, path, ld=npw + 5).cpu().numpy()[:, :npw]
    for e in d["expect"]:
        v = w[e["band"], npw - 1 if e["pw"] == "last" else e["pw"]]
        assert (v.real if e["part"] == "real" else v.imag) == e["value"]     # read_wfc_pw_test.cpp:60-65
    ref = O.read_wfc_pw(path, 0, d["nkstot"], d["nbands"], [0.0, 0.0, 0.0], d["lat0"], lat.tpiba, lat.G, *d["grid"],
                        b.ig2ixyz_k[0, :npw])
    assert np.array_equal(w, ref)


@pytest.mark.parametrize("ext", ["dat", "txt"])
def test_write_matches_reference_and_oracle(torch, ctx, case, tmp_path, ext):
    from apw import pw
    d, lat, b = case
    npw = int(b.npwk[0])
    path = os.path.join(GOLD, d["file"])
    ld = npw + 3
    w = _read(ctx, torch, d, lat, b, path, ld=ld)
    hdr = struct.unpack_from("<3i4d2i3di", open(path, "rb").read(), 0)
    out = str(tmp_path / f"g.{ext}")
    pw.write_wfc_pw(ctx, out, 0, d["nkstot"], [0.0, 0.0, 0.0], hdr[6], d["nbands"], hdr[9], d["lat0"], lat.tpiba,
                    lat.G, *d["grid"], npw, b.ig2ixyz_k[0, :npw], w, ld)
    ref = str(tmp_path / f"o.{ext}")
    O.write_wfc_pw(ref, 0, d["nkstot"], [0.0, 0.0, 0.0], hdr[6], d["nbands"], hdr[9], d["lat0"], lat.tpiba, lat.G,
                   *d["grid"], b.ig2ixyz_k[0, :npw], w.cpu().numpy()[:, :npw])
    assert open(out, "rb").read() == open(ref, "rb").read()
    if ext == "dat":
        assert open(out, "rb").read() == open(path, "rb").read()


@pytest.mark.parametrize("what", ["nbands", "kvec", "lat0", "G"])
def test_read_consistency_errors(torch, ctx, case, what):
    from apw._lib import ApwError
    d, lat, b = case
    path = os.path.join(GOLD, d["file"])
    kw = {"nbands": {"nbands": 4}, "kvec": {"kvec": (-1.0, 1.0, -1.0)}, "lat0": {"lat0": 5.0},
          "G": {"G": lat.G * 1.0101}}[what]
    msg = {"nbands": "nbands_in", "kvec": "k vector in file", "lat0": "lat0_in", "G": "G_in != G"}[what]
    with pytest.raises(ApwError, match=msg):
        _read(ctx, torch, d, lat, b, path, **kw)


def test_read_missing_and_unknown(torch, ctx, case, tmp_path):
    from apw._lib import ApwError
    d, lat, b = case
    with pytest.raises(ApwError, match="can't open file"):
        _read(ctx, torch, d, lat, b, str(tmp_path / "notfound.dat"))
    with pytest.raises(ApwError, match="unknown file type"):
        _read(ctx, torch, d, lat, b, str(tmp_path / "notfound"))


def test_round_trip_general_k(torch, ctx, tmp_path):
    """A general k-point of the 4x4x4 mesh (uneven sphere): write from HBM, read back into
    a different leading dimension, identical bits."""
    from apw import configs, pw
    s = configs.make("kpts16", ik=37)
    npw = s.npw
    kc = s.lat.gcar(s.basis.kvec_d[s.ik])
    psi = torch.from_numpy(np.ascontiguousarray(s.psi0)).cuda()
    out = str(tmp_path / "k.dat")
    pw.write_wfc_pw(ctx, out, s.ik, s.basis.nks, kc, 1.0 / s.basis.nks, s.nband, s.meta["ecutwfc"], s.lat.lat0,
                    s.lat.tpiba, s.lat.G, s.nx, s.ny, s.nz, npw, s.ig2ixyz, psi, s.npwx)
    back = torch.zeros((s.nband, npw + 7), dtype=torch.complex128, device="cuda")
    pw.read_wfc_pw(ctx, out, s.ik, s.basis.nks, s.nband, kc, s.lat.lat0, s.lat.tpiba, s.lat.G, s.nx, s.ny, s.nz,
                   npw, s.ig2ixyz, back, npw + 7)
    ctx.synchronize()
    assert np.array_equal(back.cpu().numpy()[:, :npw], s.psi0[:, :npw])
