"""The autotuner's table and cache file, on the CPU (no GPU call): load, look up, save, and
the rejection of a file that is not a tuning cache."""
import os

import pytest

from apw import _lib

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libapw.so not built")


def test_cache_round_trip(tmp_path):
    from apw import pw
    pw.tune_clear()
    src = tmp_path / "in.json"
    src.write_text('{"version": 1, "entries": [\n'
                   '{"key": "gemm f64 CN m=160 n=640 k=52224 cu=256", "choice": [6, 3], "us": 812.50},\n'
                   '{"key": "fft f32 96x96x96 npw=51627 cu=256", "choice": [36, 0], "us": 2210.00}\n'
                   ']}\n')
    assert pw.tune_load(str(src)) == 2
    st = pw.tune_stats()
    assert st["entries"] == 2 and st["tuned"] == 0
    ents = dict((k, (c, us)) for k, c, us in pw.tune_entries())
    assert ents["gemm f64 CN m=160 n=640 k=52224 cu=256"] == ((6, 3), 812.5)
    assert ents["fft f32 96x96x96 npw=51627 cu=256"] == ((36, 0), 2210.0)
    out = tmp_path / "out.json"
    assert pw.tune_save(str(out)) == 2
    pw.tune_clear()
    assert pw.tune_stats()["entries"] == 0
    assert pw.tune_load(str(out)) == 2
    assert dict((k, (c, us)) for k, c, us in pw.tune_entries()) == ents
    import json
    assert len(json.loads(out.read_text())["entries"]) == 2   # the file is valid JSON
    pw.tune_clear()


def test_rejects_other_files(tmp_path):
    from apw import pw
    bad = tmp_path / "bad.json"
    bad.write_text('{"entries": []}\n')
    with pytest.raises(_lib.ApwError, match="not a tuning cache"):
        pw.tune_load(str(bad))
    with pytest.raises(_lib.ApwError, match="cannot open"):
        pw.tune_load(str(tmp_path / "missing.json"))
    with pytest.raises(_lib.ApwError, match="mode"):
        pw.tune_set_mode(3)
    assert pw.tune_stats()["entries"] == 0


def test_committed_cache_loads():
    from apw import pw
    if not os.path.exists(pw.TUNE_CACHE_FILE):
        pytest.skip("no committed tuning cache")
    pw.tune_clear()
    n = pw.tune_load()
    assert n > 0
    keys = [k for k, _, _ in pw.tune_entries()]
    assert any(k.startswith("gemm ") for k in keys) and any(k.startswith("fft ") for k in keys)
    pw.tune_clear()
