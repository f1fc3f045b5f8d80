"""Write the reference's own kernel test vectors as JSON fixtures.

Reads the numeric initialisers (inputs and expected outputs) of the ABACUS unit
tests that pin the H|psi> kernels and stores them as data under tests/golden/:

* source/module_hamilt_pw/hamilt_pwdft/kernels/test/ekinetic_op_test.cpp
* source/module_hamilt_pw/hamilt_pwdft/kernels/test/veff_op_test.cpp
* source/module_hamilt_pw/hamilt_pwdft/kernels/test/nonlocal_op_test.cpp
* source/module_basis/module_pw/kernels/test/pw_op_test.cpp
* source/module_elecstate/kernels/test/elecstate_op_test.cpp
* source/module_hamilt_pw/hamilt_pwdft/kernels/test/vnl_op_test.cpp
* source/module_hamilt_pw/hamilt_pwdft/kernels/test/wf_op_test.cpp
* source/module_base/kernels/test/math_op_test.cpp
* source/module_io/test/read_wfc_pw_test.cpp (expected values) and the data file it
  reads, source/module_io/test/support/WAVEFUNC1.dat (copied as ref_WAVEFUNC1.dat)

Only numbers are copied (the fixtures are inputs and expected outputs); no code.
Run once in a container that has the reference tree:

    python tests/golden/extract_reference_vectors.py /root/reference
"""
from __future__ import annotations

import json
import os
import re
import sys

FILES = {
    "ekinetic_op": "source/module_hamilt_pw/hamilt_pwdft/kernels/test/ekinetic_op_test.cpp",
    "veff_op": "source/module_hamilt_pw/hamilt_pwdft/kernels/test/veff_op_test.cpp",
    "nonlocal_op": "source/module_hamilt_pw/hamilt_pwdft/kernels/test/nonlocal_op_test.cpp",
    "pw_op": "source/module_basis/module_pw/kernels/test/pw_op_test.cpp",
    "elecstate_op": "source/module_elecstate/kernels/test/elecstate_op_test.cpp",
    "vnl_op": "source/module_hamilt_pw/hamilt_pwdft/kernels/test/vnl_op_test.cpp",
    "wf_op": "source/module_hamilt_pw/hamilt_pwdft/kernels/test/wf_op_test.cpp",
    "math_op": "source/module_base/kernels/test/math_op_test.cpp",
}

_VEC = re.compile(r"std::vector<\s*(std::complex<double>|double|int)\s*>\s*(\w+)\s*=?\s*\{", re.S)
_SCALAR = re.compile(r"const\s+(int|double|bool)\s+(\w+)\s*=\s*([^;]+);")
_INT_DECL = re.compile(r"\bint\s+(\w+)\s*=\s*(-?\d+)\s*,\s*(\w+)\s*=\s*(-?\d+)\s*;")
_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
# "int a = 1, b = 2;" / "double x = 0.5, y = 1e-3;" (several declarators, literal values)
_MULTI = re.compile(r"(?<![(\w])(int|double)\s+((?:\w+\s*=\s*" + _NUM + r"\s*,\s*)*\w+\s*=\s*" + _NUM + r")\s*;")


def _braced(text: str, start: int) -> str:
    depth, i = 0, start
    while True:
        c = text[i]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                return text[start:i + 1]
        i += 1


def parse(path: str) -> dict:
    text = open(path).read()
    # drop the GPU-only sections' duplicates: keep the first definition of each name
    out: dict = {}
    for m in _VEC.finditer(text):
        kind, name = m.group(1), m.group(2)
        if name in out:
            continue
        body = _braced(text, m.end() - 1)
        if kind == "std::complex<double>":
            pairs = re.findall(r"\{\s*(" + _NUM + r")\s*,\s*(" + _NUM + r")\s*\}", body)
            out[name] = {"re": [float(a) for a, _ in pairs], "im": [float(b) for _, b in pairs]}
        else:
            nums = re.findall(_NUM, body[1:-1])
            out[name] = [int(x) for x in nums] if kind == "int" else [float(x) for x in nums]
    for m in _SCALAR.finditer(text):
        kind, name, val = m.group(1), m.group(2), m.group(3).strip()
        if name in out:
            continue
        if kind == "bool":
            out[name] = val == "true"
        elif kind == "int":
            try:
                out[name] = int(val)
            except ValueError:
                pass
        else:
            out[name] = float(val)
    for m in _MULTI.finditer(text):
        for decl in m.group(2).split(","):
            name, val = (x.strip() for x in decl.split("="))
            out.setdefault(name, int(val) if m.group(1) == "int" else float(val))
    for m in _INT_DECL.finditer(text):
        out.setdefault(m.group(1), int(m.group(2)))
        out.setdefault(m.group(3), int(m.group(4)))
    return out


# data files the reference's tests read, copied byte for byte
DATA = {"ref_WAVEFUNC1.dat": "source/module_io/test/support/WAVEFUNC1.dat"}
# expected values of read_wfc_pw_test.cpp's single-process case (EXPECT_DOUBLE_EQ lines)
WFC_TEST = "source/module_io/test/read_wfc_pw_test.cpp"
_EXPECT = re.compile(r"EXPECT_DOUBLE_EQ\(wfcatom\((\d+),\s*([^)]*?)\)\.(real|imag)\(\),\s*(" + _NUM + r")\)")


def wfc_expectations(path: str) -> list:
    text = open(path).read()
    block = text[text.index("NPROC_IN_POOL == 1"): text.index("NPROC_IN_POOL == 4")]
    out = []
    for m in _EXPECT.finditer(block):
        col = m.group(2).strip()
        out.append({"band": int(m.group(1)), "pw": "last" if "npwk" in col else int(col), "part": m.group(3),
                    "value": float(m.group(4))})
    return out


def main(root: str) -> None:
    here = os.path.dirname(os.path.abspath(__file__))
    for dst, rel in DATA.items():
        with open(os.path.join(root, rel), "rb") as f:
            blob = f.read()
        with open(os.path.join(here, dst), "wb") as f:
            f.write(blob)
        print(dst, len(blob), "bytes")
    with open(os.path.join(here, "ref_read_wfc_pw.json"), "w") as f:
        json.dump({"_source": WFC_TEST, "file": "ref_WAVEFUNC1.dat", "lat0": 5.3233,
                   "latvec": [-0.5, 0.0, 0.5, 0.0, 0.5, 0.5, -0.5, 0.5, 0.0], "grid": [12, 12, 12], "ecutwfc": 20.0,
                   "nkstot": 8, "nbands": 8, "expect": wfc_expectations(os.path.join(root, WFC_TEST))}, f, indent=1)
    for key, rel in FILES.items():
        data = parse(os.path.join(root, rel))
        data["_source"] = rel
        with open(os.path.join(here, f"ref_{key}.json"), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print(key, sorted(k for k in data if not k.startswith("_")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
