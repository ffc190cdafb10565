"""Write the golden fixtures under tests/golden/ from data the reference's own tests hold.

Run in the build container (the reference checkout is not present on GPU machines):

    python tests/golden/make_golden.py /root/reference

1. ``spgemm_cases.json`` -- the 51 dense (A, B) pairs of the reference's SpGEMM known-answer test
   (cpp_tests/test_main.cpp, ``testCases.push_back`` blocks; run with no arguments at :1548-1574). Only the numbers
   are taken (the array initialisers and the matrix shapes); the expected product is the dense product computed here
   with numpy, and the expected diagonal set is the set of nonzero diagonals of that product, which is what the
   reference test compares after pruning both sides (cpp_tests/test_checkResult.cpp:6-61).
2. ``spmv_known.json`` -- the ``spmv_2`` case (cpp_tests/test_main.cpp:1780-1846): the 8x8 matrix built from
   h = 1/sqrt(2), the x vector it sets, and the printed known answer from its comment (:1782-1794, six digits).
"""
from __future__ import annotations

import json
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _numbers(body: str, env: dict) -> list:
    out = []
    for tok in body.split(","):
        tok = tok.strip()
        if not tok:
            continue
        out.append(float(eval(tok, {"__builtins__": {}}, env)))  # tokens are literals or "-1 * h_val"
    return out


def parse_spgemm_cases(src: str):
    arrays = {}
    for m in re.finditer(r"valType\s+([ab]\d+(?:_real|_imag)?)\s*\[\s*([\d\s*]+)\]\s*=\s*\{(.*?)\};", src, re.S):
        name, size, body = m.group(1), int(eval(m.group(2), {"__builtins__": {}})), m.group(3)
        body = re.sub(r"//[^\n]*", "", body)
        vals = _numbers(body, {})
        if len(vals) > size:
            raise ValueError(f"{name}: {len(vals)} values for size {size}")
        arrays[name] = vals + [0.0] * (size - len(vals))  # C++ zero-fills a short initialiser
    shapes = {}
    for m in re.finditer(r"std::array<unsigned int,\s*2>\s+(shape[AB]\d+)\s*=\s*\{\s*(\d+)\s*,\s*(\d+)\s*\}", src):
        shapes[m.group(1)] = (int(m.group(2)), int(m.group(3)))
    mats = {}
    for m in re.finditer(r"std::shared_ptr<diaq::denseMatrix>\s+([AB]\d+)\s*=\s*std::make_shared<diaq::denseMatrix>"
                         r"\(\s*(\w+)\s*,\s*heap\((\w+)\)\s*(?:,\s*heap\((\w+)\)\s*)?\)", src, re.S):
        var, dim, re_name, im_name = m.groups()
        if dim.isdigit():
            shape = (int(dim), int(dim))
        else:
            shape = shapes[dim]
        n = shape[0] * shape[1]
        real = np.array(arrays[re_name][:n]).reshape(shape)
        imag = np.array(arrays[im_name][:n]).reshape(shape) if im_name else np.zeros(shape)
        mats[var] = real + 1j * imag
    cases = []
    for m in re.finditer(r"testCase(\d+)\.push_back\((\w+)\);\s*testCase\1\.push_back\((\w+)\);", src):
        idx, a, b = int(m.group(1)), m.group(2), m.group(3)
        cases.append((idx, mats[a], mats[b]))
    return cases


def nonzero_diagonals(P: np.ndarray):
    rows, cols = P.shape
    out = []
    for d in range(-(rows - 1), cols):
        if np.any(np.diagonal(P, offset=d) != 0):
            out.append(d)
    return out


def enc(M: np.ndarray):
    return {"shape": list(M.shape), "re": M.real.ravel().tolist(), "im": M.imag.ravel().tolist()}


def main(ref_root: str):
    src = open(os.path.join(ref_root, "cpp_tests", "test_main.cpp")).read()
    cases = parse_spgemm_cases(src)
    out = []
    for idx, A, B in cases:
        P = A @ B
        out.append({"case": idx, "A": enc(A), "B": enc(B), "product": enc(P),
                    "product_offsets_pruned": nonzero_diagonals(P),
                    "A_offsets": nonzero_diagonals(A), "B_offsets": nonzero_diagonals(B)})
    with open(os.path.join(HERE, "spgemm_cases.json"), "w") as f:
        json.dump({"source": "cpp_tests/test_main.cpp testCases 0..50 (inputs); products by numpy",
                   "cases": out}, f)
    print("spgemm cases:", len(out))

    h = 1.0 / math.sqrt(2.0)
    M = np.zeros((8, 8))
    for b in range(0, 8, 2):  # the 2x2 blocks [[h, h], [h, -h]] of spmv_2
        M[b, b], M[b, b + 1], M[b + 1, b], M[b + 1, b + 1] = h, h, h, -h
    x = np.array([0.279931, 0.279931, 0, 0, 0, 0, 0.126182, 0.126182]) \
        + 1j * np.array([0.085921, 0.085921, 0, 0, 0, 0, 0.631137, 0.631137])
    printed = np.array([0.395882, -5.55112e-17, 0, 0, 0, 0, 0.178448, -5.55112e-17]) \
        + 1j * np.array([0.121511, -2.08167e-17, 0, 0, 0, 0, 0.892563, 0])
    with open(os.path.join(HERE, "spmv_known.json"), "w") as f:
        json.dump({"source": "cpp_tests/test_main.cpp spmv_2 (:1780-1846), known answer printed at :1782-1794",
                   "A": enc(M + 0j), "x": {"re": x.real.tolist(), "im": x.imag.tolist()},
                   "y_printed": {"re": printed.real.tolist(), "im": printed.imag.tolist()},
                   "printed_digits": 6, "printed_offsets": [-1, 0, 1]}, f, indent=1)
    print("spmv known answer written")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
