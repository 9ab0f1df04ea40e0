"""Writes tests/golden/jw_tridiag_513.json from the reference's embedded data.

benchmark_scripts/vkFFT_scripts/src/sample_1_benchmark_VkFFT_double.cpp:49-2106
holds four 513-entry arrays a0, b0, c0, x0 (sub-, main-, super-diagonal and
right-hand side of a diagonally dominant tridiagonal system; a0[0] and
c0[512] lie outside the matrix).  They are inputs only -- the reference
stores no solution -- so the fixture is data, checked through its residual.
Run in a container where the reference is mounted:
    python tests/golden/extract_reference_fixture.py [/root/reference]
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "benchmark_scripts", "vkFFT_scripts", "src", "sample_1_benchmark_VkFFT_double.cpp")
text = open(src).read()
out = {}
for name, key in (("a0", "a"), ("b0", "b"), ("c0", "c"), ("x0", "x")):
    m = re.search(r"double\s+" + name + r"\s*\[\s*513\s*\]\s*=\s*\{(.*?)\}", text, re.S)
    vals = [float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
    assert len(vals) == 513, (name, len(vals))
    out[key] = vals
out["source"] = "benchmark_scripts/vkFFT_scripts/src/sample_1_benchmark_VkFFT_double.cpp:49-2106"
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "jw_tridiag_513.json")
with open(dst, "w") as f:
    json.dump(out, f)
print("wrote", dst)
