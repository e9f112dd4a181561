"""Writes tests/golden/reference_tables.json: the constant tables of the
reference's detection path, read from its sources as data.

  * kernel_dx: the 9x9 derivative kernel, float32 values
    (src/cctag/filter/cvRecode.cpp:74-83)
  * lut_thin1 / lut_thin2: the 512-entry thinning tables
    (src/cctag/filter/thinning.cpp:12-13)
  * bank3 / bank4: radius-ratio banks (src/cctag/CCTagMarkersBank.cpp:120-282)

Run once where the reference tree is available (the tests only read the JSON):
    python tests/golden/make_reference_tables.py /root/reference
The committed JSON is what the tests read.
"""
import json
import os
import re
import sys

import numpy as np


def floats(text):
    return [float(np.float32(v)) for v in re.findall(r"(-?\d+\.\d+)f", text)]


def block(src, start, end="};"):
    i = src.index(start)
    return src[i:src.index(end, i)]


def main(root):
    cc = os.path.join(root, "src", "cctag")
    rec = open(os.path.join(cc, "filter", "cvRecode.cpp")).read()
    kblock = block(rec, "kerneldX = (cv::Mat_<float>(9,9) <<", ");")
    kvals = re.findall(r"(-?\d+\.\d+f|0\.f)", kblock)
    kernel = [0.0 if v == "0.f" else float(np.float32(v[:-1])) for v in kvals]
    assert len(kernel) == 81, len(kernel)

    th = open(os.path.join(cc, "filter", "thinning.cpp")).read()
    luts = []
    for name in ("lutthin1", "lutthin2"):
        b = block(th, "static const int %s[512]" % name, "};")
        vals = [int(v) for v in re.findall(r"\b(\d+)\b", b.split("=", 1)[1])]
        assert len(vals) == 512, (name, len(vals))
        luts.append(vals)

    bk = open(os.path.join(cc, "CCTagMarkersBank.cpp")).read()
    b3 = floats(block(bk, "idThreeCrowns[32][5]"))
    b4 = floats(block(bk, "idFourCrowns[128][7]"))
    assert len(b3) == 160 and len(b4) == 896

    out = {
        "source": "alicevision/CCTag v1.0.3",
        "kernel_dx": kernel,
        "lut_thin1": luts[0],
        "lut_thin2": luts[1],
        "bank3": [b3[i:i + 5] for i in range(0, 160, 5)],
        "bank4": [b4[i:i + 7] for i in range(0, 896, 7)],
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")
    with open(dst, "w") as f:
        json.dump(out, f)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
