"""Write the committed fixtures of tests/golden/.

* fo_192.txt -- the reference phantom (field_setup, sirt-mimo.cu:470-509)
  evaluated in numpy float32 (x = -0.1f + i * 0.001f, squares as x * x),
  written in the mimo-io.cuh:22-45 text layout with 9 significant digits.
  599 pixels at 0.03 and 228 at 0.05.
* data_dynamic_kat.json -- the first rows (iter, epsilon %, delta, chosen
  scale) of the reference's convergence logs data_dynamic/*.txt, produced
  by its ground-truth line-search drivers (sirt-test-weight.cu:320-402,
  sirt-norm-test-weight.cu:407-502).  Data only; read once from the
  reference checkout when this script is run.

Usage: python tests/golden/make_fixtures.py [reference_root]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def phantom_f32(n=192):
    f32 = np.float32
    x = np.array([f32(-0.1) + f32(i) * f32(0.001) for i in range(n)], dtype=f32)
    rc, rp = f32(0.015), f32(0.005)
    ca = np.cos(f32(-30) * (f32(3.14159265) / f32(180)), dtype=f32)
    sa = np.sin(f32(30) * (f32(3.14159265) / f32(180)), dtype=f32)
    X, Y = x[None, :], x[:, None]
    fo = np.zeros((n, n), f32)
    fo[(X * X + Y * Y) <= rc * rc] = f32(0.03)
    fo[((X - rc * ca) ** 2 + (Y - rc * sa) ** 2) <= rp * rp] = f32(0.05)
    fo[((X + rc * ca) ** 2 + (Y - rc * sa) ** 2) <= rp * rp] = f32(0.05)
    fo[(X * X + (Y + rc) ** 2) <= rp * rp] = f32(0.05)
    return fo


def write_mimo(path, a2d):
    with open(path, "w") as fh:
        for row in a2d:
            fh.write("".join(f"{float(v):.9g} " for v in row) + "\n")
        fh.write("\n\n")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    fo = phantom_f32()
    assert (fo == np.float32(0.03)).sum() == 599 and (fo == np.float32(0.05)).sum() == 228
    write_mimo(os.path.join(HERE, "fo_192.txt"), fo)
    logs = {}
    ddir = os.path.join(ref, "data_dynamic")
    for name in sorted(os.listdir(ddir)):
        rows = []
        with open(os.path.join(ddir, name)) as fh:
            for line in fh:
                if line.strip():
                    it, eps, diff, scale = line.split()
                    rows.append([int(it), float(eps), float(diff), float(scale)])
                if len(rows) == 6:
                    break
        logs[name] = rows
    with open(os.path.join(HERE, "data_dynamic_kat.json"), "w") as fh:
        json.dump({"source": "data_dynamic/*.txt (first 6 rows)", "logs": logs}, fh, indent=1)


if __name__ == "__main__":
    main()
