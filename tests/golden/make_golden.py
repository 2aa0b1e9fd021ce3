"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Run in the development container (it needs the reference tree, default
/root/reference, or $WINO_REF_ROOT):

    python tests/golden/make_golden.py

It builds oracle/_ref (oracle/Makefile target `ref`):
  * libverfiy_ref.so -- the reference's CPU direct convolution
    (include/verfiy.h:5-44) compiled unchanged with ROCm clang++;
  * ref_select       -- the reference's selector (src/conv2d.cpp:24-114)
    compiled unchanged, run once per problem in a fresh process.

Outputs (data only: inputs and expected outputs):
  verfiy_cases.npz   per case: <name>_x, <name>_w (fp16 bits as uint16),
                     <name>_y (the reference conv2dcpu output, uint16 bits),
                     <name>_shape (n c h w k r s u v p q)
  selector_cases.json  [{"shape": [n c h w k r s u v p q], "algo": a}, ...]
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

REF = os.environ.get("WINO_REF_ROOT", "/root/reference")


def _f16(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).astype(np.float16))


def verfiy_cases():
    rng = np.random.default_rng(20261203)
    cases = []

    def add(name, n, c, h, w, k, r=3, s=3, u=1, v=1, p=1, q=1, dist="u11", x=None, wt=None):
        if x is None:
            if dist == "u11":
                x = rng.uniform(-1, 1, (n, c, h, w))
                wt = rng.uniform(-1, 1, (k, c, r, s))
            elif dist == "normal":
                x = np.clip(rng.standard_normal((n, c, h, w)), -4, 4)
                wt = rng.standard_normal((k, c, r, s)) * np.sqrt(2.0 / (9 * c))
            elif dist == "harness":  # src/main.cpp:36-44: rand()%255/255
                x = rng.integers(0, 255, (n, c, h, w)) / 255.0
                wt = rng.integers(0, 255, (k, c, r, s)) / 255.0
            elif dist == "subnormal":
                x = rng.uniform(-1, 1, (n, c, h, w)) * 2.0 ** -8
                wt = rng.uniform(-1, 1, (k, c, r, s)) * 2.0 ** -8
            elif dist == "large":
                x = rng.uniform(-200, 200, (n, c, h, w))
                wt = rng.uniform(-200, 200, (k, c, r, s))
        cases.append((name, (n, c, h, w, k, r, s, u, v, p, q), _f16(x), _f16(wt)))

    add("tiny_6x6", 1, 8, 6, 6, 4)
    add("odd_7x7", 2, 16, 7, 7, 8, dist="normal")
    add("c3_5x5", 1, 3, 5, 5, 2)
    add("harness_8x8", 1, 4, 8, 8, 3, dist="harness")
    add("subnormal_6x6", 1, 6, 6, 6, 3, dist="subnormal")
    add("large_5x5", 1, 3, 5, 5, 2, dist="large")
    add("strided_9x7", 1, 4, 9, 7, 3, u=2, v=2)
    add("rect_6x10", 2, 5, 6, 10, 3)
    add("k5x3_8x8", 1, 3, 8, 8, 2, r=5, s=3, p=2, q=1)
    add("f43_14x14", 1, 16, 14, 14, 8, dist="normal")
    # Product rounding probe: (1+2^-10)^2 rounds to 1+2^-9 in fp16, so the
    # sum with -(1+2^-9) is 0 (exact products would give 2^-20).
    add("probe_product", 1, 2, 1, 1, 1, r=1, s=1, p=0, q=0,
        x=_f16(np.array([1 + 2.0 ** -10, -1.0]).reshape(1, 2, 1, 1)),
        wt=_f16(np.array([1 + 2.0 ** -10, 1 + 2.0 ** -9]).reshape(1, 2, 1, 1)))
    # Final-cast probe: 1 + 2^-11 + 2^-24 rounds up to 1+2^-10 when converted
    # directly from double, and to 1.0 when it passes through float first.
    add("probe_sum", 1, 3, 1, 1, 1, r=1, s=1, p=0, q=0,
        x=_f16(np.array([1.0, 2.0 ** -11, 2.0 ** -24]).reshape(1, 3, 1, 1)),
        wt=_f16(np.ones((1, 3, 1, 1))))
    return cases


def selector_shapes():
    shapes = []
    pre = [(64, 256, 14, 14, 256), (256, 192, 14, 14, 192), (16, 256, 26, 26, 512),
           (32, 256, 14, 14, 256), (2, 1280, 16, 16, 1280), (2, 960, 64, 64, 32)]
    final = [(16, 128, 64, 64, 27), (16, 256, 32, 32, 256), (16, 64, 128, 128, 64),
             (2, 1920, 32, 32, 640), (2, 640, 64, 64, 640), (2, 320, 64, 64, 4)]
    cfgs = [(1, 64, 56, 56, 64), (64, 256, 28, 28, 256), (32, 512, 14, 14, 512)]
    for s in pre + final + cfgs:
        shapes.append(s + (3, 3, 1, 1, 1, 1))
    for (c, hw) in [(64, 56), (128, 28), (256, 14), (512, 7)]:
        for n in range(1, 33):
            shapes.append((n, c, hw, hw, c, 3, 3, 1, 1, 1, 1))
    rng = np.random.default_rng(7)
    for _ in range(40):
        n = int(rng.integers(1, 65))
        c = int(rng.choice([3, 16, 32, 64, 96, 128, 256, 384, 512, 1024]))
        hw = int(rng.choice([5, 7, 13, 14, 16, 28, 31, 56, 64]))
        k = int(rng.choice([4, 16, 27, 64, 128, 256, 512, 1024]))
        shapes.append((n, c, hw, hw, k, 3, 3, 1, 1, 1, 1))
    return shapes


def main():
    O.build(ref=True)
    cases = verfiy_cases()
    out = {}
    for name, shp, x, wt in cases:
        n, c, h, w, k, r, s, u, v, p, q = shp
        y = O.conv2d_verfiy_ref(x, wt, u, v, p, q)
        out[name + "_x"] = x.view(np.uint16)
        out[name + "_w"] = wt.view(np.uint16)
        out[name + "_y"] = y.view(np.uint16)
        out[name + "_shape"] = np.array(shp, dtype=np.int64)
    out["names"] = np.array([c[0] for c in cases])
    np.savez_compressed(os.path.join(HERE, "verfiy_cases.npz"), **out)

    sel_bin = os.path.join(ROOT, "oracle", "_ref", "ref_select")
    recs = []
    for shp in selector_shapes():
        res = subprocess.run([sel_bin] + [str(v) for v in shp], check=True,
                             capture_output=True, text=True, timeout=120)
        algo, psize = (int(t) for t in res.stdout.split()[-2:])
        recs.append({"shape": list(shp), "algo": algo, "ref_param_size": psize})
    with open(os.path.join(HERE, "selector_cases.json"), "w") as f:
        json.dump(recs, f, indent=0)
    print(f"wrote {len(cases)} conv cases, {len(recs)} selector cases")


if __name__ == "__main__":
    main()
