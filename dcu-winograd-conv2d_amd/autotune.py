"""Autotuner of the grouped GEMM tile shape and reduction split.

For every shape class (3x3 filter K x C, image H x W) and a ladder of batch
sizes N it runs the problem through the grouped launch with every tile
configuration of the GEMM table and every reduction split that divides the
work sensibly, times each with HIP events, and writes the fastest as a line

    m K C P_min P_max cfg split

of a tuning table that wino_engine_load_tuning reads (P = tiles of the
problem; consecutive ladder points split the P axis at their midpoints).
A small problem is measured as several copies in one batch (enough work for
~20 GFLOP per launch, at most 64 copies), so that the choice reflects its
throughput inside a loaded grouped launch rather than the fixed launch
latency.  The tile choice never changes results: every configuration
accumulates each output in the same order (checked by the GPU tests); the
split changes the FP32 summation order within the stated tolerance.

    python autotune.py [--out tuning/mi355x.tsv] [--classes C,K,HW ...]
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

DEFAULT_CLASSES = [(64, 64, 56), (128, 128, 28), (256, 256, 14), (512, 512, 7), (256, 256, 28)]
LADDER = [1, 2, 4, 8, 16, 32, 64]
N_CFG = 6


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "mi355x.tsv"))
    ap.add_argument("--classes", nargs="*", default=None, help="C,K,HW triples")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--precision", type=int, default=0)
    args = ap.parse_args()
    classes = DEFAULT_CLASSES if not args.classes else [
        tuple(int(v) for v in s.split(",")) for s in args.classes]

    import numpy as np
    import torch
    import winoconv as wc

    rng = np.random.default_rng(1)
    eng = wc.Engine(args.precision, 4 << 30)
    lines = ["# m K C P_min P_max cfg split   (grouped GEMM tile table, MI355X; autotune.py)"]
    try:
        for (c, k, hw) in classes:
            wt = torch.from_numpy((rng.standard_normal((k, c, 3, 3)) * 0.05).astype(np.float16)).cuda()
            fid = eng.register_filter(wt.data_ptr(), k, c)
            rows = []
            for n, algo in [(n, a) for n in LADDER for a in (wc.ALGO_F23, wc.ALGO_F43)]:
                x = torch.randn((n, c, hw, hw), device="cuda").half()
                copies = int(min(64, max(1, -(-2e10 // (2.0 * n * k * hw * hw * c * 9)))))
                ys = [torch.empty((n, k, hw, hw), device="cuda", dtype=torch.float16)
                      for _ in range(copies)]
                torch.cuda.synchronize()
                b = wc.Batch(eng, copies)
                try:
                    for y in ys:
                        b.add(fid, x.data_ptr(), y.data_ptr(), n, hw, hw, algo)
                    inf = b.info(0)
                    m, P = inf.m, inf.tiles
                    nsteps = inf.cp // (32 if args.precision == 0 else 128)
                    best = None
                    for cfg in range(N_CFG):
                        for split in (1, 2, 4):
                            # a reduction split writes FP32 partials: only
                            # for problems too small to fill the device alone
                            if split > nsteps or (split > 1 and P > 256):
                                continue
                            try:
                                for j in range(copies):
                                    b.set_tile(j, cfg, split)
                            except wc.WinoError:
                                continue  # tile not available for this K / precision
                            for _ in range(3):
                                b.launch()
                            torch.cuda.synchronize()
                            e0 = torch.cuda.Event(enable_timing=True)
                            e1 = torch.cuda.Event(enable_timing=True)
                            e0.record()
                            for _ in range(args.reps):
                                b.launch()
                            e1.record()
                            torch.cuda.synchronize()
                            us = e0.elapsed_time(e1) * 1e3 / args.reps / copies
                            if best is None or us < best[0]:
                                best = (us, cfg, split)
                    rows.append((m, P, best))
                    print(f"C={c} K={k} HW={hw} N={n} m={m} P={P}: cfg={best[1]} split={best[2]} "
                          f"{best[0]:.1f} us per problem ({copies} per launch)", flush=True)
                finally:
                    b.destroy()
            # P ranges per m: midpoints between consecutive ladder points
            for mm in sorted({r[0] for r in rows}):
                pts = sorted([(r[1], r[2]) for r in rows if r[0] == mm])
                for i, (P, best) in enumerate(pts):
                    lo = 1 if i == 0 else (pts[i - 1][0] + P) // 2 + 1
                    hi = (1 << 62) if i == len(pts) - 1 else (P + pts[i + 1][0]) // 2
                    lines.append(f"{mm} {k} {c} {lo} {hi} {best[1]} {best[2]}")
    finally:
        eng.destroy()
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.out} ({len(lines) - 1} entries)", flush=True)


if __name__ == "__main__":
    main()
