"""The command-line harness (conv2dfp16demo), run as the reference's harness
is run (src/main.cpp:7-134): 11 positional arguments, a warm-up call, timed
calls, then its own per-element CPU check (include/verfiy.h semantics,
|gpu - cpu| / gpu <= 0.01) must report error:0."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "dcu-winograd-conv2d_amd", "conv2dfp16demo")

needs_demo = pytest.mark.skipif(not os.path.exists(DEMO), reason="conv2dfp16demo not built")


def run_demo(*args, timeout=120):
    return subprocess.run([DEMO, *[str(a) for a in args]], capture_output=True, text=True,
                          timeout=timeout)


@needs_demo
def test_demo_usage_without_arguments():
    r = run_demo()
    assert r.returncode == 1
    assert "usage: conv2dfp16demo n c h w k r s u v p q" in r.stderr


@needs_demo
def test_demo_rejects_unknown_option():
    r = run_demo(1, 1, 4, 4, 1, 3, 3, 1, 1, 1, 1, "--bogus")
    assert r.returncode == 1 and "usage" in r.stderr


DEMO_CASES = [
    # (args, extra options, expected algorithm line)
    ((2, 16, 9, 9, 8, 3, 3, 1, 1, 1, 1), (), None),
    ((4, 8, 28, 28, 8, 3, 3, 1, 1, 1, 1), (), "F(2x3) fused"),        # img >= 10 flt
    ((4, 64, 14, 14, 128, 3, 3, 1, 1, 1, 1), (), "F(4x3) non-fused"),   # k <= 4 P4
    ((1, 64, 14, 14, 128, 3, 3, 1, 1, 1, 1), (), "F(2x3) non-fused"),   # k > 4 P4
    ((3, 40, 13, 17, 24, 3, 3, 1, 1, 1, 1), ("--algo", "f23"), "F(2x3) non-fused"),
    ((2, 32, 15, 15, 48, 3, 3, 1, 1, 1, 1), ("--algo", "f43"), "F(4x3) non-fused"),
    ((2, 32, 15, 15, 48, 3, 3, 1, 1, 1, 1), ("--algo", "f43", "--precision", "fp8x2"), "fp8x2"),
    # not 3x3 / stride 1 / pad 1: the direct fallback
    ((2, 8, 12, 12, 16, 5, 5, 1, 1, 2, 2), (), "direct"),
    ((3, 16, 15, 17, 24, 3, 3, 2, 2, 1, 1), (), "direct"),
    ((2, 32, 9, 9, 16, 1, 1, 1, 1, 0, 0), (), "direct"),
    ((2, 16, 9, 9, 8, 3, 3, 1, 1, 1, 1), ("--algo", "direct"), "direct"),
]


@pytest.mark.gpu
@needs_demo
@pytest.mark.parametrize("args,opts,expect", DEMO_CASES)
def test_demo_verifies(args, opts, expect):
    print("demo", args, opts, flush=True)
    r = run_demo(*args, *opts, "--iters", 5)
    assert r.returncode == 0, r.stdout + r.stderr
    assert re.search(r"paramsize:\d+", r.stdout)
    assert re.search(r"time: [0-9.]+ us", r.stdout)
    assert "================finish,error:0=========================" in r.stdout, r.stdout
    if expect:
        assert expect in r.stdout, r.stdout
