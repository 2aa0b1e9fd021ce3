"""Selector parity: the reference rule (src/conv2d.cpp:83-114).

Fixture tests/golden/selector_cases.json holds the algorithm the reference's
own compiled getkernelInfo chose for 183 problems (each in a fresh process).
Both the oracle restatement and the library's pure selector must agree.
"""
import json
import os

import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "selector_cases.json")


def _cases():
    with open(GOLD) as f:
        return json.load(f)


def test_oracle_selector_matches_reference(oracle):
    cases = _cases()
    assert len(cases) > 100
    for c in cases:
        assert oracle.select(*c["shape"])[0] == c["algo"], c


def test_library_selector_matches_reference(wc):
    for c in _cases():
        n, ch, h, w, k = c["shape"][:5]
        assert wc.select_algo(n, ch, h, w, k) == c["algo"], c


def test_selector_appendix_workspace(oracle):
    # workspace sizes of the reference's hipMalloc (src/conv2d.cpp:96-110)
    assert oracle.select(64, 256, 14, 14, 256) == (0, 42467328, 1024)
    assert oracle.select(2, 1920, 32, 32, 640) == (1, 81264640, 512)
    assert oracle.select(64, 256, 28, 28, 256)[0] == 2
    assert oracle.select(1, 64, 56, 56, 64) == (0, 2101248, 196)
    assert oracle.select(32, 512, 14, 14, 512) == (0, 56623104, 512)


def test_fused_rule_invariant(oracle):
    # fused iff n*c*h*w >= 10*k*c*r*s (32-bit, no overflow at these sizes)
    for (n, c, h, k) in [(1, 64, 56, 64), (16, 64, 56, 64), (2, 320, 64, 4), (4, 8, 5, 300)]:
        fused = n * c * h * h >= 10 * k * c * 9
        assert (oracle.select(n, c, h, h, k)[0] == 2) == fused
