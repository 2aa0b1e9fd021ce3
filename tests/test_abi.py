"""The C ABI library loads and exports every symbol the headers declare;
struct layouts match the reference's (include/conv2d.h:8-36).  No GPU calls."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for hdr in ("conv2d.h", "winoconv.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_exports_every_declared_symbol(wc):
    names = _declared_functions()
    assert {"getParamsize", "getkernelInfo", "conv2d_fp16", "wino_run"} <= names
    L = wc.lib()
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(wc._SIGS) >= names - {"wino_select_algo"} | {"wino_select_algo"}


def test_struct_layouts(wc):
    assert C.sizeof(wc.Problem) == 72
    assert C.sizeof(wc.KernelInfo) == 40


def test_param_size_and_unsupported_problems(wc):
    L = wc.lib()
    pr = wc.make_problem(0, 0, 0, 1, 8, 8, 8, 4)
    size = C.c_int(0)
    assert L.getParamsize(C.byref(pr), C.byref(size)) == 0 and size.value > 0
    blob = C.create_string_buffer(size.value)
    # non-3x3, strided and zero-extent problems are rejected before any HIP call
    for bad in (dict(r=5, s=5), dict(u=2, v=2), dict(p=0, q=0), dict(n=0)):
        f = dict(n=1, c=8, h=8, wd=8, k=4)
        f.update({("wd" if k == "w" else k): v for k, v in bad.items()})
        pb = wc.make_problem(0, 0, 0, f["n"], f["c"], f["h"], f["wd"], f["k"],
                             **{k: v for k, v in bad.items() if k in "rsuvpq"})
        assert L.getkernelInfo(C.byref(pb), None, blob) == wc.ERR_UNSUPPORTED
        assert L.wino_param_status(blob) == wc.ERR_UNSUPPORTED
        assert L.wino_run(blob) == wc.ERR_UNSUPPORTED
    # an uninitialised blob is refused, not dereferenced
    junk = C.create_string_buffer(size.value)
    assert L.wino_run(junk) == wc.ERR_INVALID
    assert L.wino_param_destroy(junk) == wc.ERR_INVALID


def test_status_strings(wc):
    for st in range(8):
        assert wc.status_string(st)
