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
    # zero extents, zero strides and filters larger than the padded image are
    # rejected before any HIP call
    for bad in (dict(n=0), dict(k=0), dict(r=0), dict(u=0), dict(r=11, p=1), dict(s=9, q=0)):
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


def test_general_problems_plan_the_direct_fallback(wc):
    """A problem other than 3x3 stride 1 pad 1 plans WINO_ALGO_DIRECT; a 1x1
    filter whose channel count is a multiple of 8 is read in place, so the
    plan has no workspace and no HIP call happens here (the other shapes are
    planned in the GPU tests).  Winograd algorithms and fp8 are refused."""
    L = wc.lib()
    size = C.c_int(0)
    L.getParamsize(None, C.byref(size))
    for dims in (dict(r=1, s=1, p=0, q=0), dict(r=1, s=1, u=2, v=3, p=0, q=1)):
        blob = C.create_string_buffer(size.value)
        pr = wc.make_problem(0, 0, 0, 2, 8, 9, 11, 4, **dims)
        info = wc.KernelInfo()
        assert L.getkernelInfo(C.byref(pr), C.byref(info), blob) == 0, dims
        inf = wc.Info()
        assert L.wino_param_info(blob, C.byref(inf)) == 0
        r, s, u, v = dims.get("r", 3), dims.get("s", 3), dims.get("u", 1), dims.get("v", 1)
        p, q = dims.get("p", 1), dims.get("q", 1)
        oh, ow = (9 + 2 * p - r) // u + 1, (11 + 2 * q - s) // v + 1
        assert inf.algo == wc.ALGO_DIRECT and inf.ws_bytes == 0
        assert (inf.oh, inf.ow) == (oh, ow)
        assert info.blockx == -(-2 * oh * ow // 64) and info.blocky == 1 and info.threadx == 256
        for algo in (wc.ALGO_F43, wc.ALGO_F23, wc.ALGO_F23_FUSED):
            assert L.wino_param_set_algo(blob, algo) == wc.ERR_UNSUPPORTED
        assert L.wino_param_set_algo(blob, wc.ALGO_DIRECT) == 0
        assert L.wino_param_set_precision(blob, wc.PREC_FP8) == wc.ERR_UNSUPPORTED
        assert L.wino_param_destroy(blob) == 0


def test_status_strings(wc):
    for st in range(8):
        assert wc.status_string(st)
