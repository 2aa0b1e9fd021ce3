"""ctypes front end of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The operator never
routes through it.

fp16 tensors cross this boundary as numpy float16 arrays (bit patterns are
reinterpreted as uint16 for the C side).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_VERFIY_PATH = os.path.join(HERE, "_ref", "libverfiy_ref.so")

_lib = None
_ref = None

_i = C.c_int
_i64 = C.c_int64
_u32 = C.c_uint32
_p = C.c_void_p


def build(ref: bool = False) -> None:
    """Compile liboracle.so (and, with ref=True, the reference-built checker)."""
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_h2d.restype = C.c_double
        L.oracle_h2d.argtypes = [C.c_uint16]
        L.oracle_d2h.restype = C.c_uint16
        L.oracle_d2h.argtypes = [C.c_double]
        L.oracle_conv2d_verfiy.argtypes = [_p, _p, _p] + [_i] * 11
        L.oracle_conv2d_f64.argtypes = [_p, _p, _p] + [_i] * 11
        L.oracle_conv2d_f64_sample.argtypes = [_p, _p] + [_i] * 11 + [_p, _i64, _p]
        L.oracle_conv2d_verfiy_sample.argtypes = [_p, _p] + [_i] * 11 + [_p, _i64, _p]
        L.oracle_wino_table.argtypes = [_i, _i, _p]
        L.oracle_wino_table.restype = _i
        L.oracle_wino_filter.argtypes = [_i, _p, _i, _i, _p]
        L.oracle_wino_input.argtypes = [_i, _p, _i, _i, _i, _i, _i, _p]
        L.oracle_wino_gemm.argtypes = [_i, _p, _p, _i, _i, _i64, _p]
        L.oracle_wino_output.argtypes = [_i, _p, _i, _i, _i, _i, _p]
        L.oracle_wino_tiles.argtypes = [_i, _i, _i, _i]
        L.oracle_wino_tiles.restype = _i64
        L.oracle_round_f16.argtypes = [_p, _i64]
        L.oracle_e4m3_decode.argtypes = [C.c_uint8]
        L.oracle_e4m3_decode.restype = C.c_double
        L.oracle_e4m3_encode.argtypes = [C.c_double]
        L.oracle_e4m3_encode.restype = C.c_uint8
        L.oracle_quant_e4m3.argtypes = [_p, _i64, C.c_double]
        L.oracle_select.argtypes = [_u32] * 11 + [_p, _p, _p]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own conv2dcpu (verfiy.h) built from the reference tree,
    or None when it has not been built (it needs the reference sources)."""
    global _ref
    if _ref is None and os.path.exists(REF_VERFIY_PATH):
        R = C.CDLL(REF_VERFIY_PATH)
        R.ref_conv2dcpu.argtypes = [_p, _p, _p] + [_i] * 11
        _ref = R
    return _ref


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def out_hw(h, w, r=3, s=3, u=1, v=1, p=1, q=1):
    return (h + 2 * p - r) // u + 1, (w + 2 * q - s) // v + 1


def conv2d_verfiy(x: np.ndarray, wt: np.ndarray, u=1, v=1, p=1, q=1) -> np.ndarray:
    """verfiy.h:5-44 restated: fp16 in, fp16 out (bit-faithful)."""
    x = np.ascontiguousarray(x, dtype=np.float16)
    wt = np.ascontiguousarray(wt, dtype=np.float16)
    n, c, h, w = x.shape
    k, c2, r, s = wt.shape
    assert c == c2
    oh, ow = out_hw(h, w, r, s, u, v, p, q)
    out = np.zeros((n, k, oh, ow), dtype=np.float16)
    lib().oracle_conv2d_verfiy(_ptr(x), _ptr(wt), _ptr(out), n, c, h, w, k, r, s, u, v, p, q)
    return out


def conv2d_verfiy_ref(x: np.ndarray, wt: np.ndarray, u=1, v=1, p=1, q=1) -> np.ndarray:
    """The reference's own compiled conv2dcpu (needs oracle/_ref)."""
    R = ref_lib()
    if R is None:
        raise FileNotFoundError(REF_VERFIY_PATH)
    x = np.ascontiguousarray(x, dtype=np.float16)
    wt = np.ascontiguousarray(wt, dtype=np.float16)
    n, c, h, w = x.shape
    k, _, r, s = wt.shape
    oh, ow = out_hw(h, w, r, s, u, v, p, q)
    out = np.zeros((n, k, oh, ow), dtype=np.float16)
    R.ref_conv2dcpu(_ptr(x), _ptr(wt), _ptr(out), n, c, h, w, k, r, s, u, v, p, q)
    return out


def conv2d_f64(x: np.ndarray, wt: np.ndarray, u=1, v=1, p=1, q=1) -> np.ndarray:
    """Exact FP64 direct correlation of fp16 inputs."""
    x = np.ascontiguousarray(x, dtype=np.float16)
    wt = np.ascontiguousarray(wt, dtype=np.float16)
    n, c, h, w = x.shape
    k, _, r, s = wt.shape
    oh, ow = out_hw(h, w, r, s, u, v, p, q)
    out = np.zeros((n, k, oh, ow), dtype=np.float64)
    lib().oracle_conv2d_f64(_ptr(x), _ptr(wt), _ptr(out), n, c, h, w, k, r, s, u, v, p, q)
    return out


def conv2d_f64_sample(x, wt, idx: np.ndarray, u=1, v=1, p=1, q=1) -> np.ndarray:
    """FP64 direct correlation at flat NKHW output indices `idx`."""
    x = np.ascontiguousarray(x, dtype=np.float16)
    wt = np.ascontiguousarray(wt, dtype=np.float16)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    n, c, h, w = x.shape
    k, _, r, s = wt.shape
    out = np.zeros(idx.shape[0], dtype=np.float64)
    lib().oracle_conv2d_f64_sample(_ptr(x), _ptr(wt), n, c, h, w, k, r, s, u, v, p, q,
                                   _ptr(idx), idx.shape[0], _ptr(out))
    return out


def conv2d_verfiy_sample(x, wt, idx: np.ndarray, u=1, v=1, p=1, q=1) -> np.ndarray:
    """verfiy.h semantics at flat NKHW output indices `idx` (fp16 result)."""
    x = np.ascontiguousarray(x, dtype=np.float16)
    wt = np.ascontiguousarray(wt, dtype=np.float16)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    n, c, h, w = x.shape
    k, _, r, s = wt.shape
    out = np.zeros(idx.shape[0], dtype=np.float16)
    lib().oracle_conv2d_verfiy_sample(_ptr(x), _ptr(wt), n, c, h, w, k, r, s, u, v, p, q,
                                      _ptr(idx), idx.shape[0], _ptr(out))
    return out


def wino_table(m: int, which: int) -> np.ndarray:
    T = m + 2
    shape = [(T, 3), (T, T), (m, T)][which]
    out = np.zeros(shape, dtype=np.float64)
    lib().oracle_wino_table(m, which, _ptr(out))
    return out


def wino_tiles(m, n, oh, ow) -> int:
    return int(lib().oracle_wino_tiles(m, n, oh, ow))


def wino_filter(m: int, wt: np.ndarray) -> np.ndarray:
    """U[t][k][c] (FP64) of fp16 filters [K][C][3][3]."""
    wt = np.ascontiguousarray(wt, dtype=np.float16)
    K, Cc = wt.shape[:2]
    T = m + 2
    U = np.zeros((T * T, K, Cc), dtype=np.float64)
    lib().oracle_wino_filter(m, _ptr(wt), K, Cc, _ptr(U))
    return U


def wino_input(m: int, x: np.ndarray, pad: int = 1) -> np.ndarray:
    """V[t][c][p] (FP64) of an fp16 NCHW input."""
    x = np.ascontiguousarray(x, dtype=np.float16)
    N, Cc, H, W = x.shape
    T = m + 2
    P = wino_tiles(m, N, H + 2 * pad - 2, W + 2 * pad - 2)
    V = np.zeros((T * T, Cc, P), dtype=np.float64)
    lib().oracle_wino_input(m, _ptr(x), N, Cc, H, W, pad, _ptr(V))
    return V


def wino_gemm(U: np.ndarray, V: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    T2, K, Cc = U.shape
    P = V.shape[2]
    M = np.zeros((T2, K, P), dtype=np.float64)
    lib().oracle_wino_gemm(T2, _ptr(U), _ptr(V), K, Cc, P, _ptr(M))
    return M


def wino_output(m: int, M: np.ndarray, N: int, oh: int, ow: int) -> np.ndarray:
    M = np.ascontiguousarray(M, dtype=np.float64)
    K = M.shape[1]
    Y = np.zeros((N, K, oh, ow), dtype=np.float64)
    lib().oracle_wino_output(m, _ptr(M), N, K, oh, ow, _ptr(Y))
    return Y


def round_f16(a: np.ndarray) -> np.ndarray:
    a = np.array(a, dtype=np.float64, order="C", copy=True)
    lib().oracle_round_f16(_ptr(a), a.size)
    return a


def e4m3_decode(b: int) -> float:
    return float(lib().oracle_e4m3_decode(b))


def e4m3_encode(x: float) -> int:
    return int(lib().oracle_e4m3_encode(x))


def quant_e4m3(a: np.ndarray, scale: float) -> np.ndarray:
    a = np.array(a, dtype=np.float64, order="C", copy=True)
    lib().oracle_quant_e4m3(_ptr(a), a.size, float(scale))
    return a


def select(n, c, h, w, k, r=3, s=3, u=1, v=1, p=1, q=1):
    """(algo, workspace bytes, tiles) of the reference selector."""
    algo = C.c_int(0)
    ws = C.c_uint64(0)
    tiles = C.c_uint64(0)
    lib().oracle_select(n, c, h, w, k, r, s, u, v, p, q, C.byref(algo), C.byref(ws), C.byref(tiles))
    return algo.value, ws.value, tiles.value


def wino_model(m: int, x: np.ndarray, wt: np.ndarray, quant_uv: str = "none",
               round_m: bool = False) -> np.ndarray:
    """Winograd pipeline in FP64 with the operand rounding of a given GPU
    precision: quant_uv 'none' (exact), 'f16' (U and V rounded to fp16)."""
    N, Cc, H, W = x.shape
    oh, ow = out_hw(H, W)
    U = wino_filter(m, wt)
    V = wino_input(m, x)
    if quant_uv == "f16":
        U = round_f16(U)
        V = round_f16(V)
    M = wino_gemm(U, V)
    if round_m:
        M = round_f16(M)
    return wino_output(m, M, N, oh, ow)


def errors(got: np.ndarray, ref: np.ndarray):
    """(max-abs, relative-L2) of got against ref, in float64."""
    g = np.asarray(got, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    d = g - r
    max_abs = float(np.max(np.abs(d))) if d.size else 0.0
    den = float(np.sqrt(np.sum(r * r)))
    rel = float(np.sqrt(np.sum(d * d)) / den) if den > 0 else float(np.sqrt(np.sum(d * d)))
    return max_abs, rel
