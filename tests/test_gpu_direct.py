"""The general fallback (WINO_ALGO_DIRECT, implicit GEMM on MFMA) against the
FP64 direct-convolution oracle, for filter sizes, strides and paddings the
Winograd paths do not compute, through the reference entry points
(getkernelInfo picks it for such problems).

Tolerance: FP32 accumulation of fp16 products and one fp16 rounding of the
result give rel-L2 vs FP64 of a few 1e-4; the bar is 1e-3.  The reference
harness's own per-element check (|gpu - cpu| / gpu <= 0.01 against the
verfiy.h restatement, positive data as src/main.cpp:35-43) must pass too."""
import numpy as np
import pytest

DIRECT_CASES = [
    # n, c, h, w, k, r, s, u, v, p, q
    (2, 5, 9, 7, 3, 5, 3, 2, 1, 2, 0),
    (1, 64, 14, 14, 32, 1, 1, 1, 1, 0, 0),
    (3, 16, 15, 17, 40, 3, 3, 2, 2, 1, 1),
    (2, 8, 8, 8, 70, 7, 7, 1, 1, 3, 3),
    (2, 33, 10, 11, 65, 2, 2, 1, 1, 0, 0),
    (1, 4, 5, 5, 2, 5, 5, 1, 1, 0, 0),          # one output pixel
    (4, 128, 28, 28, 96, 1, 1, 2, 2, 0, 0),
    (1, 3, 32, 32, 16, 3, 3, 1, 1, 1, 1),       # 3x3 s1 p1, forced direct
    (2, 24, 13, 9, 12, 3, 5, 1, 3, 0, 2),
    (2, 16, 10, 10, 200, 3, 3, 2, 2, 1, 1),     # 256-channel blocks
    (1, 5, 7, 9, 300, 1, 3, 1, 1, 0, 1),        # 256-channel blocks, C*R*S odd
]


def _run(wc, gpu, x, wt, dims, algo=None):
    n, c, h, w = x.shape
    k, _, r, s = wt.shape
    u, v, p, q = dims
    oh, ow = (h + 2 * p - r) // u + 1, (w + 2 * q - s) // v + 1
    dx, dw = gpu.to_dev(x), gpu.to_dev(wt)
    dy = gpu.empty_f16((n, k, oh, ow))
    op = wc.Conv2d(dx.data_ptr(), dw.data_ptr(), dy.data_ptr(), n, c, h, w, k,
                   algo=wc.ALGO_AUTO if algo is None else algo, r=r, s=s, u=u, v=v, p=p, q=q)
    try:
        op()
        info = op.info()
        y = gpu.to_host(dy)
    finally:
        op.destroy()
    return y, info


@pytest.mark.gpu
@pytest.mark.parametrize("case", DIRECT_CASES)
def test_direct_matches_fp64(wc, gpu, oracle, case):
    n, c, h, w, k, r, s, u, v, p, q = case
    print("direct", case, flush=True)
    rng = np.random.default_rng(sum(case))
    x = rng.standard_normal((n, c, h, w)).astype(np.float16)
    wt = (rng.standard_normal((k, c, r, s)) * np.sqrt(2.0 / (c * r * s))).astype(np.float16)
    winograd = (r, s, u, v, p, q) == (3, 3, 1, 1, 1, 1)
    y, info = _run(wc, gpu, x, wt, (u, v, p, q), algo=wc.ALGO_DIRECT if winograd else None)
    assert info.algo == wc.ALGO_DIRECT
    assert not np.isnan(y).any()
    ref = oracle.conv2d_f64(x, wt, u, v, p, q)
    assert y.shape == ref.shape
    err = np.linalg.norm(y.astype(np.float64) - ref) / np.linalg.norm(ref)
    assert err <= 1e-3, err
    # every element within fp16 rounding of the FP32 sum (|y| 2^-11) plus the
    # FP32 accumulation error bound over the reduction
    red = c * r * s
    bound = np.abs(ref) * 2.0 ** -10 + 2.0 ** -20 * red * np.abs(
        oracle.conv2d_f64(np.abs(x.astype(np.float32)).astype(np.float16),
                          np.abs(wt.astype(np.float32)).astype(np.float16), u, v, p, q)) + 1e-6
    assert np.all(np.abs(y.astype(np.float64) - ref) <= bound)


@pytest.mark.gpu
@pytest.mark.parametrize("case", DIRECT_CASES[:4])
def test_direct_passes_reference_harness_check(wc, gpu, oracle, case):
    n, c, h, w, k, r, s, u, v, p, q = case
    rng = np.random.default_rng(7)
    x = (rng.integers(0, 255, (n, c, h, w)) / 255.0).astype(np.float16)
    wt = (rng.integers(0, 255, (k, c, r, s)) / 255.0).astype(np.float16)
    y, _ = _run(wc, gpu, x, wt, (u, v, p, q))
    cpu = oracle.conv2d_verfiy(x, wt, u, v, p, q).astype(np.float32)
    g = y.astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        bad = (np.abs(g - cpu) / g > 0.01) | np.isnan(g) | np.isinf(g)
    assert not bad.any(), np.argwhere(bad)[:5]


@pytest.mark.gpu
def test_general_problems_plan_direct(wc, gpu):
    """getkernelInfo plans WINO_ALGO_DIRECT for every problem outside 3x3
    stride 1 pad 1, with the rearranged-filter workspace when the filter is
    not already [K][R*S][C] with C a multiple of 8; Winograd algorithms and
    fp8 are refused for such problems."""
    import ctypes as C
    L = wc.lib()
    size = C.c_int(0)
    L.getParamsize(None, C.byref(size))
    for dims in (dict(r=5, s=5, p=2, q=2), dict(u=2, v=2), dict(p=0, q=0),
                 dict(r=5, s=3, u=2, v=1, p=2, q=0)):
        blob = C.create_string_buffer(size.value)
        pr = wc.make_problem(0, 0, 0, 2, 5, 9, 11, 4, **dims)
        info = wc.KernelInfo()
        assert L.getkernelInfo(C.byref(pr), C.byref(info), blob) == 0, dims
        inf = wc.Info()
        assert L.wino_param_info(blob, C.byref(inf)) == 0
        r, s, u, v = dims.get("r", 3), dims.get("s", 3), dims.get("u", 1), dims.get("v", 1)
        p, q = dims.get("p", 1), dims.get("q", 1)
        oh, ow = (9 + 2 * p - r) // u + 1, (11 + 2 * q - s) // v + 1
        assert inf.algo == wc.ALGO_DIRECT and (inf.oh, inf.ow) == (oh, ow)
        assert inf.u_bytes == 4 * r * s * 8 * 2  # K x R*S x Cp fp16, Cp = 8
        assert info.blockx == -(-2 * oh * ow // 64) and info.blocky == 1
        for algo in (wc.ALGO_F43, wc.ALGO_F23, wc.ALGO_F23_FUSED):
            assert L.wino_param_set_algo(blob, algo) == wc.ERR_UNSUPPORTED
        assert L.wino_param_set_algo(blob, wc.ALGO_DIRECT) == 0
        assert L.wino_param_set_precision(blob, wc.PREC_FP8) == wc.ERR_UNSUPPORTED
        assert L.wino_param_destroy(blob) == 0


@pytest.mark.gpu
def test_direct_resident_filter_and_rebinding(wc, gpu, oracle):
    """A resident filter is rearranged once; rebinding a new filter of the
    same shape makes the next call rearrange again."""
    rng = np.random.default_rng(5)
    n, c, h, w, k, r, s = 2, 6, 11, 10, 20, 5, 3
    x = rng.standard_normal((n, c, h, w)).astype(np.float16)
    w1 = (rng.standard_normal((k, c, r, s)) * 0.2).astype(np.float16)
    w2 = (rng.standard_normal((k, c, r, s)) * 0.2).astype(np.float16)
    oh, ow = (h + 2 - r) // 2 + 1, (w + 2 - s) // 1 + 1
    dx, d1, d2 = gpu.to_dev(x), gpu.to_dev(w1), gpu.to_dev(w2)
    dy = gpu.empty_f16((n, k, oh, ow))
    op = wc.Conv2d(dx.data_ptr(), d1.data_ptr(), dy.data_ptr(), n, c, h, w, k, resident=True,
                   r=r, s=s, u=2, v=1, p=1, q=1)
    try:
        for wt, dw in ((w1, d1), (w2, d2)):
            op.set_tensors(dx.data_ptr(), dw.data_ptr(), dy.data_ptr())
            op()
            op()
            y = gpu.to_host(dy).astype(np.float64)
            ref = oracle.conv2d_f64(x, wt, 2, 1, 1, 1)
            assert np.linalg.norm(y - ref) / np.linalg.norm(ref) <= 1e-3
    finally:
        op.destroy()


@pytest.mark.gpu
def test_direct_equals_winograd_within_tolerance(wc, gpu, oracle):
    """On a 3x3 s1 p1 problem the fallback and the F(4x3) path agree with
    each other to their joint tolerance."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 32, 14, 14)).astype(np.float16)
    wt = (rng.standard_normal((48, 32, 3, 3)) * 0.08).astype(np.float16)
    yd, _ = _run(wc, gpu, x, wt, (1, 1, 1, 1), algo=wc.ALGO_DIRECT)
    yw, _ = _run(wc, gpu, x, wt, (1, 1, 1, 1), algo=wc.ALGO_F43)
    ref = oracle.conv2d_f64(x, wt)
    assert np.linalg.norm(yd - ref) / np.linalg.norm(ref) <= 1e-3
    assert np.linalg.norm(yd.astype(np.float64) - yw) / np.linalg.norm(ref) <= 1e-2


@pytest.mark.gpu
@pytest.mark.parametrize("case", [(2, 5, 9, 7, 3, 5, 3, 2, 1, 2, 0), (3, 16, 12, 10, 20, 3, 3, 1, 1, 1, 1)])
def test_gpu_fp64_error_statement_matches_oracle(wc, gpu, oracle, case):
    """wino_conv_error_f64 (whole-tensor error against an FP64 direct
    convolution computed on the GPU) equals the same statement computed from
    the CPU FP64 oracle, to FP64 summation-order noise; a planted error in
    one element moves max-abs by exactly that amount."""
    n, c, h, w, k, r, s, u, v, p, q = case
    rng = np.random.default_rng(17)
    x = rng.standard_normal((n, c, h, w)).astype(np.float16)
    wt = (rng.standard_normal((k, c, r, s)) * 0.3).astype(np.float16)
    ref = oracle.conv2d_f64(x, wt, u, v, p, q)
    y = ref.astype(np.float16)
    y.flat[7] = np.float16(float(y.flat[7]) + 0.5)
    dx, dw, dy = gpu.to_dev(x), gpu.to_dev(wt), gpu.to_dev(y)
    mx, rl = wc.conv_error_f64(dx.data_ptr(), dw.data_ptr(), dy.data_ptr(), n, c, h, w, k,
                               r, s, u, v, p, q)
    d = y.astype(np.float64) - ref
    exp_rl = np.linalg.norm(d) / np.linalg.norm(ref)
    assert abs(mx - np.abs(d).max()) <= 1e-12 * max(1.0, np.abs(d).max())
    assert abs(rl - exp_rl) <= 1e-12 * exp_rl
    assert np.abs(d).max() >= 0.49
