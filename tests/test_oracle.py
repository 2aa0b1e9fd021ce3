"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

* verfiy restatement: bit-exact against the reference's own conv2dcpu
  (include/verfiy.h, compiled from the reference tree) on the committed
  fixtures, including two probes that separate plausible mis-readings.
* FP64 direct conv: against torch.nn.functional.conv2d in float64 (a library
  routine) and a brute-force loop on a tiny case.
* Winograd stage model: the exact-arithmetic identity A^T[(G g G^T) . (B^T d B)]A
  = correlation(d, g) on random data (any wrong coefficient, sign, index or
  transposed table breaks it), plus closed-form checks of each table.
* fp16 / e4m3 codecs: numpy's float16 cast and torch's float8_e4m3fn cast.
"""
import os

import numpy as np
import pytest
import torch

from conftest import rng_inputs

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cases():
    d = np.load(os.path.join(GOLD, "verfiy_cases.npz"))
    return d, [str(n) for n in d["names"]]


def test_verfiy_restatement_matches_reference_fixtures(oracle):
    d, names = _cases()
    assert len(names) >= 12
    for name in names:
        n, c, h, w, k, r, s, u, v, p, q = (int(t) for t in d[name + "_shape"])
        x = d[name + "_x"].view(np.float16).reshape(n, c, h, w)
        wt = d[name + "_w"].view(np.float16).reshape(k, c, r, s)
        got = oracle.conv2d_verfiy(x, wt, u, v, p, q).view(np.uint16)
        exp = d[name + "_y"].reshape(got.shape)
        assert np.array_equal(got, exp), name


def test_verfiy_probes_pin_rounding(oracle):
    d, _ = _cases()
    # fp16-rounded products: (1+2^-10)^2 -> 1+2^-9, sum with -(1+2^-9) is 0
    assert d["probe_product_y"].view(np.float16).ravel()[0] == 0.0
    # final cast is a direct double->fp16 rounding: 1+2^-11+2^-24 -> 1+2^-10
    assert float(d["probe_sum_y"].view(np.float16).ravel()[0]) == 1 + 2.0 ** -10


def test_verfiy_sample_matches_full(oracle):
    x, wt = rng_inputs(5, 2, 6, 7, 9, 4)
    full = oracle.conv2d_verfiy(x, wt).ravel()
    idx = np.array([0, 3, 17, 100, full.size - 1], dtype=np.int64)
    assert np.array_equal(oracle.conv2d_verfiy_sample(x, wt, idx).view(np.uint16),
                          full[idx].view(np.uint16))


def test_fp64_direct_matches_torch(oracle):
    for seed, shp, stride, pad in [(1, (2, 5, 9, 7, 3), 1, 1), (2, (1, 3, 8, 8, 4), 2, 1),
                                   (3, (1, 4, 6, 6, 2), 1, 0)]:
        n, c, h, w, k = shp
        x, wt = rng_inputs(seed, n, c, h, w, k)
        got = oracle.conv2d_f64(x, wt, stride, stride, pad, pad)
        ref = torch.nn.functional.conv2d(torch.from_numpy(x.astype(np.float64)),
                                         torch.from_numpy(wt.astype(np.float64)),
                                         stride=stride, padding=pad).numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_fp64_direct_general_matches_torch(oracle):
    """Unequal filter sides, strides and paddings (the direct fallback's
    problems): a transposed index or a swapped stride fails this."""
    rng = np.random.default_rng(11)
    for (n, c, h, w, k, r, s, u, v, p, q) in [(2, 3, 9, 7, 4, 5, 3, 2, 1, 2, 0),
                                              (1, 2, 6, 11, 3, 1, 4, 1, 3, 0, 2),
                                              (1, 4, 8, 8, 2, 7, 7, 1, 1, 3, 3)]:
        x = rng.standard_normal((n, c, h, w)).astype(np.float16)
        wt = rng.standard_normal((k, c, r, s)).astype(np.float16)
        got = oracle.conv2d_f64(x, wt, u, v, p, q)
        ref = torch.nn.functional.conv2d(torch.from_numpy(x.astype(np.float64)),
                                         torch.from_numpy(wt.astype(np.float64)),
                                         stride=(u, v), padding=(p, q)).numpy()
        assert got.shape == ref.shape
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_fp64_sample_matches_full(oracle):
    x, wt = rng_inputs(4, 2, 6, 7, 9, 4)
    full = oracle.conv2d_f64(x, wt).ravel()
    idx = np.array([0, 5, 99, full.size - 1], dtype=np.int64)
    np.testing.assert_allclose(oracle.conv2d_f64_sample(x, wt, idx), full[idx], rtol=0, atol=0)


def test_fp64_brute_force_tiny(oracle):
    x, wt = rng_inputs(9, 1, 2, 3, 3, 1)
    xd, wd = x.astype(np.float64), wt.astype(np.float64)
    pad = np.zeros((1, 2, 5, 5))
    pad[:, :, 1:4, 1:4] = xd
    exp = np.zeros((1, 1, 3, 3))
    for i in range(3):
        for j in range(3):
            exp[0, 0, i, j] = np.sum(pad[0, :, i:i + 3, j:j + 3] * wd[0])
    np.testing.assert_allclose(oracle.conv2d_f64(x, wt), exp, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("m", [2, 4])
def test_winograd_tables_closed_form(oracle, m):
    G, BT, AT = (oracle.wino_table(m, i) for i in range(3))
    T = m + 2
    # 1-D identity in exact arithmetic: A^T[(G g) . (B^T d)] = correlate(d, g)
    rng = np.random.default_rng(m)
    for _ in range(5):
        g = rng.standard_normal(3)
        d = rng.standard_normal(T)
        y = AT @ ((G @ g) * (BT @ d))
        exp = np.array([np.dot(d[i:i + 3], g) for i in range(m)])
        np.testing.assert_allclose(y, exp, rtol=1e-12, atol=1e-12)
    # interpolation points: F(2,3) {0, 1, -1, inf}; F(4,3) {0, -1, 1, 1/2, -1/2, inf}
    assert G[0, 0] != 0 and G[-1, 2] == 1 and G[-1, 0] == 0


@pytest.mark.parametrize("m", [2, 4])
def test_winograd_stage_model_equals_direct(oracle, m):
    for seed, (n, c, h, w, k) in [(11, (2, 3, 7, 9, 4)), (12, (1, 5, 14, 14, 3)),
                                  (13, (3, 2, 5, 6, 2))]:
        x, wt = rng_inputs(seed, n, c, h, w, k)
        y = oracle.wino_model(m, x, wt, quant_uv="none")
        ref = oracle.conv2d_f64(x, wt)
        np.testing.assert_allclose(y, ref, rtol=1e-10, atol=1e-10)


def test_stage_layouts(oracle):
    """U[t][k][c], V[t][c][p] with p = (n, th, tw) row-major, M[t][k][p]."""
    x, wt = rng_inputs(21, 2, 3, 6, 10, 2)
    V = oracle.wino_input(2, x)
    assert V.shape == (16, 3, 2 * 3 * 5)
    # tile p = (n=1, th=2, tw=3): top-left input pixel (4-1, 6-1) = (3, 5)
    p = (1 * 3 + 2) * 5 + 3
    d = np.zeros((4, 4))
    xp = np.zeros((8, 12))
    xp[1:7, 1:11] = x[1, 2].astype(np.float64)
    d[:] = xp[4:8, 6:10]
    BT = oracle.wino_table(2, 1)
    np.testing.assert_allclose(V[:, 2, p].reshape(4, 4), BT @ d @ BT.T, rtol=1e-14)
    U = oracle.wino_filter(4, wt)
    G = oracle.wino_table(4, 0)
    np.testing.assert_allclose(U[:, 1, 2].reshape(6, 6),
                               G @ wt[1, 2].astype(np.float64) @ G.T, rtol=1e-14)


def test_fp16_codec_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-9, 5, 2000),
                           [65504.0, 65519.99, 65520.0, 2.0 ** -25, 3 * 2.0 ** -26, -0.0]])
    for v in vals:
        with np.errstate(over="ignore"):
            exp = np.array(v, np.float64).astype(np.float16).view(np.uint16)
        assert oracle.lib().oracle_d2h(float(v)) == exp, v
    for b in range(0, 65536, 7):
        h = np.array(b, np.uint16).view(np.float16)
        if np.isfinite(h):
            assert oracle.lib().oracle_h2d(b) == float(h)


def test_e4m3_codec(oracle):
    # every finite code decodes and re-encodes to itself; values match torch
    codes = [b for b in range(256) if (b & 0x7f) != 0x7f]
    dec = np.array([oracle.e4m3_decode(b) for b in codes])
    tref = torch.tensor(codes, dtype=torch.uint8).view(torch.float8_e4m3fn).to(torch.float64).numpy()
    np.testing.assert_array_equal(dec, tref)
    for b, v in zip(codes, dec):
        assert oracle.e4m3_encode(v) == (b if v != 0 or b == 0 else oracle.e4m3_encode(v))
    assert max(dec) == 448.0
    # round-to-nearest-even against torch's cast, inside the finite range
    rng = np.random.default_rng(8)
    xs = np.concatenate([rng.uniform(-448, 448, 4000), rng.uniform(-1, 1, 4000) * 2.0 ** -6,
                         (dec[:-1] + np.diff(dec) / 2)[np.isfinite(dec[:-1])]])
    xs = xs[np.abs(xs) < 448]
    tq = torch.from_numpy(xs).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    oq = np.array([oracle.e4m3_encode(v) for v in xs], dtype=np.uint8)
    # torch keeps the sign of zero; compare values
    np.testing.assert_array_equal(
        np.array([oracle.e4m3_decode(b) for b in oq]),
        torch.from_numpy(tq).view(torch.float8_e4m3fn).to(torch.float64).numpy())
    # saturation
    assert oracle.e4m3_encode(1e6) == 0x7e and oracle.e4m3_encode(-500.0) == 0xfe
