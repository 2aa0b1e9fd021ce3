"""The device e4m3 quantizer (hardware v_cvt_pk_fp8_f32 behind a +-448
clamp, used by every fp8 transform kernel) against the oracle's integer
quantizer, on every code, every rounding tie, its neighbours, saturation and
random values.  Bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _codes(oracle):
    vals = [oracle.e4m3_decode(b) for b in range(256) if (b & 0x7f) != 0x7f]
    return np.array(sorted(set(vals)), dtype=np.float64)


def test_quantizer_matches_oracle(wc, gpu, oracle):
    import torch
    codes = _codes(oracle)
    pos = codes[codes >= 0]
    mids = (pos[:-1] + pos[1:]) / 2  # exact ties: round to even
    f32 = lambda a: np.asarray(a, dtype=np.float32)
    nxt = np.nextafter(f32(mids), f32(np.inf))
    prv = np.nextafter(f32(mids), f32(-np.inf))
    rng = np.random.default_rng(8)
    rnd = np.exp2(rng.uniform(-14, 9.5, 20000)).astype(np.float32)
    sat = f32([448, 448.5, 463.9, 464, 465, 480, 1e4, 3e38])
    x = np.concatenate([f32(codes), f32(mids), nxt, prv, rnd, sat, f32([0.0, 2.0 ** -10, 2.0 ** -9 * 0.5])])
    x = np.concatenate([x, -x]).astype(np.float32)
    dx = torch.from_numpy(x).to("cuda:0")
    dq = torch.zeros(x.size, dtype=torch.uint8, device="cuda:0")
    torch.cuda.synchronize()
    wc.check(wc.lib().wino_quantize_e4m3(dx.data_ptr(), dq.data_ptr(), x.size), "quantize")
    got = dq.cpu().numpy()
    want = np.array([oracle.e4m3_encode(float(v)) for v in x], dtype=np.uint8)
    bad = np.nonzero(got != want)[0]
    print(f"{x.size} values, {bad.size} mismatches", flush=True)
    assert bad.size == 0, [(float(x[i]), int(got[i]), int(want[i])) for i in bad[:10]]
