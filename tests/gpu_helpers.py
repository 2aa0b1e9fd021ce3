"""Device-memory plumbing for the GPU tests (torch tensors on cuda:0)."""
import numpy as np
import torch


def to_dev(a: np.ndarray) -> torch.Tensor:
    a = np.ascontiguousarray(a)
    if a.dtype == np.float16:
        return torch.from_numpy(a.view(np.int16)).view(torch.float16).to("cuda:0")
    return torch.from_numpy(a).to("cuda:0")


def empty_f16(shape) -> torch.Tensor:
    return torch.full(tuple(shape), float("nan"), dtype=torch.float16, device="cuda:0")


def to_host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    if t.dtype == torch.float16:
        return t.detach().cpu().view(torch.int16).numpy().view(np.float16)
    return t.detach().cpu().numpy()


def read_bytes(ptr: int, nbytes: int) -> np.ndarray:
    """Copies nbytes of device memory at ptr to a host uint8 array."""
    import winoconv
    torch.cuda.synchronize()
    out = np.empty(nbytes, dtype=np.uint8)
    winoconv.check(winoconv.lib().wino_memcpy_d2h(out.ctypes.data, ptr, nbytes), "d2h")
    return out


def run_conv(wc, x: np.ndarray, wt: np.ndarray, algo, precision=0, resident=False, repeat=1):
    """Runs the operator through the C ABI; returns (y fp16 host, info)."""
    n, c, h, w = x.shape
    k = wt.shape[0]
    dx, dw = to_dev(x), to_dev(wt)
    dy = empty_f16((n, k, h, w))
    op = wc.Conv2d(dx.data_ptr(), dw.data_ptr(), dy.data_ptr(), n, c, h, w, k, algo=algo,
                   precision=precision, resident=resident)
    try:
        for _ in range(repeat):
            op()
        y = to_host(dy)
        info = op.info()
    finally:
        op.destroy()
    return y, info
