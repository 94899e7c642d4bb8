"""Device-memory plumbing: PyTorch-ROCm tensors hold the HBM buffers handed
to the C ABI, and torch's current HIP stream orders the calls."""
import numpy as np


def torch():
    import torch as _t
    return _t


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("gala_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
    return t


def stream_ptr():
    t = require_gpu()
    return t.cuda.current_stream().cuda_stream


def to_dev(a, dtype=None):
    """numpy / torch -> contiguous device tensor (float64 unless dtype given)."""
    t = require_gpu()
    if isinstance(a, t.Tensor):
        x = a.to(device="cuda", dtype=dtype or a.dtype)
        return x.contiguous()
    np_dtype = {None: np.float64, t.float64: np.float64, t.float32: np.float32,
                t.int64: np.int64, t.int32: np.int32}[dtype]
    a = np.ascontiguousarray(a, dtype=np_dtype)
    if not a.flags.writeable:
        a = a.copy()
    return t.from_numpy(a).to("cuda")


def empty(shape, dtype=None):
    t = require_gpu()
    return t.empty(shape, dtype=dtype or t.float64, device="cuda")


def ptr(x):
    return None if x is None else x.data_ptr()


def to_host(x):
    return x.detach().cpu().numpy()
