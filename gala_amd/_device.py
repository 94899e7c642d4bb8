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
    """numpy / torch -> contiguous device tensor, float64 unless dtype is given
    (a torch tensor of another dtype is converted, as a numpy array is: the
    kernels read these buffers as the C ABI's types)."""
    t = require_gpu()
    if isinstance(a, t.Tensor):
        x = a.to(device="cuda", dtype=dtype or t.float64)
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


def _shape(a):
    return tuple(a.shape) if hasattr(a, "shape") else np.shape(a)


def check_w0(w0):
    """Raise ValueError unless w0 (numpy or torch, any device) is (6, N)."""
    shp = _shape(w0)
    if len(shp) != 2 or shp[0] != 6:
        raise ValueError(f"w0 must have shape (6, N), got {shp}")
    return shp[1]


def check_per_orbit(v, N, name):
    """Raise ValueError unless v is a scalar or has shape (N,)."""
    if v is None:
        return
    shp = _shape(v)
    if shp not in ((), (N,)):
        raise ValueError(f"{name} must be a scalar or have length {N}, got shape {shp}")


def check_permutation(order, N):
    """Raise ValueError unless order (numpy or torch, any device) is a
    permutation of 0..N-1: a short, out-of-range or repeated order would
    leave orbits unvisited or read past the state arrays."""
    shp = _shape(order)
    if shp != (N,):
        raise ValueError(f"order must have length {N}, got shape {shp}")
    if N == 0:
        return
    if hasattr(order, "device") and hasattr(order, "bincount"):
        t = torch()
        o = order.to(t.int64)
        lo, hi = int(o.min().item()), int(o.max().item())
        ok = lo >= 0 and hi < N and bool((t.bincount(o, minlength=N) == 1).all().item())
    else:
        o = np.asarray(order).astype(np.int64)
        lo, hi = int(o.min()), int(o.max())
        ok = lo >= 0 and hi < N and bool(np.all(np.bincount(o, minlength=N) == 1))
    if not ok:
        raise ValueError("order must be a permutation of 0..N-1 (entries out of range or "
                         "repeated)")


def phase_space(w0):
    """w0 as a contiguous float64 (6, N) device tensor (ValueError otherwise)."""
    check_w0(w0)
    return to_dev(w0)


def per_orbit(v, N, name, dtype=None):
    """A per-orbit array of length N (scalars broadcast) as a device tensor."""
    t = torch()
    if v is None:
        return None
    check_per_orbit(v, N, name)
    if isinstance(v, t.Tensor):
        return to_dev(v.expand(N) if v.dim() == 0 else v, dtype=dtype)
    v = np.asarray(v)
    return to_dev(np.broadcast_to(v, (N,)) if v.ndim == 0 else v, dtype=dtype)


def permutation(order, N):
    """order as an int32 device tensor, checked by check_permutation."""
    check_permutation(order, N)
    return to_dev(order, dtype=torch().int32)
