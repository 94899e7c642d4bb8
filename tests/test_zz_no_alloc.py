"""gamd_dop853_noalloc against gamd_dop853 through the C ABI: the same
kernels with the work-queue counter in the library's fixed device array
instead of a per-call allocation, so results must agree bit for bit.  Two
calls in a row, the second on a smaller input (the array keeps the first
call's counter, which the second must clear), with and without a queue
order and dense output."""
import ctypes

import numpy as np
import pytest

from tests.helpers import mw_oracle, mw_product, vcirc_oracle

pytestmark = pytest.mark.gpu


def _call(fn, pot, w0, T, order=None, t_out=None):
    import torch
    from gala_amd import _device as dev
    from gala_amd import _lib
    lib = _lib.load()
    N = w0.shape[1]
    w0d = dev.to_dev(w0)
    t0 = torch.zeros(N, dtype=torch.float64, device="cuda")
    t1 = dev.to_dev(T)
    h0 = torch.ones(N, dtype=torch.float64, device="cuda")
    wfin = torch.full((6, N), np.nan, dtype=torch.float64, device="cuda")
    status = torch.full((N,), 99, dtype=torch.int32, device="cuda")
    stats = torch.full((N, 4), -1, dtype=torch.int64, device="cuda")
    ordd = None if order is None else dev.to_dev(order, dtype=torch.int32)
    tout = None if t_out is None else dev.to_dev(t_out)
    wout = None if t_out is None else torch.full((6, len(t_out), N), np.nan,
                                                  dtype=torch.float64, device="cuda")
    rc = getattr(lib, fn)(pot.c_handle, w0d.data_ptr(), N, t0.data_ptr(), t1.data_ptr(),
                          h0.data_ptr(), 1e-8, 1e-10, 0, 0, 0.0, dev.ptr(tout),
                          0 if tout is None else len(t_out), dev.ptr(wout), wfin.data_ptr(),
                          status.data_ptr(), stats.data_ptr(), dev.ptr(ordd), 0, 0,
                          dev.stream_ptr())
    assert rc == 0, (fn, rc)
    torch.cuda.synchronize()
    out = [dev.to_host(x) for x in (wfin, status, stats)]
    if wout is not None:
        out.append(dev.to_host(wout))
    return out


def test_noalloc_matches_allocating_entry_point():
    from gala_amd import workloads
    print("gamd_dop853_noalloc vs gamd_dop853: 3000 then 700 orbits", flush=True)
    pot = mw_product()
    w0, T = workloads.composite_dop853(vcirc_oracle(mw_oracle()), n=3000, seed=31)
    T = T * 0.05
    for n, order in ((3000, None), (700, "shuffled"), (700, None)):
        w, t = np.ascontiguousarray(w0[:, :n]), np.ascontiguousarray(T[:n])
        o = (np.random.default_rng(n).permutation(n).astype(np.int32)
             if order == "shuffled" else None)
        ref = _call("gamd_dop853", pot, w, t, order=o)
        got = _call("gamd_dop853_noalloc", pot, w, t, order=o)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
        assert np.all(got[1] == 1)
    # dense output, smaller again
    t_out = np.linspace(0.0, 50.0, 11)
    w, t = np.ascontiguousarray(w0[:, :300]), np.full(300, 50.0)
    ref = _call("gamd_dop853", pot, w, t, t_out=t_out)
    got = _call("gamd_dop853_noalloc", pot, w, t, t_out=t_out)
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_noalloc_rejects_what_gamd_dop853_rejects():
    from gala_amd import _lib
    lib = _lib.load()
    assert lib.gamd_dop853_noalloc(None, None, 1, None, None, None, 1e-8, 1e-10, 0, 0, 0.0, None,
                                   0, None, None, None, None, None, 0, 0, None) == -1
    fake = ctypes.c_void_p(8)  # never dereferenced: the block check fails first
    assert lib.gamd_dop853_noalloc(fake, fake, 1, fake, fake, None, 1e-8, 1e-10, 0, 0, 0.0, None,
                                   0, None, fake, None, None, None, 0, 96, None) == -1
