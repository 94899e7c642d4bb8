"""ctypes binding of libgala_amd.so (the C ABI declared in include/gala_amd.h).

The library must be built (``python -m gala_amd.build``); there is no CPU
fallback: every entry point here runs the HIP kernels, and using them without
the library or without a GPU raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgala_amd.so")

MAX_COMPONENTS = 8
MAX_BODIES = 64   # GAMD_MAX_BODIES: bodies of one DOP853 N-body system
MAX_MASSIVE = 8   # GAMD_MAX_MASSIVE: bodies carrying a potential
SCF_MAX_NMAX = 20
SCF_MAX_LMAX = 10

NULL, HERNQUIST, PLUMMER, MIYAMOTONAGAI, NFW, SCF = range(6)

STATUS_NAMES = {0: "GAMD_OK", -1: "GAMD_ERR_INVALID", -2: "GAMD_ERR_HIP",
                -3: "GAMD_ERR_NOMEM", -4: "GAMD_ERR_UNSUPPORTED"}


class Component(C.Structure):
    """gamd_component (include/gala_amd.h)."""
    _fields_ = [("type", C.c_int32), ("shift_rotate", C.c_int32), ("nmax", C.c_int32),
                ("lmax", C.c_int32), ("pars", C.c_double * 8), ("q0", C.c_double * 3),
                ("R", C.c_double * 9), ("S", C.c_void_p), ("T", C.c_void_p)]


class MLPNorm(C.Structure):
    """gamd_mlp_norm (include/gala_amd.h)."""
    _fields_ = [("mu", C.c_double), ("inv_sigma", C.c_double),
                ("lnr0", C.c_double), ("inv_dlnr", C.c_double), ("pad_r", C.c_double),
                ("scale", C.c_void_p), ("gm", C.c_double), ("rs", C.c_double),
                ("phi_c", C.c_double)]


MLP_KNOTS = 32  # GAMD_MLP_KNOTS


# every symbol include/gala_amd.h declares, with its signature
_vp, _sz, _d, _i64, _i32, _int = C.c_void_p, C.c_size_t, C.c_double, C.c_int64, C.c_int32, C.c_int
_f, _nm = C.c_float, C.POINTER(MLPNorm)
SIGNATURES = {
    "gamd_version": (C.c_char_p, []),
    "gamd_potential_create": (_int, [C.POINTER(Component), _int, C.POINTER(_vp)]),
    "gamd_potential_destroy": (_int, [_vp]),
    "gamd_potential_value": (_int, [_vp, _d, _vp, _sz, _vp, _vp]),
    "gamd_potential_density": (_int, [_vp, _d, _vp, _sz, _vp, _vp]),
    "gamd_potential_gradient": (_int, [_vp, _d, _vp, _sz, _vp, _vp]),
    "gamd_potential_hessian": (_int, [_vp, _d, _vp, _sz, _vp, _vp]),
    "gamd_potential_d2_dr2": (_int, [_vp, _d, _vp, _sz, _vp, _vp]),
    "gamd_hamiltonian_energy": (_int, [_vp, _d, _vp, _sz, _vp, _vp]),
    "gamd_leapfrog": (_int, [_vp, _vp, _sz, _d, _d, _i64, _vp, _i64, _vp, _vp, _vp]),
    "gamd_leapfrog_resume": (_int, [_vp, _vp, _sz, _d, _i64, _vp, _vp, _vp]),
    "gamd_leapfrog_stream_workspace_bytes": (_i64, [_sz, _i64]),
    "gamd_leapfrog_stream": (_int, [_vp, _vp, _sz, _d, _d, _i64, _i64, _i64, _int, _vp, _vp, _vp,
                                    _sz, _vp]),
    "gamd_ruth4": (_int, [_vp, _vp, _sz, _d, _d, _i64, _i64, _vp, _vp, _vp]),
    "gamd_dop853": (_int, [_vp, _vp, _sz, _vp, _vp, _vp, _d, _d, _i64, _i64, _d, _vp, _sz,
                           _vp, _vp, _vp, _vp, _vp, _i64, _int, _vp]),
    "gamd_dop853_noalloc": (_int, [_vp, _vp, _sz, _vp, _vp, _vp, _d, _d, _i64, _i64, _d, _vp,
                                   _sz, _vp, _vp, _vp, _vp, _vp, _i64, _int, _vp]),
    "gamd_dop853_dense_workspace_bytes": (_i64, [_sz, _i64]),
    "gamd_dop853_dense_stream": (_int, [_vp, _vp, _sz, _vp, _vp, _vp, _d, _d, _i64, _i64, _d,
                                        _vp, _sz, _i64, _int, _vp, _vp, _vp, _vp, _vp, _i64,
                                        _int, _vp, _sz, _vp]),
    "gamd_dop853_batched": (_int, [_vp, _vp, _sz, _d, _d, _d, _d, _d, _i64, _i64, _d, _int, _vp,
                                   _sz, _vp, _vp, _vp, _vp, _vp]),
    "gamd_nbody_leapfrog": (_int, [_vp, _vp, _int, _vp, _sz, _d, _d, _i64, _i64, _vp, _vp,
                                   _vp]),
    "gamd_nbody_ruth4": (_int, [_vp, _vp, _int, _vp, _sz, _d, _d, _i64, _i64, _vp, _vp, _vp]),
    "gamd_nbody_dop853": (_int, [_vp, _vp, _int, _vp, _sz, _d, _d, _d, _d, _d, _i64, _i64, _d,
                                 _vp, _sz, _vp, _vp, _vp, _vp, _vp]),
    "gamd_mockstream_leapfrog": (_int, [_vp, _vp, _int, _vp, _sz, _d, _d, _vp, _vp, _sz, _d,
                                        _vp, _vp]),
    "gamd_mockstream_dop853": (_int, [_vp, _vp, _int, _vp, _sz, _vp, _vp, _vp, _vp, _d, _d, _d,
                                      _d, _i64, _vp, _vp, _vp, _vp, _vp]),
    "gamd_mlp_param_count": (_i64, []),
    "gamd_mlp_bf16_count": (_i64, []),
    "gamd_mlp_workspace_bytes": (_i64, [_i64]),
    "gamd_mlp_sync_bf16": (_int, [_vp, _vp, _vp]),
    "gamd_mlp_forward": (_int, [_vp, _vp, _nm, _vp, _sz, _vp, _vp, _vp]),
    "gamd_mlp_make_targets": (_int, [_nm, _vp, _vp, _vp, _sz, _vp, _vp, _vp]),
    "gamd_mlp_gradients": (_int, [_vp, _vp, _vp, _vp, _sz, _vp, _vp, _vp, _sz, _vp]),
    "gamd_mlp_adam": (_int, [_vp, _vp, _vp, _vp, _i64, _f, _f, _f, _f, _vp, _vp]),
    "gamd_surrogate_leapfrog_step": (_int, [_vp, _vp, _nm, _vp, _vp, _vp, _vp, _vp, _vp, _sz,
                                            _d, _int, _i32, _d, _vp]),
}

_LIB = None


def load():
    """Load libgala_amd.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m gala_amd.build` (hipcc, gfx950)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _LIB = lib
    return _LIB


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {STATUS_NAMES.get(rc, rc)}")


def make_components(components):
    """components: list of dicts with type, pars, q0, R, shift_rotate and, for
    SCF, S, T (numpy, (nmax+1, lmax+1, lmax+1)).  Returns (ctypes array, keepalive)."""
    n = len(components)
    if n > MAX_COMPONENTS:
        raise ValueError(f"at most {MAX_COMPONENTS} components, got {n}")
    arr = (Component * max(n, 1))()
    keep = []
    for i, c in enumerate(components):
        cc = arr[i]
        cc.type = c["type"]
        cc.shift_rotate = int(c["shift_rotate"])
        for k, v in enumerate(c["pars"]):
            cc.pars[k] = float(v)
        for k in range(3):
            cc.q0[k] = float(c["q0"][k])
        for k, v in enumerate(np.asarray(c["R"], dtype=np.float64).ravel()):
            cc.R[k] = float(v)
        if c["type"] == SCF:
            S = np.ascontiguousarray(c["S"], dtype=np.float64)
            T = np.ascontiguousarray(c["T"], dtype=np.float64)
            cc.nmax, cc.lmax = S.shape[0] - 1, S.shape[1] - 1
            cc.S, cc.T = S.ctypes.data, T.ctypes.data
            keep += [S, T]
    return arr, keep


def body_array(handles):
    """gamd_potential* array (host) for the body arguments of the N-body entry
    points: one handle per body, None for a massless body."""
    arr = (C.c_void_p * max(len(handles), 1))()
    for i, h in enumerate(handles):
        arr[i] = h
    return arr
