"""ctypes binding of libphrike_hip.so (include/phrike_hip.h).

The library is built in-tree by ``python -m phrike_amd.build``.  There is no
fallback: if the library is missing or cannot be loaded, every entry point
raises, so a HIP-backend call never silently runs anything else.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_size_t, c_void_p

LIB_PATH = os.environ.get("PHRIKE_HIP_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                           "libphrike_hip.so")

PH_F64 = 0
PH_F32 = 1
PH_STEP_BANDLIMITED = 1
PH_STEP_SOLENOIDAL = 2
PH_STEP_CHAIN = 4
PH_STEP_CHAINED = 8
# device-resident run loop scalars (ph_run3d / ph_run2d)
PH_RS_T, PH_RS_DT, PH_RS_TEND, PH_RS_TSTOP, PH_RS_DTNEXT, PH_RS_DTUSED, PH_RS_N = range(7)
PH_RF_ACTIVE, PH_RF_STEPS, PH_RF_N = range(3)
# monitoring reductions (ph_monitor3d / ph_monitor2d / ph_monitor_mhd3d)
(PH_MON_MASS, PH_MON_RV2, PH_MON_VMAX, PH_MON_VMIN, PH_MON_RMAX, PH_MON_RMIN, PH_MON_RS1, PH_MON_RS2,
 PH_MON_PMAX, PH_MON_PMIN, PH_MON_PSUM, PH_MON_B2, PH_MON_SHIFT, PH_MON_N) = range(14)

# symbol -> (restype, argtypes); mirrors include/phrike_hip.h
_SIGNATURES = {
    "ph_abi_version": (c_int, []),
    "ph_last_error": (ctypes.c_char_p, []),
    "ph_timing_enable": (c_int, [c_int]),
    "ph_timing_collect": (c_int, [ctypes.c_char_p, c_size_t]),
    "ph_plan3d_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_int, c_double, c_double, c_double,
                                 c_int, c_int, c_int, c_int, c_double, c_double]),
    "ph_plan3d_destroy": (c_int, [c_void_p]),
    "ph_plan3d_workspace_bytes": (c_int, [c_void_p, POINTER(c_size_t)]),
    "ph_rhs3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_deriv3d": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_step3d": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p]),
    "ph_step3d_ex": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_int, c_void_p, c_void_p]),
    "ph_run3d": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_double, c_int, c_void_p,
                         c_void_p]),
    "ph_max_wave_speed3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "ph_conserved_sums3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_fft3d": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_ifft3d": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_monitor3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_monitor2d": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_monitor_mhd3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_plan2d_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_double, c_double, c_int, c_int,
                                 c_int, c_int, c_double, c_double, c_int]),
    "ph_plan2d_destroy": (c_int, [c_void_p]),
    "ph_plan2d_workspace_bytes": (c_int, [c_void_p, POINTER(c_size_t)]),
    "ph_rhs2d": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_deriv2d": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_step2d": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_int,
                          c_void_p, c_void_p]),
    "ph_step2d_noalloc": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_int,
                                  c_void_p]),
    "ph_max_wave_speed2d": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                                    c_void_p, c_void_p]),
    "ph_run2d": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_double, c_void_p,
                         c_void_p]),
    "ph_conserved_sums2d": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_runs_clip": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_runs_advance": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_gll_apply": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p]),
    "ph_planmhd3d_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_int, c_double, c_double, c_double,
                                    c_int, c_int, c_int, c_int, c_double, c_double]),
    "ph_planmhd3d_destroy": (c_int, [c_void_p]),
    "ph_planmhd3d_workspace_bytes": (c_int, [c_void_p, POINTER(c_size_t)]),
    "ph_rhs_mhd3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_project_mhd3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_step_mhd3d": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "ph_step_mhd3d_ex": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_int, c_void_p, c_void_p]),
    "ph_max_wave_speed_mhd3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "ph_conserved_sums_mhd3d": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_planh_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_double, c_double, c_int, c_int,
                                c_int, c_int, c_double, c_double, c_int, c_void_p, c_void_p, c_int, c_int]),
    "ph_planh_destroy": (c_int, [c_void_p]),
    "ph_planh_workspace_bytes": (c_int, [c_void_p, POINTER(c_size_t)]),
    "ph_rhs2dh": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph_step2dh": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_int,
                           c_void_p, c_void_p]),
    "ph_max_wave_speed2dh": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                                     c_void_p]),
}

PH_GLL_F64 = 0
PH_GLL_SPLIT_F32 = 1
PH_GLL_SPLIT_F64IO = 2
PH_GLL_OFFSET = 8

_lib = None


class HipBackendError(RuntimeError):
    """A failed C-ABI call (the library's message is attached)."""


def load(path: str = LIB_PATH):
    """Load the library once; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HipBackendError(
            f"{path} is missing: build it with `python -m phrike_amd.build` "
            "(the hip backend has no fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def timing_enable(on: bool) -> None:
    call("ph_timing_enable", int(bool(on)))


def timing_collect() -> dict:
    """{kernel tag: (launches, total ms)} since the last collect."""
    buf = ctypes.create_string_buffer(1 << 16)
    call("ph_timing_collect", buf, len(buf))
    out = {}
    for line in buf.value.decode().splitlines():
        tag, n, ms = line.split()
        out[tag] = (int(n), float(ms))
    return out


def symbols():
    return list(_SIGNATURES)


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().ph_last_error()
        msg = msg.decode() if msg else ""
        raise HipBackendError(f"{what} failed ({rc}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def ptr(t) -> c_void_p:
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return c_void_p(0)
    return c_void_p(t.data_ptr())


def stream_handle(torch, device=None) -> c_void_p:
    return c_void_p(torch.cuda.current_stream(device).cuda_stream)
