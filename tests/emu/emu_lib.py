"""ctypes access to the CPU-emulated backend (tests/emu/build_emu.py) with
numpy arrays standing in for device buffers."""
import ctypes
import os

import numpy as np

from phrike_amd import _lib as _real

HERE = os.path.dirname(os.path.abspath(__file__))


def load(asan=False, defines=(), tag=""):
    from build_emu import build
    path = build(asan, defines, tag)
    lib = ctypes.CDLL(path)
    for name, (res, args) in _real._SIGNATURES.items():
        if name.startswith("ph_gll"):
            continue
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


def P(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


class Emu3D:
    def __init__(self, lib, nx, ny, nz, L=(2 * np.pi,) * 3, dtype=np.float64, dealias=True,
                 filt=(1, 16, 36.0), gamma=1.4):
        self.lib, self.dtype = lib, dtype
        self.shape = (nz, ny, nx)
        h = ctypes.c_void_p()
        code = 0 if dtype == np.float64 else 1
        self._chk(lib.ph_plan3d_create(ctypes.byref(h), nx, ny, nz, *L, code, int(dealias), *filt, gamma))
        self.h = h
        n = ctypes.c_size_t()
        self._chk(lib.ph_plan3d_workspace_bytes(h, ctypes.byref(n)))
        self.ws = np.zeros(n.value // 8 + 64, dtype=np.float64)

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.ph_last_error().decode())

    def deriv(self, axis, f):
        f = np.ascontiguousarray(f, dtype=self.dtype)
        out = np.zeros_like(f)
        nf = f.size // int(np.prod(self.shape))
        self._chk(self.lib.ph_deriv3d(self.h, axis, nf, P(f), P(out), P(self.ws), None))
        return out

    def rhs(self, U):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        out = np.zeros_like(U)
        self._chk(self.lib.ph_rhs3d(self.h, P(U), P(out), P(self.ws), None))
        return out

    def step(self, U, dt, flags=0, dt_dev=None):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        out = np.zeros_like(U)
        self._chk(self.lib.ph_step3d_ex(self.h, P(U), P(out), dt, P(dt_dev), flags, P(self.ws), None))
        return out


class Emu2D:
    def __init__(self, lib, nx, ny, batch=1, L=(1.0, 1.0), dtype=np.float64, dealias=True,
                 filt=(1, 16, 36.0), gamma=1.4):
        self.lib, self.dtype = lib, dtype
        h = ctypes.c_void_p()
        code = 0 if dtype == np.float64 else 1
        self._chk(lib.ph_plan2d_create(ctypes.byref(h), nx, ny, *L, code, int(dealias), *filt, gamma, batch))
        self.h = h
        n = ctypes.c_size_t()
        self._chk(lib.ph_plan2d_workspace_bytes(h, ctypes.byref(n)))
        self.ws = np.zeros(n.value // 8 + 64, dtype=np.float64)

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.ph_last_error().decode())

    def step(self, U, dt, filt=True, active=None, out=None, noalloc=False, dt_dev=None):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        out = np.zeros_like(U) if out is None else out
        nb = U.shape[0] if U.ndim == 4 else 1
        if noalloc:
            self._chk(self.lib.ph_step2d_noalloc(self.h, nb, P(U), P(out), dt, P(dt_dev), P(active), int(filt),
                                                 None))
        else:
            self._chk(self.lib.ph_step2d(self.h, nb, P(U), P(out), dt, P(dt_dev), P(active), int(filt),
                                         P(self.ws), None))
        return out

    def rhs(self, U):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        out = np.zeros_like(U)
        nb = U.shape[0] if U.ndim == 4 else 1
        self._chk(self.lib.ph_rhs2d(self.h, nb, P(U), P(out), P(self.ws), None))
        return out

    def dt(self, U, cfl, active=None):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        nb = U.shape[0] if U.ndim == 4 else 1
        m, d = np.zeros(nb), np.zeros(nb)
        self._chk(self.lib.ph_max_wave_speed2d(self.h, nb, P(U), P(m), P(d), cfl, P(active), P(self.ws), None))
        return m, d


class EmuHybrid:
    def __init__(self, lib, nx, ny, D, Fn, dy_min, batch=1, lx=1.0, dtype=np.float64, dealias=True,
                 filt=(1, 16, 36.0), gamma=1.4, mode=0, ybc=1):
        self.lib, self.dtype = lib, dtype
        self.D = np.ascontiguousarray(D, dtype=np.float64)
        self.Fn = np.ascontiguousarray(Fn, dtype=np.float64)
        h = ctypes.c_void_p()
        code = 0 if dtype == np.float64 else 1
        self._chk(lib.ph_planh_create(ctypes.byref(h), nx, ny, lx, dy_min, code, int(dealias), *filt, gamma,
                                      batch, self.D.ctypes.data, self.Fn.ctypes.data, mode, ybc))
        self.h = h
        n = ctypes.c_size_t()
        self._chk(lib.ph_planh_workspace_bytes(h, ctypes.byref(n)))
        self.ws = np.zeros(n.value // 8 + 64, dtype=np.float64)

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.ph_last_error().decode())

    def rhs(self, U):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        out = np.zeros_like(U)
        nb = U.shape[0] if U.ndim == 4 else 1
        self._chk(self.lib.ph_rhs2dh(self.h, nb, P(U), P(out), P(self.ws), None))
        return out

    def step(self, U, dt, filt=True, active=None, out=None):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        out = np.zeros_like(U) if out is None else out
        nb = U.shape[0] if U.ndim == 4 else 1
        self._chk(self.lib.ph_step2dh(self.h, nb, P(U), P(out), dt, None, P(active), int(filt), P(self.ws), None))
        return out

    def dt(self, U, cfl):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        nb = U.shape[0] if U.ndim == 4 else 1
        out = np.zeros(nb, dtype=np.float64)
        self._chk(self.lib.ph_max_wave_speed2dh(self.h, nb, P(U), None, P(out), cfl, P(self.ws), None))
        return out


class EmuMHD3D:
    """The 3D ideal-MHD plan (ph_planmhd3d_*) on the CPU emulation."""

    def __init__(self, lib, nx, ny, nz, L=(2 * np.pi,) * 3, dtype=np.float64, dealias=True,
                 filt=(1, 16, 36.0), gamma=5.0 / 3.0):
        self.lib, self.dtype = lib, dtype
        self.shape = (nz, ny, nx)
        h = ctypes.c_void_p()
        code = 0 if dtype == np.float64 else 1
        self._chk(lib.ph_planmhd3d_create(ctypes.byref(h), nx, ny, nz, *L, code, int(dealias), *filt, gamma))
        self.h = h
        n = ctypes.c_size_t()
        self._chk(lib.ph_planmhd3d_workspace_bytes(h, ctypes.byref(n)))
        self.ws = np.zeros(n.value // 8 + 64, dtype=np.float64)
        self.sc = np.zeros(8, dtype=np.float64)

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.ph_last_error().decode())

    def rhs(self, U):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        out = np.zeros_like(U)
        self._chk(self.lib.ph_rhs_mhd3d(self.h, P(U), P(out), P(self.ws), None))
        return out

    def project(self, U):
        V = np.array(U, dtype=self.dtype, copy=True, order="C")
        self._chk(self.lib.ph_project_mhd3d(self.h, P(V), P(self.ws), None))
        return V

    def step(self, U, dt, flags=0):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        out = np.zeros_like(U)
        self._chk(self.lib.ph_step_mhd3d_ex(self.h, P(U), P(out), dt, flags, P(self.ws), None))
        return out

    def max_speed_dt(self, U, cfl):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        m, d = np.zeros(1), np.zeros(1)
        self._chk(self.lib.ph_max_wave_speed_mhd3d(self.h, P(U), P(m), P(d), cfl, P(self.ws), None))
        return float(m[0]), float(d[0])

    def sums(self, U):
        U = np.ascontiguousarray(U, dtype=self.dtype)
        self._chk(self.lib.ph_conserved_sums_mhd3d(self.h, P(U), P(self.sc), P(self.ws), None))
        return self.sc[:7].copy()
