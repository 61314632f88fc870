"""Grid objects of the ``hip`` backend.

Same constructor arguments, attributes and operator methods as the reference's
Grid3D / Grid2D (phrike/grid.py:775-1354, 1357-1744), with arrays held as
PyTorch-ROCm tensors on one GPU.  Operators call libphrike_hip.so through the
C ABI; nothing here computes a transform on the host.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from . import _lib

try:
    import torch
except Exception:  # pragma: no cover
    torch = None


def _require_hip(backend: str) -> None:
    if str(backend).lower() != "hip":
        raise ValueError(
            f"phrike_amd provides the 'hip' backend only (got backend={backend!r}); "
            "the 'numpy' and 'torch' backends are the reference's own")
    if torch is None or not torch.cuda.is_available():
        raise RuntimeError("backend='hip' needs a ROCm GPU visible to PyTorch")
    _lib.load()


def _filter_cfg(filter_params: Optional[Dict]):
    if filter_params and bool(filter_params.get("enabled", False)):
        return 1, int(filter_params.get("p", 8)), float(filter_params.get("alpha", 36.0))
    return 0, 8, 36.0


def _dtype_code(precision: str):
    if precision == "single":
        return _lib.PH_F32, torch.float32
    return _lib.PH_F64, torch.float64


def _axis_k(n: int, length: float) -> np.ndarray:
    return 2.0 * np.pi * np.fft.fftfreq(n, d=length / n)


def _mask_1d(n: int, dealias: bool) -> np.ndarray:
    if not dealias:
        return np.ones(n)
    return (np.abs(np.fft.fftfreq(n) * n) <= n // 3).astype(float)


def _sigma_1d(n: int, on: int, p: int, alpha: float) -> np.ndarray:
    if not on:
        return np.ones(n)
    m = np.abs(np.fft.fftfreq(n) * n)
    return np.exp(-alpha * (m / max(np.max(m), 1.0)) ** p)


class Grid3D:
    """Uniform periodic 3D grid; arrays are (..., Nz, Ny, Nx) on the GPU."""

    def __init__(self, Nx: int, Ny: int, Nz: int, Lx: float, Ly: float, Lz: float,
                 dealias: bool = True, filter_params: Optional[Dict] = None,
                 fft_workers: int = 1, backend: str = "hip",
                 torch_device: Optional[str] = None, precision: str = "double",
                 debug: bool = False):
        _require_hip(backend)
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.Lx, self.Ly, self.Lz = float(Lx), float(Ly), float(Lz)
        self.dealias = bool(dealias)
        self.filter_params = filter_params
        self.fft_workers = fft_workers
        self.backend = "hip"
        self.precision = "single" if precision == "single" else "double"
        self.debug = debug
        self.torch_device = torch_device or "cuda"
        self.device = torch.device(self.torch_device)
        self._dtype_code, self.dtype = _dtype_code(self.precision)
        self._use_torch = True  # arrays are torch tensors
        self.dx = self.Lx / self.Nx
        self.dy = self.Ly / self.Ny
        self.dz = self.Lz / self.Nz
        dev, dt = self.device, self.dtype
        self.x = torch.tensor(np.linspace(0.0, self.Lx, self.Nx, endpoint=False), dtype=dt, device=dev)
        self.y = torch.tensor(np.linspace(0.0, self.Ly, self.Ny, endpoint=False), dtype=dt, device=dev)
        self.z = torch.tensor(np.linspace(0.0, self.Lz, self.Nz, endpoint=False), dtype=dt, device=dev)
        self._kx_np = _axis_k(self.Nx, self.Lx)
        self._ky_np = _axis_k(self.Ny, self.Ly)
        self._kz_np = _axis_k(self.Nz, self.Lz)
        self.kx = torch.tensor(self._kx_np, dtype=dt, device=dev)
        self.ky = torch.tensor(self._ky_np, dtype=dt, device=dev)
        self.kz = torch.tensor(self._kz_np, dtype=dt, device=dev)
        ct = torch.complex64 if dt == torch.float32 else torch.complex128
        self.ikx = (1j * self.kx.to(ct))[None, None, :]
        self.iky = (1j * self.ky.to(ct))[None, :, None]
        self.ikz = (1j * self.kz.to(ct))[:, None, None]
        fon, fp, fa = _filter_cfg(filter_params)
        self._filter = (fon, fp, fa)
        self._gamma = 1.4
        self._plans = {}
        self._ws = None

    # ---- plan / workspace ----
    def _plan(self, gamma: Optional[float] = None):
        g = float(self._gamma if gamma is None else gamma)
        if g not in self._plans:
            h = _lib.c_void_p()
            fon, fp, fa = self._filter
            with torch.cuda.device(self.device):
                _lib.call("ph_plan3d_create", _lib.ctypes.byref(h), self.Nx, self.Ny, self.Nz,
                          self.Lx, self.Ly, self.Lz, self._dtype_code, int(self.dealias), fon, fp, fa, g)
            self._plans[g] = h
        return self._plans[g]

    def workspace(self, gamma: Optional[float] = None):
        plan = self._plan(gamma)
        n = _lib.c_size_t()
        _lib.call("ph_plan3d_workspace_bytes", plan, _lib.ctypes.byref(n))
        if self._ws is None or self._ws.numel() < n.value:
            self._ws = torch.empty(n.value, dtype=torch.uint8, device=self.device)
        return self._ws

    def __del__(self):
        try:
            lib = _lib.load()
            for h in getattr(self, "_plans", {}).values():
                lib.ph_plan3d_destroy(h)
            if getattr(self, "_fplan", None) is not None:
                lib.ph_plan3d_destroy(self._fplan)
        except Exception:
            pass

    # ---- helpers ----
    def _as_fields(self, f):
        if not isinstance(f, torch.Tensor):
            f = torch.as_tensor(np.asarray(f), device=self.device)
        if f.shape[-3:] != (self.Nz, self.Ny, self.Nx):
            raise ValueError(f"expected (..., {self.Nz}, {self.Ny}, {self.Nx}), got {tuple(f.shape)}")
        f = f.to(device=self.device, dtype=self.dtype).contiguous()
        return f

    def _apply(self, axis: int, f):
        f = self._as_fields(f)
        lead = f.shape[:-3]
        flat = f.reshape(-1, self.Nz, self.Ny, self.Nx)
        out = torch.empty_like(flat)
        ws = self.workspace()
        st = _lib.stream_handle(torch, self.device)
        plan = self._plan()
        for s in range(0, flat.shape[0], 5):
            e = min(s + 5, flat.shape[0])
            _lib.call("ph_deriv3d", plan, axis, e - s, _lib.ptr(flat[s:e]), _lib.ptr(out[s:e]),
                      _lib.ptr(ws), st)
        return out.reshape(*lead, self.Nz, self.Ny, self.Nx)

    # ---- reference operators (phrike/grid.py:1705-1744) ----
    def dx1(self, f):
        return self._apply(0, f)

    def dy1(self, f):
        return self._apply(1, f)

    def dz1(self, f):
        return self._apply(2, f)

    def apply_spectral_filter(self, f):
        return self._apply(-1, f)

    def xyz_mesh(self):
        Z, Y, X = torch.meshgrid(self.z, self.y, self.x, indexing="ij")
        return X, Y, Z

    # ---- spectral helpers of the reference's Grid3D (phrike/grid.py:1396-1401,
    # 1497-1518) ----
    @property
    def k2(self):
        """kx^2 + ky^2 + kz^2 on (Nz, Ny, Nx) (phrike/grid.py:1396-1400)."""
        return (self.kx[None, None, :] ** 2 + self.ky[None, :, None] ** 2) + self.kz[:, None, None] ** 2

    @property
    def k2_safe(self):
        """k2 with its zero (the mean mode) replaced by 1 (phrike/grid.py:1401)."""
        k2 = self.k2
        return torch.where(k2 == 0.0, torch.ones_like(k2), k2)

    def _apply_masks(self, F):
        """F *= dealias mask; F *= filter sigma, in place (phrike/grid.py:
        1515-1518); the tensor-product factors are applied axis by axis."""
        fon, fp, fa = self._filter
        for axis, n in ((-1, self.Nx), (-2, self.Ny), (-3, self.Nz)):
            f = _mask_1d(n, self.dealias) * _sigma_1d(n, fon, fp, fa)
            shape = [1, 1, 1]
            shape[axis] = n
            F *= torch.as_tensor(f, dtype=F.real.dtype, device=F.device).reshape(shape)
        return F

    def _full_plan(self):
        """A plan that keeps every mode and has no filter (for fftn/ifftn)."""
        if getattr(self, "_fplan", None) is None:
            h = _lib.c_void_p()
            with torch.cuda.device(self.device):
                _lib.call("ph_plan3d_create", _lib.ctypes.byref(h), self.Nx, self.Ny, self.Nz, self.Lx, self.Ly,
                          self.Lz, self._dtype_code, 0, 0, 8, 36.0, self._gamma)
            n = _lib.c_size_t()
            _lib.call("ph_plan3d_workspace_bytes", h, _lib.ctypes.byref(n))
            self._fplan = h
            self._fws = torch.empty(n.value, dtype=torch.uint8, device=self.device)
        return self._fplan, self._fws

    def _ctype(self):
        return torch.complex64 if self.dtype == torch.float32 else torch.complex128

    def fftn(self, f):
        """Complex FFT over the last three axes (phrike/grid.py:1497-1504):
        the library's forward transform gives the kx >= 0 half; the rest
        follows from Hermitian symmetry (the fields are real)."""
        f = self._as_fields(f)
        lead = f.shape[:-3]
        flat = f.reshape(-1, self.Nz, self.Ny, self.Nx)
        mx = self.Nx // 2 + 1
        half = torch.empty((flat.shape[0], self.Nz, self.Ny, mx), dtype=self._ctype(), device=self.device)
        plan, ws = self._full_plan()
        st = _lib.stream_handle(torch, self.device)
        for s in range(0, flat.shape[0], 5):
            e = min(s + 5, flat.shape[0])
            _lib.call("ph_fft3d", plan, e - s, _lib.ptr(flat[s:e]), _lib.ptr(half[s:e]), _lib.ptr(ws), st)
        full = torch.empty((flat.shape[0], self.Nz, self.Ny, self.Nx), dtype=half.dtype, device=self.device)
        full[..., :mx] = half
        if self.Nx > 2:
            # F[kz, ky, kx] = conj(F[-kz, -ky, -kx]) for kx > Nx/2
            neg = torch.roll(torch.flip(half[..., 1:self.Nx - mx + 1], dims=(-3, -2)), shifts=(1, 1), dims=(-3, -2))
            full[..., mx:] = torch.flip(neg, dims=(-1,)).conj()
        return full.reshape(*lead, self.Nz, self.Ny, self.Nx)

    def ifftn(self, F):
        """Real part of the inverse complex FFT over the last three axes
        (phrike/grid.py:1506-1513): that is the inverse of the Hermitian part
        (F(k) + conj F(-k))/2, whose kx >= 0 half the library transforms."""
        F = torch.as_tensor(F, device=self.device).to(self._ctype())
        if F.shape[-3:] != (self.Nz, self.Ny, self.Nx):
            raise ValueError(f"expected (..., {self.Nz}, {self.Ny}, {self.Nx}), got {tuple(F.shape)}")
        lead = F.shape[:-3]
        flat = F.reshape(-1, self.Nz, self.Ny, self.Nx)
        mx = self.Nx // 2 + 1
        rev = torch.roll(torch.flip(flat, dims=(-3, -2, -1)), shifts=(1, 1, 1), dims=(-3, -2, -1))  # F(-k)
        herm = (0.5 * (flat + rev.conj()))[..., :mx].contiguous()
        out = torch.empty((flat.shape[0], self.Nz, self.Ny, self.Nx), dtype=self.dtype, device=self.device)
        plan, ws = self._full_plan()
        st = _lib.stream_handle(torch, self.device)
        for s in range(0, flat.shape[0], 5):
            e = min(s + 5, flat.shape[0])
            _lib.call("ph_ifft3d", plan, e - s, _lib.ptr(herm[s:e]), _lib.ptr(out[s:e]), _lib.ptr(ws), st)
        return out.reshape(*lead, self.Nz, self.Ny, self.Nx)

    # full 3D tables, materialised on request only (phrike/grid.py:1404-1415)
    @property
    def dealias_mask(self):
        mx, my, mz = (_mask_1d(n, self.dealias) for n in (self.Nx, self.Ny, self.Nz))
        m = mz[:, None, None] * my[None, :, None] * mx[None, None, :]
        return torch.tensor(m, dtype=self.dtype, device=self.device)

    @property
    def filter_sigma(self):
        fon, fp, fa = self._filter
        sx, sy, sz = (_sigma_1d(n, fon, fp, fa) for n in (self.Nx, self.Ny, self.Nz))
        s = sz[:, None, None] * sy[None, :, None] * sx[None, None, :]
        return torch.tensor(s, dtype=self.dtype, device=self.device)


class Grid2D:
    """Uniform periodic 2D grid (Fourier x Fourier); arrays (..., Ny, Nx).

    ``batch`` sizes the plan for that many same-shape runs stacked as
    (batch, 4, Ny, Nx) (used by the ensemble server)."""

    def __init__(self, Nx: int, Ny: int, Lx: float, Ly: float, dealias: bool = True,
                 filter_params: Optional[Dict] = None, fft_workers: int = 1,
                 backend: str = "hip", torch_device: Optional[str] = None,
                 precision: str = "double", debug: bool = False,
                 basis_x: str = "fourier", basis_y: str = "fourier",
                 bc: str = "dirichlet", bc_config=None, batch: int = 1):
        _require_hip(backend)
        if basis_x != "fourier" or basis_y != "fourier":
            raise ValueError("Grid2D(backend='hip') is Fourier x Fourier; use "
                             "phrike_amd.hybrid.HybridGrid2D for Legendre y")
        self.Nx, self.Ny = int(Nx), int(Ny)
        self.Lx, self.Ly = float(Lx), float(Ly)
        self.dealias = bool(dealias)
        self.filter_params = filter_params
        self.fft_workers = fft_workers
        self.backend = "hip"
        self.basis_x, self.basis_y = basis_x, basis_y
        self.bc, self.bc_config = bc, bc_config
        self.precision = "single" if precision == "single" else "double"
        self.debug = debug
        self.torch_device = torch_device or "cuda"
        self.device = torch.device(self.torch_device)
        self._dtype_code, self.dtype = _dtype_code(self.precision)
        self._use_torch = True
        self.batch = int(batch)
        self.dx = self.Lx / self.Nx
        self.dy = self.Ly / self.Ny
        self.dx_min = None
        self.dy_min = None
        dev, dt = self.device, self.dtype
        self.x = torch.tensor(np.linspace(0.0, self.Lx, self.Nx, endpoint=False), dtype=dt, device=dev)
        self.y = torch.tensor(np.linspace(0.0, self.Ly, self.Ny, endpoint=False), dtype=dt, device=dev)
        self.kx = torch.tensor(_axis_k(self.Nx, self.Lx), dtype=dt, device=dev)
        self.ky = torch.tensor(_axis_k(self.Ny, self.Ly), dtype=dt, device=dev)
        ct = torch.complex64 if dt == torch.float32 else torch.complex128
        self.ikx = (1j * self.kx.to(ct))[None, :]
        self.iky = (1j * self.ky.to(ct))[:, None]
        self._filter = _filter_cfg(filter_params)
        self._gamma = 1.4
        self._plans = {}
        self._ws = None

    def _plan(self, gamma: Optional[float] = None):
        g = float(self._gamma if gamma is None else gamma)
        if g not in self._plans:
            h = _lib.c_void_p()
            fon, fp, fa = self._filter
            with torch.cuda.device(self.device):
                _lib.call("ph_plan2d_create", _lib.ctypes.byref(h), self.Nx, self.Ny, self.Lx, self.Ly,
                          self._dtype_code, int(self.dealias), fon, fp, fa, g, self.batch)
            self._plans[g] = h
        return self._plans[g]

    def workspace(self, gamma: Optional[float] = None):
        plan = self._plan(gamma)
        n = _lib.c_size_t()
        _lib.call("ph_plan2d_workspace_bytes", plan, _lib.ctypes.byref(n))
        if self._ws is None or self._ws.numel() < n.value:
            self._ws = torch.empty(n.value, dtype=torch.uint8, device=self.device)
        return self._ws

    def __del__(self):
        try:
            lib = _lib.load()
            for h in getattr(self, "_plans", {}).values():
                lib.ph_plan2d_destroy(h)
        except Exception:
            pass

    def _as_runs(self, f):
        """-> (runs, 4, Ny, Nx) contiguous; accepts (4,Ny,Nx), (B,4,Ny,Nx) or
        (Ny,Nx) / (k,Ny,Nx) fields (padded into 4-field groups)."""
        if not isinstance(f, torch.Tensor):
            f = torch.as_tensor(np.asarray(f), device=self.device)
        if f.shape[-2:] != (self.Ny, self.Nx):
            raise ValueError(f"expected (..., {self.Ny}, {self.Nx}), got {tuple(f.shape)}")
        return f.to(device=self.device, dtype=self.dtype).contiguous()

    def _apply(self, axis: int, f):
        f = self._as_runs(f)
        lead = f.shape[:-2]
        flat = f.reshape(-1, self.Ny, self.Nx)
        nf = flat.shape[0]
        groups = (nf + 3) // 4
        pad = torch.zeros((groups * 4, self.Ny, self.Nx), dtype=self.dtype, device=self.device)
        pad[:nf] = flat
        out = torch.empty_like(pad)
        ws = self.workspace()
        st = _lib.stream_handle(torch, self.device)
        plan = self._plan()
        for s in range(0, groups, self.batch):
            e = min(s + self.batch, groups)
            _lib.call("ph_deriv2d", plan, e - s, axis, _lib.ptr(pad[4 * s:4 * e]),
                      _lib.ptr(out[4 * s:4 * e]), _lib.ptr(ws), st)
        return out[:nf].reshape(*lead, self.Ny, self.Nx)

    def dx1(self, f):
        return self._apply(0, f)

    def dy1(self, f):
        return self._apply(1, f)

    def apply_spectral_filter(self, f):
        return self._apply(-1, f)

    def apply_boundary_conditions(self, U, equations=None):
        # Fourier x Fourier: no boundary conditions (phrike/grid.py:1082-1084)
        return U

    def xy_mesh(self):
        Y, X = torch.meshgrid(self.y, self.x, indexing="ij")
        return X, Y

    @property
    def dealias_mask(self):
        m = _mask_1d(self.Ny, self.dealias)[:, None] * _mask_1d(self.Nx, self.dealias)[None, :]
        return torch.tensor(m, dtype=self.dtype, device=self.device)

    @property
    def filter_sigma(self):
        fon, fp, fa = self._filter
        s = _sigma_1d(self.Ny, fon, fp, fa)[:, None] * _sigma_1d(self.Nx, fon, fp, fa)[None, :]
        return torch.tensor(s, dtype=self.dtype, device=self.device)
