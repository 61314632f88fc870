"""ph_step2d_noalloc (static scratch, no workspace, no allocation) against
ph_step2d through the C ABI on the GPU: identical bits on two calls in a row
(the second with a smaller batch, grid and precision), device memory unchanged
by the call, and a batch beyond the static scratch rejected before launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP = {"enabled": True, "p": 16, "alpha": 36.0}


def _case(dev, n, nb, prec, seed0):
    import torch
    from phrike_amd.grid import Grid2D
    from phrike_amd.initial_conditions import khi_2d_periodic
    g = Grid2D(n, n, 1.0, 1.0, filter_params=FP, precision=prec, batch=nb, torch_device=str(dev))
    U = torch.as_tensor(np.stack([khi_2d_periodic(n, n, seed=seed0 + s) for s in range(nb)]),
                        dtype=g.dtype, device=dev)
    active = torch.ones(nb, dtype=torch.int32, device=dev)
    active[1] = 0
    dt = torch.tensor([2e-4 * (1 + 0.1 * s) for s in range(nb)], dtype=torch.float64, device=dev)
    return g, U, active, dt


def test_step2d_noalloc_matches_workspace_step(dev):
    import torch
    from phrike_amd import _lib
    st = _lib.stream_handle(torch, dev)
    for n, nb, prec, seed0 in ((256, 6, "single", 10), (64, 2, "double", 30)):
        print(f"noalloc case {n}^2 x {nb} {prec}", flush=True)
        g, U, active, dt = _case(dev, n, nb, prec, seed0)
        plan, ws = g._plan(1.4), g.workspace(1.4)
        a = torch.full_like(U, 3.0)
        b = torch.full_like(U, 3.0)
        _lib.call("ph_step2d", plan, nb, _lib.ptr(U), _lib.ptr(a), 0.0, _lib.ptr(dt), _lib.ptr(active), 1,
                  _lib.ptr(ws), st)
        torch.cuda.synchronize(dev)
        free0 = torch.cuda.mem_get_info(dev)[0]
        _lib.call("ph_step2d_noalloc", plan, nb, _lib.ptr(U), _lib.ptr(b), 0.0, _lib.ptr(dt),
                  _lib.ptr(active), 1, st)
        torch.cuda.synchronize(dev)
        free1 = torch.cuda.mem_get_info(dev)[0]
        assert torch.equal(a, b)
        assert torch.all(b[1] == 3.0)  # the inactive run is untouched
        assert not torch.equal(b[0], U[0])
        assert free1 == free0, (free0, free1)


def test_step2d_noalloc_rejects_oversized_batch(dev):
    import torch
    from phrike_amd import _lib
    g, U, active, dt = _case(dev, 1024, 2, "double", 50)
    out = torch.empty_like(U)
    with pytest.raises(_lib.HipBackendError, match="static scratch"):
        _lib.call("ph_step2d_noalloc", g._plan(1.4), 2, _lib.ptr(U), _lib.ptr(out), 1e-4, None, None, 1,
                  _lib.stream_handle(torch, dev))
