"""CPU oracle for the pseudo-spectral Euler hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package (``phrike_amd``)
imports this package: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Contents
--------
``spectral_np``  numpy/scipy restatement of the reference's Fourier operators,
                 Euler fluxes, RHS and RK4 step in 3D and 2D
                 (phrike/grid.py, phrike/equations.py, phrike/solver.py).
``gll_np``       Legendre-Gauss-Lobatto nodes, weights, differentiation and
                 nodal-filter matrices (phrike/basis/legendre.py).
``ics``          harness-defined initial conditions (TGV with the
                 p = 100 + (1/16)(...) pressure, decaying turbulence, KHI).

Parity pinning of each function is listed in the module headers.
"""
