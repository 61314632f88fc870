"""Properties of the decaying-turbulence initial state (the FP32-vs-FP64
turbulence workload: random-phase solenoidal velocity, E(k) ~ k^4 exp(-2 (k/4)^2), rms Mach 0.3,
rho = 1 + 1% Gaussian noise, uniform p) checked from the mathematics: the
spectral divergence vanishes, the rms Mach number is 0.3, the shell spectrum
peaks near k = 4 and follows the k^4 exp(-2 (k/4)^2) shape, and the seeds
make it reproducible."""
import numpy as np


def _fields(U, gamma=1.4):
    rho = U[0]
    u = U[1:4] / rho
    p = (gamma - 1.0) * (U[4] - 0.5 * rho * (u * u).sum(axis=0))
    return rho, u, p


def test_turbulence_ic_properties():
    from phrike_amd.initial_conditions import decaying_turbulence_3d
    n = 64
    U = decaying_turbulence_3d(n, seed=7, density_seed=9)
    rho, u, p = _fields(U)
    # rms Mach 0.3 against c0 = sqrt(gamma p0 / rho0) = 1
    mach = np.sqrt(np.mean((u * u).sum(axis=0)))
    assert abs(mach - 0.3) < 1e-12
    # uniform p0 = 1/gamma (up to the conservative round trip)
    assert np.abs(p - 1.0 / 1.4).max() < 1e-13
    # rho: 1 + 0.01 N(0, 1)
    assert abs(rho.mean() - 1.0) < 5e-4 and abs(rho.std() - 0.01) < 5e-4
    # solenoidal: spectral divergence of u (computed here with numpy FFTs)
    k = np.fft.fftfreq(n, d=1.0 / n)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    uh = [np.fft.fftn(c) for c in u]
    div = np.fft.ifftn(1j * (KX * uh[0] + KY * uh[1] + KZ * uh[2])).real
    grad = np.fft.ifftn(1j * KX * uh[0]).real
    assert np.abs(div).max() < 1e-12 * np.abs(grad).max()
    # shell spectrum: peak near k = 4 and the shape k^4 exp(-2 (k/4)^2) in
    # the well-populated shells
    kmag = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    e = 0.5 * sum(np.abs(c) ** 2 for c in uh)
    shells = np.rint(kmag).astype(int)
    E = np.bincount(shells.ravel(), weights=e.ravel())[:n // 2]
    kk = np.arange(len(E))
    assert 3 <= int(np.argmax(E[1:]) + 1) <= 5
    model = kk ** 4 * np.exp(-2.0 * (kk / 4.0) ** 2)
    sel = slice(2, 9)
    ratio = E[sel] / model[sel]
    assert ratio.std() / ratio.mean() < 0.25  # random phases: shell sums scatter
    # seeded: the same seeds give the same state, another seed another one
    assert np.array_equal(U, decaying_turbulence_3d(n, seed=7, density_seed=9))
    assert not np.array_equal(U, decaying_turbulence_3d(n, seed=8, density_seed=9))
