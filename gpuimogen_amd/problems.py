"""Initial conditions of the benchmark and test problems (host numpy, FP64).

States are shaped (5, nz, ny, nx) with slabs (rho, E, px, py, pz), x fastest.
"""
from __future__ import annotations

import numpy as np


def conserved(rho, vx, vy, vz, P, gamma):
    q = np.empty((5,) + np.shape(rho))
    q[0] = rho
    q[1] = P / (gamma - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    q[2] = rho * vx
    q[3] = rho * vy
    q[4] = rho * vz
    return q


def sedov3d(n=256, gamma=5.0 / 3.0, energy=1.0, p_background=1e-5, r2max=9):
    """Sedov-Taylor blast on the unit cube, n^3 cells, dx = 1/n.

    rho = 1, v = 0, P = p_background; the total energy `energy` is deposited as
    internal energy, uniformly, in every cell whose integer offset (i,j,k) from
    the centre cell satisfies i^2+j^2+k^2 <= r2max (123 cells for r2max = 9).
    The centre cell is 0-based index floor(n/2 + 0.5) - 1 on each axis, the
    reference's 1-based floor(N/2+0.5) convention
    (experiment/SedovTaylorBlastWave/SedovTaylorBlastWaveInitializer.m:95).
    Returns (state, h, deposit_cell_count)."""
    dx = 1.0 / n
    c = int(np.floor(n / 2 + 0.5)) - 1
    q = np.zeros((5, n, n, n))
    q[0] = 1.0
    q[1] = p_background / (gamma - 1)
    r = int(np.floor(np.sqrt(r2max)))
    offs = [(i, j, k) for i in range(-r, r + 1) for j in range(-r, r + 1) for k in range(-r, r + 1)
            if i * i + j * j + k * k <= r2max]
    e_cell = energy / (len(offs) * dx ** 3)
    for (i, j, k) in offs:
        q[1, c + k, c + j, c + i] = e_cell
    return q, (dx, dx, dx), len(offs)


def sod1d(n=1024, gamma=1.4):
    """Sod shock tube on [0,1]: (rho, v, P) = (1,0,1) | (0.125,0,0.1) at x = 0.5."""
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    rho = np.where(x < 0.5, 1.0, 0.125)
    P = np.where(x < 0.5, 1.0, 0.1)
    z = np.zeros(n)
    return conserved(rho, z, z, z, P, gamma).reshape(5, 1, 1, n), (dx, 1.0, 1.0)


def gaussian_field(n, kmax, rng):
    """Unit-variance Gaussian random field on an n^3 periodic grid whose
    Fourier modes are white for |k| <= kmax (integer wavenumbers) and zero
    above: FFT-filtered white noise."""
    white = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n) * n
    kz, ky, kx = np.meshgrid(k, k, np.fft.rfftfreq(n) * n, indexing="ij")
    spec = np.fft.rfftn(white)
    spec[(kx * kx + ky * ky + kz * kz) > kmax * kmax] = 0.0
    f = np.fft.irfftn(spec, s=(n, n, n))
    f -= f.mean()
    return f / f.std()


def gasdust3d(n=192, seed=0, gamma=5.0 / 3.0, eps_median=0.1, eps_sigma_ln=1.0, mach=0.3, kmax=8,
              dust_cs=0.3, dust_gamma=1.01):
    """Two-fluid gas + dust box, n^3 periodic cells on the unit cube.

    Gas: rho = 1, P = 1, gamma = 5/3. Dust-to-gas ratio lognormal with median
    eps_median and log-width eps_sigma_ln, from a Gaussian random field with
    |k| <= kmax. Gas and dust velocities: independent Gaussian random fields
    (|k| <= kmax) per component with standard deviation mach * c_s. Both
    fluids are swept with HLLC, as the reference's gas-dust profile runs them
    (run/profile_3DXYZ_gasdust.m:55-57). The dust is a gamma = 1.01 fluid
    (fluidDetailModel('10um_iron_balls'), astroParams/fluidDetailModel.m:148)
    with sound speed dust_cs: the drag source leaves the dust energy slab
    unchanged (cudaSource2FluidDrag.cu:1678-1680, the dust is taken as
    pressureless), so the dust's internal energy absorbs its kinetic energy
    changes and must start large enough to stay positive.
    Returns (gas, dust, h, dust_cs)."""
    rng = np.random.default_rng(seed)
    cs = np.sqrt(gamma * 1.0 / 1.0)
    rho = np.ones((n, n, n))
    vg = [mach * cs * gaussian_field(n, kmax, rng) for _ in range(3)]
    gas = conserved(rho, vg[0], vg[1], vg[2], np.ones((n, n, n)), gamma)
    del vg
    eps = eps_median * np.exp(eps_sigma_ln * gaussian_field(n, kmax, rng))
    rd = eps * rho
    vd = [mach * cs * gaussian_field(n, kmax, rng) for _ in range(3)]
    dust = conserved(rd, vd[0], vd[1], vd[2], rd * dust_cs ** 2 / dust_gamma, dust_gamma)
    del vd
    dx = 1.0 / n
    ciso = dust_cs
    return gas, dust, (dx, dx, dx), ciso
