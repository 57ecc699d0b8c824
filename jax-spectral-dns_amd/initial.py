"""Seeded synthetic initial conditions for the channel-flow configurations.

The generator draws Fourier-Chebyshev coefficients for the wall-normal
velocity v and vorticity omega_y with amplitude exp(-(|kx_i|+|kz_i|)/2) 0.8^n
(kx_i, kz_i integer mode indices, n the Chebyshev degree), makes the spectrum
Hermitian, evaluates it at the Gauss-Lobatto points, multiplies v by
(1-y^2)^2 and omega_y by (1-y^2) so that v = v' = omega_y = 0 at the walls (v uses degrees
n <= Ny-5 and omega_y degrees n <= Ny-3, so the multiplied profiles stay in
the Ny-point collocation space and the wall conditions hold to rounding),
recovers u and w from continuity (the same relations as the per-wavenumber
step, navier_stokes.py:1506-1527) and scales the field to a target energy.
Output layout is the reference's: complex128 (3, Nx_modes, Ny, Nz_modes).
"""

from __future__ import annotations

import numpy as np

from .domain_math import Grid


def _hermitian(c: np.ndarray) -> np.ndarray:
    """Make c[kx, kz, :] = conj(c[-kx, -kz, :]) on the stored index set."""
    m, mz = c.shape[0], c.shape[1]
    out = c.copy()
    for i in range(m):
        for j in range(mz):
            pi, pj = (m - i) % m, (mz - j) % mz
            key = (i, j) < (pi, pj)
            if (i, j) == (pi, pj):
                out[i, j] = c[i, j].real
            elif not key:
                out[i, j] = np.conj(c[pi, pj])
    return out


def spectral_energy(grid: Grid, vel_hat: np.ndarray) -> float:
    """Energy 0.5 <u.u> / volume from the spectrum (Parseval on the x-z grid,
    the y quadrature of the reference's definite integral)."""
    m, mz = vel_hat.shape[1], vel_hat.shape[3]
    pi = (m - np.arange(m)) % m
    pj = (mz - np.arange(mz)) % mz
    herm = 0.5 * (vel_hat + np.conj(vel_hat[:, pi][:, :, :, pj]))
    s2 = grid.scaling**2 / (grid.px * grid.pz)
    per_y = (np.abs(herm) ** 2).sum(axis=(0, 1, 3)) * s2
    return float(np.dot(grid.y_weights / 4.0, per_y))


def perturbation(grid: Grid, seed: int = 0, energy: float = 1e-6, decay: float = 0.8) -> np.ndarray:
    rng = np.random.default_rng(seed)
    m, ny, mz = grid.nxm, grid.ny, grid.nzm
    ix = np.concatenate([np.arange(m // 2 + 1), np.arange(-(m // 2), 0)])
    iz = np.concatenate([np.arange(mz // 2 + 1), np.arange(-(mz // 2), 0)])
    amp = np.exp(-(np.abs(ix)[:, None] + np.abs(iz)[None, :]) / 2.0)
    cheb = decay ** np.arange(ny)

    def draw() -> np.ndarray:
        re = rng.standard_normal((m, mz, ny))
        im = rng.standard_normal((m, mz, ny))
        return (re + 1j * im) / np.sqrt(2.0) * amp[:, :, None] * cheb[None, None, :]

    cv, cw = _hermitian(draw()), _hermitian(draw())
    cv[:, :, ny - 4 :] = 0.0
    cw[:, :, ny - 2 :] = 0.0
    cv[0, 0, :] = 0.0
    T = np.cos(np.outer(np.arange(ny), np.arange(ny)) * np.pi / (ny - 1))  # T_n(y_j) [j, n]
    y = grid.y
    v = np.einsum("jn,abn->abj", T, cv) * ((1 - y**2) ** 2)[None, None, :]
    om = np.einsum("jn,abn->abj", T, cw) * (1 - y**2)[None, None, :]
    dv = np.einsum("lj,abj->abl", grid.D, v)
    kx = grid.kx[:, None, None]
    kz = grid.kz[None, :, None]
    k2 = -(kx**2 + kz**2)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = (-1j * kx * dv + 1j * kz * om) / k2
        w = (-1j * kz * dv - 1j * kx * om) / k2
    u[0, 0] = 0.0
    w[0, 0] = 0.0
    vel = np.stack([u, v, w]).transpose(0, 1, 3, 2).copy()  # (3, kx, y, kz)
    vel *= np.sqrt(energy / spectral_energy(grid, vel))
    return vel


def laminar(grid: Grid, u_max: float = 1.0) -> np.ndarray:
    """Spectral representation of U = u_max (1 - y^2): only the (0,0) mode,
    scaled like the reference's forward transform (sqrt(Px Pz) / prod(L/2pi))."""
    vel = np.zeros((3, grid.nxm, grid.ny, grid.nzm), dtype=np.complex128)
    vel[0, 0, :, 0] = u_max * (1 - grid.y**2) * np.sqrt(grid.px * grid.pz) / grid.scaling
    return vel
