"""Gradient parameterisation of the optimisation problems: the initial
velocity as a function of (omega_y, v, u_00, w_00).

``vort_yvel_to_vel`` restates navier_stokes.py:1063-1203: v is projected in
the least-squares sense onto Chebyshev combinations phi_n = T_{n+4} -
a T_{n+2} - b T_n with phi(+-1) = phi'(+-1) = 0 (smooth_and_enforce_bc_vel_y,
navier_stokes.py:1027-1061; here one precomputed N x N projector A A^+
instead of a least-squares solve per pencil), u and w follow from continuity
and the definition of omega_y, the (0,0) pencil takes u_00 and w_00.  It is
written in torch so that autograd composes it with the discrete adjoint of
the step (``vjp.run``): the reference differentiates J with respect to these
parameters with jax.grad (optimiser.py:115-121).  It runs once per gradient
evaluation, on the device of its inputs.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .domain_math import Grid


def clamped_basis(y: np.ndarray, M: Optional[int] = None) -> np.ndarray:
    """Columns phi_n at the grid points, n < M = N - 4 (the Chebyshev
    combination of navier_stokes.py:1033-1048, T_n(y_j) evaluated by the
    Chebyshev series)."""
    N = len(y)
    M_ = N - 4 if M is None else M
    A = np.zeros((N, M_))
    for n in range(M_):
        c = np.zeros(n + 5)
        c[n + 4] = 1.0
        c[n + 2] = -(2.0 * (n + 2)) / (n + 1)
        c[n] = (n + 3) / (n + 1)
        A[:, n] = np.polynomial.chebyshev.chebval(y, c)
    return A


def clamped_projector(y: np.ndarray, M: Optional[int] = None) -> np.ndarray:
    """Orthogonal projector onto the span of the clamped basis (the map
    v -> A lstsq(A, v))."""
    A = clamped_basis(y, M)
    Q, _ = np.linalg.qr(A)
    return Q @ Q.T


class VelocityParameterisation:
    def __init__(self, grid: Grid, device: torch.device, M: Optional[int] = None):
        self.grid = grid
        c128 = torch.complex128
        self.P = torch.from_numpy(clamped_projector(grid.y, M)).to(device=device, dtype=c128)
        self.D = torch.from_numpy(np.ascontiguousarray(grid.D)).to(device=device, dtype=c128)
        kx = torch.from_numpy(grid.kx).to(device)[:, None, None]
        kz = torch.from_numpy(grid.kz).to(device)[None, None, :]
        self.kx, self.kz = kx, kz
        mk2 = -(kx**2 + kz**2)
        self.mk2 = torch.where(mk2 == 0, torch.ones_like(mk2), mk2)

    def __call__(
        self,
        vort: torch.Tensor,
        vel_y: torch.Tensor,
        vel_x_00: torch.Tensor,
        vel_z_00: torch.Tensor,
        two_d: bool = False,
    ) -> torch.Tensor:
        """(Nx, Ny, Nz) complex inputs -> (3, Nx, Ny, Nz) velocity spectrum."""
        if two_d:
            vort = torch.zeros_like(vel_y)
            vel_z_00 = torch.zeros_like(vel_x_00)
        vs = torch.matmul(self.P, vel_y)
        vy = torch.matmul(self.D, vs)
        u = (-1j * self.kx * vy + 1j * self.kz * vort) / self.mk2
        w = torch.zeros_like(u) if two_d else (-1j * self.kz * vy - 1j * self.kx * vort) / self.mk2
        pencil = torch.zeros_like(u)
        pencil[0, :, 0] = 1.0
        u = u * (1 - pencil) + pencil * vel_x_00.to(u.dtype)[None, :, None]
        w = w * (1 - pencil) + pencil * vel_z_00.to(w.dtype)[None, :, None]
        return torch.stack([u, vs, w])


def vort_yvel_to_vel(
    grid: Grid,
    vort: torch.Tensor,
    vel_y: torch.Tensor,
    vel_x_00: torch.Tensor,
    vel_z_00: torch.Tensor,
    two_d: bool = False,
) -> torch.Tensor:
    return VelocityParameterisation(grid, vel_y.device)(vort, vel_y, vel_x_00, vel_z_00, two_d)
