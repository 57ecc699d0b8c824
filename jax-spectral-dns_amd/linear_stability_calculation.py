"""Linear stability of plane channel flow and optimal transient growth: the
host-side initial-condition generator of the optimisation runs.

Mirrors the interface of LinearStabilityCalculation
(linear_stability_calculation.py:34-638): eigenvalues of the linearised
operator about the base profile, the largest energy gain G(T) of a single
(alpha, beta) wave over a time T from the least stable modes
(calculate_transient_growth_svd / _max_energy, :424-557) and the initial
condition that attains it (calculate_transient_growth_initial_condition,
:559-638).  The reference discretises the primitive-variable equations on
symmetric / antisymmetric Chebyshev bases; this module uses the equivalent
Orr-Sommerfeld / Squire form on Gauss-Lobatto collocation points, which has
the same spectrum and the same energy gains (only the discretisation error
differs).  For a perturbation q(y) exp(i(alpha x + beta z) + lambda t),
k^2 = alpha^2 + beta^2, L = D^2 - k^2:

    lambda L v   = -i alpha U L v + i alpha U'' v + L^2 v / Re,
    lambda eta   = -i beta U' v - i alpha U eta + L eta / Re,

with v = v' = 0 and eta = 0 at the walls, imposed by replacing boundary rows
(their infinite eigenvalues are discarded).  The kinetic energy per unit
area is (1 / (2 k^2)) int (|Dv|^2 + k^2 |v|^2 + |eta|^2) dy (Clenshaw-Curtis
weights), the gain G(T) = ||F exp(Lambda T) F^-1||_2^2 with M = F^H F the
energy matrix of the retained modes, and the optimal initial condition is
F^-1 times the first right singular vector.  u and w follow from
continuity and the definition of eta:
u = (i / k^2)(alpha Dv - beta eta), w = (i / k^2)(beta Dv + alpha eta).
Host NumPy / SciPy, run once per optimisation; nothing here is on the GPU
step path.
"""

from __future__ import annotations

from typing import List, Optional, Tuple

import numpy as np
import scipy.linalg

from . import domain_math as DM


def clenshaw_curtis_weights(n: int) -> np.ndarray:
    """Quadrature weights on the n Gauss-Lobatto points y_j = cos(pi j / (n-1))
    (Clenshaw & Curtis 1960; the closed form of Trefethen, Spectral Methods in
    MATLAB, 2000, program clencurt)."""
    N = n - 1
    theta = np.pi * np.arange(n) / N
    w = np.zeros(n)
    v = np.ones(N - 1)
    inner = theta[1:-1]
    if N % 2 == 0:
        w[0] = w[-1] = 1.0 / (N * N - 1)
        for k in range(1, N // 2):
            v -= 2.0 * np.cos(2 * k * inner) / (4 * k * k - 1)
        v -= np.cos(N * inner) / (N * N - 1)
    else:
        w[0] = w[-1] = 1.0 / (N * N)
        for k in range(1, (N - 1) // 2 + 1):
            v -= 2.0 * np.cos(2 * k * inner) / (4 * k * k - 1)
    w[1:-1] = 2.0 * v / N
    return w


class LinearStabilityCalculation:
    def __init__(
        self,
        Re: float = 180.0,
        alpha: float = 3.25,
        beta: float = 0.0,
        n: int = 50,
        U_max: Optional[float] = None,
        U_base: Optional[np.ndarray] = None,
    ):
        self.Re = float(Re)
        self.alpha = float(alpha)
        self.beta = float(beta)
        self.n = int(n)
        self.ys = DM.cheb_grid(self.n)
        self.D = DM.cheb_diff(self.ys)
        if U_base is not None:
            self.U_base = np.asarray(U_base, dtype=np.float64)
        else:
            self.U_base = (1.0 if U_max is None else float(U_max)) * (1.0 - self.ys**2)
        self.eigenvalues: Optional[np.ndarray] = None
        self.eigenvectors: Optional[np.ndarray] = None  # columns [v; eta], sorted like eigenvalues
        self.S: Optional[np.ndarray] = None
        self.coeffs: Optional[np.ndarray] = None
        self.velocity_field_ = None

    @property
    def k2(self) -> float:
        return self.alpha**2 + self.beta**2

    def assemble_matrix(self) -> Tuple[np.ndarray, np.ndarray]:
        """A q = lambda B q for q = [v; eta] with the wall conditions in rows
        0, 1, n-2, n-1 (v, v') and n, 2n-1 (eta)."""
        n, a, b, Re = self.n, self.alpha, self.beta, self.Re
        D = self.D
        I = np.eye(n)
        L = D @ D - self.k2 * I
        U = np.diag(self.U_base)
        dU = np.diag(D @ self.U_base)
        d2U = np.diag(D @ (D @ self.U_base))
        A = np.zeros((2 * n, 2 * n), dtype=complex)
        B = np.zeros((2 * n, 2 * n), dtype=complex)
        A[:n, :n] = -1j * a * U @ L + 1j * a * d2U + (L @ L) / Re
        B[:n, :n] = L
        A[n:, :n] = -1j * b * dU
        A[n:, n:] = -1j * a * U + L / Re
        B[n:, n:] = I
        for r, row in ((0, np.eye(n)[0]), (n - 1, np.eye(n)[n - 1]), (1, D[0]), (n - 2, D[n - 1])):
            A[r, :] = 0.0
            B[r, :] = 0.0
            A[r, :n] = row
        for r in (n, 2 * n - 1):
            A[r, :] = 0.0
            B[r, :] = 0.0
            A[r, r] = 1.0
        return A, B

    def calculate_eigenvalues(self) -> Tuple[np.ndarray, np.ndarray]:
        """Finite eigenvalues sorted by decreasing growth rate Re(lambda), and
        their eigenvectors [v; eta] (columns)."""
        A, B = self.assemble_matrix()
        lam, vec = scipy.linalg.eig(A, B)
        ok = np.isfinite(lam) & (np.abs(lam) < 1e6)
        lam, vec = lam[ok], vec[:, ok]
        order = np.argsort(-lam.real)
        self.eigenvalues, self.eigenvectors = lam[order], vec[:, order]
        return self.eigenvalues, self.eigenvectors

    def _velocity(self, q: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        n = self.n
        v, eta = q[:n], q[n:]
        dv = self.D @ v
        u = 1j / self.k2 * (self.alpha * dv - self.beta * eta)
        w = 1j / self.k2 * (self.beta * dv + self.alpha * eta)
        return u, v, w

    def energy_matrix(self, Q: np.ndarray) -> np.ndarray:
        """M_jk = (1/(2k^2)) int (conj(Dv_j) Dv_k + k^2 conj(v_j) v_k + conj(eta_j) eta_k) dy."""
        n = self.n
        w = clenshaw_curtis_weights(n)
        V, E = Q[:n], Q[n:]
        DV = self.D @ V
        M = (DV.conj().T * w) @ DV + self.k2 * (V.conj().T * w) @ V + (E.conj().T * w) @ E
        M = 0.5 * (M + M.conj().T) / (2.0 * self.k2)
        return M

    def calculate_transient_growth_svd(
        self, T: float, number_of_modes: int, Ts: Optional[List[float]] = None
    ) -> Tuple[np.ndarray, np.ndarray]:
        """Singular values of F exp(Lambda T) F^-1 over the number_of_modes least
        stable modes and the mode coefficients of the optimal initial condition
        (linear_stability_calculation.py:424-538).  With Ts, G(t) = S[0]^2 for
        each t is stored in self.Gs."""
        if self.eigenvalues is None:
            self.calculate_eigenvalues()
        m = min(int(number_of_modes), len(self.eigenvalues))
        lam, Q = self.eigenvalues[:m], self.eigenvectors[:, :m]
        M = self.energy_matrix(Q)
        F = scipy.linalg.cholesky(M, lower=False)
        F_inv = np.linalg.inv(F)
        if Ts is not None:
            self.Gs = [float(np.linalg.svd(F @ np.diag(np.exp(lam * t)) @ F_inv, compute_uv=False)[0] ** 2) for t in Ts]
        _, S, Vh = np.linalg.svd(F @ np.diag(np.exp(lam * T)) @ F_inv)
        coeffs = F_inv @ Vh.conj().T[:, 0]
        self.S, self.coeffs = S, coeffs
        return S, coeffs

    def calculate_transient_growth_max_energy(self, T: float, number_of_modes: int, Ts: Optional[List[float]] = None) -> float:
        """G(T), the largest energy gain over time T (:540-557)."""
        S, _ = self.calculate_transient_growth_svd(T, number_of_modes, Ts=Ts)
        return float(S[0] ** 2)

    def optimal_profiles(self, T: float, number_of_modes: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """u(y), v(y), w(y) of the optimal initial condition on the n collocation points."""
        _, coeffs = self.calculate_transient_growth_svd(T, number_of_modes)
        q = self.eigenvectors[:, : len(coeffs)] @ coeffs
        return self._velocity(q)

    # ---- fields on a DNS grid -------------------------------------------
    def _check_grid(self, domain) -> None:
        if len(domain.grid[1]) != self.n:
            raise ValueError("the stability calculation must use the domain's Ny = %d points" % len(domain.grid[1]))

    def y_slice_to_3d_field(self, domain, y_slice_i: np.ndarray, factor: complex = 1.0) -> np.ndarray:
        """Re{factor * q(y) exp(i(alpha x + beta z))} on the domain's physical
        grid (linear_stability_calculation.py:313-335)."""
        self._check_grid(domain)
        x, z = np.asarray(domain.grid[0]), np.asarray(domain.grid[2])
        phase = np.exp(1j * self.alpha * x)[:, None, None] * np.exp(1j * self.beta * z)[None, None, :]
        return (factor * phase * np.asarray(y_slice_i)[None, :, None]).real

    def velocity_field_from_y_slice(self, domain, y_slice, factor: complex = 1.0):
        """Physical VectorField of the three profiles
        (linear_stability_calculation.py:337-364)."""
        import torch

        from .field import PhysicalField, VectorField

        names = ("velocity_pert_x", "velocity_pert_y", "velocity_pert_z")
        comps = [
            PhysicalField(domain, torch.from_numpy(self.y_slice_to_3d_field(domain, q, factor)), name=nm)
            for q, nm in zip(y_slice, names)
        ]
        self.velocity_field_ = VectorField(comps)
        return self.velocity_field_

    def velocity_field(self, domain, evec: np.ndarray, factor: complex = 1.0):
        """Physical VectorField of one eigenvector (linear_stability_calculation.py:268-311).
        Here evec is [v; eta] on the collocation points (the reference's is
        the primitive (u, v, w, p) coefficient vector of its own basis)."""
        return self.velocity_field_from_y_slice(domain, self._velocity(np.asarray(evec)), factor)

    def velocity_field_single_mode(
        self,
        domain,
        mode: int = 0,
        factor: complex = 1.0,
        recompute_partial: bool = False,
        recompute_full: bool = False,
        save: bool = False,
    ):
        """Physical VectorField of eigenmode `mode` (0: least stable), e.g. a
        Tollmien-Schlichting wave (linear_stability_calculation.py:225-266).
        The reference first tries fields cached on disk; this computes them
        (recompute_partial / recompute_full accepted for the same call
        sites).  save writes the three components as field files
        (u/v/w_Re_<Re>_n_<n>_mode_<mode>, field.save_to_file)."""
        if recompute_full or self.eigenvalues is None:
            self.calculate_eigenvalues()
        u = self.velocity_field(domain, self.eigenvectors[:, mode], factor)
        if save:
            for c, comp in zip("uvw", u):
                comp.save_to_file("%s_Re_%s_n_%d_mode_%d" % (c, self.Re, self.n, mode))
        return u

    def energy_over_time(self, domain, mode: int = 0, eps: float = 1.0):
        """(E(t, dim=None), lambda): the linear theory's energy of eps times
        eigenmode `mode` at time t, E(0) exp(2 Re(lambda) t), all components or
        one (linear_stability_calculation.py:366-422)."""
        if self.eigenvalues is None:
            self.calculate_eigenvalues()
        if getattr(self, "velocity_field_", None) is None:
            self.velocity_field(domain, self.eigenvectors[:, mode])
        lam = self.eigenvalues[mode]
        comp_energy = [float(c.energy()) for c in self.velocity_field_]

        def out(t: float, dim: Optional[int] = None) -> float:
            e = sum(comp_energy) if dim is None else comp_energy[dim]
            return float(eps**2 * e * np.exp(2.0 * lam.real * t))

        return out, lam

    def calculate_transient_growth_initial_condition(
        self, domain, T: float, number_of_modes: int, energy: Optional[float] = None
    ):
        """Spectral velocity field (complex128, (3, Nx, Ny, Nz) in the reference
        layout) of the physical field Re{(u, v, w)(y) exp(i(alpha x + beta z))}
        on the domain's grid (:559-638, velocity_field / y_slice_to_3d_field
        :268-364); the collocation points must be the domain's (n = Ny).
        With energy, the field is scaled to that energy."""
        import torch

        self._check_grid(domain)
        u, v, w = self.optimal_profiles(T, number_of_modes)
        phys = np.stack([self.y_slice_to_3d_field(domain, c) for c in (u, v, w)])
        hat = torch.stack([domain.field_hat(torch.from_numpy(phys[c])) for c in range(3)])
        if energy is not None:
            from .initial import spectral_energy

            e = spectral_energy(domain.grid_math, hat.numpy())
            hat = hat * (energy / e) ** 0.5
        return hat
