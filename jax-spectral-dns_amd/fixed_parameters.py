"""Immutable parameter bundles of the equations (reference
fixed_parameters.py:14-36): same classes, field names and defaults.

``NavierStokesVelVortFixedParameters`` carries the reference's operator
matrix fields.  The reference fills them with complex128 per-wavenumber
matrices (or None when they are built on the fly, navier_stokes.py:437-502);
here the operators live on the GPU inside the step plan in factored form
(shared eigenvectors of the interior D^2 and per-|k| wall-correction tables),
so the matrix fields are None, as in the reference's on-the-fly mode.
"""

from __future__ import annotations

import dataclasses
from typing import Any, Optional


@dataclasses.dataclass(frozen=True)
class FixedParameters:
    domain: Any
    dt: float


@dataclasses.dataclass(frozen=True)
class NavierStokesVelVortFixedParameters:
    physical_domain: Any
    poisson_mat: Optional[Any]
    rk_mats_rhs: Optional[Any]
    rk_mats_lhs_inv: Optional[Any]
    rk_rhs_inhom: Optional[Any]
    rk_mats_lhs_inv_inhom: Optional[Any]
    rk_mats_rhs_ns: Optional[Any]
    rk_mats_lhs_inv_ns: Optional[Any]
    Re_tau: float
    max_cfl: float = 0.3
    dt_update_frequency: int = 10  # update the time step every dt_update_frequency steps
    u_max_over_u_tau: float = 1e0
    number_of_rk_steps: int = 3
