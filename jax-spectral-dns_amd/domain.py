"""Domain surface of the reference (jax_spectral_dns/domain.py) for the
channel geometry (x, z periodic, y Chebyshev collocation).

``PhysicalDomain.create`` / ``.hat()`` / ``FourierDomain`` keep the
reference's names, shapes and grids (domain.py:130-317, 791-848).  The
stand-alone transforms ``field_hat`` / ``field_no_hat`` and ``diff`` are
input-preparation and post-processing utilities on torch tensors (they follow
domain.py:716-764, 1167-1242 and 891-918 on the reference's physical grid);
inside the time step the same transforms run fused in the HIP kernels.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any, Optional, Sequence, Tuple

import numpy as np
import torch

from . import domain_math as DM

C128 = torch.complex128


def _t(a: np.ndarray, like: Optional[torch.Tensor] = None, dtype=torch.float64) -> torch.Tensor:
    out = torch.as_tensor(np.ascontiguousarray(a), dtype=dtype)
    return out.to(like.device) if like is not None else out


@dataclass(frozen=True)
class PhysicalDomain:
    """Physical-space domain (domain.py:683-764)."""

    number_of_dimensions: int
    periodic_directions: Tuple[bool, ...]
    scale_factors: Tuple[float, ...]
    shape: Tuple[int, ...]
    grid: Tuple[np.ndarray, ...]
    diff_mats: Tuple[Optional[np.ndarray], ...]
    aliasing: float = 3 / 2
    dealias_nonperiodic: bool = False
    grid_math: Any = field(default=None, compare=False, repr=False)

    @classmethod
    def create(
        cls,
        shape: Tuple[int, ...],
        periodic_directions: Tuple[bool, ...],
        scale_factors: Optional[Tuple[float, ...]] = None,
        aliasing: float = 3 / 2,
        dealias_nonperiodic: bool = False,
    ) -> "PhysicalDomain":
        """domain.py:201-317 for the channel (True, False, True)."""
        if tuple(periodic_directions) != (True, False, True):
            raise NotImplementedError("the MI355X step supports the channel (x, z periodic, y Chebyshev) only")
        if dealias_nonperiodic:
            raise NotImplementedError("dealias_nonperiodic is not supported by the MI355X step")
        if aliasing != 3 / 2:
            raise NotImplementedError("the MI355X step dealiases with the 3/2 rule only")
        sf = list(scale_factors) if scale_factors is not None else [2 * np.pi, 1.0, 2 * np.pi]
        g = DM.make_grid(tuple(int(s) for s in shape), sf[0], sf[2])
        x = np.linspace(0.0, sf[0], g.px + 1)[:-1]
        z = np.linspace(0.0, sf[2], g.pz + 1)[:-1]
        return cls(
            3,
            (True, False, True),
            (float(sf[0]), 1.0, float(sf[2])),
            (g.nxm, g.ny, g.nzm),
            (x, g.y, z),
            (None, g.D, None),
            aliasing,
            False,
            g,
        )

    def get_shape(self) -> Tuple[int, ...]:
        return self.shape

    def get_shape_aliasing(self) -> Tuple[int, ...]:
        return tuple(len(g) for g in self.grid)

    def number_of_cells(self, direction: int) -> int:
        return len(self.grid[direction])

    def all_dimensions(self) -> Sequence[int]:
        return range(self.number_of_dimensions)

    def is_periodic(self, direction: int) -> bool:
        return self.periodic_directions[direction]

    def all_periodic_dimensions(self):
        return [d for d in self.all_dimensions() if self.is_periodic(d)]

    def all_nonperiodic_dimensions(self):
        return [d for d in self.all_dimensions() if not self.is_periodic(d)]

    def get_extent(self, i: int) -> float:
        return self.scale_factors[i] if self.is_periodic(i) else 2.0

    def hat(self) -> "FourierDomain":
        return FourierDomain.FromPhysicalDomain(self)

    def field_hat(self, f: torch.Tensor) -> torch.Tensor:
        """fftn over (x, z) with norm='ortho' / prod(L/2pi), then per axis
        [0:ks], [ks], conj([ks]), [N-ks+1:N] (domain.py:716-764)."""
        g = self.grid_math
        out = torch.fft.fftn(f.to(C128), dim=(0, 2), norm="ortho") / g.scaling
        for axis in (0, 2):
            n = f.shape[axis]
            ks = int((n - n * (1 - 1 / self.aliasing)) / 2)
            o1, o2, o3 = out.narrow(axis, 0, ks), out.narrow(axis, ks, 1), out.narrow(axis, n - ks + 1, ks - 1)
            out = torch.cat([o1, o2, torch.conj(o2), o3], dim=axis)
        return out

    def diff(self, f: torch.Tensor, direction: int, order: int = 1) -> torch.Tensor:
        """Physical-space derivative along y (domain.py:355-373)."""
        if direction != 1:
            raise NotImplementedError("physical-space periodic derivatives are not part of the step")
        Dm = _t(np.linalg.matrix_power(self.grid_math.D, order), f)
        return torch.einsum("ijk,lj->ilk", f, Dm.to(f.dtype))

    def mgrid(self):
        return np.meshgrid(*self.grid, indexing="ij")


@dataclass(frozen=True)
class FourierDomain:
    """Spectral domain (domain.py:769-1242)."""

    number_of_dimensions: int
    periodic_directions: Tuple[bool, ...]
    scale_factors: Tuple[float, ...]
    shape: Tuple[int, ...]
    grid: Tuple[np.ndarray, ...]
    physical_domain: PhysicalDomain
    aliasing: float = 3 / 2

    @classmethod
    def FromPhysicalDomain(cls, pd: PhysicalDomain) -> "FourierDomain":
        g = pd.grid_math
        return cls(3, pd.periodic_directions, pd.scale_factors, pd.shape, (g.kx, g.y, g.kz), pd, pd.aliasing)

    @property
    def grid_math(self):
        return self.physical_domain.grid_math

    def get_shape(self) -> Tuple[int, ...]:
        return self.shape

    def get_shape_aliasing(self) -> Tuple[int, ...]:
        return self.physical_domain.get_shape_aliasing()

    def number_of_cells(self, direction: int) -> int:
        return len(self.grid[direction])

    def all_dimensions(self) -> Sequence[int]:
        return range(self.number_of_dimensions)

    def all_periodic_dimensions(self):
        return [0, 2]

    def all_nonperiodic_dimensions(self):
        return [1]

    def field_no_hat(self, fh: torch.Tensor) -> torch.Tensor:
        """Zero-pad [0:ks+1], zeros(ceil(N(a-1))), [N-ks:N] per axis, then
        ifftn(ortho).real * prod(L/2pi) (domain.py:1167-1242)."""
        g = self.grid_math
        for axis in (0, 2):
            n = fh.shape[axis]
            ks = int(n / 2)
            nz = math.ceil(n * (self.aliasing - 1))
            zs = list(fh.shape)
            zs[axis] = nz
            fh = torch.cat(
                [fh.narrow(axis, 0, ks + 1), torch.zeros(zs, dtype=fh.dtype, device=fh.device), fh.narrow(axis, n - ks, ks)],
                dim=axis,
            )
        return torch.fft.ifftn(fh, dim=(0, 2), norm="ortho").real / (1 / g.scaling)

    def diff(self, fh: torch.Tensor, direction: int, order: int = 1) -> torch.Tensor:
        """(i k)^order f in x/z; matrix_power(D, order) along y (domain.py:891-918)."""
        g = self.grid_math
        if direction == 0:
            return _t((1j * g.kx) ** order, fh, C128)[:, None, None] * fh
        if direction == 2:
            return _t((1j * g.kz) ** order, fh, C128)[None, None, :] * fh
        Dm = _t(np.linalg.matrix_power(g.D, order), fh, C128)
        return torch.einsum("ijk,lj->ilk", fh, Dm)

    def curl(self, vh: torch.Tensor) -> torch.Tensor:
        """domain.py:920-933."""
        u_y, u_z = self.diff(vh[0], 1), self.diff(vh[0], 2)
        v_x, v_z = self.diff(vh[1], 0), self.diff(vh[1], 2)
        w_x, w_y = self.diff(vh[2], 0), self.diff(vh[2], 1)
        return torch.stack([w_y - v_z, u_z - w_x, v_x - u_y])

    def divergence(self, vh: torch.Tensor) -> torch.Tensor:
        return self.diff(vh[0], 0) + self.diff(vh[1], 1) + self.diff(vh[2], 2)
