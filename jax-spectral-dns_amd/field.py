"""Field surface of the reference (jax_spectral_dns/field.py), numerics only.

Fields wrap torch tensors (complex128 spectral, float64 physical) on any
device.  ``PhysicalField.FromFunc`` / ``.hat()`` / ``.energy()``,
``FourierField.no_hat()`` and ``VectorField`` keep the reference's names and
semantics (field.py:294-611, 1154-1473, 2764-3091); plotting and HDF5 I/O are
not part of this build.
"""

from __future__ import annotations

from typing import Callable, Generic, List, Optional, Sequence, TypeVar, Union

import numpy as np
import torch

from .domain import FourierDomain, PhysicalDomain

T = TypeVar("T", "PhysicalField", "FourierField")


class Field:
    """Common arithmetic of physical and spectral fields (field.py:81-291)."""

    def __init__(self, data: torch.Tensor, name: str = "field"):
        self.data = data
        self.name = name
        self.time_step = 0

    def get_data(self) -> torch.Tensor:
        return self.data

    def get_name(self) -> str:
        return self.name

    def set_name(self, name: str) -> None:
        self.name = name

    def set_time_step(self, t: int) -> None:
        self.time_step = t

    def get_time_step(self) -> int:
        return self.time_step

    def _new(self, data: torch.Tensor):
        raise NotImplementedError

    def __add__(self, other):
        return self._new(self.data + (other.data if isinstance(other, Field) else other))

    def __sub__(self, other):
        return self._new(self.data - (other.data if isinstance(other, Field) else other))

    def __mul__(self, other):
        return self._new(self.data * (other.data if isinstance(other, Field) else other))

    __rmul__ = __mul__

    def __neg__(self):
        return self._new(-self.data)

    def __abs__(self) -> float:
        """norm / number of dofs (field.py:235-236)."""
        return float(torch.linalg.norm(self.data.reshape(-1)) / self.data.numel())


class PhysicalField(Field):
    def __init__(self, physical_domain: PhysicalDomain, data: torch.Tensor, name: str = "field"):
        super().__init__(data, name)
        self.physical_domain = physical_domain

    def _new(self, data):
        return PhysicalField(self.physical_domain, data, self.name)

    @classmethod
    def FromFunc(cls, domain: PhysicalDomain, func: Callable, name: str = "field", device=None) -> "PhysicalField":
        X = np.meshgrid(*domain.grid, indexing="ij")
        return cls(domain, torch.as_tensor(func(X), dtype=torch.float64, device=device), name)

    @classmethod
    def Zeros(cls, domain: PhysicalDomain, name: str = "field", device=None) -> "PhysicalField":
        return cls(domain, torch.zeros(domain.get_shape_aliasing(), dtype=torch.float64, device=device), name)

    def hat(self) -> "FourierField":
        return FourierField(self.physical_domain, self.physical_domain.field_hat(self.data), self.name + "_hat")

    def volume_integral(self) -> torch.Tensor:
        """z and x: L/N sum; y: the definite-integral quadrature
        (field.py:1456-1466, 2618-2711)."""
        g = self.physical_domain.grid_math
        w = torch.as_tensor(g.y_weights, device=self.data.device)
        fz = self.physical_domain.scale_factors[2] / self.data.shape[2] * self.data.sum(dim=2)
        fy = (fz * w[None, :]).sum(dim=1)
        return self.physical_domain.scale_factors[0] / self.data.shape[0] * fy.sum()

    def energy(self) -> float:
        vol = 2.0 * self.physical_domain.scale_factors[0] * self.physical_domain.scale_factors[2]
        return float((0.5 * self * self).volume_integral() / vol)

    def energy_with_other(self, other: "PhysicalField") -> float:
        vol = 2.0 * self.physical_domain.scale_factors[0] * self.physical_domain.scale_factors[2]
        return float((0.5 * self * other).volume_integral() / vol)

    def diff(self, direction: int, order: int = 1) -> "PhysicalField":
        return PhysicalField(self.physical_domain, self.physical_domain.diff(self.data, direction, order), self.name)


class FourierField(Field):
    def __init__(self, physical_domain: PhysicalDomain, data: torch.Tensor, name: str = "field_hat"):
        super().__init__(data, name)
        self.physical_domain = physical_domain
        self.fourier_domain: FourierDomain = physical_domain.hat()

    def _new(self, data):
        return FourierField(self.physical_domain, data, self.name)

    @classmethod
    def Zeros(cls, domain: PhysicalDomain, name: str = "field_hat", device=None) -> "FourierField":
        return cls(domain, torch.zeros(domain.get_shape(), dtype=torch.complex128, device=device), name)

    def no_hat(self) -> PhysicalField:
        return PhysicalField(self.physical_domain, self.fourier_domain.field_no_hat(self.data), self.name.replace("_hat", ""))

    def diff(self, direction: int, order: int = 1) -> "FourierField":
        return FourierField(self.physical_domain, self.fourier_domain.diff(self.data, direction, order), self.name)

    def energy(self) -> float:
        return self.no_hat().energy()


class VectorField(Generic[T]):
    """List of scalar fields with vector operations (field.py:294-611)."""

    def __init__(self, elements: Sequence[T], name: Optional[str] = None):
        self.elements: List[T] = list(elements)
        self.name = name or (self.elements[0].name if self.elements else "vector")
        self.time_step = 0

    @classmethod
    def FromData(cls, field_cls, domain: PhysicalDomain, data: torch.Tensor, name: str = "field") -> "VectorField":
        return cls([field_cls(domain, data[i], name + "_" + "xyz"[i]) for i in range(data.shape[0])], name)

    def __getitem__(self, i: int) -> T:
        return self.elements[i]

    def __len__(self) -> int:
        return len(self.elements)

    def __iter__(self):
        return iter(self.elements)

    def get_data(self) -> torch.Tensor:
        return torch.stack([f.data for f in self.elements])

    def get_name(self) -> str:
        return self.name

    def set_name(self, name: str) -> None:
        self.name = name

    def set_time_step(self, t: int) -> None:
        self.time_step = t

    def get_time_step(self) -> int:
        return self.time_step

    def hat(self) -> "VectorField":
        return VectorField([f.hat() for f in self.elements], self.name + "_hat")  # type: ignore[union-attr]

    def no_hat(self) -> "VectorField":
        return VectorField([f.no_hat() for f in self.elements], self.name.replace("_hat", ""))  # type: ignore[union-attr]

    def energy(self) -> float:
        return float(sum(f.energy() for f in self.elements))

    def energy_with_other(self, other: "VectorField") -> float:
        return float(sum(a.energy_with_other(b) for a, b in zip(self.elements, other.elements)))  # type: ignore[union-attr]

    def _combine(self, other, op) -> "VectorField":
        if isinstance(other, VectorField):
            return VectorField([op(a, b) for a, b in zip(self.elements, other.elements)], self.name)
        return VectorField([op(a, other) for a in self.elements], self.name)

    def __add__(self, other):
        return self._combine(other, lambda a, b: a + b)

    def __sub__(self, other):
        return self._combine(other, lambda a, b: a - b)

    def __mul__(self, other):
        return self._combine(other, lambda a, b: a * b)

    __rmul__ = __mul__

    def __neg__(self):
        return VectorField([-a for a in self.elements], self.name)

    def curl(self) -> "VectorField":
        fd = self.elements[0].fourier_domain  # type: ignore[union-attr]
        c = fd.curl(self.get_data())
        return VectorField.FromData(FourierField, self.elements[0].physical_domain, c, "vort_hat")  # type: ignore[union-attr]

    def div(self):
        fd = self.elements[0].fourier_domain  # type: ignore[union-attr]
        return FourierField(self.elements[0].physical_domain, fd.divergence(self.get_data()), "div_hat")  # type: ignore[union-attr]
