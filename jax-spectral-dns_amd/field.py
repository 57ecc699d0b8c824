"""Field surface of the reference (jax_spectral_dns/field.py), numerics only.

Fields wrap torch tensors (complex128 spectral, float64 physical) on any
device.  ``PhysicalField.FromFunc`` / ``.hat()`` / ``.energy()``,
``FourierField.no_hat()`` and ``VectorField`` keep the reference's names and
semantics (field.py:294-611, 1154-1473, 2764-3091).  Field files are HDF5 in
the reference's layout (save_to_file / FromFile, through the package's own
reader and writer); plotting is not part of this build.
"""

from __future__ import annotations

import os
from typing import Callable, Generic, List, Optional, Sequence, Tuple, TypeVar, Union

import numpy as np
import torch

from . import hdf5
from .domain import FourierDomain, PhysicalDomain

T = TypeVar("T", "PhysicalField", "FourierField")


def _field_path(filename: str) -> str:
    """A name not starting with "." or "/" lives in Field.field_dir
    (field.py:153-160)."""
    if filename[0] in "./":
        return filename
    os.makedirs(Field.field_dir, exist_ok=True)
    return Field.field_dir + "/" + filename


def _numpy(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def _read_group(filename: str, name: str, time_step: int) -> Tuple[hdf5.Dataset, int]:
    """Dataset `name` of group str(time_step); a negative time_step counts from
    the last group in h5py's iteration order (field.py:409-431)."""
    tree = hdf5.read(filename)
    if time_step < 0:
        time_step = int(hdf5.names(tree)[-1]) + time_step + 1
    grp = tree.get(str(time_step))
    assert isinstance(grp, dict), "time step %d not found, only %s found." % (time_step, hdf5.names(tree))
    dset = grp.get(name)
    assert isinstance(dset, hdf5.Dataset), "dataset %s not found, only %s found." % (name, hdf5.names(grp))
    return dset, time_step


class Field:
    """Common arithmetic of physical and spectral fields (field.py:81-291)."""

    field_dir = os.environ.get("JAX_SPECTRAL_DNS_FIELD_DIR", "./fields/")

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

    def save_to_hdf_file(self, filename: str) -> None:
        """Group str(time_step) with a gzip-9 dataset named after the field and
        its int64 attribute time_step (field.py:144-151)."""
        ds = hdf5.Dataset(_numpy(self.data), {"time_step": np.int64(self.time_step)})
        hdf5.write(filename, {str(self.get_time_step()): {self.name: ds}}, compression=9)

    def save_to_file(self, filename: str) -> None:
        """field.py:153-165."""
        self.save_to_hdf_file(_field_path(filename))


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

    @classmethod
    def read_hdf(cls, filename: str, name: str, time_step: int) -> torch.Tensor:
        """field.py:1352-1373."""
        return torch.from_numpy(np.array(_read_group(filename, name, time_step)[0].data))

    @classmethod
    def FromFile(
        cls, domain: PhysicalDomain, filename: str, name: str = "field", time_step: int = -1, device=None
    ) -> "PhysicalField":
        """A field saved by save_to_file (field.py:1375-1420); its shape must be
        the domain's."""
        data = cls.read_hdf(_field_path(filename), name, time_step)
        if tuple(data.shape) != tuple(domain.get_shape_aliasing()):
            raise ValueError("field of shape %s does not match the domain %s" % (tuple(data.shape), domain.get_shape_aliasing()))
        return cls(domain, data.to(device=device), name)


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

    def field_2d(self, direction: int, wavenumber: int = 0) -> "FourierField":
        """Keep the modes of index wavenumber (and N - wavenumber) along
        direction, zero the rest (field.py:2912-2935)."""
        n = self.data.shape[direction]
        keep = [wavenumber] if wavenumber == 0 else [wavenumber, n - wavenumber]
        out = torch.zeros_like(self.data)
        idx = [slice(None)] * self.data.dim()
        for k in keep:
            idx[direction] = slice(k, k + 1)
            out[tuple(idx)] = out[tuple(idx)] + self.data[tuple(idx)]
        return FourierField(self.physical_domain, out, self.name)

    @classmethod
    def FromFile(
        cls, domain: PhysicalDomain, filename: str, name: str = "field_hat", time_step: int = -1, device=None
    ) -> "FourierField":
        data = torch.from_numpy(np.array(_read_group(_field_path(filename), name, time_step)[0].data))
        if tuple(data.shape) != tuple(domain.get_shape()):
            raise ValueError("field of shape %s does not match the domain %s" % (tuple(data.shape), domain.get_shape()))
        return cls(domain, data.to(device=device), name)


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

    def field_2d(self, direction: int, wavenumber: int = 0) -> "VectorField":
        """field.py:636-639."""
        return VectorField([f.field_2d(direction, wavenumber) for f in self.elements], self.name)  # type: ignore[union-attr]

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

    def save_to_hdf_file(self, filename: str) -> None:
        """Group str(time_step) with one gzip-9 dataset (components stacked on
        axis 0) named after the field, no attribute (field.py:718-736)."""
        hdf5.write(filename, {str(self.get_time_step()): {self.name: hdf5.Dataset(_numpy(self.get_data()))}}, compression=9)

    def save_to_file(self, filename: str) -> None:
        """field.py:740-762."""
        self.save_to_hdf_file(_field_path(filename))

    @classmethod
    def read_hdf(cls, filename: str, name: str, time_step: int) -> Tuple[torch.Tensor, int]:
        """field.py:409-431."""
        dset, ts = _read_group(filename, name, time_step)
        return torch.from_numpy(np.array(dset.data)), ts

    @classmethod
    def FromFile(
        cls, domain: PhysicalDomain, filename: str, name: str = "field", time_step: int = -1, device=None
    ) -> "VectorField":
        """A vector field saved by save_to_file (field.py:433-470): physical
        components when the stored arrays have the domain's physical shape,
        spectral components when they have its mode shape."""
        data, ts = cls.read_hdf(_field_path(filename), name, time_step)
        data = data.to(device=device)
        if tuple(data.shape[1:]) == tuple(domain.get_shape_aliasing()) and not data.is_complex():
            out = cls.FromData(PhysicalField, domain, data, name)
        elif tuple(data.shape[1:]) == tuple(domain.get_shape()):
            out = cls.FromData(FourierField, domain, data, name)
        else:
            raise ValueError("field of shape %s does not match the domain" % (tuple(data.shape),))
        out.set_time_step(ts)
        return out

    def curl(self) -> "VectorField":
        fd = self.elements[0].fourier_domain  # type: ignore[union-attr]
        c = fd.curl(self.get_data())
        return VectorField.FromData(FourierField, self.elements[0].physical_domain, c, "vort_hat")  # type: ignore[union-attr]

    def div(self):
        fd = self.elements[0].fourier_domain  # type: ignore[union-attr]
        return FourierField(self.elements[0].physical_domain, fd.divergence(self.get_data()), "div_hat")  # type: ignore[union-attr]
