"""Shared helpers for the tests: seeded inputs and relative errors."""

from __future__ import annotations

import numpy as np
import torch

from jax_spectral_dns_amd import domain_math as DM
from jax_spectral_dns_amd import initial
from oracle import jsd_oracle as O


def rel(a, b) -> float:
    a = a.detach().cpu() if isinstance(a, torch.Tensor) else torch.as_tensor(a)
    b = b.detach().cpu() if isinstance(b, torch.Tensor) else torch.as_tensor(b)
    return float(torch.linalg.norm((a - b).reshape(-1)) / torch.linalg.norm(b.reshape(-1)))


def pair(shape, lx=2 * np.pi, lz=2 * np.pi):
    """(product grid, oracle domain) for the same box."""
    return DM.make_grid(shape, lx, lz), O.make_domain(shape, lx, lz)


def perturbation_field(g, seed=0, energy=1e-6):
    return torch.from_numpy(initial.perturbation(g, seed=seed, energy=energy))


def ns_field(g, seed=0, frac=0.05):
    """Laminar Poiseuille flow plus a seeded perturbation at frac of its energy."""
    base = initial.laminar(g)
    e_base = initial.spectral_energy(g, base)
    return torch.from_numpy(base + initial.perturbation(g, seed=seed, energy=frac * e_base))
