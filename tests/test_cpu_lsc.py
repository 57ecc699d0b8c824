"""Host-side linear stability / optimal transient growth
(linear_stability_calculation.py:34-638), pinned on published answers:
Orszag's (1971) least stable Orr-Sommerfeld eigenvalue at Re = 10000,
alpha = 1, and the Reddy & Henningson (1993) transient-growth curve at
Re = 3000, alpha = 1, beta = 0 that the reference ships
(rh93_transient_growth.csv, digitised; copied as a fixture)."""

import os

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd.domain import PhysicalDomain
from jax_spectral_dns_amd.linear_stability_calculation import LinearStabilityCalculation, clenshaw_curtis_weights
from oracle import jsd_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("n", [9, 10, 33, 64])
def test_clenshaw_curtis_is_exact_on_polynomials(n):
    y = np.cos(np.pi * np.arange(n) / (n - 1))
    w = clenshaw_curtis_weights(n)
    for p in range(0, n - 1):
        exact = 0.0 if p % 2 else 2.0 / (p + 1)
        assert abs((w * y**p).sum() - exact) < 1e-13


def test_orszag_eigenvalue():
    """Least stable mode of plane Poiseuille flow at Re = 10000, alpha = 1:
    c = 0.23752649 + 0.00373967 i (Orszag, J. Fluid Mech. 50, 1971)."""
    lsc = LinearStabilityCalculation(Re=10000.0, alpha=1.0, beta=0.0, n=120)
    lam, _ = lsc.calculate_eigenvalues()
    c = 1j * lam[0] / lsc.alpha
    assert abs(c.real - 0.23752649) < 1e-8
    assert abs(c.imag - 0.00373967) < 1e-8


def test_transient_growth_matches_reddy_henningson_1993():
    """G(t) at Re = 3000, alpha = 1, beta = 0 against the digitised curve: within
    3 % on t in [2, 44] (digitisation), peak 20.36 near t = 15."""
    data = np.genfromtxt(os.path.join(HERE, "golden", "rh93_transient_growth.csv"), delimiter=",").T
    lsc = LinearStabilityCalculation(Re=3000.0, alpha=1.0, beta=0.0, n=100)
    lsc.calculate_transient_growth_max_energy(15.0, 60, Ts=list(data[0]))
    G = np.array(lsc.Gs)
    assert abs(G[0] - 1.0) < 1e-12
    sel = (data[0] > 2.0) & (data[0] < 44.0)
    assert np.all(np.abs(G[sel] - data[1][sel]) / data[1][sel] < 0.03)
    assert abs(max(G) - 20.36) < 0.05


def test_optimal_initial_condition_is_divergence_free_and_scaled():
    """The optimal initial condition on a DNS grid (n = Ny): a single (alpha, beta)
    wave, divergence free, at the requested energy."""
    dom = PhysicalDomain.create((8, 33, 8), (True, False, True), scale_factors=(2 * np.pi, 1.0, 2 * np.pi))
    lsc = LinearStabilityCalculation(Re=1000.0, alpha=1.0, beta=1.0, n=33)
    u = lsc.calculate_transient_growth_initial_condition(dom, 10.0, 30, energy=1e-4)
    d = O.make_domain((8, 33, 8))
    div = O.diff(d, u[0], 0, 1) + O.diff(d, u[1], 1, 1) + O.diff(d, u[2], 2, 1)
    assert torch.linalg.norm(div) < 1e-9 * torch.linalg.norm(u)
    assert abs(float(O.energy(d, u)) - 1e-4) < 1e-12
    mags = torch.linalg.norm(u, dim=(0, 2))  # per (kx, kz)
    nz = [tuple(ij) for ij in torch.nonzero(mags > 1e-12 * mags.max()).tolist()]
    M, Z = mags.shape
    assert sorted(nz) == [(1, 1), (M - 1, Z - 1)]  # (alpha, beta) and its conjugate only


def test_neutral_point_of_plane_poiseuille():
    """Critical point of plane Poiseuille flow, Re = 5772.22, alpha = 1.02056:
    neutral least stable mode with phase speed c_r = 0.26400174 (Orszag 1971),
    the case of the reference's Tollmien-Schlichting growth-rate test
    (tests/test_project.py:1320-1500)."""
    lsc = LinearStabilityCalculation(Re=5772.22, alpha=1.02056, beta=0.0, n=100)
    lam, _ = lsc.calculate_eigenvalues()
    c = 1j * lam[0] / lsc.alpha
    assert abs(c.real - 0.26400174) < 1e-7
    assert abs(lam[0].real) < 1e-7


def test_single_mode_field_and_energy_over_time():
    """velocity_field_single_mode: one (alpha, 0) wave on the DNS grid,
    divergence free, no-slip; energy_over_time(0) is its energy and the
    energy grows as exp(2 Re(lambda) t)."""
    shape = (4, 33, 4)
    dom = PhysicalDomain.create(shape, (True, False, True), scale_factors=(2 * np.pi / 1.02056, 1.0, 1e-3))
    lsc = LinearStabilityCalculation(Re=6000.0, alpha=1.02056, beta=0.0, n=33)
    u = lsc.velocity_field_single_mode(dom)
    phys = torch.stack([c.get_data() for c in u])
    assert phys.shape == (3,) + tuple(dom.get_shape_aliasing())
    # walls
    assert float(phys[:, :, 0].abs().max()) < 1e-12 * float(phys.abs().max())
    assert float(phys[:, :, -1].abs().max()) < 1e-12 * float(phys.abs().max())
    hat = torch.stack([dom.field_hat(phys[c]) for c in range(3)])
    d = O.make_domain(shape, 2 * np.pi / 1.02056, 1e-3)
    div = O.diff(d, hat[0], 0, 1) + O.diff(d, hat[1], 1, 1) + O.diff(d, hat[2], 2, 1)
    assert torch.linalg.norm(div) < 1e-9 * torch.linalg.norm(hat)
    mags = torch.linalg.norm(hat, dim=(0, 2))
    nz = sorted(tuple(ij) for ij in torch.nonzero(mags > 1e-12 * mags.max()).tolist())
    assert nz == [(1, 0), (hat.shape[1] - 1, 0)]
    fn, lam = lsc.energy_over_time(dom, eps=0.1)
    e0 = sum(float(c.energy()) for c in u)
    assert abs(fn(0.0) - 0.01 * e0) < 1e-14 * e0
    assert abs(fn(2.0) / fn(0.0) - np.exp(4.0 * lam.real)) < 1e-14
    assert abs(float(O.energy(d, hat)) - e0) < 1e-12 * e0


def test_single_mode_fields_saved_as_field_files(tmp_path, monkeypatch):
    """save=True writes the three components as field files
    (linear_stability_calculation.py:262-265) that FromFile reads back."""
    from jax_spectral_dns_amd.field import Field, PhysicalField

    monkeypatch.setattr(Field, "field_dir", str(tmp_path))
    dom = PhysicalDomain.create((4, 33, 4), (True, False, True), scale_factors=(2 * np.pi, 1.0, 1e-3))
    lsc = LinearStabilityCalculation(Re=6000.0, alpha=1.0, beta=0.0, n=33)
    u = lsc.velocity_field_single_mode(dom, save=True)
    for c, comp in zip("uvw", u):
        back = PhysicalField.FromFile(dom, "%s_Re_%s_n_%d_mode_%d" % (c, lsc.Re, lsc.n, 0), name=comp.name)
        assert torch.equal(back.get_data(), comp.get_data())
