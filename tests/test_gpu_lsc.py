"""End to end against a published answer: the optimal initial condition of
the host linear stability calculation, advanced by the HIP linearised
perturbation step, attains the optimal transient growth of plane Poiseuille
flow at Re = 3000, alpha = 1, beta = 0 over T = 15: G = 20.36 (Reddy &
Henningson 1993, the curve the reference ships as rh93_transient_growth.csv)."""

import numpy as np
import pytest

from jax_spectral_dns_amd import KIND_PERTURBATION
from jax_spectral_dns_amd.domain import PhysicalDomain
from jax_spectral_dns_amd.linear_stability_calculation import LinearStabilityCalculation
from jax_spectral_dns_amd.plan import StepPlan

pytestmark = pytest.mark.gpu


def test_linearised_step_attains_optimal_transient_growth():
    re, dt, T = 3000.0, 0.01, 15.0
    dom = PhysicalDomain.create((8, 65, 4), (True, False, True), scale_factors=(2 * np.pi, 1.0, 2 * np.pi))
    lsc = LinearStabilityCalculation(Re=re, alpha=1.0, beta=0.0, n=65)
    G = lsc.calculate_transient_growth_max_energy(T, 60)
    u0 = lsc.calculate_transient_growth_initial_condition(dom, T, 60, energy=1e-6).to("cuda")
    plan = StepPlan(dom.grid_math, re, dt, KIND_PERTURBATION, linearise=True)
    e0 = plan.energy(u0)
    u = u0.clone()
    plan.step(u, 0.0, int(round(T / dt)))
    gain = plan.energy(u) / e0
    print("DNS gain %.6f, stability calculation G(T) %.6f" % (gain, G))
    assert abs(G - 20.36) < 0.02
    assert abs(gain - G) < 2e-3 * G
