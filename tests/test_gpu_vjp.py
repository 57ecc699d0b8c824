"""Discrete adjoint (VJP) of the linearised perturbation step.

* transposition identity <A x, y> = <x, A^T y> between the GPU forward step
  and the GPU adjoint (real-pair inner product), independent of the oracle;
* the gain gradient dJ/du0, J = E(u_T)/E(u_0), against torch.autograd of
  the oracle (the stand-in for jax.grad of the reference), <= 1e-9;
* the same gradient against centred finite differences along 16 seeded
  random directions, <= 1e-8 relative: numerator and denominator energies
  are exact quadratics along a line, so their centred differences carry no
  truncation error and only ~1e-14 rounding.
"""

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import KIND_PERTURBATION
from jax_spectral_dns_amd.plan import StepPlan
from jax_spectral_dns_amd.vjp import gain_and_grad
from oracle import jsd_oracle as O
from tests.helpers import pair, perturbation_field, rel

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _dot(a, b):
    return float((torch.view_as_real(a) * torch.view_as_real(b)).sum())


@pytest.mark.parametrize("shape,lx,lz", [((24, 65, 4), 2 * np.pi, 2 * np.pi), ((24, 33, 16), 2 * np.pi, np.pi)])
def test_transposition_identity(shape, lx, lz):
    g, _ = pair(shape, lx, lz)
    plan = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=True)
    x = perturbation_field(g, seed=21, energy=1.0).to(DEV)
    y = perturbation_field(g, seed=22, energy=1.0).to(DEV)
    # arbitrary (non-solenoidal, non-Hermitian) directions exercise every path
    rng = np.random.default_rng(3)
    x = x + 0.1 * torch.from_numpy(rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape)).to(DEV)
    x[:, :, 0] = 0
    x[:, :, -1] = 0
    for n in (1, 3):
        ax = x.clone()
        plan.step(ax, 0.0, n)
        aty = y.clone()
        plan.vjp(aty, n)
        lhs, rhs = _dot(ax, y), _dot(x, aty)
        print("n=%d <Ax,y>=%.15e <x,A^T y>=%.15e" % (n, lhs, rhs))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1e-30) + 1e-13 * float(ax.abs().max() * y.abs().max())


def test_gain_gradient_against_oracle_autograd():
    shape = (24, 65, 4)
    g, d = pair(shape)
    re, dt, n = 3000.0, 0.01, 20
    u0 = perturbation_field(g, seed=0, energy=1e-6)
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=True)
    J, grad = gain_and_grad(u0.to(DEV), plan, n, checkpoint_every=5)
    sp = O.make_perturbation(d, re, dt, True)
    x = torch.view_as_real(u0).clone().requires_grad_(True)
    u = torch.view_as_complex(x)
    Jr = O.energy(d, O.run(sp, u, n)) / O.energy(d, u)
    Jr.backward()
    gref = torch.view_as_complex(x.grad)
    err = rel(grad, gref)
    print("J %.15f vs %.15f, grad rel %.3e" % (J, float(Jr), err))
    assert abs(J - float(Jr)) / float(Jr) < 1e-11
    assert err < 1e-9


def test_gain_gradient_against_finite_differences():
    shape = (48, 97, 32)
    g, _ = pair(shape)
    re, dt, n = 3000.0, 0.01, 100
    u0 = perturbation_field(g, seed=0, energy=1e-6).to(DEV)
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=True)
    J, grad = gain_and_grad(u0, plan, n, checkpoint_every=50)

    def energies(u):
        v = u.clone()
        plan.step(v, 0.0, n)
        return plan.energy(v), plan.energy(u)

    # The step is linear and the energy quadratic, so N(e) = E(A(u0 + e d)) and
    # D(e) = E(u0 + e d) are quadratics in e: their centred differences are
    # exact for any h (no truncation error), and h of the size of u0 keeps the
    # rounding noise at ~1e-14 relative.  J' = (N' D - N D') / D^2.
    h = 0.5
    N0, D0 = energies(u0)
    assert abs(N0 / D0 - J) <= 1e-12 * abs(J)
    worst = 0.0
    for k in range(16):
        rng = np.random.default_rng(1 + k)
        dirn = torch.from_numpy(rng.standard_normal(u0.shape) + 1j * rng.standard_normal(u0.shape)).to(DEV)
        dirn[:, :, 0] = 0
        dirn[:, :, -1] = 0
        dirn *= torch.linalg.norm(u0) / torch.linalg.norm(dirn)
        Np, Dp = energies(u0 + h * dirn)
        Nm, Dm = energies(u0 - h * dirn)
        dN, dD = (Np - Nm) / (2 * h), (Dp - Dm) / (2 * h)
        fd = (dN * D0 - N0 * dD) / (D0 * D0)
        ad = _dot(grad, dirn)
        e = abs(fd - ad) / max(abs(ad), 1e-300)
        worst = max(worst, e)
        print("dir %2d fd %.12e ad %.12e rel %.2e" % (k, fd, ad, e))
    assert worst < 1e-8


def test_linearised_1500_step_gradient_against_finite_differences():
    """The differentiation workload at full length: 48 x 97 x 32, Re 3000,
    dt 0.01, T = 15 (1500 steps), checkpoint every 50; the exact-quadratic
    centred differences along 16 seeded directions (<= 1e-8)."""
    shape = (48, 97, 32)
    g, _ = pair(shape)
    re, dt, n = 3000.0, 0.01, 1500
    u0 = perturbation_field(g, seed=0, energy=1e-6).to(DEV)
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=True)
    J, grad = gain_and_grad(u0, plan, n, checkpoint_every=50)

    def energies(u):
        v = u.clone()
        plan.step(v, 0.0, n)
        return plan.energy(v), plan.energy(u)

    h = 0.5
    N0, D0 = energies(u0)
    assert abs(N0 / D0 - J) <= 1e-12 * abs(J)
    worst = 0.0
    for k in range(16):
        rng = np.random.default_rng(1 + k)
        dirn = torch.from_numpy(rng.standard_normal(u0.shape) + 1j * rng.standard_normal(u0.shape)).to(DEV)
        dirn[:, :, 0] = 0
        dirn[:, :, -1] = 0
        dirn *= torch.linalg.norm(u0) / torch.linalg.norm(dirn)
        Np, Dp = energies(u0 + h * dirn)
        Nm, Dm = energies(u0 - h * dirn)
        dN, dD = (Np - Nm) / (2 * h), (Dp - Dm) / (2 * h)
        fd = (dN * D0 - N0 * dD) / (D0 * D0)
        ad = _dot(grad, dirn)
        e = abs(fd - ad) / max(abs(ad), 1e-300)
        worst = max(worst, e)
        print("dir %2d fd %.12e ad %.12e rel %.2e" % (k, fd, ad, e), flush=True)
    print("J(T = 15) = %.12f" % J)
    assert worst < 1e-8


def test_parameter_gradient_against_oracle_autograd():
    """dJ/d(omega_y, v, u_00, w_00) through vort_yvel_to_vel (the reference's
    optimisation parameters) against autograd of the oracle, <= 1e-9."""
    from jax_spectral_dns_amd.optimiser import gain_and_grad_parameters

    shape = (24, 33, 16)
    g, d = pair(shape, 2 * np.pi, np.pi)
    re, dt, n = 3000.0, 0.01, 8
    rng = np.random.default_rng(9)
    sh = (g.nxm, g.ny, g.nzm)
    vort = torch.from_numpy(rng.standard_normal(sh) + 1j * rng.standard_normal(sh)) * 1e-3
    vy = torch.from_numpy(rng.standard_normal(sh) + 1j * rng.standard_normal(sh)) * 1e-3
    ux = torch.from_numpy(rng.standard_normal(g.ny) + 0j) * 1e-3
    wz = torch.from_numpy(rng.standard_normal(g.ny) + 0j) * 1e-3
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=True)
    J, grads = gain_and_grad_parameters(plan, [t.to(DEV) for t in (vort, vy, ux, wz)], n, checkpoint_every=3)
    sp = O.make_perturbation(d, re, dt, True)
    leaves = [torch.view_as_real(t.clone()).requires_grad_(True) for t in (vort, vy, ux, wz)]
    u0 = O.vort_yvel_to_vel(d, *[torch.view_as_complex(x) for x in leaves])
    Jr = O.energy(d, O.run(sp, u0, n)) / O.energy(d, u0)
    ref = torch.autograd.grad(Jr, leaves)
    print("J %.15f vs %.15f" % (J, float(Jr)))
    assert abs(J - float(Jr)) / float(Jr) < 1e-11
    for name, a, b in zip(("vort", "vel_y", "vel_x_00", "vel_z_00"), grads, ref):
        err = rel(a, torch.view_as_complex(b))
        print(name, "grad rel", err)
        assert err < 1e-9, name


@pytest.mark.parametrize("shape", [(24, 65, 4), (24, 48, 16), (24, 33, 16)])
def test_eigen_adjoint_solve_matches_table_adjoint(monkeypatch, shape):
    """The eigen-form transpose of the solve with the h_v / h_g transpose folded
    in (k_solve_T8p, the default) against the table-form transpose and
    k_ygemm<GEMM_HVT> (JSD_SOLVET=table): the same adjoint to rounding over 1
    and 3 steps, odd and even Ny, linearised and nonlinear; and the
    transposition identity of the eigen form holds also for inputs with
    nonzero wall values, which enter the forward step through the wall
    columns of D and D^2."""
    from jax_spectral_dns_amd import KIND_NS

    g, _ = pair(shape, 2 * np.pi, np.pi)
    rng = np.random.default_rng(5)

    def rnd():
        return torch.from_numpy(rng.standard_normal((3, g.nxm, g.ny, g.nzm)) + 1j * rng.standard_normal((3, g.nxm, g.ny, g.nzm))).to(DEV)

    monkeypatch.delenv("JSD_SOLVET", raising=False)
    pe = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=True)
    monkeypatch.setenv("JSD_SOLVET", "table")
    pt = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=True)
    monkeypatch.delenv("JSD_SOLVET", raising=False)
    x = perturbation_field(g, seed=23, energy=1.0).to(DEV) + 0.1 * rnd()
    y = rnd()
    for n in (1, 3):
        a, b = y.clone(), y.clone()
        pe.vjp(a, n)
        pt.vjp(b, n)
        err = rel(a, b)
        print("linearised, n=%d: eigen vs table adjoint rel %.3e" % (n, err))
        assert err < 1e-12 * n
        ax = x.clone()
        pe.step(ax, 0.0, n)
        lhs, rhs = _dot(ax, y), _dot(x, a)
        print("  <Ax,y>=%.15e <x,A^T y>=%.15e (walls of x nonzero)" % (lhs, rhs))
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs) + 1e-13 * float(ax.abs().max() * y.abs().max())
    # nonlinear NS step, adjoint about a forward state
    u = perturbation_field(g, seed=24, energy=1e-2).to(DEV)
    monkeypatch.delenv("JSD_SOLVET", raising=False)
    ne = StepPlan(g, 3000.0, 0.005, KIND_NS, adjoint=True)
    monkeypatch.setenv("JSD_SOLVET", "table")
    nt = StepPlan(g, 3000.0, 0.005, KIND_NS, adjoint=True)
    monkeypatch.delenv("JSD_SOLVET", raising=False)
    a, b = y.clone(), y.clone()
    ne.vjp_state(u, a, -2.0 / 3000.0)
    nt.vjp_state(u, b, -2.0 / 3000.0)
    err = rel(a, b)
    print("NS: eigen vs table adjoint rel %.3e" % err)
    assert err < 1e-11
