"""Discrete adjoint of the nonlinear steps (NS and nonlinear perturbation),
linearised about the recomputed forward state (jsd_step_vjp_state).

* the gain gradient dJ/du0, J = E(u_T)/E(u_0), against torch.autograd of the
  oracle (the stand-in for jax.grad of the reference), <= 1e-9;
* against centred finite differences along seeded random directions: J is
  smooth, the h^2 truncation term is ~1e-10 of the derivative at h = 1e-4
  (relative to |u0|), so the bound is 1e-6;
* checkpoint-and-recompute with different chunk lengths gives the same
  gradient (the recompute is deterministic).
"""

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import KIND_NS, KIND_PERTURBATION
from jax_spectral_dns_amd.plan import StepPlan
from jax_spectral_dns_amd.vjp import gain_and_grad
from oracle import jsd_oracle as O
from tests.helpers import ns_field, pair, perturbation_field, rel

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _dot(a, b):
    return float((torch.view_as_real(a) * torch.view_as_real(b)).sum())


def _oracle_grad(sp, d, u0, n):
    x = torch.view_as_real(u0.clone()).requires_grad_(True)
    u = torch.view_as_complex(x)
    Jr = O.energy(d, O.run(sp, u, n)) / O.energy(d, u)
    Jr.backward()
    return float(Jr), torch.view_as_complex(x.grad)


@pytest.mark.parametrize("shape,lz", [((24, 33, 16), np.pi), ((24, 65, 16), np.pi), ((24, 33, 48), np.pi)])
def test_nonlinear_perturbation_gradient_against_oracle_autograd(shape, lz):
    """(24, 33, 48): z transform length 128, the register z passes of the
    forward and the adjoint."""
    g, d = pair(shape, 2 * np.pi, lz)
    re, dt, n = 3000.0, 0.005, 4
    u0 = perturbation_field(g, seed=0, energy=1e-2)
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=False, adjoint=True)
    J, grad = gain_and_grad(u0.to(DEV), plan, n, checkpoint_every=2)
    sp = O.make_perturbation(d, re, dt, False)
    Jr, gref = _oracle_grad(sp, d, u0, n)
    err = rel(grad, gref)
    print("J %.15f vs %.15f, grad rel %.3e" % (J, Jr, err))
    assert abs(J - Jr) / abs(Jr) < 1e-11
    assert err < 1e-9


def test_ns_gradient_against_oracle_autograd():
    g, d = pair((24, 33, 16), 2 * np.pi, np.pi)
    re, dt, n = 5000.0, 0.005, 3
    dpdx = -2.0 / re
    u0 = ns_field(g, seed=2)
    plan = StepPlan(g, re, dt, KIND_NS, adjoint=True)
    J, grad = gain_and_grad(u0.to(DEV), plan, n, checkpoint_every=3, dPdx=dpdx)
    sp = O.make_ns(d, re, dt, dpdx)
    Jr, gref = _oracle_grad(sp, d, u0, n)
    err = rel(grad, gref)
    print("J %.15f vs %.15f, grad rel %.3e" % (J, Jr, err))
    assert abs(J - Jr) / abs(Jr) < 1e-11
    assert err < 1e-9


def test_nonlinear_gradient_against_finite_differences():
    g, _ = pair((24, 33, 16), 2 * np.pi, np.pi)
    re, dt, n = 3000.0, 0.005, 6
    u0 = perturbation_field(g, seed=3, energy=1e-2).to(DEV)
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=False, adjoint=True)
    J, grad = gain_and_grad(u0, plan, n, checkpoint_every=4)

    def gain(u):
        v = u.clone()
        plan.step(v, 0.0, n)
        return plan.energy(v) / plan.energy(u)

    h = 1e-4
    worst = 0.0
    for k in range(8):
        rng = np.random.default_rng(1 + k)
        dirn = torch.from_numpy(rng.standard_normal(u0.shape) + 1j * rng.standard_normal(u0.shape)).to(DEV)
        dirn[:, :, 0] = 0
        dirn[:, :, -1] = 0
        dirn *= torch.linalg.norm(u0) / torch.linalg.norm(dirn)
        fd = (gain(u0 + h * dirn) - gain(u0 - h * dirn)) / (2 * h)
        ad = _dot(grad, dirn)
        e = abs(fd - ad) / max(abs(ad), 1e-300)
        worst = max(worst, e)
        print("dir %d fd %.12e ad %.12e rel %.2e" % (k, fd, ad, e))
    assert worst < 1e-6


def test_recompute_chunks_give_the_same_gradient():
    g, _ = pair((24, 33, 16), 2 * np.pi, np.pi)
    u0 = perturbation_field(g, seed=4, energy=1e-2).to(DEV)
    plan = StepPlan(g, 3000.0, 0.005, KIND_PERTURBATION, linearise=False, adjoint=True)
    J1, g1 = gain_and_grad(u0, plan, 6, checkpoint_every=1)
    J2, g2 = gain_and_grad(u0, plan, 6, checkpoint_every=6)
    J3, g3 = gain_and_grad(u0, plan, 6, checkpoint_every=4)
    assert J1 == J2 == J3
    assert rel(g1, g2) < 1e-14 and rel(g1, g3) < 1e-14


@pytest.mark.parametrize("shape", [(24, 65, 16), (48, 33, 8), (96, 17, 8)])
def test_register_adjoint_x_pass_matches_lds_kernels(monkeypatch, shape):
    """The merged register transpose of the inverse x pass (k_adj_x_fwd_r,
    the default) against the two-launch LDS kernels (JSD_XPASS=0): the same
    nonlinear-step adjoint to rounding, x lengths 40, 75 and 150."""
    from jax_spectral_dns_amd import KIND_NS

    g, _ = pair(shape, 2 * np.pi, np.pi)
    rng = np.random.default_rng(9)
    y = torch.from_numpy(rng.standard_normal((3, g.nxm, g.ny, g.nzm)) + 1j * rng.standard_normal((3, g.nxm, g.ny, g.nzm))).to("cuda")
    u = perturbation_field(g, seed=25, energy=1e-2).to("cuda")
    out = []
    for mode in ("1", "0"):
        monkeypatch.setenv("JSD_XPASS", mode)
        plan = StepPlan(g, 3000.0, 0.005, KIND_NS, adjoint=True)
        a = y.clone()
        plan.vjp_state(u, a, -2.0 / 3000.0)
        out.append(a)
    monkeypatch.delenv("JSD_XPASS", raising=False)
    err = rel(out[0], out[1])
    print(shape, "register vs LDS adjoint x pass rel", err)
    assert err < 1e-13


@pytest.mark.parametrize("kind", [KIND_NS, KIND_PERTURBATION])
@pytest.mark.parametrize("shape", [(24, 33, 64), (24, 32, 96)])
def test_register_z_passes_match_lds_kernels(monkeypatch, shape, kind):
    """The register z passes (k_fft_z_mid_r, k_adj_z_r; z lengths 128 and
    256, the default) against the LDS kernels (JSD_ZPASS=lds): the same
    nonlinear step and the same step adjoint to rounding."""
    g, _ = pair(shape, 2 * np.pi, np.pi)
    assert g.ppz == (256 if shape[2] >= 96 else 128)
    rng = np.random.default_rng(10)
    y = torch.from_numpy(rng.standard_normal((3, g.nxm, g.ny, g.nzm)) + 1j * rng.standard_normal((3, g.nxm, g.ny, g.nzm))).to(DEV)
    u = perturbation_field(g, seed=26, energy=1e-2).to(DEV)
    dpdx = -2.0 / 3000.0 if kind == KIND_NS else 0.0
    steps, adjs = [], []
    for mode in ("reg", "lds"):
        if mode == "lds":
            monkeypatch.setenv("JSD_ZPASS", "lds")
        kw = {} if kind == KIND_NS else {"linearise": False}
        plan = StepPlan(g, 3000.0, 0.005, kind, adjoint=True, **kw)
        s = u.clone()
        plan.step(s, dpdx, 2)
        steps.append(s)
        a = y.clone()
        plan.vjp_state(u, a, dpdx)
        adjs.append(a)
    monkeypatch.delenv("JSD_ZPASS", raising=False)
    es, ea = rel(steps[0], steps[1]), rel(adjs[0], adjs[1])
    print(shape, kind, "register vs LDS z passes: step rel %.3e, adjoint rel %.3e" % (es, ea))
    assert es < 1e-13
    assert ea < 1e-13


@pytest.mark.parametrize("kind", [KIND_NS, KIND_PERTURBATION])
@pytest.mark.parametrize("shape", [(24, 65, 16), (24, 33, 48), (24, 32, 96)])
def test_recorded_steps_gradient_matches_recompute(shape, kind):
    """checkpoint_every=0: every step recorded by jsd_step_record (substep
    states and kept transforms) and its adjoint taken by jsd_step_vjp_recorded
    with nothing recomputed, against the checkpoint-and-recompute gradient
    (the recorded substep states come from the forward's chained y-derivatives,
    the recomputed ones from fresh ones: equal to rounding); z lengths 32
    (LDS z pass), 128 and 256 (the five-row layout)."""
    g, _ = pair(shape, 2 * np.pi, np.pi)
    kw = {} if kind == KIND_NS else {"linearise": False}
    dpdx = -2.0 / 3000.0 if kind == KIND_NS else 0.0
    plan = StepPlan(g, 3000.0, 0.005, kind, adjoint=True, **kw)
    u0 = (ns_field(g, seed=4) if kind == KIND_NS else perturbation_field(g, seed=27, energy=1e-2)).to(DEV)
    J0, g0 = gain_and_grad(u0, plan, 5, checkpoint_every=2, dPdx=dpdx)
    J1, g1 = gain_and_grad(u0, plan, 5, checkpoint_every=0, dPdx=dpdx)
    err = rel(g1, g0)
    print(shape, kind, "recorded vs recompute: J %.3e, grad rel %.3e" % (abs(J1 - J0) / abs(J0), err))
    assert abs(J1 - J0) <= 1e-13 * abs(J0)
    assert err < 1e-11


def test_step_record_rejects_linearised_plans_and_short_records():
    g, _ = pair((24, 33, 16), 2 * np.pi, np.pi)
    lin = StepPlan(g, 3000.0, 0.005, KIND_PERTURBATION, linearise=True, adjoint=True)
    with pytest.raises(Exception):
        lin.record_elems()
    plan = StepPlan(g, 3000.0, 0.005, KIND_NS, adjoint=True)
    u = ns_field(g, seed=5).to(DEV)
    short = torch.empty(plan.record_elems() - 1, dtype=torch.complex128, device=DEV)
    with pytest.raises(Exception):
        plan.step_record(u, short)
