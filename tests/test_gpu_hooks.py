"""The step's plug-in points and options on the GPU against the oracle:

* the source term of get_source_term (navier_stokes.py:1697-1704,
  1461-1490) through jsd_step_source, forward and dual;
* enforce_base (navier_stokes.py:1581-1590);
* the nonlinear_update_fn plug-in (navier_stokes.py:324-332) through
  jsd_substep_terms, directly and through the equation class;
* a base flow other than the laminar profile (velocity_base_hat,
  navier_stokes_perturbation.py:351-380);
* the time-dependent linearise_switch / coupling_term_switch (:385-416);
* the dual's objectives gain_3d, gain_no_00 and dissipation
  (navier_stokes_perturbation_dual.py:903-1104) and their source term;
* the stages they use: laplacian, dissipation sum, volume integral.
Tolerances: 1e-11 relative per step (the step's parity bar), 1e-12 for
stages.
"""

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import KIND_DUAL, KIND_NS, KIND_PERTURBATION
from jax_spectral_dns_amd.domain import PhysicalDomain
from jax_spectral_dns_amd.field import FourierField, VectorField
from jax_spectral_dns_amd.navier_stokes import NavierStokesVelVort
from jax_spectral_dns_amd.navier_stokes_perturbation import NavierStokesVelVortPerturbation
from jax_spectral_dns_amd.navier_stokes_perturbation_dual import NavierStokesVelVortPerturbationDual
from jax_spectral_dns_amd.plan import StepPlan
from oracle import jsd_oracle as O
from tests.helpers import ns_field, pair, perturbation_field, rel

pytestmark = pytest.mark.gpu
DEV = "cuda"
SHAPE, LX, LZ = (16, 33, 16), 2 * np.pi, np.pi


def _source(g, seed):
    """A smooth source field with a (0,0) part (the perturbation generator's
    spectrum, so it vanishes at the walls like the reference's laplacian
    sources of wall-bounded states need not; any field is allowed)."""
    s = perturbation_field(g, seed=seed, energy=1.0)
    s[0, 0, :, 0] = torch.from_numpy(0.3 * (1 - g.y**2) + 0j)
    s[2, 0, :, 0] = torch.from_numpy(0.1 * g.y * (1 - g.y**2) + 0j)
    return s


@pytest.mark.parametrize("kind,linearise", [(KIND_NS, False), (KIND_PERTURBATION, False), (KIND_PERTURBATION, True)])
def test_step_with_source_term(kind, linearise):
    g, d = pair(SHAPE, LX, LZ)
    if kind == KIND_NS:
        re, dt, dpdx = 5000.0, 0.005, -2.0 / 5000.0
        u = ns_field(g, seed=40)
        sp = O.make_ns(d, re, dt, dpdx)
    else:
        re, dt, dpdx = 3000.0, 0.01, 0.0
        u = perturbation_field(g, seed=41, energy=1e-6 if linearise else 1e-2)
        sp = O.make_perturbation(d, re, dt, linearise)
    src = _source(g, 42) * 1e-2
    sp.source = lambda t: src
    plan = StepPlan(g, re, dt, kind, linearise=linearise)
    out, ref = u.to(DEV).clone(), u.clone()
    s_dev = src.to(DEV)
    for n in range(3):
        ref = O.rk_step(sp, ref, n)
        plan.step_source(out, dpdx, s_dev)
        err = rel(out, ref)
        print("step", n, "rel", err)
        assert err < 1e-11 * (n + 1)
    # the source changes the result (the test is not vacuous)
    sp.source = None
    assert rel(ref, O.run(sp, u, 3)) > 1e-6


def test_enforce_base():
    g, d = pair(SHAPE, LX, LZ)
    re, dt, dpdx = 5000.0, 0.005, -2.0 / 5000.0
    u = ns_field(g, seed=43)
    sp = O.make_ns(d, re, dt, dpdx)
    sp.enforce_base = True
    src = _source(g, 44) * 1e-2
    plan = StepPlan(g, re, dt, KIND_NS)
    out, ref = u.to(DEV).clone(), u.clone()
    for n in range(3):
        ref = O.rk_step(sp, ref, n)
        plan.step_source(out, dpdx, None, enforce_base=True)
        assert rel(out, ref) < 1e-11 * (n + 1)
    # with a source as well
    sp.source = lambda t: src
    s_dev = src.to(DEV)
    for n in range(2):
        ref = O.rk_step(sp, ref, n)
        plan.step_source(out, dpdx, s_dev, enforce_base=True)
        assert rel(out, ref) < 1e-11 * (n + 4)


def test_dual_step_with_source_term():
    g, d = pair(SHAPE, LX, LZ)
    re, dt = 3000.0, 0.01
    fwd_o = O.make_perturbation(d, re, dt, False)
    uf = perturbation_field(g, seed=45, energy=1e-2)
    v = perturbation_field(g, seed=46, energy=1e-2)
    src = _source(g, 47) * 1e-2
    dual_o = O.make_dual(fwd_o, lambda t: uf, False)
    dual_o.source = lambda t: src
    fwd = StepPlan(g, re, dt, KIND_PERTURBATION)
    dual = StepPlan(g, re, dt, KIND_DUAL, ops=fwd.ops, explicit_dt=-dt, plan_re=-re)
    out, ref = v.to(DEV).clone(), v.clone()
    for n in range(2):
        ref = O.rk_step(dual_o, ref, n)
        dual.step_source(out, 0.0, src.to(DEV), fwd_hat=uf.to(DEV))
        assert rel(out, ref) < 1e-11 * (n + 1)


def test_substep_terms_with_the_oracle_terms():
    """jsd_substep_terms with h_v, h_g, conv_ns computed by the oracle in
    every substep reproduces the oracle's step."""
    g, d = pair(SHAPE, LX, LZ)
    re, dt, dpdx = 5000.0, 0.005, -2.0 / 5000.0
    u = ns_field(g, seed=48)
    sp = O.make_ns(d, re, dt, dpdx)
    ref = O.rk_step(sp, u, 0)
    plan = StepPlan(g, re, dt, KIND_NS)
    out = u.to(DEV).clone()
    for s in range(3):
        h_v, h_g, _, conv = O.nl_ns_skew(d, out.cpu())
        plan.substep_terms(out, s, h_v.to(DEV), h_g.to(DEV), conv.to(DEV).contiguous(), dpdx)
    assert rel(out, ref) < 1e-11


def _pd(shape=SHAPE, lx=LX, lz=LZ):
    return PhysicalDomain.create(shape, (True, False, True), scale_factors=(lx, 1.0, lz))


def test_nonlinear_update_fn_plugin_in_the_equation():
    """A user's nonlinear_update_fn (here: half the library's term) replaces
    the term in every substep of perform_time_step / solve."""
    g, d = pair(SHAPE, LX, LZ)
    re, dt = 5000.0, 0.005
    u0 = ns_field(g, seed=49)
    pd = _pd()
    vh = VectorField.FromData(FourierField, pd, u0.to(DEV), "velocity_hat")
    nse = NavierStokesVelVort(vh, Re=re, dt=dt, end_time=2 * dt)
    nse.dPdx = -2.0 / re
    default = nse.nonlinear_update_fn

    def half(vel, t):
        h_v, h_g, vort, conv = default(vel, t)
        return 0.5 * h_v, 0.5 * h_g, vort, 0.5 * conv

    nse.nonlinear_update_fn = half
    out, _ = nse.perform_time_step(u0.to(DEV), -2.0 / re, 0)
    sp = O.make_ns(d, re, dt, -2.0 / re)

    def half_o(vh, t):
        h_v, h_g, vort, conv = O.nl_ns_skew(d, vh)
        return 0.5 * h_v, 0.5 * h_g, vort, 0.5 * conv

    sp.nonlinear = half_o
    ref = O.rk_step(sp, u0, 0)
    assert rel(out, ref) < 1e-11
    sp.nonlinear = lambda vh, t: O.nl_ns_skew(d, vh)
    assert rel(ref, O.rk_step(sp, u0, 0)) > 1e-8


@pytest.mark.parametrize("linearise", [True, False])
def test_custom_base_flow(linearise):
    """velocity_base_hat with the profile U = 0.7 (1 - y^2) + 0.2 y^3 - 0.1 y."""
    g, d = pair(SHAPE, LX, LZ)
    re, dt = 3000.0, 0.01
    y = g.y
    prof = 0.7 * (1 - y**2) + 0.2 * y**3 - 0.1 * y
    X = np.meshgrid(d.x, d.y, d.z, indexing="ij")
    ub = torch.from_numpy(0.7 * (1 - X[1] ** 2) + 0.2 * X[1] ** 3 - 0.1 * X[1])
    zero = torch.zeros_like(ub)
    base_hat = torch.stack([O.field_hat(d, ub), O.field_hat(d, zero), O.field_hat(d, zero)])
    u0 = perturbation_field(g, seed=50, energy=1e-6 if linearise else 1e-2)
    pd = _pd()
    vh = VectorField.FromData(FourierField, pd, u0.to(DEV), "velocity_hat")
    bh = VectorField.FromData(FourierField, pd, base_hat.to(DEV), "velocity_base_hat")
    nse = NavierStokesVelVortPerturbation(vh, Re=re, dt=dt, end_time=3 * dt, linearise=linearise, velocity_base_hat=bh)
    assert np.abs(nse.base_profile - prof).max() < 1e-13
    sp = O.make_perturbation(d, re, dt, linearise, base_hat=base_hat)
    out, ref = u0.to(DEV).clone(), u0.clone()
    for n in range(3):
        out, _ = nse.perform_time_step(out, 0.0, n)
        ref = O.rk_step(sp, ref, n)
        assert rel(out, ref) < 1e-11 * (n + 1)
    # a base flow with x dependence is rejected
    bad = base_hat.clone()
    bad[0, 1, :, 0] = 1.0
    with pytest.raises(NotImplementedError):
        NavierStokesVelVortPerturbation(vh, Re=re, dt=dt, end_time=3 * dt, velocity_base_hat=bad.to(DEV))


def test_linearise_and_coupling_switches():
    """linearise_switch = 0.5 over 4 steps: steps 0, 1 linearised, 2, 3
    nonlinear; coupling_term_switch = 0.25: step 0 without the coupling
    term (navier_stokes_perturbation.py:385-416)."""
    g, d = pair(SHAPE, LX, LZ)
    re, dt = 3000.0, 0.01
    u0 = perturbation_field(g, seed=51, energy=1e-2)
    pd = _pd()
    vh = VectorField.FromData(FourierField, pd, u0.to(DEV), "velocity_hat")
    nse = NavierStokesVelVortPerturbation(
        vh, Re=re, dt=dt, end_time=4 * dt, linearise_switch=0.5, coupling_term_switch=0.25
    )
    assert [nse.linearise(t) for t in range(4)] == [True, True, False, False]
    assert [nse.coupling_term(t) for t in range(4)] == [False, True, True, True]
    traj, _, n = nse.solve_scan()
    assert n == 4
    ops = O.build_operators(d, re, dt)
    ref = u0.clone()
    for t in range(4):
        sp = O.make_perturbation(d, re, dt, nse.linearise(t), nse.coupling_term(t), ops=ops)
        ref = O.rk_step(sp, ref, t)
    out = nse.get_latest_field("velocity_hat").get_data()
    assert rel(out, ref) < 4e-11


@pytest.mark.parametrize("mode", ["gain_3d", "gain_no_00", "dissipation"])
@pytest.mark.parametrize("linearise", [True, False])
def test_dual_objectives(mode, linearise):
    g, d = pair(SHAPE, LX, LZ)
    re, dt, n = 2000.0, 0.01, 4
    u0 = perturbation_field(g, seed=52, energy=1e-6 if linearise else 1e-3)
    u0[0, 0, :, 0] = torch.from_numpy(1e-3 * (1 - g.y**2) * g.y + 0j)  # a mean part for the objectives to see
    pd = _pd()
    vh = VectorField.FromData(FourierField, pd, u0.to(DEV), "velocity_hat")
    nse = NavierStokesVelVortPerturbation(vh, Re=re, dt=dt, end_time=n * dt, linearise=linearise)
    dual = NavierStokesVelVortPerturbationDual.FromNavierStokesVelVortPerturbation(
        nse, checkpointing=False, optimisation_mode=mode
    )
    grad = dual.get_grad()
    obj = dual.get_objective_fun()
    fwd = O.make_perturbation(d, re, dt, linearise)
    rg, robj, _ = O.dual_gradient(fwd, u0, n, linearise, mode=mode, re_tau=re)
    print(mode, "grad rel %.3e obj %.15e vs %.15e" % (rel(grad, rg), obj, float(robj)))
    assert rel(grad, rg) < 1e-10
    assert abs(obj - float(robj)) <= 1e-11 * abs(float(robj))


def test_dual_objectives_with_checkpointing_match_stored():
    g, _ = pair(SHAPE, LX, LZ)
    re, dt, n = 2000.0, 0.01, 6
    u0 = perturbation_field(g, seed=53, energy=1e-3)
    pd = _pd()
    out = []
    for ck in (False, True):
        vh = VectorField.FromData(FourierField, pd, u0.to(DEV), "velocity_hat")
        nse = NavierStokesVelVortPerturbation(vh, Re=re, dt=dt, end_time=n * dt, linearise=False)
        dual = NavierStokesVelVortPerturbationDual(nse, checkpointing=ck, checkpoint_every=3, optimisation_mode="dissipation")
        out.append((dual.get_grad(), dual.get_objective_fun()))
    assert torch.equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


def test_laplacian_dissipation_and_volume_integral_stages():
    g, d = pair(SHAPE, LX, LZ)
    plan = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION)
    u = perturbation_field(g, seed=54, energy=1.0)
    u[0, 0, :, 0] = torch.from_numpy(0.5 * (1 - g.y**2) + 0j)
    lap = plan.laplacian(u.to(DEV), -0.25)
    ref = -0.25 * torch.stack([O.laplacian(d, u[i]) for i in range(3)])
    assert rel(lap, ref) < 1e-12
    acc = torch.zeros((g.nxm, g.ny, g.nzm), dtype=torch.complex128, device=DEV)
    plan.dissipation_accumulate(u.to(DEV), acc)
    plan.dissipation_accumulate(u.to(DEV), acc, 0.5)
    racc = torch.zeros_like(u[0])
    for i in range(3):
        for j in range(3):
            racc = racc + O.diff(d, u[i], j) ** 2
    assert rel(acc, 1.5 * racc) < 1e-12
    vi = plan.volume_integral(acc)
    rvi = float(O.volume_integral(d, O.field_no_hat(d, 1.5 * racc)))
    assert abs(vi - rvi) < 1e-12 * abs(rvi)


def test_fixed_parameters():
    g, _ = pair(SHAPE, LX, LZ)
    pd = _pd()
    vh = VectorField.FromData(FourierField, pd, ns_field(g, seed=55).to(DEV), "velocity_hat")
    nse = NavierStokesVelVort(vh, Re=5000.0, dt=0.005, end_time=0.01, max_cfl=0.4)
    fp = nse.get_fixed_params()
    assert fp.Re_tau == 5000.0 and fp.max_cfl == 0.4 and fp.number_of_rk_steps == 3
    assert fp.physical_domain is pd and fp.dt_update_frequency == 10 and fp.u_max_over_u_tau == 1.0
    assert fp.poisson_mat is None and fp.rk_mats_lhs_inv is None
