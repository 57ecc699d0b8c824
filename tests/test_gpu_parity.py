"""Parity of the HIP path (through the C ABI) with the oracle, stage by stage
and step by step, on seeded inputs.  Tolerances: <= 1e-11 relative L2 per
step (the FP64 parity tolerance of the step), nonlinear terms and
derivatives <= 1e-12."""

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import KIND_DUAL, KIND_NS, KIND_PERTURBATION
from jax_spectral_dns_amd.plan import StepPlan
from oracle import jsd_oracle as O
from tests.helpers import ns_field, pair, perturbation_field, rel

pytestmark = pytest.mark.gpu
DEV = "cuda"

SHAPES = [((24, 65, 4), 2 * np.pi, 2 * np.pi), ((24, 33, 16), 2 * np.pi, np.pi)]

# The y-derivatives run on the even/odd blocks of the centro-(anti)symmetrised
# D and D^2, which differ from numpy's D and D @ D by rounding; D^2's entries
# grow like Ny^4, so its products differ by up to a few 1e-13 relative.
YDIFF_TOL = {1: 1e-13, 2: 1e-12}


@pytest.mark.parametrize("shape,lx,lz", SHAPES)
def test_ydiff(shape, lx, lz):
    g, d = pair(shape, lx, lz)
    plan = StepPlan(g, 3000.0, 0.01, KIND_NS)
    u = perturbation_field(g, seed=1, energy=1.0)
    for order in (1, 2):
        out = plan.ydiff(u.to(DEV), order).cpu()
        ref = torch.stack([O.diff(d, u[c], 1, order) for c in range(3)])
        assert rel(out, ref) < YDIFF_TOL[order], order


@pytest.mark.parametrize("shape", [(16, 48, 16), (8, 130, 8), (8, 24, 8), (8, 127, 8), (8, 128, 8)])
def test_ydiff_even_and_largest_ny(shape):
    """Even Ny (no centre point in the even/odd split), the largest Ny the
    parity blocks hold, Ny = 24 (fewer k-steps than the streaming kernel's
    prefetch distance: the blocked kernel runs), and Ny = 127 / 128 (every
    MFMA tile row stored; 127 has its centre row inside the tiles)."""
    g, d = pair(shape)
    plan = StepPlan(g, 3000.0, 0.01, KIND_NS)
    rng = np.random.default_rng(11)
    u = torch.from_numpy(rng.standard_normal(plan.shape) + 1j * rng.standard_normal(plan.shape))
    for order in (1, 2):
        out = plan.ydiff(u.to(DEV), order).cpu()
        ref = torch.stack([O.diff(d, u[c], 1, order) for c in range(3)])
        assert rel(out, ref) < YDIFF_TOL[order], order


@pytest.mark.parametrize("shape,lx,lz", SHAPES)
def test_energy(shape, lx, lz):
    g, d = pair(shape, lx, lz)
    plan = StepPlan(g, 3000.0, 0.01, KIND_NS)
    u = ns_field(g, seed=2)
    e = plan.energy(u.to(DEV))
    assert abs(e - float(O.energy(d, u))) / float(O.energy(d, u)) < 1e-12


@pytest.mark.parametrize("shape,lx,lz", SHAPES)
def test_nonlinear_ns(shape, lx, lz):
    g, d = pair(shape, lx, lz)
    plan = StepPlan(g, 5000.0, 0.005, KIND_NS)
    u = ns_field(g, seed=3)
    hv, hg, conv = plan.nonlinear(u.to(DEV))
    rv, rg, _, rc = O.nl_ns_skew(d, u)
    assert rel(hv, rv) < 1e-12
    assert rel(hg, rg) < 1e-12
    assert rel(conv, rc) < 1e-12


@pytest.mark.parametrize("shape", [(16, 33, 160), (16, 33, 96), (8, 33, 128), (16, 33, 80), (8, 32, 48), (16, 33, 64)])
def test_nonlinear_ns_z256(shape):
    """z transform lengths 256 (16 x 16) and 128 (8 x 16) of the register
    kernel against the oracle; at length 256, 97 and 129 stored kz take its
    pruned first stage (inputs n2 = 5..11 zero)."""
    g, d = pair(shape)
    assert g.ppz == (256 if shape[2] >= 96 else 128)
    plan = StepPlan(g, 5000.0, 0.005, KIND_NS)
    u = ns_field(g, seed=12)
    hv, hg, conv = plan.nonlinear(u.to(DEV))
    rv, rg, _, rc = O.nl_ns_skew(d, u)
    assert rel(hv, rv) < 1e-12
    assert rel(hg, rg) < 1e-12
    assert rel(conv, rc) < 1e-12


@pytest.mark.parametrize("shape", [(8, 33, 8), (16, 33, 16)])
def test_nonlinear_ns_runtime_lengths(shape):
    """x transform lengths 15 and 25 have no compile-time kernel: the
    run-time-length FFT kernels run here."""
    g, d = pair(shape)
    assert g.ppx in (15, 25)
    plan = StepPlan(g, 5000.0, 0.005, KIND_NS)
    u = ns_field(g, seed=7)
    hv, hg, conv = plan.nonlinear(u.to(DEV))
    rv, rg, _, rc = O.nl_ns_skew(d, u)
    assert rel(hv, rv) < 1e-12
    assert rel(hg, rg) < 1e-12
    assert rel(conv, rc) < 1e-12


@pytest.mark.parametrize("shape", [(24, 33, 16), (16, 33, 160), (16, 33, 128)])
@pytest.mark.parametrize("kind", [KIND_NS, KIND_DUAL])
def test_compile_time_and_runtime_fft_agree(monkeypatch, kind, shape):
    """The compile-time-length FFT kernels (z length 256: the 16 x 16
    register kernel) and the run-time-length ones (forced with
    JSD_FFT_GENERIC=1 at plan creation) agree to rounding."""
    g, _ = pair(shape, 2 * np.pi, np.pi)
    u = perturbation_field(g, seed=8, energy=1e-2).to(DEV)
    uf = perturbation_field(g, seed=9, energy=1e-2).to(DEV)

    def make():
        if kind == KIND_NS:
            return StepPlan(g, 5000.0, 0.005, KIND_NS)
        fwd_ops = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION).ops
        return StepPlan(g, 3000.0, 0.01, KIND_DUAL, ops=fwd_ops, explicit_dt=-0.01, plan_re=-3000.0)

    monkeypatch.delenv("JSD_FFT_GENERIC", raising=False)
    a = make().nonlinear(u, None if kind == KIND_NS else uf)
    monkeypatch.setenv("JSD_FFT_GENERIC", "1")
    b = make().nonlinear(u, None if kind == KIND_NS else uf)
    for x, y in zip(a, b):
        assert rel(x, y) < 1e-13


@pytest.mark.parametrize("linearise", [True, False])
@pytest.mark.parametrize("shape,lx,lz", SHAPES)
def test_nonlinear_perturbation(shape, lx, lz, linearise):
    g, d = pair(shape, lx, lz)
    plan = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=linearise)
    u = perturbation_field(g, seed=4, energy=1e-2)
    hv, hg, conv = plan.nonlinear(u.to(DEV))
    base = O.base_flow_hat(d)
    rv, rg, _, rc = O.nl_perturbation_skew(d, u, base, linearise)
    assert rel(hv, rv) < 1e-12
    assert rel(hg, rg) < 1e-12
    assert rel(conv, rc) < 1e-12


@pytest.mark.parametrize("coupling_term", [True, False])
@pytest.mark.parametrize("shape", [(16, 33, 96), (16, 33, 160), (16, 32, 96), (16, 33, 80), (16, 32, 64)])
def test_nonlinear_perturbation_z256(shape, coupling_term):
    """The perturbation product on the register z transform, lengths 256 and
    128 (two y lines per workgroup, base flow looked up per line, the second
    line masked when Ny is odd); 161 stored kz take the unpruned first stage."""
    g, d = pair(shape)
    assert g.ppz == (256 if shape[2] >= 96 else 128)
    plan = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=False, coupling_term=coupling_term)
    u = perturbation_field(g, seed=14, energy=1e-2)
    hv, hg, conv = plan.nonlinear(u.to(DEV))
    base = O.base_flow_hat(d)
    rv, rg, _, rc = O.nl_perturbation_skew(d, u, base, False, coupling_term)
    assert rel(hv, rv) < 1e-12
    assert rel(hg, rg) < 1e-12
    assert rel(conv, rc) < 1e-12


@pytest.mark.parametrize("shape", [(16, 33, 96), (16, 33, 160), (16, 33, 80), (8, 32, 48)])
def test_nonlinear_dual_z256(shape):
    """The dual product (12 packed transforms) on the register z transform,
    lengths 256 and 128."""
    g, d = pair(shape)
    assert g.ppz == (256 if shape[2] >= 96 else 128)
    fwd = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION)
    plan = StepPlan(g, 3000.0, 0.01, KIND_DUAL, ops=fwd.ops, explicit_dt=-0.01, plan_re=-3000.0)
    v = perturbation_field(g, seed=15, energy=1e-2)
    uf = perturbation_field(g, seed=16, energy=1e-2)
    hv, hg, conv = plan.nonlinear(v.to(DEV), uf.to(DEV))
    rv, rg, _, rc = O.nl_dual_skew(d, v, O.base_flow_hat(d), uf, False)
    assert rel(hv, rv) < 1e-12
    assert rel(hg, rg) < 1e-12
    assert rel(conv, rc) < 1e-12


@pytest.mark.parametrize("linearise", [True, False])
@pytest.mark.parametrize("shape,lx,lz", SHAPES)
def test_nonlinear_dual(shape, lx, lz, linearise):
    g, d = pair(shape, lx, lz)
    fwd = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=linearise)
    plan = StepPlan(g, 3000.0, 0.01, KIND_DUAL, linearise=linearise, ops=fwd.ops, explicit_dt=-0.01, plan_re=-3000.0)
    v = perturbation_field(g, seed=5, energy=1e-2)
    uf = perturbation_field(g, seed=6, energy=1e-2)
    hv, hg, conv = plan.nonlinear(v.to(DEV), None if linearise else uf.to(DEV))
    base = O.base_flow_hat(d)
    rv, rg, _, rc = O.nl_dual_skew(d, v, base, uf, linearise)
    assert rel(hv, rv) < 1e-12
    assert rel(hg, rg) < 1e-12
    assert rel(conv, rc) < 1e-12


@pytest.mark.parametrize("shape,lx,lz", SHAPES)
def test_step_ns(shape, lx, lz):
    g, d = pair(shape, lx, lz)
    re, dt = 5000.0, 0.005
    dpdx = -2.0 / re
    plan = StepPlan(g, re, dt, KIND_NS)
    u = ns_field(g, seed=7)
    sp = O.make_ns(d, re, dt, dpdx)
    ref = u.clone()
    out = u.to(DEV).clone()
    for n in range(3):
        ref = O.rk_step(sp, ref, n)
        plan.step(out, dpdx, 1)
        err = rel(out, ref)
        print("step", n, "rel", err)
        assert err < 1e-11 * (n + 1)


@pytest.mark.parametrize("linearise", [True, False])
def test_step_perturbation(linearise):
    g, d = pair((24, 65, 4))
    re, dt = 3000.0, 0.01
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=linearise)
    u = perturbation_field(g, seed=0, energy=1e-6 if linearise else 1e-2)
    sp = O.make_perturbation(d, re, dt, linearise)
    ref = u.clone()
    out = u.to(DEV).clone()
    for n in range(5):
        ref = O.rk_step(sp, ref, n)
        plan.step(out, 0.0, 1)
        err = rel(out, ref)
        print("step", n, "rel", err)
        assert err < 1e-11 * (n + 1)


@pytest.mark.parametrize("shape", [(24, 65, 16), (8, 130, 8)])
@pytest.mark.parametrize("kind", [KIND_NS, KIND_PERTURBATION, KIND_DUAL])
def test_eigen_solve_matches_table_solve(kind, shape):
    """The implicit solve in eigen form (one shared E, E^-1 per parity half,
    the default) and with the per-|k| Helmholtz / Poisson inverse tables
    agree to rounding over three steps, for odd and even Ny."""
    g, _ = pair(shape, 2 * np.pi, np.pi)
    if kind == KIND_NS:
        u = ns_field(g, seed=21)
        make = lambda eig: StepPlan(g, 5000.0, 0.005, KIND_NS, eigen_solve=eig)
    elif kind == KIND_PERTURBATION:
        u = perturbation_field(g, seed=22, energy=1e-2)
        make = lambda eig: StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, eigen_solve=eig)
    else:
        u = perturbation_field(g, seed=23, energy=1e-2)
        fwd_ops = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=True, eigen_solve=False).ops  # full tables
        make = lambda eig: StepPlan(
            g, 3000.0, 0.01, KIND_DUAL, linearise=True, ops=fwd_ops, explicit_dt=-0.01, plan_re=-3000.0, eigen_solve=eig
        )
    a, b = u.to(DEV).clone(), u.to(DEV).clone()
    pe, pt = make(True), make(False)
    for n in range(3):
        if kind == KIND_DUAL:
            pe.dual_step(a, None)
            pt.dual_step(b, None)
        else:
            pe.step(a, -2.0 / 5000.0 if kind == KIND_NS else 0.0, 1)
            pt.step(b, -2.0 / 5000.0 if kind == KIND_NS else 0.0, 1)
        assert rel(a, b) < 1e-12 * (n + 1), n


@pytest.mark.parametrize("kind", [KIND_NS, KIND_PERTURBATION])
def test_multistep_call_chains_derivatives(kind):
    """In eigen form the y-derivatives of the state in substeps 2 and 3 come
    from the previous substep's solve instead of the y-derivative kernel.
    Three steps in one call match the oracle and are bit-identical to three
    single-step calls (the chain never crosses a step boundary)."""
    g, d = pair((24, 65, 16), 2 * np.pi, np.pi)
    if kind == KIND_NS:
        re, dt, dpdx = 5000.0, 0.005, -2.0 / 5000.0
        u = ns_field(g, seed=31)
        sp = O.make_ns(d, re, dt, dpdx)
    else:
        re, dt, dpdx = 3000.0, 0.01, 0.0
        u = perturbation_field(g, seed=32, energy=1e-2)
        sp = O.make_perturbation(d, re, dt, False)
    plan = StepPlan(g, re, dt, kind)
    ref = u.clone()
    for n in range(3):
        ref = O.rk_step(sp, ref, n)
    a = u.to(DEV).clone()
    plan.step(a, dpdx, 3)
    b = u.to(DEV).clone()
    for _ in range(3):
        plan.step(b, dpdx, 1)
    assert rel(a, ref) < 3e-11
    assert torch.equal(a, b)


@pytest.mark.parametrize("shape", [(24, 65, 16), (8, 130, 8), (16, 33, 4), (128, 129, 128)])
@pytest.mark.parametrize("kind,linearise", [(KIND_NS, False), (KIND_PERTURBATION, False), (KIND_PERTURBATION, True),
                                            (KIND_DUAL, True)])
def test_eight_pencil_solve_matches_four_pencil_solve(monkeypatch, kind, linearise, shape):
    """The solves with eight pencils per workgroup (the default with eight
    waves, one parity each; and with four waves) perform the four-pencil
    kernel's arithmetic in the same order, with reciprocals in place of two
    divisions: three steps agree to rounding
    (<= 1e-14 per step, measured <= 3e-15), for odd and even Ny, a last group
    that is partly empty (16 x 4 gives 17 x 5 = 85 pencils), every
    derivative-output mode and the full size."""
    if shape == (128, 129, 128) and kind != KIND_NS:
        pytest.skip("full size: the plain step only")
    g, _ = pair(shape, 2 * np.pi, np.pi)
    if kind == KIND_NS:
        u = ns_field(g, seed=33)
        make = lambda: StepPlan(g, 5000.0, 0.005, KIND_NS)
    elif kind == KIND_PERTURBATION:
        u = perturbation_field(g, seed=34, energy=1e-6 if linearise else 1e-2)
        make = lambda: StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=linearise)
    else:
        u = perturbation_field(g, seed=35, energy=1e-6)
        fwd_ops = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=True).ops
        make = lambda: StepPlan(g, 3000.0, 0.01, KIND_DUAL, linearise=True, ops=fwd_ops, explicit_dt=-0.01, plan_re=-3000.0)
    monkeypatch.delenv("JSD_SOLVE4", raising=False)
    monkeypatch.delenv("JSD_SOLVE", raising=False)
    p8w = make()  # default: 8 pencils, 8 waves (one parity per wave)
    monkeypatch.setenv("JSD_SOLVE", "8")
    p8 = make()
    monkeypatch.setenv("JSD_SOLVE", "4")
    p4 = make()
    monkeypatch.delenv("JSD_SOLVE", raising=False)
    a, b, c = u.to(DEV).clone(), u.to(DEV).clone(), u.to(DEV).clone()
    dp = -2.0 / 5000.0 if kind == KIND_NS else 0.0
    for n in range(3):
        for plan, x in ((p8, a), (p4, b), (p8w, c)):
            if kind == KIND_DUAL:
                plan.dual_step(x, None)
            else:
                plan.step(x, dp, 1)
        err = rel(a, b)
        err_w = rel(c, a)
        print("step", n, "8 vs 4 pencils rel", err, "8 pencils 8 vs 4 waves rel", err_w)
        assert err <= 1e-14 * (n + 1), (n, err)
        assert err_w <= 1e-14 * (n + 1), (n, err_w)


# x transform lengths of the compile-time kernels: 40 (8 5), 75 (5 5 3),
# 100 (5 5 4), 150 (5 5 3 2), 160 (8 5 4), 200 (8 5 5)
X_SHAPES = [(24, 33, 16), (48, 33, 8), (64, 17, 8), (96, 17, 8), (104, 17, 8), (128, 33, 8)]


@pytest.mark.parametrize("shape", X_SHAPES)
@pytest.mark.parametrize("kind", [KIND_NS, KIND_PERTURBATION, KIND_DUAL])
def test_register_x_passes_match_lds_x_passes(monkeypatch, kind, shape):
    """The x passes of jsd_fft_x.inc (the default: the inverse's last stage
    stored straight to T, the forward's first stage in registers, twiddle
    powers; and, with JSD_XPASS=2, the inverse's first stage in registers
    too) agree with the LDS x passes (JSD_XPASS=0) to rounding on every
    compile-time x length, for the NS, perturbation and dual products; and
    match the oracle's nonlinear term."""
    g, d = pair(shape, 2 * np.pi, np.pi)
    u = perturbation_field(g, seed=51, energy=1e-2)
    uf = perturbation_field(g, seed=52, energy=1e-2)

    def make():
        if kind == KIND_NS:
            return StepPlan(g, 5000.0, 0.005, KIND_NS)
        if kind == KIND_PERTURBATION:
            return StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION, linearise=False)
        fwd_ops = StepPlan(g, 3000.0, 0.01, KIND_PERTURBATION).ops
        return StepPlan(g, 3000.0, 0.01, KIND_DUAL, linearise=False, ops=fwd_ops, explicit_dt=-0.01, plan_re=-3000.0)

    out = {}
    for mode in ("0", "1", "2"):
        monkeypatch.setenv("JSD_XPASS", mode)
        out[mode] = make().nonlinear(u.to(DEV), None if kind != KIND_DUAL else uf.to(DEV))
    monkeypatch.delenv("JSD_XPASS", raising=False)
    a = out["1"]
    for mode in ("1", "2"):
        for x, y in zip(out[mode], out["0"]):
            err = rel(x, y)
            print(shape, kind, "x-pass mode", mode, "vs LDS kernels rel", err)
            assert err < 1e-13
    if kind == KIND_NS:
        ref = O.nl_ns_skew(d, u)
        for x, r in zip((a[0], a[1], a[2]), (ref[0], ref[1], ref[3])):
            assert rel(x, r) < 1e-12
