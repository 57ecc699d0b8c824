"""Pins of the oracle (the CPU restatement) against closed forms, the
reference's own known-answer tests (tests/test_project.py of the reference)
and brute-force evaluation.  Each check is chosen so that a plausible mistake
in the restatement (a dropped term, a wrong sign or index, a transposed
operand, a wrong scaling) fails it."""

import math
import os

import numpy as np
import pytest
import torch
from numpy.polynomial import chebyshev as C

from oracle import jsd_oracle as O


def _abs(f: torch.Tensor) -> float:
    """Field.__abs__: norm / number of dofs (field.py:235-236)."""
    return float(torch.linalg.norm(f.reshape(-1)) / f.numel())


def _phys(dom, fn):
    X = np.meshgrid(dom.x, dom.y, dom.z, indexing="ij")
    return torch.from_numpy(fn(X))


# --------------------------------------------------------------------------
# grids and derivative matrices (domain.py:130-180)
# --------------------------------------------------------------------------


@pytest.mark.parametrize("n", [9, 33, 65, 129])
def test_cheb_diff_exact_on_polynomials(n):
    """D T_k(y_j) = T_k'(y_j) for every degree k < n (collocation is exact)."""
    y = O.cheb_grid(n)
    D = O.cheb_diff_mat(y)
    err = 0.0
    for k in range(n):
        c = np.zeros(k + 1)
        c[k] = 1.0
        f = C.chebval(y, c)
        df = C.chebval(y, C.chebder(c))
        err = max(err, np.abs(D @ f - df).max() / max(1.0, np.abs(df).max()))
    assert err < 5e-12 * n


def test_grids_and_wavenumbers():
    d = O.make_domain((128, 129, 128), lx=4 * np.pi, lz=2 * np.pi)
    assert (d.nxm, d.ny, d.nzm, d.px, d.pz) == (129, 129, 129, 194, 194)
    k = np.round(d.kx * d.lx / (2 * np.pi)).astype(int)
    assert list(k[:3]) == [0, 1, 2] and k[64] == 64 and k[65] == -64 and k[-1] == -1
    assert np.allclose(d.kx * d.lx / (2 * np.pi), k, atol=1e-12)
    d0 = O.make_domain((24, 65, 4))
    assert (d0.nxm, d0.ny, d0.nzm, d0.px, d0.pz) == (25, 65, 5, 38, 8)
    assert abs(d0.y[0] - 1.0) < 1e-15 and abs(d0.y[-1] + 1.0) < 1e-15


# --------------------------------------------------------------------------
# transforms (domain.py:716-764, 1167-1242)
# --------------------------------------------------------------------------


def test_field_hat_closed_form_single_modes():
    """sin(2 a x) cos(3 b z) (1 - y^2): four stored coefficients with the
    ortho scaling sqrt(Px Pz) / prod(L / 2 pi), all others zero."""
    lx, lz = 2 * np.pi / 1.3, 2 * np.pi / 0.7
    d = O.make_domain((24, 17, 16), lx, lz)
    ax, bz = 2 * np.pi / lx, 2 * np.pi / lz
    f = _phys(d, lambda X: np.sin(2 * ax * X[0]) * np.cos(3 * bz * X[2]) * (1 - X[1] ** 2))
    fh = O.field_hat(d, f).numpy()
    scale = math.sqrt(d.px * d.pz) / d.scaling
    expect = np.zeros_like(fh)
    prof = 1 - d.y**2
    for sx in (1, -1):
        for sz in (1, -1):
            i = 2 if sx > 0 else d.nxm - 2
            j = 3 if sz > 0 else d.nzm - 3
            expect[i, :, j] = sx / (4j) * prof * scale
    assert np.abs(fh - expect).max() < 1e-12 * scale


def test_field_hat_plus_minus_K_convention():
    """cos(K x + 2 z): the true spectrum has (K, 2) and (-K, -2).  The
    reference keeps conj of the +K row as the -K row without mirroring kz,
    so the stored (-K, +2) entry holds the true (-K, -2) value and the
    stored (-K, -2) entry is zero; kx != -K rows are untouched."""
    d = O.make_domain((24, 17, 16))
    K = d.nxm // 2
    f = _phys(d, lambda X: np.cos(K * X[0] + 2 * X[2]) + 0 * X[1])
    fh = O.field_hat(d, f).numpy()
    scale = math.sqrt(d.px * d.pz) / d.scaling
    iK, imK = K, K + 1
    j2, jm2 = 2, d.nzm - 2
    assert abs(fh[iK, 3, j2] - 0.5 * scale) < 1e-12 * scale
    assert abs(fh[imK, 3, j2] - 0.5 * scale) < 1e-12 * scale
    assert abs(fh[imK, 3, jm2]) < 1e-12 * scale
    assert abs(fh[iK, 3, jm2]) < 1e-12 * scale
    # and along z: cos(2 x + KZ z) -> the stored (+2, -KZ) entry holds the
    # true (-2, -KZ) value (conj of the +KZ column, kx not mirrored)
    KZ = d.nzm // 2
    g = _phys(d, lambda X: np.cos(2 * X[0] + KZ * X[2]) + 0 * X[1])
    gh = O.field_hat(d, g).numpy()
    assert abs(gh[2, 3, KZ + 1] - 0.5 * scale) < 1e-12 * scale
    assert abs(gh[d.nxm - 2, 3, KZ + 1]) < 1e-12 * scale
    assert abs(gh[2, 3, KZ] - 0.5 * scale) < 1e-12 * scale


def test_field_no_hat_closed_form():
    """A single stored mode X[3, :, 1] = 1 maps to cos(3x + z) s / sqrt(Px Pz)
    (the real part of the full complex inverse transform)."""
    d = O.make_domain((24, 9, 16))
    X = torch.zeros((d.nxm, d.ny, d.nzm), dtype=torch.complex128)
    X[3, :, 1] = 1.0
    f = O.field_no_hat(d, X).numpy()
    Xg = np.meshgrid(d.x, d.y, d.z, indexing="ij")
    expect = np.cos(3 * Xg[0] + Xg[2]) * d.scaling / math.sqrt(d.px * d.pz)
    assert np.abs(f - expect).max() < 1e-14


def test_transform_round_trip_band_limited():
    d = O.make_domain((24, 17, 16), 2 * np.pi, np.pi)
    f = _phys(d, lambda X: np.cos(3 * X[0] - 4 * X[2]) * X[1] + np.sin(X[0]) ** 2 * (1 - X[1] ** 2))
    g = O.field_no_hat(d, O.field_hat(d, f))
    assert float((g - f).abs().max()) < 1e-13


def test_fourier_derivatives():
    """d/dx and d/dz of sin(2x)cos(3z)y against the analytic derivative
    (test_fourier_simple_3D, tests/test_project.py:492-575, tolerance 1e-9)."""
    d = O.make_domain((24, 17, 16))
    fn = lambda X: np.sin(2 * X[0]) * np.cos(3 * X[2]) * X[1] ** 3
    fh = O.field_hat(d, _phys(d, fn))
    dx = O.field_no_hat(d, O.diff(d, fh, 0)).numpy()
    dz = O.field_no_hat(d, O.diff(d, fh, 2, 2)).numpy()
    dy = O.field_no_hat(d, O.diff(d, fh, 1)).numpy()
    X = np.meshgrid(d.x, d.y, d.z, indexing="ij")
    assert np.abs(dx - 2 * np.cos(2 * X[0]) * np.cos(3 * X[2]) * X[1] ** 3).max() < 1e-9
    assert np.abs(dz + 9 * np.sin(2 * X[0]) * np.cos(3 * X[2]) * X[1] ** 3).max() < 1e-9
    assert np.abs(dy - 3 * np.sin(2 * X[0]) * np.cos(3 * X[2]) * X[1] ** 2).max() < 1e-9


# --------------------------------------------------------------------------
# integrals and energy (domain.py:485-628, field.py:1456-1473, 2618-2711)
# --------------------------------------------------------------------------


def test_definite_integral_golden_1d_cheb():
    """tests/test_project.py:679-691: int_{-1}^{1} exp(sin(2 pi y)) - 1 dy on
    60 Gauss-Lobatto points = 0.5321317555 (tolerance 1e-10)."""
    d = O.make_domain((4, 60, 4))
    w = O.y_int_weights(d)
    val = float(np.dot(w, np.exp(np.sin(d.y * 2 * np.pi)) - 1))
    assert abs(val - 0.5321317555) < 1e-10


def test_y_weights_clenshaw_curtis_on_polynomials():
    """The quadrature integrates polynomials of degree <= N-2 exactly."""
    d = O.make_domain((4, 33, 4))
    w = O.y_int_weights(d)
    for k in range(0, 31):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, d.y**k) - exact) < 1e-13


def test_energy_closed_form():
    """U = 1 - y^2: E = 0.5 int int int U^2 / (2 Lx Lz) = 4/15."""
    d = O.make_domain((24, 33, 8), 4 * np.pi, 2 * np.pi)
    base = O.base_flow_hat(d)
    assert abs(float(O.energy(d, base)) - 4.0 / 15.0) < 1e-13


def test_energy_parseval_against_spectral_formula():
    from jax_spectral_dns_amd import domain_math as DM
    from jax_spectral_dns_amd import initial

    g = DM.make_grid((24, 33, 16), 2 * np.pi, np.pi)
    d = O.make_domain((24, 33, 16), 2 * np.pi, np.pi)
    u = initial.perturbation(g, seed=11, energy=0.37)
    u[0, g.nxm // 2 + 1] *= 1.7  # break Hermitian symmetry on the -K row
    assert abs(float(O.energy(d, torch.from_numpy(u))) - initial.spectral_energy(g, u)) < 1e-13


# --------------------------------------------------------------------------
# operators (domain.py:857-889, navier_stokes.py:920-1022)
# --------------------------------------------------------------------------


@pytest.mark.parametrize("shape,lx,lz", [((24, 28, 20), 1.0, 1.0), ((20, 28, 24), 1.0, 1.5)])
def test_poisson_analytic(shape, lx, lz):
    """test_poisson_slices / test_poisson_no_slices (tests/test_project.py:957-1073):
    sin(2 pi x/Lx) sin(2 pi (z+1)/Lz) cos(pi y / 2), tolerance 1e-8."""
    d = O.make_domain(shape, lx, lz)
    ops = O.build_operators(d, 1.0, 1.0)
    fac = (2 * np.pi / lx) ** 2 + np.pi**2 / 4 + (2 * np.pi / lz) ** 2
    ua = lambda X: np.sin(X[0] * 2 * np.pi / lx) * np.sin((X[2] + 1.0) * 2 * np.pi / lz) * np.cos(X[1] * np.pi / 2)
    rhs = _phys(d, lambda X: -fac * ua(X))
    rh = O.field_hat(d, rhs).permute(0, 2, 1)
    sol = O.apply_k(d, ops.poisson, rh, "poisson").permute(0, 2, 1).clone()
    sol[0] = 0
    sol[:, :, 0] = 0
    out = O.field_no_hat(d, sol)
    assert _abs(out - _phys(d, ua)) < 1e-8


def test_helmholtz_inverse_residual():
    """lhs_inv really inverts hom_dir(I - beta dt L(k)) (residual at rounding)."""
    d = O.make_domain((8, 33, 8))
    ops = O.build_operators(d, 500.0, 0.01)
    D2 = np.linalg.matrix_power(d.D, 2)
    kx, kz = d.kx[3], d.kz[2]
    L = D2 / 500.0 - np.eye(d.ny) * (kx**2 + kz**2) / 500.0
    A = O.hom_dirichlet(np.eye(d.ny) - O.RK_BETA[1] * 0.01 * L)
    assert np.abs(A @ ops.lhs_inv[1, 3, 2] - np.eye(d.ny)).max() < 1e-12


# --------------------------------------------------------------------------
# nonlinear terms
# --------------------------------------------------------------------------


def test_linear_perturbation_terms_closed_form():
    """For u = (u, v, w) about U = 1 - y^2 the linearised term is
    H0 = -(U u_x + v U') + (div u) U / 2, H1 = -U v_x, H2 = -U w_x; check the
    oracle's literal FFT-based evaluation against these products evaluated
    directly in physical space for a smooth band-limited field."""
    d = O.make_domain((24, 17, 16))
    base = O.base_flow_hat(d)
    X = np.meshgrid(d.x, d.y, d.z, indexing="ij")
    y = X[1]
    u = np.sin(2 * X[0]) * np.cos(X[2]) * (1 - y**2) * y
    v = np.cos(X[0] + 2 * X[2]) * (1 - y**2) ** 2
    w = np.sin(3 * X[2] - X[0]) * (1 - y**2)
    vh = torch.stack([O.field_hat(d, torch.from_numpy(a)) for a in (u, v, w)])
    hv, hg, vort, conv = O.nl_perturbation_skew(d, vh, base, linearise=True)
    U = 1 - y**2
    dU = -2 * y
    ux = 2 * np.cos(2 * X[0]) * np.cos(X[2]) * (1 - y**2) * y
    vx = -np.sin(X[0] + 2 * X[2]) * (1 - y**2) ** 2
    vy = np.cos(X[0] + 2 * X[2]) * 2 * (1 - y**2) * (-2 * y)
    wx = -np.cos(3 * X[2] - X[0]) * (1 - y**2)
    wz = 3 * np.cos(3 * X[2] - X[0]) * (1 - y**2)
    div = ux + vy + wz
    H = [-(U * ux + v * dU) + 0.5 * div * U, -U * vx, -U * wx]
    for c in range(3):
        got = O.field_no_hat(d, -conv[c]).numpy()
        assert np.abs(got - H[c]).max() < 1e-11


def test_transform_grid_invariance():
    """Any transform grid P >= 3K+1 gives the same truncated spectrum of the
    quadratic nonlinear term up to rounding (the product kernels rely on it)."""
    from tests.helpers import ns_field, pair

    g, d = pair((24, 33, 16), 2 * np.pi, np.pi)
    u = ns_field(g, seed=3)
    d2 = O.make_domain((24, 33, 16), 2 * np.pi, np.pi, physical_override=(40, 25))
    a = O.nl_ns_skew(d, u)
    b = O.nl_ns_skew(d2, u)
    for x, y in zip(a, b):
        assert float(torch.linalg.norm(x - y) / torch.linalg.norm(x)) < 1e-13
    with pytest.raises(ValueError):
        O.make_domain((24, 33, 16), physical_override=(36, 25))


# --------------------------------------------------------------------------
# full step: the reference's known answers
# --------------------------------------------------------------------------


def _laminar_ic(d, re, perturbation_factor=0.01):
    """solve_navier_stokes_laminar initial field (navier_stokes.py:2022-2083)."""
    um = re / 2.0
    X = np.meshgrid(d.x, d.y, d.z, indexing="ij")
    ana = um * (1 - X[1] ** 2) + 0.0 * X[0] * X[2]
    u = np.pi / 3 * um * (perturbation_factor * np.cos(X[1] * np.pi / 2) * (np.cos(3 * X[0]) ** 2 * np.cos(4 * X[2]) ** 2)) + (
        1 - perturbation_factor
    ) * ana
    v = 0.1 * perturbation_factor * um * (np.pi / 3 * (1 - X[1] ** 2) ** 2 * np.cos(3 * X[0]) * np.cos(4 * X[2]))
    w = 0.1 * np.pi / 3 * perturbation_factor * um * (np.cos(X[1] * np.pi / 2) * np.cos(5 * X[0]) * np.cos(3 * X[2]))
    return torch.stack([O.field_hat(d, torch.from_numpy(a)) for a in (u, v, w)]), ana


def test_navier_stokes_laminar_converges():
    """test_navier_stokes_laminar (tests/test_project.py:1075-1134): Re = 1.5,
    4 x 96 x 4, aliasing 1, dt = 0.04, T = 1 converges to Re/2 (1 - y^2)
    within 6e-5 in every component."""
    re = 1.5
    d = O.make_domain((4, 96, 4), 1.87, 0.93, aliasing=1.0)
    vh, ana = _laminar_ic(d, re)
    sp = O.make_ns(d, re, 0.04, -1.0)
    n = O.number_of_steps(1.0, 0.04)
    assert n == 25
    vh = O.run(sp, vh, n)
    vel = [O.field_no_hat(d, vh[c]) for c in range(3)]
    assert _abs(vel[0] - torch.from_numpy(ana)) < 6e-5
    assert _abs(vel[1]) < 6e-5
    assert _abs(vel[2]) < 6e-5


def test_perturbation_laminar_decays():
    """test_perturbation_laminar (tests/test_project.py:1239-1273): Re = 1.5,
    16 x 48 x 16, aliasing 1, dt = 0.04, T = 2: the perturbation decays below
    5e-7 in every component."""
    re = 1.5
    pf = 0.01
    d = O.make_domain((16, 48, 16), 1.87, 0.93, aliasing=1.0)
    X = np.meshgrid(d.x, d.y, d.z, indexing="ij")
    a = 0.1 * pf * (np.pi / 3 * np.cos(X[1] * np.pi / 2) * np.cos(3 * X[0]) * np.cos(4 * X[2]))
    c = 0.1 * np.pi / 3 * pf * (np.cos(X[1] * np.pi / 2) * np.cos(5 * X[0]) * np.cos(3 * X[2]))
    vh = torch.stack([O.field_hat(d, torch.from_numpy(f)) for f in (a, a, c)])
    sp = O.make_perturbation(d, re, 0.04, linearise=False)
    vh = O.run(sp, vh, O.number_of_steps(2.0, 0.04))
    for k in range(3):
        assert _abs(O.field_no_hat(d, vh[k])) < 5e-7


def test_step_keeps_continuity_and_wall_conditions():
    """After a step: i kx u + D v + i kz w = 0 and v = dv/dy = 0 at the walls
    (the influence-matrix correction), to rounding."""
    from tests.helpers import pair, perturbation_field

    g, d = pair((24, 33, 16), 2 * np.pi, np.pi)
    u = perturbation_field(g, seed=2, energy=1e-3)
    sp = O.make_perturbation(d, 2000.0, 0.01, linearise=False)
    u1 = O.rk_step(sp, u, 0)
    div = O.divergence(d, u1)
    assert float(div.abs().max()) < 1e-10 * float(u1.abs().max()) * d.ny**2
    dv = O.diff(d, u1[1], 1)
    assert float(dv[:, 0].abs().max()) < 1e-11 and float(dv[:, -1].abs().max()) < 1e-11
    assert float(u1[:, :, 0].abs().max()) == 0.0 and float(u1[:, :, -1].abs().max()) == 0.0


def test_checkpointed_adjoint_identical():
    """test_checkpointing (tests/test_project.py:1505-1555): the adjoint with
    checkpoint-and-recompute equals the adjoint with the full history (the
    reference asserts an energy of the difference < 1e-30)."""
    from tests.helpers import pair, perturbation_field

    g, d = pair((24, 17, 8), 2 * np.pi, 2 * np.pi)
    u0 = perturbation_field(g, seed=1, energy=1e-4)
    fwd = O.make_perturbation(d, 1000.0, 0.02, linearise=False)
    g1, gain1, v1 = O.dual_gradient(fwd, u0, 6, linearise=False)
    g2, gain2, v2 = O.dual_gradient(fwd, u0, 6, linearise=False, checkpoint=3)
    assert float(O.energy(d, v1 - v2)) < 1e-30
    assert float(gain1) == float(gain2)


def test_parameterisation_pins():
    """vort_yvel_to_vel (navier_stokes.py:1027-1203): the clamped projection
    is idempotent, keeps profiles already in the clamped span, gives
    v = v' = 0 at both walls; the velocity satisfies continuity and has the
    prescribed omega_y and (0,0) profiles."""
    shape = (8, 33, 8)
    d = O.make_domain(shape, 2 * np.pi, np.pi)
    rng = np.random.default_rng(0)
    sh = (d.nxm, d.ny, d.nzm)
    vort = torch.from_numpy(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    vy = torch.from_numpy(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    ux = torch.from_numpy(rng.standard_normal(d.ny) + 0j)
    wz = torch.from_numpy(rng.standard_normal(d.ny) + 0j)
    vel = O.vort_yvel_to_vel(d, vort, vy, ux, wz)
    u, v, w = vel[0], vel[1], vel[2]
    D = torch.from_numpy(d.D).to(torch.complex128)
    dv = torch.matmul(D, v)
    scale = float(v.abs().max())
    assert float(v[:, 0, :].abs().max()) < 1e-12 * scale and float(v[:, -1, :].abs().max()) < 1e-12 * scale
    assert float(dv[:, 0, :].abs().max()) < 1e-10 * float(dv.abs().max())
    assert float(dv[:, -1, :].abs().max()) < 1e-10 * float(dv.abs().max())
    # idempotent
    v2 = O.smooth_and_enforce_bc_vel_y(d, v.permute(0, 2, 1)).permute(0, 2, 1)
    assert float((v2 - v).abs().max()) < 1e-12 * scale
    # a profile in the span is kept
    A = torch.from_numpy(O.clamped_cheb_basis(d.y)).to(torch.complex128)
    c = torch.from_numpy(rng.standard_normal(A.shape[1]) + 0j)
    prof = A @ c
    kept = O.smooth_and_enforce_bc_vel_y(d, prof[None, :])[0]
    assert float((kept - prof).abs().max()) < 1e-12 * float(prof.abs().max())
    # continuity i kx u + D v + i kz w = 0 and omega_y = i kz u - i kx w away from (0,0)
    kx = torch.from_numpy(d.kx)[:, None, None]
    kz = torch.from_numpy(d.kz)[None, None, :]
    div = 1j * kx * u + dv + 1j * kz * w
    div[0, :, 0] = 0
    assert float(div.abs().max()) < 1e-11 * float(dv.abs().max())
    om = 1j * kz * u - 1j * kx * w
    mask = torch.ones_like(om)
    mask[0, :, 0] = 0
    assert float(((om - vort) * mask).abs().max()) < 1e-11 * float(vort.abs().max())
    assert float((u[0, :, 0] - ux).abs().max()) == 0.0 and float((w[0, :, 0] - wz).abs().max()) == 0.0


def test_golden_fixtures_match_their_generators():
    """The committed oracle fixtures start from the initial conditions that
    tests/golden/make_golden.py names: the linearised fixture's stored u0 is the seeded field and
    one oracle step from it reproduces the stored u1 exactly; the NS fixture's stored
    energy history starts at the energy of its seeded initial field, has one
    entry per step and stays finite and positive."""
    from jax_spectral_dns_amd import domain_math as DM
    from jax_spectral_dns_amd import initial

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    fx0 = np.load(os.path.join(here, "linearised_25x65x5_1500.npz"))
    re, dt, n = fx0["params"][:3]
    g0 = DM.make_grid((24, 65, 4))
    assert np.array_equal(fx0["u0"], initial.perturbation(g0, seed=0, energy=1e-6))
    d0 = O.make_domain((24, 65, 4))
    sp = O.make_perturbation(d0, float(re), float(dt), linearise=True)
    u1 = O.rk_step(sp, torch.from_numpy(fx0["u0"]), 0)
    assert np.array_equal(u1.numpy(), fx0["u1"])
    assert len(fx0["energy"]) == int(n) + 1

    fx1 = np.load(os.path.join(here, "ns_128x129x128_200_energy.npz"))
    g1 = DM.make_grid((128, 129, 128), 4 * np.pi, 2 * np.pi)
    base = initial.laminar(g1)
    u0 = base + initial.perturbation(g1, seed=0, energy=0.05 * initial.spectral_energy(g1, base))
    d1 = O.make_domain((128, 129, 128), 4 * np.pi, 2 * np.pi)
    e0 = float(O.energy(d1, torch.from_numpy(u0)))
    assert abs(e0 - fx1["energy"][0]) <= 1e-14 * e0
    assert len(fx1["energy"]) == int(fx1["params"][2]) + 1
    assert np.all(np.isfinite(fx1["energy"])) and np.all(fx1["energy"] > 0)
