"""CPU-only checks of the product's host side: the C-ABI library loads and
exports every symbol include/jsd.h declares (no compute call without a GPU),
the host precompute matches the oracle's operator assembly, the transform
plans are valid, and the seeded generator has the documented properties."""

import ctypes
import os
import re

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import _lib
from jax_spectral_dns_amd import domain_math as DM
from jax_spectral_dns_amd import initial
from oracle import jsd_oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    header = open(os.path.join(REPO, "include", "jsd.h")).read()
    declared = set(re.findall(r"^\s*(?:int|const char\*|void)\s+(jsd_\w+)\s*\(", header, re.M))
    assert declared == set(_lib.EXPORTS)
    lib = _lib.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.jsd_version() >= 1


def test_library_rejects_bad_plans_without_gpu():
    """Argument validation happens before any device call."""
    lib = _lib.load()
    desc = _lib.PlanDesc()
    desc.nxm, desc.ny, desc.nzm = 24, 65, 5  # even mode count
    h = ctypes.c_void_p()
    assert lib.jsd_plan_create(ctypes.byref(desc), ctypes.byref(h)) == -1
    assert b"odd" in lib.jsd_last_error()
    desc.nxm, desc.ny = 25, 400
    assert lib.jsd_plan_create(ctypes.byref(desc), ctypes.byref(h)) == -1
    assert lib.jsd_plan_create(None, ctypes.byref(h)) == -1


@pytest.mark.parametrize("shape,lx,lz", [((24, 65, 4), 2 * np.pi, 2 * np.pi), ((128, 129, 128), 4 * np.pi, 2 * np.pi)])
def test_grid_matches_oracle(shape, lx, lz):
    g = DM.make_grid(shape, lx, lz)
    d = O.make_domain(shape, lx, lz)
    assert (g.nxm, g.ny, g.nzm, g.px, g.pz) == (d.nxm, d.ny, d.nzm, d.px, d.pz)
    assert np.array_equal(g.y, d.y)
    assert np.array_equal(g.kx, d.kx) and np.array_equal(g.kz, d.kz)
    assert np.abs(g.D - d.D).max() <= 1e-13 * np.abs(d.D).max()
    assert np.abs(g.y_weights - O.y_int_weights(d)).max() < 1e-14
    assert g.scaling == d.scaling
    assert g.ppx >= 3 * (g.nxm // 2) + 1 and DM.factorize(g.ppx)
    assert g.ppz >= 3 * (g.nzm // 2) + 1 and DM.factorize(g.ppz)


def test_operator_tables_match_oracle():
    shape = (24, 33, 8)
    g = DM.make_grid(shape)
    d = O.make_domain(shape)
    ops = DM.build_operators(g, 1234.0, 0.013)
    ref = O.build_operators(d, 1234.0, 0.013)
    for s in range(3):
        assert np.abs(ops.lhs_inv[s] - ref.lhs_inv[s]).max() < 1e-12
        # R_s = rhs_base + alpha dt (-k^2)/Re on the interior diagonal
        for a in range(g.ka):
            for b in range(g.kb):
                k2 = g.kx[a] ** 2 + g.kz[b] ** 2
                R = ops.rhs_base[s].copy()
                idx = np.arange(1, g.ny - 1)
                R[idx, idx] += DM.RK_ALPHA[s] * 0.013 * (-k2) / 1234.0
                assert np.abs(R - ref.rhs[s, a, b]).max() < 1e-13
    assert np.abs(ops.poisson - ref.poisson).max() < 1e-10 * np.abs(ref.poisson).max()
    assert np.abs(ops.rhs_ns - ref.rhs_ns).max() < 1e-15
    assert np.abs(ops.lhs_inv_ns - ref.lhs_inv_ns).max() < 1e-12


@pytest.mark.parametrize("ny", [33, 48, 129, 130])
def test_interior_eigen_reproduces_inverse_tables(ny):
    """The eigen form of the implicit solve: on each parity half of the
    interior rows, E diag(1/((1 + c k^2) - c lam)) E^-1 is the interior block
    of lhs_inv[s] (c = beta_s dt/Re) and E diag(1/(lam - k^2)) E^-1 that of
    the Poisson inverse, for every |k| class tested; E E^-1 = I and
    D^2 E = E diag(lam) to rounding.  A sign slip in c or k^2, a swapped
    parity block or a dropped centre column fails here."""
    shape = (8, ny, 8)
    g = DM.make_grid(shape, 2 * np.pi, np.pi)
    re, dt = 2500.0, 0.004
    ops = DM.build_operators(g, re, dt)
    n = g.ny
    ne, no = n // 2 - 1 + n % 2, n // 2 - 1
    f = ops.y_eig
    sizes = [ne * ne, no * no, ne * ne, no * no, ne, no]
    parts = np.split(f, np.cumsum(sizes)[:-1])
    Ee, Eo = parts[0].reshape(ne, ne), parts[1].reshape(no, no)
    Xe, Xo = parts[2].reshape(ne, ne), parts[3].reshape(no, no)
    le, lo = parts[4], parts[5]
    Be, Bo = DM.parity_blocks(g.D2)
    for B, E, X, lam in ((Be, Ee, Xe, le), (Bo, Eo, Xo, lo)):
        assert np.all(lam < 0)
        assert np.abs(E @ X - np.eye(len(lam))).max() < 1e-13
        assert np.abs(B @ E - E * lam).max() < 1e-12 * np.abs(B).max()
    # parity basis of the interior: x_int = T_e e + T_o o
    Te = np.zeros((n, ne))
    To = np.zeros((n, no))
    for h in range(ne):
        Te[h + 1, h] = 1.0
        Te[n - 2 - h, h] = 1.0
    for h in range(no):
        To[h + 1, h] = 1.0
        To[n - 2 - h, h] = -1.0
    Pe = np.linalg.pinv(Te)
    Po = np.linalg.pinv(To)
    for a, b in ((0, 0), (0, 1), (1, 0), (g.ka - 1, g.kb - 1), (2, 3)):
        k2 = g.kx[a] ** 2 + g.kz[b] ** 2
        for s in range(3):
            c = DM.RK_BETA[s] * dt / re
            Li = ops.lhs_inv[s, a, b]
            P = ops.poisson[a, b]
            for T, Pr, E, X, lam in ((Te, Pe, Ee, Xe, le), (To, Po, Eo, Xo, lo)):
                dl = 1.0 / ((1.0 + c * k2) - c * lam)
                helm = T @ (E * dl) @ X @ Pr
                pois = T @ (E * (dl / (lam - k2))) @ X @ Pr
                ref_h = Li @ T @ Pr
                ref_p = P @ Li @ T @ Pr
                assert np.linalg.norm(helm - ref_h) < 1e-11 * np.linalg.norm(ref_h)
                assert np.linalg.norm(pois - ref_p) < 1e-10 * np.linalg.norm(ref_p)


def test_wall_correction_solution_satisfies_neumann():
    """v_p + a v_a + b flip(v_a) has zero wall derivative for any v_p."""
    g = DM.make_grid((24, 33, 8))
    ops = DM.build_operators(g, 3000.0, 0.01)
    rng = np.random.default_rng(0)
    for s in range(3):
        for a, b in ((0, 0), (3, 2), (g.ka - 1, g.kb - 1)):
            vp = rng.standard_normal(g.ny)
            vp[0] = vp[-1] = 0.0
            dvp = g.D @ vp
            mi = ops.infl_inv[s, a, b].reshape(2, 2)
            ab = mi @ np.array([-dvp[0], -dvp[-1]])
            va = ops.v_a[s, a, b]
            v = vp + ab[0] * va + ab[1] * va[::-1]
            dv = g.D @ v
            assert abs(dv[0]) < 1e-8 * np.abs(dvp).max() and abs(dv[-1]) < 1e-8 * np.abs(dvp).max()


@pytest.mark.parametrize("p", [7, 8, 25, 40, 49, 75, 125, 147, 196])
def test_fft_plan_is_a_permutation_and_exact(p):
    perm, tw, rad = DM.dit_tables(p, inverse=False)
    assert sorted(perm.tolist()) == list(range(p))
    assert int(np.prod(rad)) == p
    x = np.random.default_rng(p).standard_normal(p) + 1j
    # execute the in-place DIT plan exactly as the kernel does
    buf = x[perm].astype(np.complex128)
    L = 1
    for r in rad:
        Lp, L = L, L * r
        for g0 in range(p // L):
            for k1 in range(Lp):
                pos = g0 * L + k1 + np.arange(r) * Lp
                v = buf[pos] * tw[(np.arange(r) * k1 * (p // L)) % p]
                buf[pos] = np.array([sum(v[m] * tw[(q * m * (p // r)) % p] for m in range(r)) for q in range(r)])
    assert np.abs(buf - np.fft.fft(x)).max() < 1e-13 * p


def test_generator_properties():
    g = DM.make_grid((24, 33, 16), 2 * np.pi, np.pi)
    u = initial.perturbation(g, seed=5, energy=1e-6)
    assert u.shape == (3, g.nxm, g.ny, g.nzm)
    assert abs(initial.spectral_energy(g, u) - 1e-6) < 1e-18
    # walls: u = v = w = 0 and dv/dy = 0
    assert np.abs(u[:, :, 0]).max() < 1e-14 and np.abs(u[:, :, -1]).max() < 1e-14
    dv = np.einsum("lj,ijk->ilk", g.D, u[1])
    assert np.abs(dv[:, 0]).max() < 1e-10 * np.abs(u[1]).max() * g.ny**2
    # Hermitian: u[-k] = conj u[k]
    pi = (g.nxm - np.arange(g.nxm)) % g.nxm
    pj = (g.nzm - np.arange(g.nzm)) % g.nzm
    assert np.abs(u - np.conj(u[:, pi][:, :, :, pj])).max() < 1e-15
    # solenoidal: i kx u + D v + i kz w = 0
    div = 1j * g.kx[:, None, None] * u[0] + dv + 1j * g.kz[None, None, :] * u[2]
    assert np.abs(div).max() < 1e-10 * np.abs(u).max() * g.ny**2
    # deterministic
    assert np.array_equal(u, initial.perturbation(g, seed=5, energy=1e-6))


def test_product_requires_gpu():
    """Without a GPU the plan refuses to build: no CPU fallback exists."""
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from jax_spectral_dns_amd.plan import StepPlan

    with pytest.raises(_lib.JsdError):
        StepPlan(DM.make_grid((24, 65, 4)), 3000.0, 0.01, _lib.KIND_PERTURBATION)


def test_parameterisation_matches_oracle():
    """Package form (one precomputed projector, torch) against the oracle's
    per-pencil least squares."""
    from jax_spectral_dns_amd.parameterisation import vort_yvel_to_vel

    shape = (8, 33, 8)
    g = DM.make_grid(shape, 2 * np.pi, np.pi)
    d = O.make_domain(shape, 2 * np.pi, np.pi)
    rng = np.random.default_rng(1)
    sh = (g.nxm, g.ny, g.nzm)
    vort = torch.from_numpy(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    vy = torch.from_numpy(rng.standard_normal(sh) + 1j * rng.standard_normal(sh))
    ux = torch.from_numpy(rng.standard_normal(g.ny) + 0j)
    wz = torch.from_numpy(rng.standard_normal(g.ny) + 0j)
    a = vort_yvel_to_vel(g, vort, vy, ux, wz)
    b = O.vort_yvel_to_vel(d, vort, vy, ux, wz)
    assert float((a - b).abs().max() / b.abs().max()) < 1e-13


def test_sampled_operator_tables_match_full_ones():
    """build_operators(full=False) (the eigen-form plans' build) equals the
    full build at the checked |k| entries, and its wall-correction tables
    (from LU solves) equal the full build's to rounding."""
    g = DM.make_grid((24, 33, 16), 2 * np.pi, np.pi)
    full = DM.build_operators(g, 2000.0, 0.01)
    samp = DM.build_operators(g, 2000.0, 0.01, full=False)
    assert full.full and not samp.full
    nk = g.ka * g.kb
    ck = DM.check_indices(nk)
    assert ck[0] == 0 and ck[-1] == nk - 1 and all(k % 13 == 0 or k == nk - 1 for k in ck)
    li_f = full.lhs_inv.reshape(3, nk, g.ny, g.ny)
    li_s = samp.lhs_inv.reshape(3, nk, g.ny, g.ny)
    assert np.array_equal(li_f[:, ck], li_s[:, ck])
    po_f = full.poisson.reshape(nk, g.ny, g.ny)
    po_s = samp.poisson.reshape(nk, g.ny, g.ny)
    assert np.array_equal(po_f[ck], po_s[ck])
    assert np.abs(samp.v_a - full.v_a).max() < 1e-13 * np.abs(full.v_a).max()
    assert np.abs(samp.infl_inv - full.infl_inv).max() < 1e-10 * np.abs(full.infl_inv).max()
    for name in ("rhs_base", "rhs_ns", "lhs_inv_ns", "y_eig"):
        assert np.array_equal(getattr(samp, name), getattr(full, name)), name
