"""Pins of the oracle's source term, enforce_base switch and the dual's
objectives (gain_3d, gain_no_00, dissipation) and their inputs.

* The (0,0) source enters the mean-flow block exactly like a pressure
  gradient (navier_stokes.py:1461-1490), which the laminar-flow pin fixes.
* The h_v / h_g addends follow NumPy's broadcasting of the reference's
  expression (navier_stokes.py:1697-1704) evaluated on the reference's
  meshgrid (domain.py:827), independently of the oracle's diff helper; a
  curl-free source without a (0,0) mode changes nothing.
* enforce_base drops exactly the (0,0) convective term.
* The dissipation average, taken of the squared spectral coefficients, sees
  only the mean-flow shear: closed form for a parabolic mean profile.
"""

import numpy as np
import pytest
import torch

from oracle import jsd_oracle as O
from tests.helpers import ns_field, pair, perturbation_field, rel


def _mode_scale(d):
    return np.sqrt(d.px * d.pz) / d.scaling


def test_source_00_is_a_pressure_gradient():
    g, d = pair((16, 33, 8), 2 * np.pi, np.pi)
    u = ns_field(g, seed=3)
    c = -2.0 / 1000.0
    a = O.rk_step(O.make_ns(d, 1000.0, 0.01, c), u, 0)
    sp = O.make_ns(d, 1000.0, 0.01, 0.0)
    src = torch.zeros_like(u)
    # dpdx_spectral = dPdx sqrt(Px (2pi/Lx)^2 Pz (2pi/Lz)^2): the (0,0) coefficient of a constant field
    src[0, 0, :, 0] = -c * ((d.px * (2 * np.pi / d.lx) ** 2) * (d.pz * (2 * np.pi / d.lz) ** 2)) ** 0.5
    sp.source = lambda t: src
    b = O.rk_step(sp, u, 0)
    assert rel(b, a) < 1e-15


def test_source_addends_follow_numpy_broadcasting():
    d = O.make_domain((16, 17, 8), 2 * np.pi, np.pi)
    rng = np.random.default_rng(4)
    sh = (3, d.nxm, d.ny, d.nzm)
    s = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    # the reference's FourierDomain mgrid and diff: (1j * mgrid[direction]) ** order * field
    mg = np.meshgrid(d.kx, d.y, d.kz, indexing="ij")
    curl_1 = (1j * mg[2]) * s[0] - (1j * mg[0]) * s[2]
    want_hv = ((1j * mg[2]) * curl_1[0] - (1j * mg[0]) * curl_1[2])[1]
    want_hg = curl_1[1]
    z = torch.zeros((d.nxm, d.ny, d.nzm), dtype=torch.complex128)
    hv, hg = O.source_step_terms(d, torch.from_numpy(s), z, z)
    for k in range(d.nxm):
        assert np.abs(hv[k].numpy() - want_hv).max() < 1e-12 * np.abs(want_hv).max()
        assert np.abs(hg[k].numpy() - want_hg).max() < 1e-12 * np.abs(want_hg).max()


def test_curl_free_source_changes_nothing():
    g, d = pair((16, 33, 8), 2 * np.pi, np.pi)
    u = perturbation_field(g, seed=6, energy=1e-3)
    phi = perturbation_field(g, seed=7, energy=1.0)[1]
    phi[0, :, 0] = 0
    src = torch.stack([O.diff(d, phi, 0), O.diff(d, phi, 1), O.diff(d, phi, 2)])
    sp = O.make_perturbation(d, 2000.0, 0.01, linearise=False)
    a = O.rk_step(sp, u, 0)
    sp.source = lambda t: src
    b = O.rk_step(sp, u, 0)
    # phi has no (0,0) mode, so s_x and s_z vanish there too: nothing changes
    assert rel(b, a) < 1e-14


def test_enforce_base_drops_only_the_mean_convection():
    """enforce_base equals a nonlinear term whose conv_ns has its (0,0) line
    zeroed (the new and the carried old term), and nothing else: the other
    pencils change only through the mean flow's feedback in later substeps."""
    g, d = pair((16, 33, 8), 2 * np.pi, np.pi)
    u = ns_field(g, seed=8)
    sp = O.make_ns(d, 2000.0, 0.01, -2.0 / 2000.0)
    plain = O.rk_step(sp, u, 0)
    sp.enforce_base = True
    enf = O.rk_step(sp, u, 0)
    sp.enforce_base = False
    base_nl = sp.nonlinear

    def no_mean_conv(vh, t):
        h_v, h_g, vort, conv = base_nl(vh, t)
        conv = conv.clone()
        conv[:, 0, :, 0] = 0
        return h_v, h_g, vort, conv

    sp.nonlinear = no_mean_conv
    wrapped = O.rk_step(sp, u, 0)
    assert rel(enf, wrapped) < 1e-15
    assert rel(enf[:, 0, :, 0], plain[:, 0, :, 0]) > 1e-8


def test_laplacian_closed_form():
    d = O.make_domain((16, 17, 8))
    X = np.meshgrid(d.x, d.y, d.z, indexing="ij")
    y = X[1]
    f = np.sin(2 * X[0]) * np.cos(X[2]) * (1 - y**2) * y
    lap = -5 * f + np.sin(2 * X[0]) * np.cos(X[2]) * (-6 * y)
    got = O.field_no_hat(d, O.laplacian(d, O.field_hat(d, torch.from_numpy(f)))).numpy()
    assert np.abs(got - lap).max() < 1e-11


def test_dissipation_average_closed_form():
    """A mean profile U = c (1 - y^2) (coefficient C = c sqrt(PxPz)/s) gives
    sum_ij (d_j u_i)^2 = 4 C^2 y^2 on the (0,0) line; the volume integral of
    its no_hat over n_states identical states / |Re n_states| is
    Lx Lz (8/3) C^2 s / (sqrt(PxPz) |Re|).  Non-mean modes add nothing (the
    volume integral sees the (0,0) coefficient only)."""
    g, d = pair((16, 17, 8), 2 * np.pi, np.pi)
    c, re = 0.7, 1234.0
    C = c * _mode_scale(d)
    u = perturbation_field(g, seed=9, energy=1e-2)
    u[:, 0, :, 0] = 0
    u[0, 0, :, 0] = torch.from_numpy(C * (1 - d.y**2) + 0j)
    hist = [u, u, u]
    got = float(O.objective(d, "dissipation", u, u, hist, -re))
    want = d.lx * d.lz * 8.0 / 3.0 * C**2 * d.scaling / (np.sqrt(d.px * d.pz) * re)
    assert abs(got - want) < 1e-11 * abs(want)


def test_gain_3d_and_no_00_objectives():
    g, d = pair((16, 17, 8), 2 * np.pi, np.pi)
    u = perturbation_field(g, seed=10, energy=1e-2)
    only_kx0 = torch.zeros_like(u)
    only_kx0[:, 0] = u[:, 0]
    assert float(O.objective(d, "gain_3d", u, only_kx0, None, 1.0)) < 1e-28
    only_00 = torch.zeros_like(u)
    only_00[:, 0, :, 0] = u[:, 0, :, 0] + 0.1
    assert float(O.objective(d, "gain_no_00", u, only_00, None, 1.0)) < 1e-28
    rest = u - only_kx0
    g3 = float(O.objective(d, "gain_3d", u, u, None, 1.0))
    assert abs(g3 - float(O.energy(d, rest) / O.energy(d, u))) < 1e-15
    ic = O.dual_initial_condition(u, "gain_3d")
    assert float(ic[:, 0].abs().max()) == 0.0 and torch.equal(ic[:, 1:], -u[:, 1:])
    ic = O.dual_initial_condition(u, "gain_no_00")
    assert float(ic[:, 0, :, 0].abs().max()) == 0.0 and torch.equal(ic[:, 1:], -u[:, 1:])
    assert float(O.dual_initial_condition(u, "dissipation").abs().max()) == 0.0
