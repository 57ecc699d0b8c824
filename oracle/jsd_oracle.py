"""CPU restatement of the velocity-vorticity channel-flow step of
dakling/jax-spectral-dns (reference package ``jax_spectral_dns``).

TEST INFRASTRUCTURE ONLY.  This module is the checker for the HIP product
path: only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it.  The product (``jax_spectral_dns_amd``)
never imports, links or falls back to anything in ``oracle/``.

Every function restates one piece of the reference line by line, in torch
FP64 on the CPU (so that ``torch.autograd`` through it stands in for
``jax.grad`` of the reference), with NumPy for the host precompute.  The
reference file and line range each function follows is given in its
docstring.  Arrays follow the reference layout: spectral fields are
complex128 of shape (Nx_modes, Ny, Nz_modes) with wavenumbers stored in the
order [0..K, -K..-1]; physical fields are float64 of shape (Px, Ny, Pz).

How each part is pinned (so that a plausible mistake fails a check) is
listed in ``oracle/PINNING.md``; the checks themselves live in
``tests/test_oracle_*.py``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from functools import reduce
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np
import torch

C128 = torch.complex128
F64 = torch.float64

# RK3/CN low-storage coefficients, navier_stokes.py:898-918
RK_ALPHA = (29 / 96, -3 / 40, 1 / 6)
RK_BETA = (37 / 160, 5 / 24, 1 / 6)
RK_GAMMA = (8 / 15, 5 / 12, 3 / 4)
RK_XI = (0, -17 / 60, -5 / 12)


# ---------------------------------------------------------------------------
# grids and operators (domain.py)
# ---------------------------------------------------------------------------


def mode_and_physical_shape(
    shape: Tuple[int, int, int], aliasing: float = 3 / 2
) -> Tuple[Tuple[int, int, int], Tuple[int, int, int]]:
    """Stored-mode shape and physical grid size for periodic (x, z) and a
    Chebyshev y direction.  domain.py:217-258 (shape rule) together with
    get_fourier_grid domain.py:143-152 (an odd argument drops one point)."""
    periodic = (True, False, True)
    modes = tuple(
        int(n) if (not p or int(n) % 2 != 0) else int(n) + 1
        for n, p in zip(shape, periodic)
    )
    phys = []
    for m, p in zip(modes, periodic):
        if not p:
            phys.append(m)
            continue
        c = math.ceil(m * aliasing)
        arg = c if c % 2 != 0 else c + 1
        # get_fourier_grid: odd N -> N-1 points
        phys.append(arg - 1 if arg % 2 != 0 else arg)
    return modes, tuple(phys)  # type: ignore[return-value]


def cheb_grid(n: int) -> np.ndarray:
    """Gauss-Lobatto points y_j = cos(pi/(n-1) * j).  domain.py:130-140."""
    return np.array([np.cos(np.pi / (n - 1) * i) for i in np.arange(n)])


def fourier_grid(n_arg: int, length: float) -> np.ndarray:
    """Equidistant periodic grid.  domain.py:143-152."""
    n = int(n_arg - 1) if n_arg % 2 != 0 else int(n_arg)
    if n % 2 != 0:
        raise Exception("Fourier discretization points must be even!")
    return np.linspace(start=0.0, stop=length, num=int(n + 1))[:-1]


def cheb_diff_mat(xs: np.ndarray) -> np.ndarray:
    """Collocation first-derivative matrix with the negative-sum diagonal.
    domain.py:155-163 (order 1)."""
    n = len(xs)
    c = np.block([2.0, np.ones((1, n - 2)) * 1.0, 2.0]) * (-1) ** np.arange(0, n)
    x_rep = np.repeat(xs, n).reshape(n, n)
    dx = x_rep - np.transpose(x_rep)
    d_ = np.transpose((np.transpose(1 / c) @ c)) / (dx + np.eye(n))
    return np.linalg.matrix_power(d_ - np.diag(sum(np.transpose(d_))), 1)


def fourier_wavenumbers(m: int, length: float) -> np.ndarray:
    """Stored wavenumbers [0..K, -K..-1] * 2 pi / L.  domain.py:796-827."""
    inp = np.linspace(0, m - 1, int(m))
    return (np.block([inp[m // 2 :], inp[: m // 2]]) - m // 2) * (2 * np.pi) / length


def hom_dirichlet(mat: np.ndarray) -> np.ndarray:
    """First/last row and column -> unit.  domain.py:375-397."""
    n = mat.shape[0]
    out = np.block([[1, np.zeros((1, n - 1))], [np.zeros((n - 1, 1)), mat[1:, 1:]]])
    n = out.shape[0]
    return np.block([[out[:-1, :-1], np.zeros((n - 1, 1))], [np.zeros((1, n - 1)), 1]])


def inhom_dirichlet(
    mat: np.ndarray, rhs: np.ndarray, bc_left: float, bc_right: float
) -> Tuple[np.ndarray, np.ndarray]:
    """First/last row -> unit, rhs ends -> (bc_right, bc_left).
    domain.py:425-448."""
    n = mat.shape[0]
    out = np.block([[1, np.zeros((1, n - 1))], [mat[1:, :]]])
    n = out.shape[0]
    out = np.block([[out[:-1, :]], [np.zeros((1, n - 1)), 1]])
    out_rhs = np.block([bc_right, rhs[1:-1], bc_left])
    return out, out_rhs


@dataclass
class Dom:
    """Physical + Fourier domain pair for a channel (x, z periodic, y
    Chebyshev).  Mirrors PhysicalDomain.create / FourierDomain.FromPhysicalDomain
    (domain.py:201-317, domain.py:791-848)."""

    nxm: int
    ny: int
    nzm: int
    px: int
    pz: int
    lx: float
    lz: float
    aliasing: float
    y: np.ndarray
    x: np.ndarray
    z: np.ndarray
    D: np.ndarray
    kx: np.ndarray
    kz: np.ndarray
    scaling: float
    ppx: int = 0  # transform grid actually used (== px unless overridden)
    ppz: int = 0
    cache: Dict[str, object] = field(default_factory=dict)

    @property
    def overridden(self) -> bool:
        return (self.ppx, self.ppz) != (self.px, self.pz)

    @property
    def D2(self) -> np.ndarray:
        if "D2" not in self.cache:
            self.cache["D2"] = np.linalg.matrix_power(self.D, 2)
        return self.cache["D2"]  # type: ignore[return-value]

    def t_(self, name: str, mat: np.ndarray, dtype=C128) -> torch.Tensor:
        key = "t_" + name + str(dtype)
        if key not in self.cache:
            self.cache[key] = torch.from_numpy(np.ascontiguousarray(mat)).to(dtype)
        return self.cache[key]  # type: ignore[return-value]


def make_domain(
    shape: Tuple[int, int, int],
    lx: float = 2 * np.pi,
    lz: float = 2 * np.pi,
    aliasing: float = 3 / 2,
    physical_override: Optional[Tuple[int, int]] = None,
) -> Dom:
    """Build the domain pair.  ``physical_override`` replaces the physical
    (Px, Pz) sizes (any P >= 3K+1 gives the same truncated spectrum of a
    quadratic product up to rounding; used only to prove that property)."""
    modes, phys = mode_and_physical_shape(shape, aliasing)
    px, pz = phys[0], phys[2]
    ppx, ppz = (px, pz) if physical_override is None else physical_override
    for m, pp in ((modes[0], ppx), (modes[2], ppz)):
        if physical_override is not None and pp < 3 * (m // 2) + 1:
            raise ValueError("transform grid %d cannot dealias %d modes" % (pp, m))
    y = cheb_grid(modes[1])
    D = cheb_diff_mat(y)
    scaling = 1.0
    for length in (lx, lz):  # domain.py:718-720 (periodic dims 0 then 2)
        scaling *= length / (2 * np.pi)
    return Dom(
        nxm=modes[0],
        ny=modes[1],
        nzm=modes[2],
        px=px,
        pz=pz,
        lx=lx,
        lz=lz,
        aliasing=aliasing,
        y=y,
        x=np.linspace(0.0, lx, ppx + 1)[:-1],
        z=np.linspace(0.0, lz, ppz + 1)[:-1],
        D=D,
        kx=fourier_wavenumbers(modes[0], lx),
        kz=fourier_wavenumbers(modes[2], lz),
        scaling=scaling,
        ppx=ppx,
        ppz=ppz,
    )


# ---------------------------------------------------------------------------
# transforms (domain.py:716-764, domain.py:1167-1242)
# ---------------------------------------------------------------------------


def field_hat(dom: Dom, f: torch.Tensor) -> torch.Tensor:
    """Physical (Px, Ny, Pz) -> stored spectrum (Nxm, Ny, Nzm).
    fftn over (0, 2) with norm='ortho', divided by prod(L/2pi); per axis keep
    bins [0:ks], [ks], conj([ks]), [N-ks+1:N].  domain.py:716-764."""
    out = torch.fft.fftn(f.to(C128), dim=(0, 2), norm="ortho") / dom.scaling
    if dom.overridden:
        # ortho scaling of the reference grid: sqrt(Px Pz) c_k
        out = out * math.sqrt((dom.px * dom.pz) / (dom.ppx * dom.ppz))
    for axis, n, m in ((0, f.shape[0], dom.nxm), (2, f.shape[2], dom.nzm)):
        ks = int((n - n * (1 - 1 / dom.aliasing)) / 2) if not dom.overridden else m // 2
        o1 = out.narrow(axis, 0, ks)
        o2 = out.narrow(axis, ks, 1)
        o3 = out.narrow(axis, n - ks + 1, ks - 1)
        out = torch.cat([o1, o2, torch.conj(o2), o3], dim=axis)
    return out.to(C128)


def field_no_hat(dom: Dom, fh: torch.Tensor) -> torch.Tensor:
    """Stored spectrum -> physical real field: zero-pad per axis
    [0:ks+1], zeros(ceil(N*(a-1))), [N-ks:N]; ifftn(ortho).real * prod(L/2pi).
    domain.py:1167-1242.  With an overridden transform grid the padding
    fills that grid and the reference's ortho scaling is kept."""
    for axis in (0, 2):
        n = fh.shape[axis]
        ks = int(n / 2)
        f1 = fh.narrow(axis, 0, ks + 1)
        if not dom.overridden:
            nz = math.ceil(n * (dom.aliasing - 1))
        else:
            nz = (dom.ppx if axis == 0 else dom.ppz) - n
        f2 = fh.narrow(axis, n - ks, ks)
        if nz == 0:
            f2 = f2.narrow(axis, 1, f2.shape[axis] - 1)
        zshape = list(f1.shape)
        zshape[axis] = nz
        fh = torch.cat([f1, torch.zeros(zshape, dtype=fh.dtype), f2], dim=axis)
    out = torch.fft.ifftn(fh, dim=(0, 2), norm="ortho").real / (1 / dom.scaling)
    if dom.overridden:
        out = out * math.sqrt((dom.ppx * dom.ppz) / (dom.px * dom.pz))
    return out


def diff(dom: Dom, fh: torch.Tensor, direction: int, order: int = 1) -> torch.Tensor:
    """Spectral derivative: (i k)^order * f in x/z; einsum with
    matrix_power(D, order) along y.  domain.py:891-918."""
    if direction == 0:
        fac = (1j * dom.kx) ** order
        return torch.from_numpy(fac).to(C128)[:, None, None] * fh
    if direction == 2:
        fac = (1j * dom.kz) ** order
        return torch.from_numpy(fac).to(C128)[None, None, :] * fh
    Dm = dom.t_("Dpow%d" % order, np.linalg.matrix_power(dom.D, order))
    return torch.einsum("ijk,lj->ilk", fh, Dm)


def diff_phys_y(dom: Dom, f: torch.Tensor, order: int = 1) -> torch.Tensor:
    """Physical-space y derivative.  domain.py:355-373."""
    Dm = dom.t_("Dpow%d" % order, np.linalg.matrix_power(dom.D, order), F64)
    return torch.einsum("ijk,lj->ilk", f, Dm)


def divergence(dom: Dom, vh: torch.Tensor) -> torch.Tensor:
    """domain.py:676-678."""
    return diff(dom, vh[0], 0) + diff(dom, vh[1], 1) + diff(dom, vh[2], 2)


def curl(dom: Dom, vh: torch.Tensor) -> torch.Tensor:
    """FourierDomain.curl, domain.py:920-933."""
    u_y = diff(dom, vh[0], 1)
    u_z = diff(dom, vh[0], 2)
    v_x = diff(dom, vh[1], 0)
    v_z = diff(dom, vh[1], 2)
    w_x = diff(dom, vh[2], 0)
    w_y = diff(dom, vh[2], 1)
    return torch.stack([w_y - v_z, u_z - w_x, v_x - u_y])


def update_bc(f: torch.Tensor, axis: int = 1) -> torch.Tensor:
    """Zero the two y-end values.  domain.py:630-647, domain.py:1148-1165."""
    out = f.clone()
    idx = [slice(None)] * f.dim()
    idx[axis] = 0
    out[tuple(idx)] = 0
    idx[axis] = -1
    out[tuple(idx)] = 0
    return out


# ---------------------------------------------------------------------------
# integrals and energy (domain.py:485-628, field.py:1456-1473, 2618-2711)
# ---------------------------------------------------------------------------


def _yint_mat(dom: Dom) -> np.ndarray:
    """D with first row and column replaced by e0 (set_first_mat_row_and_col_to_unit,
    domain.py:509-519)."""
    n = dom.ny
    mat = np.linalg.matrix_power(dom.D, 1)
    first = np.concatenate([np.ones(1), np.zeros(n - 1)])
    return np.vstack([first, np.hstack([np.zeros((n - 1, 1)), mat[1:, 1:]])])


def y_int_weights(dom: Dom) -> np.ndarray:
    """Weights w with definite_integral_y(f) = sum_j w_j f_j: integrate with
    bc_right=0 solves D' F = f with row/col 0 -> e0 and f_0 -> 0; the
    definite value is F[0] - F[N-1].  domain.py:577-628, field.py:2628-2665."""
    n = dom.ny
    inv = np.linalg.inv(_yint_mat(dom))
    w = inv[0, :] - inv[n - 1, :]
    w[0] = 0.0  # the field's first entry is replaced by bc_right = 0
    return w


def volume_integral(dom: Dom, f: torch.Tensor) -> torch.Tensor:
    """z (L/N * sum), then y (inverse-D quadrature), then x.  field.py:1456-1466."""
    fz = dom.lz / f.shape[2] * reduce(lambda a, b: a + b, torch.moveaxis(f, 2, 0))
    # y: literal solve with the bc-modified D, then first minus last entry
    n = dom.ny
    inv = dom.t_("yint_inv", np.linalg.inv(_yint_mat(dom)), F64)
    b = fz.clone()
    b[:, 0] = 0.0
    out = torch.einsum("ij,lj->il", b, inv)
    fy = out[:, 0] - out[:, n - 1]
    return dom.lx / f.shape[0] * torch.sum(fy)


def energy_phys(dom: Dom, u: torch.Tensor) -> torch.Tensor:
    """VectorField.energy: sum_i 0.5 * int(u_i^2) / (2 Lx Lz).
    field.py:607-611, field.py:1468-1473."""
    vol = 2.0 * dom.lx * dom.lz
    e = torch.zeros((), dtype=F64)
    for i in range(u.shape[0]):
        e = e + volume_integral(dom, 0.5 * u[i] * u[i]) / vol
    return e


def energy(dom: Dom, vh: torch.Tensor) -> torch.Tensor:
    """Energy of a spectral velocity (no_hat then energy)."""
    return energy_phys(dom, torch.stack([field_no_hat(dom, vh[i]) for i in range(vh.shape[0])]))


def energy_with_other(dom: Dom, vh: torch.Tensor, wh: torch.Tensor) -> torch.Tensor:
    """field.py:613-620 / field.py:1475-1480."""
    vol = 2.0 * dom.lx * dom.lz
    e = torch.zeros((), dtype=F64)
    for i in range(vh.shape[0]):
        e = e + volume_integral(dom, 0.5 * field_no_hat(dom, vh[i]) * field_no_hat(dom, wh[i])) / vol
    return e


# ---------------------------------------------------------------------------
# nonlinear terms
# ---------------------------------------------------------------------------


def helicity_to_nonlinear_terms(dom: Dom, hel: torch.Tensor, vh: torch.Tensor):
    """h_v = -d_y(d_x H0 + d_z H2) + d_xx H1 + d_zz H1; h_g = d_z H0 - d_x H2;
    conv_ns = -H; vort = curl(u).  navier_stokes.py:48-73."""
    h_v = -diff(dom, diff(dom, hel[0], 0) + diff(dom, hel[2], 2), 1) + diff(dom, hel[1], 0, 2) + diff(dom, hel[1], 2, 2)
    h_g = diff(dom, hel[0], 2) - diff(dom, hel[2], 0)
    return h_v, h_g, curl(dom, vh), -hel


def vel_1_nabla_vel_2(dom: Dom, vel_1: torch.Tensor, vel_2_hat: torch.Tensor) -> torch.Tensor:
    """(a . grad) b with b spectral.  navier_stokes.py:76-90."""
    out = torch.zeros_like(vel_1)
    for i in range(3):
        for j in range(3):
            out[i] = out[i] + vel_1[j] * field_no_hat(dom, diff(dom, vel_2_hat[i], j))
    return out


def div_vel_1_vel_2(dom: Dom, vel_1_hat: torch.Tensor, vel_2: torch.Tensor) -> torch.Tensor:
    """(div a) b_i.  navier_stokes.py:216-224."""
    d = field_no_hat(dom, divergence(dom, vel_1_hat))
    return torch.stack([d * vel_2[i] for i in range(3)])


def _hat3(dom: Dom, f: torch.Tensor) -> torch.Tensor:
    return torch.stack([field_hat(dom, f[i]) for i in range(3)])


def _nohat3(dom: Dom, fh: torch.Tensor) -> torch.Tensor:
    return torch.stack([field_no_hat(dom, fh[i]) for i in range(3)])


def nl_ns_skew(dom: Dom, vh: torch.Tensor):
    """Skew-symmetric NS term: H = hat(-u.grad u) + 0.5 hat((div u) u).
    navier_stokes.py:236-261 (with :93-126)."""
    vel = _nohat3(dom, vh)
    dvv_hat = _hat3(dom, div_vel_1_vel_2(dom, vh, vel))
    hel = _hat3(dom, -vel_1_nabla_vel_2(dom, vel, vh)) + 0.5 * dvv_hat
    return helicity_to_nonlinear_terms(dom, hel, vh)


def nl_perturbation_skew(
    dom: Dom, vh: torch.Tensor, base_hat: torch.Tensor, linearise: bool, coupling_term: bool = True
):
    """Perturbation skew-symmetric term about base flow U.
    navier_stokes_perturbation.py:160-239 (with :43-71)."""
    c = 0.0 if linearise else 1.0
    vel = _nohat3(dom, vh)
    vel_base = _nohat3(dom, base_hat)
    dvv = (
        c * div_vel_1_vel_2(dom, vh, vel)
        + div_vel_1_vel_2(dom, base_hat, vel)
        + div_vel_1_vel_2(dom, vh, vel_base)
    )
    dvv_hat = _hat3(dom, dvv)
    hel_phys = -(
        c * vel_1_nabla_vel_2(dom, vel, vh)
        + vel_1_nabla_vel_2(dom, vel_base, vh)
        + vel_1_nabla_vel_2(dom, vel, base_hat)
    )
    hel = _hat3(dom, hel_phys) + 0.5 * dvv_hat
    h_v, h_g, vort, conv = helicity_to_nonlinear_terms(dom, hel, vh)
    if not coupling_term:
        comb = field_hat(dom, diff_phys_y(dom, vel_base[0]) * vel[1])
        h_g = h_g + diff(dom, comb, 2)
    return h_v, h_g, vort, conv


def nl_dual_skew(
    dom: Dom,
    vh: torch.Tensor,
    base_hat: torch.Tensor,
    u_hat: torch.Tensor,
    linearise: bool,
    coupling_term: bool = True,
):
    """Dual (adjoint) skew-symmetric term, W = c u + U:
    H = -hat((W.grad) v) + hat([sum_j v_j d_i W_j]_i)
        + 0.5 hat(0.5((div W) v + (div v) W)).
    navier_stokes_perturbation_dual.py:214-294 (with :61-90); the Reynolds
    stress term re_ijj is zero (navier_stokes_perturbation_dual.py:479-484)."""
    c = 0.0 if linearise else 1.0
    vel_v = _nohat3(dom, vh)
    w_hat = c * u_hat + base_hat
    vel_w = _nohat3(dom, w_hat)
    dvv = 0.5 * (div_vel_1_vel_2(dom, w_hat, vel_v) + div_vel_1_vel_2(dom, vh, vel_w))
    dvv_hat = _hat3(dom, dvv)
    t1 = _hat3(dom, vel_1_nabla_vel_2(dom, vel_w, vh))
    t2 = []
    for i in range(3):
        acc = torch.zeros_like(vel_v[0])
        for j in range(3):
            acc = acc + vel_v[j] * field_no_hat(dom, diff(dom, w_hat[j], i))
        t2.append(field_hat(dom, acc))
    hel = -t1 + torch.stack(t2) + 0.5 * dvv_hat
    h_v, h_g, vort, conv = helicity_to_nonlinear_terms(dom, hel, vh)
    if not coupling_term:
        vb0 = field_no_hat(dom, base_hat[0])
        comb_hy = field_hat(dom, diff_phys_y(dom, vb0) * vel_v[0])
        h_v = h_v + diff(dom, comb_hy, 0, 2) + diff(dom, comb_hy, 2, 2)
    return h_v, h_g, vort, conv


# ---------------------------------------------------------------------------
# operator tables (navier_stokes.py:920-1022, domain.py:857-889)
# ---------------------------------------------------------------------------


@dataclass
class Operators:
    """Per distinct (|kx|, |kz|) real operator tables.  The reference stores
    the same values per signed (kx, kz) as complex128; the values depend on
    kx^2 + kz^2 only, and -kx is computed as an exact negation, so the tables
    are bitwise identical for both signs."""

    re_tau: float
    dt: float
    rhs: np.ndarray  # (3, Ka, Kb, Ny, Ny)
    lhs_inv: np.ndarray  # (3, Ka, Kb, Ny, Ny)
    lhs_inv_inh: np.ndarray  # (3, Ka, Kb, Ny, Ny)
    rhs_inh: np.ndarray  # (3, Ka, Kb, Ny)
    poisson: np.ndarray  # (Ka, Kb, Ny, Ny)
    rhs_ns: np.ndarray  # (3, 2Ny, 2Ny)
    lhs_inv_ns: np.ndarray  # (3, 2Ny, 2Ny)


def _hom_dir_stack(mat: np.ndarray) -> np.ndarray:
    """hom_dirichlet applied to every matrix of a stack (domain.py:375-397)."""
    out = mat.copy()
    out[..., 0, :] = 0.0
    out[..., :, 0] = 0.0
    out[..., -1, :] = 0.0
    out[..., :, -1] = 0.0
    out[..., 0, 0] = 1.0
    out[..., -1, -1] = 1.0
    return out


def build_operators(dom: Dom, re_tau: float, dt: float) -> Operators:
    """prepare_assemble_rk_matrices (navier_stokes.py:920-1022) and
    assemble_poisson_matrix (domain.py:857-889), restricted to |kx|, |kz|."""
    n = dom.ny
    ka, kb = dom.nxm // 2 + 1, dom.nzm // 2 + 1
    D2 = np.linalg.matrix_power(dom.D, 2)
    Ly = 1 / re_tau * D2
    I = np.eye(n)
    Z = np.zeros((n, n))
    D2h = hom_dirichlet(np.linalg.matrix_power(dom.D, 2))
    L_ns_y = 1 / re_tau * np.block([[D2h, Z], [Z, D2h]])
    rhs = np.zeros((3, ka, kb, n, n))
    lhs_inv = np.zeros((3, ka, kb, n, n))
    lhs_inv_inh = np.zeros((3, ka, kb, n, n))
    rhs_inh = np.zeros((3, ka, kb, n))
    rhs_ns = np.zeros((3, 2 * n, 2 * n))
    lhs_inv_ns = np.zeros((3, 2 * n, 2 * n))
    kx2 = dom.kx[:ka, None] ** 2
    kz2 = dom.kz[None, :kb] ** 2
    ksq = (kx2 + kz2)[:, :, None, None]  # kx**2 + kz**2 per (a, b), as in the loop body
    L = Ly[None, None] + I[None, None] * (-(ksq)) / re_tau
    for s in range(3):
        # stacked over (a, b); np.linalg.inv runs the same LAPACK solve per matrix
        rhs[s] = _hom_dir_stack(I + RK_ALPHA[s] * dt * L)
        lhs_inv[s] = np.linalg.inv(_hom_dir_stack(I - RK_BETA[s] * dt * L))
        inh = (I - RK_BETA[s] * dt * L).copy()
        inh[..., 0, :] = 0.0
        inh[..., 0, 0] = 1.0
        inh[..., -1, :] = 0.0
        inh[..., -1, -1] = 1.0
        lhs_inv_inh[s] = np.linalg.inv(inh)
        rhs_inh[s] = inhom_dirichlet(I, np.zeros(n), 0.0, 1.0)[1]
        I_ns = np.eye(2 * n)
        L_ns = L_ns_y + I_ns * (-(0**2 + 0**2)) / re_tau
        rhs_ns[s] = hom_dirichlet(I_ns + RK_ALPHA[s] * dt * L_ns)
        lhs_inv_ns[s] = np.linalg.inv(hom_dirichlet(I_ns - RK_BETA[s] * dt * L_ns))
    # Poisson: inv(-(kx^2+kz^2) * eye_bc + hom_dir(D2)), domain.py:857-889
    eye_bc = np.eye(n)
    eye_bc[0, 0] = 0.0
    eye_bc[-1, -1] = 0.0
    poisson = np.linalg.inv((-(ksq)) * eye_bc[None, None] + D2h[None, None])
    return Operators(re_tau, dt, rhs, lhs_inv, lhs_inv_inh, rhs_inh, poisson, rhs_ns, lhs_inv_ns)


def _abs_index(m: int) -> np.ndarray:
    """Stored index -> |k| table index (i for i <= K, m - i otherwise)."""
    i = np.arange(m)
    return np.where(i <= m // 2, i, m - i)


def apply_k(dom: Dom, table: np.ndarray, x: torch.Tensor, key: str) -> torch.Tensor:
    """Per-pencil matvec table[|kx|,|kz|] @ x[kx, kz, :] for complex pencils
    x of shape (Nxm, Nzm, Ny) with a real table (Ka, Kb, Ny, Ny).  The four
    sign quadrants (+-kx, +-kz) each map one-to-one onto the table; pencils on
    kx = 0 or kz = 0 are computed twice with identical values."""
    T = dom.t_(key, table, F64)
    ka, kb = T.shape[0], T.shape[1]
    out = torch.empty_like(x)
    a = torch.arange(ka)
    b = torch.arange(kb)
    for ia in (a, (dom.nxm - a) % dom.nxm):
        for ib in (b, (dom.nzm - b) % dom.nzm):
            xs = x[ia][:, ib]
            re = torch.matmul(T, xs.real.unsqueeze(-1)).squeeze(-1)
            im = torch.matmul(T, xs.imag.unsqueeze(-1)).squeeze(-1)
            out[ia[:, None], ib[None, :]] = torch.complex(re, im)
    return out


def apply_k_vec(dom: Dom, table: np.ndarray, key: str) -> torch.Tensor:
    ai = torch.from_numpy(_abs_index(dom.nxm))
    bi = torch.from_numpy(_abs_index(dom.nzm))
    T = dom.t_(key, table, F64)
    return T[ai][:, bi].to(C128)


# ---------------------------------------------------------------------------
# one RK3/CN time step (navier_stokes.py:1219-1775)
# ---------------------------------------------------------------------------


@dataclass
class StepParams:
    dom: Dom
    ops: Operators
    dt: float  # the explicit-term time step (negative for the dual)
    dPdx: float = 0.0
    nonlinear: Optional[Callable[[torch.Tensor, int], Tuple[torch.Tensor, ...]]] = None
    # get_source_term(time_step) -> None | complex (3, Nx, Ny, Nz) (navier_stokes.py:1216-1217)
    source: Optional[Callable[[int], Optional[torch.Tensor]]] = None
    # enforce_base: the (0,0) convective terms are dropped (navier_stokes.py:1581-1590)
    enforce_base: bool = False


def source_step_terms(dom: Dom, source_hat: torch.Tensor, h_v: torch.Tensor, h_g: torch.Tensor):
    """The source term's addends to h_v, h_g as the reference writes them
    (navier_stokes.py:1697-1704): curl_source = curl(source)[1] is the y
    component (Nx, Ny, Nz); curl_source[0] and curl_source[2] index its kx
    rows 0 and 2, whose (Ny, Nz) derivatives broadcast over kx, and the
    trailing [1] keeps kx row 1 of their difference, which again broadcasts
    over every kx row of h_v; h_g gains kx row 1 of curl_source."""
    curl_source = curl(dom, source_hat)[1]
    h_v = h_v + (diff(dom, curl_source[0], 2) - diff(dom, curl_source[2], 0))[1]
    h_g = h_g + curl_source[1]
    return h_v, h_g


def _bc_pencil(x: torch.Tensor) -> torch.Tensor:
    return update_bc(x, axis=-1)


def solve_wavenumbers(
    sp: StepParams,
    step: int,
    v_1_lap_hat: torch.Tensor,
    vort_1: torch.Tensor,
    h_v: torch.Tensor,
    h_g: torch.Tensor,
    h_v_old: torch.Tensor,
    h_g_old: torch.Tensor,
    conv: torch.Tensor,
    conv_old: torch.Tensor,
    vel_hat: torch.Tensor,
    source_hat: Optional[torch.Tensor] = None,
) -> torch.Tensor:
    """All per-wavenumber solves of one RK substep, vectorised over (kx, kz).
    perform_single_rk_step_for_single_wavenumber, navier_stokes.py:1228-1548,
    mapped by get_new_vel_field_map navier_stokes.py:1558-1677."""
    dom, ops = sp.dom, sp.ops
    dt = sp.dt
    gamma, xi = RK_GAMMA[step], RK_XI[step]
    P = lambda f: f.permute(0, 2, 1)  # (Nx, Ny, Nz) -> pencils (Nx, Nz, Ny)
    s = str(step)
    Dm = dom.t_("Dpow1", np.linalg.matrix_power(dom.D, 1))
    Dv = lambda v: torch.matmul(v, Dm.T)

    # wall-normal velocity, p-part
    rhs_p = apply_k(dom, ops.rhs[step], P(v_1_lap_hat), "rhs" + s) + (dt * gamma) * P(h_v) + (dt * xi) * P(h_v_old)
    rhs_p = _bc_pencil(rhs_p)
    phi_new = apply_k(dom, ops.lhs_inv[step], rhs_p, "lhs_inv" + s)
    phi_new = _bc_pencil(phi_new)
    v_p = _bc_pencil(apply_k(dom, ops.poisson, phi_new, "poisson"))
    # a-part (navier_stokes.py:1321-1364)
    r_inh = apply_k_vec(dom, ops.rhs_inh[step], "rhs_inh" + s)
    phi_a = apply_k(dom, ops.lhs_inv_inh[step], r_inh, "lhs_inv_inh" + s)
    v_a = _bc_pencil(apply_k(dom, ops.poisson, phi_a, "poisson"))
    v_b = torch.flip(v_a, dims=(-1,))
    dvp, dva, dvb = Dv(v_p), Dv(v_a), Dv(v_b)
    M = torch.stack(
        [torch.stack([dva[..., 0], dvb[..., 0]], -1), torch.stack([dva[..., -1], dvb[..., -1]], -1)], -2
    )
    R = torch.stack([-dvp[..., 0], -dvp[..., -1]], -1).unsqueeze(-1)
    ab = torch.linalg.lstsq(M, R).solution.squeeze(-1)
    v_new = v_p + ab[..., 0:1] * v_a + ab[..., 1:2] * v_b
    v_new = _bc_pencil(v_new)
    # vorticity
    rhs_w = apply_k(dom, ops.rhs[step], P(vort_1), "rhs" + s) + (dt * gamma) * P(h_g) + (dt * xi) * P(h_g_old)
    rhs_w = _bc_pencil(rhs_w)
    vort_new = _bc_pencil(apply_k(dom, ops.lhs_inv[step], rhs_w, "lhs_inv" + s))
    # u, w from continuity (navier_stokes.py:1506-1527)
    kx = torch.from_numpy(dom.kx)[:, None, None]
    kz = torch.from_numpy(dom.kz)[None, :, None]
    j_kx = 1j * kx
    j_kz = 1j * kz
    minus_k2 = -(kx**2 + kz**2)
    # the (0,0) entry is replaced by the mean-flow solve below; a unit
    # denominator there keeps the value unused and the derivative finite
    minus_k2 = torch.where(minus_k2 == 0, torch.ones_like(minus_k2), minus_k2)
    v_y = Dv(v_new)
    u_new = (-j_kx * v_y + j_kz * vort_new) / minus_k2
    w_new = (-j_kz * v_y - j_kx * vort_new) / minus_k2
    # (0,0): mean-flow block solve (navier_stokes.py:1417-1504)
    n = dom.ny
    dpdx = sp.dPdx * ((dom.px * (2 * np.pi / dom.lx) ** 2) * (dom.pz * (2 * np.pi / dom.lz) ** 2)) ** 0.5
    v_hat00 = torch.cat([vel_hat[0][0, :, 0], vel_hat[2][0, :, 0]])
    ones = torch.ones(n, dtype=C128)
    # source_x_00, source_z_00 (navier_stokes.py:1461-1490): the same current
    # source enters the new and the old term
    if source_hat is None:
        sx00 = torch.zeros(n, dtype=C128)
        sz00 = torch.zeros(n, dtype=C128)
    else:
        sx00 = source_hat[0][0, :, 0]
        sz00 = source_hat[2][0, :, 0]
    c00, c20 = conv[0][0, :, 0], conv[2][0, :, 0]
    c00o, c20o = conv_old[0][0, :, 0], conv_old[2][0, :, 0]
    if sp.enforce_base:  # navier_stokes.py:1581-1590
        c00, c20, c00o, c20o = (torch.zeros_like(c00) for _ in range(4))
    n_new = torch.cat([-c00, -c20]) + torch.cat([-dpdx * ones + sx00, 0 * ones + sz00])
    n_old = torch.cat([-c00o, -c20o]) + torch.cat([-dpdx * ones + sx00, 0 * ones + sz00])
    lhs00 = dom.t_("lhs_inv_ns" + s, ops.lhs_inv_ns[step])
    rhs00 = dom.t_("rhs_ns" + s, ops.rhs_ns[step])
    v00 = lhs00 @ (rhs00 @ v_hat00 + (dt * gamma) * n_new + (dt * xi) * n_old)
    u_new = u_new.clone()
    w_new = w_new.clone()
    u_new[0, 0] = v00[:n]
    w_new[0, 0] = v00[n:]
    out = torch.stack([u_new, v_new, w_new]).permute(0, 1, 3, 2)
    return out


def rk_step(sp: StepParams, vel_hat: torch.Tensor, time_step: int) -> torch.Tensor:
    """One full RK3/CN step.  perform_runge_kutta_step, navier_stokes.py:1679-1775."""
    dom = sp.dom
    h_v_old = h_g_old = conv_old = None
    for step in range(3):
        h_v, h_g, vort, conv = sp.nonlinear(vel_hat, time_step)  # type: ignore[misc]
        source_hat = sp.source(time_step) if sp.source is not None else None
        if source_hat is not None:  # navier_stokes.py:1697-1704
            h_v, h_g = source_step_terms(dom, source_hat, h_v, h_g)
        if step == 0:
            h_v_old = torch.zeros_like(h_v)
            h_g_old = torch.zeros_like(h_g)
            conv_old = torch.zeros_like(conv)
        v1 = vel_hat[1]
        v_1_lap = diff(dom, v1, 0, 2) + diff(dom, v1, 1, 2) + diff(dom, v1, 2, 2)
        new = solve_wavenumbers(
            sp, step, v_1_lap, vort[1], h_v, h_g, h_v_old, h_g_old, conv, conv_old, vel_hat, source_hat
        )
        h_v_old, h_g_old, conv_old = h_v, h_g, conv
        vel_hat = torch.stack([update_bc(new[i]) for i in range(3)])
    return vel_hat


# ---------------------------------------------------------------------------
# equations (navier_stokes.py:313-384, navier_stokes_perturbation.py:347-520,
# navier_stokes_perturbation_dual.py:379-589)
# ---------------------------------------------------------------------------


def base_flow_hat(dom: Dom, u_max: float = 1.0) -> torch.Tensor:
    """Laminar base flow U = u_max (1 - y^2) as a spectral vector field.
    navier_stokes_perturbation.py:351-380."""
    X0, X1, X2 = np.meshgrid(dom.x, dom.y, dom.z, indexing="ij")
    ub = torch.from_numpy(u_max * (1 - X1**2) + 0.0 * X0 * X2)
    zero = torch.from_numpy(0.0 * X0 * X1 * X2)
    return torch.stack([field_hat(dom, ub), field_hat(dom, zero), field_hat(dom, zero)])


def instationary_base_profile(
    y: np.ndarray, t: float, re: float, kappa: float = 1.0, accelerate: bool = True, n_max: int = 100
) -> np.ndarray:
    """Time-dependent base flow U(y, t) between walls moving with +-g_w(t)
    (navier_stokes_perturbation_instationary.py:39-80, profile_fn):
    g_w = 1 - exp(-kappa t) (accelerating) or exp(-kappa t), and
    U = g_w y + Re/6 g_w'(t) (y^3 - y) + sum_{n=1}^{n_max-1} [(e^{-kappa t} - e^{-a_n t})
    (-2 Re (-1)^n / (pi n)^3) kappa^2 s / (a_n - kappa) + e^{-a_n t} c_n] sin(n pi y)
    with a_n = (n pi)^2 / Re, s = -1 (accelerating) or +1, c_n =
    (-2 Re (-1)^n / (pi n)^3) g_w'(0); g_w' is the exact derivative (the
    reference's jax.grad)."""
    def g_w(tt):
        return 1.0 - math.exp(-kappa * tt) if accelerate else math.exp(-kappa * tt)

    def d_g_w(tt):
        return kappa * math.exp(-kappa * tt) if accelerate else -kappa * math.exp(-kappa * tt)

    out = g_w(t) * y + re / 6 * d_g_w(t) * (y**3 - y)
    coeff = -1.0 if accelerate else 1.0
    for n in range(1, n_max):
        pre = -2 * re * (-1) ** n / (math.pi * n) ** 3
        c_1_n = pre * d_g_w(0.0)
        a_n = (math.pi * n) ** 2 / re
        integ = kappa**2 * coeff / (a_n - kappa)
        out = out + (math.exp(-kappa * t) - math.exp(-a_n * t)) * (pre * integ * np.sin(n * math.pi * y))
        out = out + math.exp(-a_n * t) * c_1_n * np.sin(n * math.pi * y)
    return out


def make_perturbation_instationary(
    dom: Dom, re: float, dt: float, kappa: float = 1.0, accelerate: bool = True, n_max: int = 100
) -> StepParams:
    """NavierStokesVelVortPerturbationInstationary
    (navier_stokes_perturbation_instationary.py:29-108): the nonlinear
    perturbation step about U(y, t) at t = time_step * dt (the same for the
    three substeps, navier_stokes.py:1695), with coupling_term = False
    (h_g gains d_z hat(U' v), navier_stokes_perturbation.py:214-224) and no
    pressure gradient."""
    ops = build_operators(dom, re, dt)
    sp = StepParams(dom, ops, dt, 0.0)
    X0, X1, X2 = np.meshgrid(dom.x, dom.y, dom.z, indexing="ij")
    zero = field_hat(dom, torch.from_numpy(0.0 * X0 * X1 * X2))

    def base(t: int) -> torch.Tensor:
        U = instationary_base_profile(dom.y, t * dt, re, kappa, accelerate, n_max)
        ub = torch.from_numpy(U[None, :, None] + 0.0 * X0 * X2)
        return torch.stack([field_hat(dom, ub), zero, zero])

    sp.nonlinear = lambda vh, t: nl_perturbation_skew(dom, vh, base(t), False, coupling_term=False)
    return sp


def make_dual_instationary(
    dom: Dom,
    re: float,
    dt: float,
    n_steps: int,
    history: Callable[[int], torch.Tensor],
    kappa: float = 1.0,
    accelerate: bool = True,
    n_max: int = 100,
) -> StepParams:
    """NavierStokesVelVortPerturbationInstationaryDual
    (navier_stokes_perturbation_instationary_dual.py:70-119): the nonlinear
    dual step (linearise = False, coupling_term = False: h_v gains
    (d_x^2 + d_z^2) hat(U' v_x), navier_stokes_perturbation_dual.py:269-278)
    of dual time step t about the forward base flow at time step n - t
    (vel_base_fn, :96-101) and the forward state history(t) = u at step n - t
    (navier_stokes_perturbation_dual.py:591-602); the forward's operators,
    explicit terms with -dt, no pressure gradient (:83-89)."""
    ops = build_operators(dom, re, dt)
    sp = StepParams(dom, ops, -dt, 0.0)
    X0, X1, X2 = np.meshgrid(dom.x, dom.y, dom.z, indexing="ij")
    zero = field_hat(dom, torch.from_numpy(0.0 * X0 * X1 * X2))

    def base(k: int) -> torch.Tensor:
        U = instationary_base_profile(dom.y, k * dt, re, kappa, accelerate, n_max)
        ub = torch.from_numpy(U[None, :, None] + 0.0 * X0 * X2)
        return torch.stack([field_hat(dom, ub), zero, zero])

    sp.nonlinear = lambda vh, t: nl_dual_skew(dom, vh, base(n_steps - t), history(t), False, coupling_term=False)
    return sp


def make_ns(dom: Dom, re: float, dt: float, dPdx: float, u_max_over_u_tau: float = 1.0) -> StepParams:
    re_tau = re / u_max_over_u_tau
    ops = build_operators(dom, re_tau, dt)
    sp = StepParams(dom, ops, dt, dPdx)
    sp.nonlinear = lambda vh, t: nl_ns_skew(dom, vh)
    return sp


def make_perturbation(
    dom: Dom,
    re: float,
    dt: float,
    linearise: bool,
    coupling_term: bool = True,
    u_max_over_u_tau: float = 1.0,
    ops: Optional[Operators] = None,
    base_hat: Optional[torch.Tensor] = None,
) -> StepParams:
    """velocity_base_hat defaults to the laminar profile
    (navier_stokes_perturbation.py:351-380); any base_hat may be given."""
    re_tau = re / u_max_over_u_tau
    ops = ops if ops is not None else build_operators(dom, re_tau, dt)
    base = base_flow_hat(dom, u_max_over_u_tau) if base_hat is None else base_hat
    sp = StepParams(dom, ops, dt, 0.0)
    sp.nonlinear = lambda vh, t: nl_perturbation_skew(dom, vh, base, linearise, coupling_term)
    sp.base_hat = base  # type: ignore[attr-defined]
    return sp


def make_dual(fwd: StepParams, history: Callable[[int], torch.Tensor], linearise: bool, coupling_term: bool = True) -> StepParams:
    """Dual equation: same operators as the forward (dt*L unchanged because
    both Re and dt change sign), explicit terms with -dt.
    navier_stokes_perturbation_dual.py:470-471, :542-589."""
    dom = fwd.dom
    base = fwd.base_hat  # type: ignore[attr-defined]
    sp = StepParams(dom, fwd.ops, -fwd.dt, 0.0)
    sp.nonlinear = lambda vh, t: nl_dual_skew(dom, vh, base, history(t), linearise, coupling_term)
    sp.base_hat = base  # type: ignore[attr-defined]
    return sp


def median_factor(n: int) -> int:
    """navier_stokes.py:1853-1861."""
    factors = reduce(list.__add__, ([i, n // i] for i in range(1, int(n**0.5) + 1) if n % i == 0))
    factors.sort()
    return factors[len(factors) // 2]


def number_of_steps(end_time: float, dt: float) -> int:
    """n = int(T/dt) (navier_stokes.py:1864-1871)."""
    return max(0, int(end_time / dt))


def run(sp: StepParams, vel_hat: torch.Tensor, n_steps: int, keep: bool = False, t0: int = 0):
    """Advance n_steps; optionally keep every state (write_entire_output)."""
    hist = [vel_hat] if keep else None
    for t in range(n_steps):
        vel_hat = rk_step(sp, vel_hat, t0 + t)
        if keep:
            hist.append(vel_hat)  # type: ignore[union-attr]
    return (vel_hat, hist) if keep else vel_hat


def laplacian(dom: Dom, fh: torch.Tensor) -> torch.Tensor:
    """Domain.laplacian (domain.py:671-673)."""
    return diff(dom, fh, 0, 2) + diff(dom, fh, 1, 2) + diff(dom, fh, 2, 2)


def field_2d(fh: torch.Tensor, axis: int) -> torch.Tensor:
    """FourierField.field_2d(direction) with wavenumber 0 (field.py:2912-2935):
    keep index 0 along ``axis`` (0: kx, 2: kz of a (Nx, Ny, Nz) field),
    zero the rest."""
    out = torch.zeros_like(fh)
    idx = [slice(None)] * fh.dim()
    idx[axis] = slice(0, 1)
    out[tuple(idx)] = fh[tuple(idx)]
    return out


def dual_initial_condition(uT: torch.Tensor, mode: str) -> torch.Tensor:
    """navier_stokes_perturbation_dual.py:903-935."""
    if mode == "gain":
        return -1 * uT
    if mode == "gain_3d":
        return -1 * (uT - torch.stack([field_2d(uT[i], 0) for i in range(3)]))
    if mode == "gain_no_00":
        return -1 * (uT - torch.stack([field_2d(field_2d(uT[i], 0), 2) for i in range(3)]))
    if mode == "dissipation":
        return 0 * uT
    raise ValueError(mode)


def objective(dom: Dom, mode: str, u0: torch.Tensor, uT: torch.Tensor, hist: Optional[List[torch.Tensor]], re_tau: float):
    """get_objective_fun (navier_stokes_perturbation_dual.py:941-957, 1058-1104):
    the gains E(u(T) - part)/E(u(0)); the dissipation average
    volume_integral(no_hat(sum_states sum_ij diff(u_i, j)**2)) / |Re_tau n_states|
    with the squares of the spectral coefficients, as there."""
    if mode == "gain":
        return energy(dom, uT) / energy(dom, u0)
    if mode == "gain_3d":
        return energy(dom, uT - torch.stack([field_2d(uT[i], 0) for i in range(3)])) / energy(dom, u0)
    if mode == "gain_no_00":
        return energy(dom, uT - torch.stack([field_2d(field_2d(uT[i], 0), 2) for i in range(3)])) / energy(dom, u0)
    assert hist is not None
    acc = torch.zeros_like(u0[0])
    for u in hist:
        for i in range(3):
            for j in range(3):
                acc = acc + diff(dom, u[i], j) ** 2
    acc = acc / abs(re_tau * len(hist))
    return volume_integral(dom, field_no_hat(dom, acc))


def dual_gradient(
    fwd: StepParams,
    u0_hat: torch.Tensor,
    n_steps: int,
    linearise: bool,
    checkpoint: Optional[int] = None,
    mode: str = "gain",
    re_tau: Optional[float] = None,
):
    """Adjoint gradient (u0 - v0)/E0 with the dual started from
    dual_initial_condition(u(T), mode) (v(T) = -u(T) for the gain).
    run_forward_calculation / solve_scan / get_grad,
    navier_stokes_perturbation_dual.py:690-1120.  With ``checkpoint`` set,
    only every checkpoint-th forward state is stored and each chunk is
    recomputed before its adjoint sweep (navier_stokes_perturbation_dual.py:730-757,
    :990-1035).  The dissipation mode adds the source
    -laplacian(u(t)) / |Re_dual T| at dual step t (get_source_term,
    navier_stokes_perturbation_dual.py:962-988; Re_dual = -re_tau,
    T = n_steps |dt|) and returns the dissipation average as its objective."""
    dom = fwd.dom
    if mode == "dissipation" and checkpoint is not None:
        raise ValueError("the oracle's dissipation mode keeps the whole history")
    if checkpoint is None:
        uT, hist = run(fwd, u0_hat, n_steps, keep=True)
        dual = make_dual(fwd, lambda t: hist[-1 - t], linearise)
        if mode == "dissipation":
            assert re_tau is not None
            scale = -1 / abs(-re_tau * (-n_steps * abs(fwd.dt)))
            dual.source = lambda t: scale * torch.stack([laplacian(dom, hist[-1 - t][i]) for i in range(3)])
        v = dual_initial_condition(uT, mode)
        for t in range(n_steps):
            v = rk_step(dual, v, t)
        e0 = energy(dom, u0_hat)
        obj = objective(dom, mode, u0_hat, uT, hist, re_tau if re_tau is not None else 1.0)
        return (u0_hat - v) / e0, obj, v
    else:
        inner = checkpoint
        outer = n_steps // inner
        ckpt = [u0_hat]
        u = u0_hat
        for o in range(outer):
            u = run(fwd, u, inner)
            ckpt.append(u)
        uT = u
        v = dual_initial_condition(uT, mode)
        cur: List[torch.Tensor] = []
        start = [0]

        def hist_fn(t: int) -> torch.Tensor:
            return cur[-1 - (t - start[0] * inner)]

        dual = make_dual(fwd, hist_fn, linearise)
        for o in range(outer):
            _, h = run(fwd, ckpt[-1 - (o + 1)], inner, keep=True)
            cur[:] = h
            start[0] = o
            for t in range(o * inner, (o + 1) * inner):
                v = rk_step(dual, v, t)
    e0 = energy(dom, u0_hat)
    return (u0_hat - v) / e0, objective(dom, mode, u0_hat, uT, None, 1.0), v


# ---------------------------------------------------------------------------
# gradient parameterisation (navier_stokes.py:1027-1203)
# ---------------------------------------------------------------------------


def clamped_cheb_basis(y: np.ndarray, M: Optional[int] = None) -> np.ndarray:
    """Columns phi_n(y_j) = T_{n+4} - a T_{n+2} - b T_n, a = 2(n+2)/(n+1),
    b = -(n+3)/(n+1), n < M = N-4: Chebyshev combinations with
    phi(+-1) = phi'(+-1) = 0.  smooth_and_enforce_bc_vel_y.phi_n,
    navier_stokes.py:1033-1048, evaluated at the grid as in :1054-1058."""
    N = len(y)
    M_ = N - 4 if M is None else M
    A = np.zeros((N, M_))
    for i in range(M_):
        a = +(2 * (i + 2)) / (i + 1)
        b = -(i + 3) / (i + 1)
        arr = np.eye(4 + i + 1)[4 + i] - a * np.eye(4 + i + 1)[2 + i] - b * np.eye(4 + i + 1)[i]
        A[:, i] = np.polynomial.chebyshev.chebval(y, arr)
    return A


def smooth_and_enforce_bc_vel_y(dom: Dom, v: torch.Tensor, M: Optional[int] = None) -> torch.Tensor:
    """Least-squares projection of the y-profiles v[..., :] (last axis y)
    onto the clamped basis.  navier_stokes.py:1027-1061 (jnp.linalg.lstsq)."""
    A = torch.from_numpy(clamped_cheb_basis(dom.y, M)).to(v.dtype)
    flat = v.reshape(-1, v.shape[-1]).T  # (N, batch)
    coeffs = torch.linalg.lstsq(A, flat).solution
    return (A @ coeffs).T.reshape(v.shape)


def vort_yvel_to_vel(
    dom: Dom,
    vort: torch.Tensor,
    vel_y: torch.Tensor,
    vel_x_00: torch.Tensor,
    vel_z_00: torch.Tensor,
    two_d: bool = False,
) -> torch.Tensor:
    """(omega_y, v, u_00, w_00) -> (u, v, w), arrays (Nx, Ny, Nz): v smoothed
    with the clamped projection, u and w from continuity and the definition
    of omega_y per (kx, kz); the (0,0) pencil takes u_00, w_00 and the
    smoothed v_00.  navier_stokes.py:1063-1203 (rk_00 :1087-1101,
    rk_not_00 :1103-1130)."""
    if two_d:
        vort = torch.zeros_like(vel_y)
        vel_z_00 = torch.zeros_like(vel_x_00)
    vs = smooth_and_enforce_bc_vel_y(dom, vel_y.permute(0, 2, 1)).permute(0, 2, 1)
    Dm = dom.t_("Dpow1", dom.D)
    vy = torch.matmul(Dm, vs)  # diff along y: (Ny,Ny) @ (Nx, Ny, Nz)
    kx = torch.from_numpy(dom.kx)[:, None, None]
    kz = torch.from_numpy(dom.kz)[None, None, :]
    mk2 = -(kx**2 + kz**2)
    mk2 = torch.where(mk2 == 0, torch.ones_like(mk2), mk2)
    u = (-1j * kx * vy + 1j * kz * vort) / mk2
    w = (-1j * kz * vy - 1j * kx * vort) / mk2
    if two_d:
        w = torch.zeros_like(u)
    u = u.clone()
    w = w.clone()
    u[0, :, 0] = vel_x_00.to(u.dtype)
    w[0, :, 0] = vel_z_00.to(w.dtype)
    return torch.stack([u, vs.to(u.dtype), w])
