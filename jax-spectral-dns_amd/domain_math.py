"""Host-side grid and operator precompute for the channel-flow step.

Everything here runs once per (shape, Re, dt) on the host in NumPy FP64 and
is handed to the HIP library as plain arrays.  The formulas are those of the
reference's Domain and operator assembly:

* Gauss-Lobatto grid and collocation derivative matrix D
  (reference domain.py:130-163),
* the odd-mode / physical-size rule (domain.py:217-258, 143-152),
* stored wavenumbers [0..K, -K..-1] 2 pi / L (domain.py:796-827),
* RK3/CN operator tables: R_s = hom_dir(I + alpha_s dt L), the inverse of
  hom_dir(I - beta_s dt L), the Poisson inverse
  inv(-(kx^2+kz^2) eye_bc + hom_dir(D^2)), the mean-flow 2Ny block operators
  (navier_stokes.py:920-1022, domain.py:857-889),
* the y quadrature of the definite integral (domain.py:577-628).

The tables depend on kx^2 + kz^2 only, so they are stored once per distinct
(|kx|, |kz|) pair, in real FP64 (the reference stores the same values per
signed pair as complex128).  Two derived per-(|k|, substep) quantities are
precomputed as well: the homogeneous-Neumann correction solution
v_a = BC(P (L_inh^-1 e_0)) and the inverse of the 2x2 influence matrix built
from D v_a and D flip(v_a) at the walls (navier_stokes.py:1321-1389).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

RK_ALPHA = (29 / 96, -3 / 40, 1 / 6)
RK_BETA = (37 / 160, 5 / 24, 1 / 6)
RK_GAMMA = (8 / 15, 5 / 12, 3 / 4)
RK_XI = (0.0, -17 / 60, -5 / 12)

# radices the HIP FFT kernel implements
FFT_RADICES = (8, 7, 5, 4, 3, 2)


def modes_and_physical(n: int, aliasing: float = 1.5) -> Tuple[int, int]:
    """Stored odd mode count and the reference's physical grid size for a
    periodic direction (domain.py:217-258 with get_fourier_grid dropping the
    last point of an odd argument, domain.py:143-152)."""
    m = int(n) if int(n) % 2 != 0 else int(n) + 1
    c = math.ceil(m * aliasing)
    arg = c if c % 2 != 0 else c + 1
    return m, (arg - 1 if arg % 2 != 0 else arg)


def factorize(n: int) -> List[int]:
    """Radix list (innermost first) using the kernel's radices, or [] if n has
    another prime factor."""
    out: List[int] = []
    rest = n
    for r in FFT_RADICES:
        while rest % r == 0 and rest > 1:
            out.append(r)
            rest //= r
    return out if rest == 1 else []


def transform_size(m: int, reference_p: int, direction: str = "x") -> int:
    """Transform grid P' >= 3K+1 (K = m // 2).  Any such grid computes the same
    |k| <= K spectrum of a product of two fields with |k| <= K as the
    reference's grid, up to rounding.  In x the intermediate fields have P'
    rows in HBM, so the smallest size with cheap radices (2, 3, 4, 5, 8) is
    taken; the z transform lives entirely in LDS, so a power of two (the
    cheapest butterflies) is taken there."""
    k = m // 2
    p = max(3 * k + 1, 2)
    if direction == "z":
        q = 1
        while q < p:
            q *= 2
        return q
    while True:
        f = factorize(p)
        if f and 7 not in f:
            return p
        p += 1


def cheb_grid(n: int) -> np.ndarray:
    return np.array([np.cos(np.pi / (n - 1) * i) for i in np.arange(n)])


def cheb_diff(xs: np.ndarray) -> np.ndarray:
    n = len(xs)
    c = np.block([2.0, np.ones((1, n - 2)) * 1.0, 2.0]) * (-1) ** np.arange(0, n)
    xm = np.repeat(xs, n).reshape(n, n)
    dx = xm - xm.T
    d_ = ((1 / c).T @ c).T / (dx + np.eye(n))
    return d_ - np.diag(sum(d_.T))


def wavenumbers(m: int, length: float) -> np.ndarray:
    inp = np.linspace(0, m - 1, int(m))
    return (np.block([inp[m // 2 :], inp[: m // 2]]) - m // 2) * (2 * np.pi) / length


def hom_dir(mat: np.ndarray) -> np.ndarray:
    out = mat.copy()
    out[0, :] = 0.0
    out[:, 0] = 0.0
    out[-1, :] = 0.0
    out[:, -1] = 0.0
    out[0, 0] = 1.0
    out[-1, -1] = 1.0
    return out


@dataclass
class Grid:
    nxm: int
    ny: int
    nzm: int
    px: int  # reference physical sizes (they fix the spectral scaling)
    pz: int
    lx: float
    lz: float
    ppx: int  # transform grid used by the kernels
    ppz: int
    y: np.ndarray
    D: np.ndarray
    kx: np.ndarray
    kz: np.ndarray
    scaling: float
    y_weights: np.ndarray
    cache: Dict[str, object] = field(default_factory=dict)

    @property
    def D2(self) -> np.ndarray:
        return np.linalg.matrix_power(self.D, 2)

    @property
    def ka(self) -> int:
        return self.nxm // 2 + 1

    @property
    def kb(self) -> int:
        return self.nzm // 2 + 1


def make_grid(
    shape: Tuple[int, int, int],
    lx: float = 2 * np.pi,
    lz: float = 2 * np.pi,
    transform_grid: Tuple[int, int] = None,  # type: ignore[assignment]
) -> Grid:
    nxm, px = modes_and_physical(shape[0])
    nzm, pz = modes_and_physical(shape[2])
    ny = int(shape[1])
    if nxm % 4 != 1 or nzm % 4 != 1:
        # the reference pads ceil(m/2) zeros but its grid has one point less
        # when m = 3 (mod 4); that case is inconsistent upstream
        raise ValueError("stored mode counts must be 1 mod 4 (got %d, %d)" % (nxm, nzm))
    if transform_grid is None:
        ppx, ppz = transform_size(nxm, px, "x"), transform_size(nzm, pz, "z")
    else:
        ppx, ppz = transform_grid
    y = cheb_grid(ny)
    D = cheb_diff(y)
    scaling = 1.0
    for length in (lx, lz):
        scaling *= length / (2 * np.pi)
    # y quadrature of the definite integral (integrate with bc_right = 0)
    mat = D.copy()
    mat[0, :] = 0.0
    mat[:, 0] = 0.0
    mat[0, 0] = 1.0
    inv = np.linalg.inv(mat)
    w = inv[0, :] - inv[ny - 1, :]
    w[0] = 0.0
    return Grid(nxm, ny, nzm, px, pz, lx, lz, ppx, ppz, y, D, wavenumbers(nxm, lx), wavenumbers(nzm, lz), scaling, w)


@dataclass
class OperatorTables:
    lhs_inv: np.ndarray  # (3, Ka, Kb, Ny, Ny) row-major
    poisson: np.ndarray  # (Ka, Kb, Ny, Ny)
    v_a: np.ndarray  # (3, Ka, Kb, Ny)
    infl_inv: np.ndarray  # (3, Ka, Kb, 4)  inverse 2x2 influence matrix, row-major
    rhs_base: np.ndarray  # (3, Ny, Ny) hom_dir(I + alpha_s dt D2 / Re)
    rhs_ns: np.ndarray  # (3, 2Ny, 2Ny)
    lhs_inv_ns: np.ndarray  # (3, 2Ny, 2Ny)
    y_eig: np.ndarray  # interior_eigen(D^2), flat
    full: bool = True  # False: lhs_inv / poisson hold only the |k| entries of check_indices


def parity_blocks(A: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Even and odd blocks of the interior rows/columns of a centro-symmetric
    matrix (A[N-1-i][N-1-j] = A[i][j], after symmetrising (A + J A J)/2).

    With e_l = (x_l + x_l')/2, o_l = (x_l - x_l')/2 (l' = N-1-l) on the
    interior rows l = 1..N-2, A maps e -> e by the block
    B_e[h][g] = A[h+1][g+1] + A[h+1][N-2-g] (the centre column once) and
    o -> o by B_o[h][g] = A[h+1][g+1] - A[h+1][N-2-g]; half sizes
    nE = N//2 - 1 + N%2, nO = N//2 - 1.
    """
    n = A.shape[0]
    S = 0.5 * (A + A[::-1, ::-1])
    ne, no = n // 2 - 1 + (n & 1), n // 2 - 1
    rows = np.arange(1, ne + 1)
    cols = np.arange(1, ne + 1)
    mirror = n - 1 - cols
    Be = S[np.ix_(rows, cols)] + np.where(cols == mirror, 0.0, 1.0)[None, :] * S[np.ix_(rows, mirror)]
    ro = np.arange(1, no + 1)
    Bo = S[np.ix_(ro, ro)] - S[np.ix_(ro, n - 1 - ro)]
    return Be, Bo


def _refined_eig(B: np.ndarray, sweeps: int = 3) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Real eigen-decomposition B = E diag(lam) E^-1 of a block of the interior
    Chebyshev D^2 (real, distinct, negative eigenvalues).  LAPACK's pairs are
    refined by Newton steps on the bordered system
    [[B - lam I, -e], [e^T, 0]] [de; dlam] = [-(B e - lam e); 0] with the
    residual in extended precision, and E^-1 by Newton-Schulz steps
    X <- X + X (I - E X), so that E, E^-1 and lam are accurate to rounding."""
    n = B.shape[0]
    lam, E = np.linalg.eig(B)
    if np.abs(lam.imag).max() > 0 or np.abs(E.imag).max() > 0:
        raise ValueError("interior D^2 has complex eigenvalues")
    lam, E = lam.real, E.real
    order = np.argsort(-lam)
    lam, E = lam[order], E[:, order]
    Bl = B.astype(np.longdouble)
    El = E.astype(np.longdouble)
    lal = lam.astype(np.longdouble)
    M = np.zeros((n + 1, n + 1))
    for j in range(n):
        e = El[:, j] / np.sqrt((El[:, j] ** 2).sum())
        for _ in range(sweeps):
            r = Bl @ e - lal[j] * e
            M[:n, :n] = B - float(lal[j]) * np.eye(n)
            M[:n, n] = -e.astype(np.float64)
            M[n, :n] = e.astype(np.float64)
            M[n, n] = 0.0
            rhs = np.concatenate([-r.astype(np.float64), [0.0]])
            d = np.linalg.solve(M, rhs)
            e = e + d[:n].astype(np.longdouble)
            lal[j] = lal[j] + np.longdouble(d[n])
        El[:, j] = e
    E = El.astype(np.float64)
    X = np.linalg.inv(E).astype(np.longdouble)
    El = E.astype(np.longdouble)
    eye = np.eye(n, dtype=np.longdouble)
    for _ in range(2):
        X = X + X @ (eye - El @ X)
    return E, X.astype(np.float64), lal.astype(np.float64)


def interior_eigen(D2: np.ndarray) -> np.ndarray:
    """Eigen-decompositions of the even and odd interior blocks of D^2
    (parity_blocks).  The interior block of every implicit y operator of the
    step is a function of it: the Helmholtz inverse
    (hom_dir(I - beta_s dt L(k)))^-1 restricted to the interior is
    E diag(1 / (1 + c k^2 - c lam)) E^-1 with c = beta_s dt / Re, and the
    Poisson inverse E diag(1 / (lam - k^2)) E^-1 (navier_stokes.py:920-1022),
    so one shared pair (E, E^-1) per parity replaces the per-|k| inverse
    tables on the device.  Flat layout: E_e (nE x nE), E_o (nO x nO),
    E_e^-1, E_o^-1 (row-major), lam_e (nE), lam_o (nO)."""
    Be, Bo = parity_blocks(D2)
    Ee, Xe, le = _refined_eig(Be)
    Eo, Xo, lo = _refined_eig(Bo)
    return np.concatenate([Ee.ravel(), Eo.ravel(), Xe.ravel(), Xo.ravel(), le, lo])


def check_indices(nk: int) -> np.ndarray:
    """The |k| table entries at which plan creation checks the eigen form of
    the solve against the inverse tables (every 13th and the last)."""
    k = np.arange(nk)
    return k[(k % 13 == 0) | (k == nk - 1)]


def build_operators(g: Grid, re_tau: float, dt: float, full: bool = True) -> OperatorTables:
    """The per-|k| operator tables of navier_stokes.py:920-1022 and
    domain.py:857-889.  full = False (the eigen-form solve): the Helmholtz
    and Poisson inverses are formed only at check_indices (the plan's
    cross-check of its eigen factorisation), the other entries stay zero,
    and the wall-correction solution v_a is solved in the eigen coordinates
    of D^2's even / odd interior halves, where both of its implicit problems
    are diagonal (the same values to rounding, O(Ny^2) per |k| instead of
    two Ny x Ny inversions)."""
    if not full:
        return _build_operators_sampled(g, re_tau, dt)
    n = g.ny
    ka, kb = g.ka, g.kb
    D2 = np.linalg.matrix_power(g.D, 2)
    Ly = 1 / re_tau * D2
    I = np.eye(n)
    kxa = g.kx[:ka]
    kzb = g.kz[:kb]
    k2 = (kxa[:, None] ** 2 + kzb[None, :] ** 2).reshape(-1)
    nk = ka * kb
    # L(k) = Ly + I * (-(kx^2+kz^2)) / Re, stacked over |k|
    L = Ly[None, :, :] + I[None, :, :] * (-(k2))[:, None, None] / re_tau
    lhs_inv = np.empty((3, nk, n, n))
    v_a = np.empty((3, nk, n))
    infl = np.empty((3, nk, 4))
    rhs_base = np.empty((3, n, n))
    eye_bc = np.eye(n)
    eye_bc[0, 0] = 0.0
    eye_bc[-1, -1] = 0.0
    D2h = hom_dir(D2)
    poisson = np.linalg.inv((-(k2))[:, None, None] * eye_bc[None] + D2h[None])
    for s in range(3):
        A = I[None] - RK_BETA[s] * dt * L
        lhs = A.copy()
        lhs[:, 0, :] = 0.0
        lhs[:, :, 0] = 0.0
        lhs[:, -1, :] = 0.0
        lhs[:, :, -1] = 0.0
        lhs[:, 0, 0] = 1.0
        lhs[:, -1, -1] = 1.0
        lhs_inv[s] = np.linalg.inv(lhs)
        inh = A.copy()
        inh[:, 0, :] = 0.0
        inh[:, -1, :] = 0.0
        inh[:, 0, 0] = 1.0
        inh[:, -1, -1] = 1.0
        phi_a = np.linalg.inv(inh)[:, :, 0]  # inv(L_inh) @ [1, 0, ..., 0]
        va = np.einsum("kij,kj->ki", poisson, phi_a)
        va[:, 0] = 0.0
        va[:, -1] = 0.0
        vb = va[:, ::-1]
        dva = va @ g.D.T
        dvb = vb @ g.D.T
        M = np.stack([np.stack([dva[:, 0], dvb[:, 0]], -1), np.stack([dva[:, -1], dvb[:, -1]], -1)], -2)
        infl[s] = np.linalg.inv(M).reshape(nk, 4)
        v_a[s] = va
        rhs_base[s] = hom_dir(I + RK_ALPHA[s] * dt * Ly)
    Z = np.zeros((n, n))
    L_ns = 1 / re_tau * np.block([[D2h, Z], [Z, D2h]])
    I2 = np.eye(2 * n)
    rhs_ns = np.stack([hom_dir(I2 + RK_ALPHA[s] * dt * L_ns) for s in range(3)])
    lhs_inv_ns = np.stack([np.linalg.inv(hom_dir(I2 - RK_BETA[s] * dt * L_ns)) for s in range(3)])
    return OperatorTables(
        lhs_inv=lhs_inv.reshape(3, ka, kb, n, n),
        poisson=poisson.reshape(ka, kb, n, n),
        v_a=v_a.reshape(3, ka, kb, n),
        infl_inv=infl.reshape(3, ka, kb, 4),
        rhs_base=rhs_base,
        rhs_ns=rhs_ns,
        lhs_inv_ns=lhs_inv_ns,
        y_eig=interior_eigen(D2),
    )


def _build_operators_sampled(g: Grid, re_tau: float, dt: float) -> OperatorTables:
    n = g.ny
    ka, kb = g.ka, g.kb
    D2 = np.linalg.matrix_power(g.D, 2)
    Ly = 1 / re_tau * D2
    I = np.eye(n)
    kxa = g.kx[:ka]
    kzb = g.kz[:kb]
    k2 = (kxa[:, None] ** 2 + kzb[None, :] ** 2).reshape(-1)
    nk = ka * kb
    ck = check_indices(nk)
    L = Ly[None, :, :] + I[None, :, :] * (-(k2))[:, None, None] / re_tau
    lhs_inv = np.zeros((3, nk, n, n))  # untouched pages: no memory beyond the checked entries
    poisson = np.zeros((nk, n, n))
    v_a = np.empty((3, nk, n))
    infl = np.empty((3, nk, 4))
    rhs_base = np.empty((3, n, n))
    eye_bc = np.eye(n)
    eye_bc[0, 0] = 0.0
    eye_bc[-1, -1] = 0.0
    D2h = hom_dir(D2)
    P = (-(k2))[:, None, None] * eye_bc[None] + D2h[None]
    poisson[ck] = np.linalg.inv(P[ck])
    y_eig = interior_eigen(D2)
    Ee, Eo, Xe, Xo, le, lo = _split_eigen(y_eig, n)
    ne, no = len(le), len(lo)
    # eigen coordinates of the interior of D^2's wall column (the right-hand
    # side of the inhomogeneous Helmholtz problem, times c)
    b0 = D2[1 : n - 1, 0]
    h = np.arange(no)
    be = 0.5 * (b0[:ne] + b0[::-1][:ne])
    bo = 0.5 * (b0[h] - b0[n - 3 - h])
    ze, zo = Xe @ be, Xo @ bo
    for s in range(3):
        A = I[None] - RK_BETA[s] * dt * L
        lhs = A[ck].copy()
        lhs[:, 0, :] = 0.0
        lhs[:, :, 0] = 0.0
        lhs[:, -1, :] = 0.0
        lhs[:, :, -1] = 0.0
        lhs[:, 0, 0] = 1.0
        lhs[:, -1, -1] = 1.0
        lhs_inv[s][ck] = np.linalg.inv(lhs)
        # phi_a = inv(L_inh) e_0: phi_0 = 1, phi_{N-1} = 0 and on the interior
        # ((1 + c k^2) I - c D^2) phi = c D^2[:, 0]; v_a = (D^2 - k^2)^-1 phi on
        # the interior: both diagonal in the eigen coordinates of D^2's halves
        c = RK_BETA[s] * dt / re_tau
        de = ((1.0 + c * k2)[None, :] - c * le[:, None]) * (le[:, None] - k2[None, :])
        do = ((1.0 + c * k2)[None, :] - c * lo[:, None]) * (lo[:, None] - k2[None, :])
        ve = Ee @ (c * ze[:, None] / de)  # (ne, nk)
        vo = Eo @ (c * zo[:, None] / do)  # (no, nk)
        vint = np.empty((n - 2, nk))
        vint[:ne] = ve
        vint[h] += vo
        vint[n - 3 - h] = ve[h] - vo
        va = np.zeros((nk, n))
        va[:, 1 : n - 1] = vint.T
        vb = va[:, ::-1]
        dva = va @ g.D.T
        dvb = vb @ g.D.T
        M = np.stack([np.stack([dva[:, 0], dvb[:, 0]], -1), np.stack([dva[:, -1], dvb[:, -1]], -1)], -2)
        infl[s] = np.linalg.inv(M).reshape(nk, 4)
        v_a[s] = va
        rhs_base[s] = hom_dir(I + RK_ALPHA[s] * dt * Ly)
    Z = np.zeros((n, n))
    L_ns = 1 / re_tau * np.block([[D2h, Z], [Z, D2h]])
    I2 = np.eye(2 * n)
    rhs_ns = np.stack([hom_dir(I2 + RK_ALPHA[s] * dt * L_ns) for s in range(3)])
    lhs_inv_ns = np.stack([np.linalg.inv(hom_dir(I2 - RK_BETA[s] * dt * L_ns)) for s in range(3)])
    return OperatorTables(
        lhs_inv=lhs_inv.reshape(3, ka, kb, n, n),
        poisson=poisson.reshape(ka, kb, n, n),
        v_a=v_a.reshape(3, ka, kb, n),
        infl_inv=infl.reshape(3, ka, kb, 4),
        rhs_base=rhs_base,
        rhs_ns=rhs_ns,
        lhs_inv_ns=lhs_inv_ns,
        y_eig=y_eig,
        full=False,
    )


def _split_eigen(y_eig: np.ndarray, n: int):
    """(E_e, E_o, E_e^-1, E_o^-1, lam_e, lam_o) of interior_eigen's flat layout."""
    ne, no = n // 2 - 1 + (n & 1), n // 2 - 1
    o = 0
    out = []
    for r in (ne, no, ne, no):
        out.append(y_eig[o : o + r * r].reshape(r, r))
        o += r * r
    out.append(y_eig[o : o + ne])
    out.append(y_eig[o + ne : o + ne + no])
    return out


def dit_tables(p: int, inverse: bool) -> Tuple[np.ndarray, np.ndarray, List[int]]:
    """In-place mixed-radix decimation-in-time FFT plan for length p:
    (input index for each LDS position, twiddle table w_p^m, radix list with
    the innermost radix first)."""
    radices = factorize(p)
    if not radices:
        raise ValueError("unsupported FFT length %d" % p)

    # recursive split, outermost radix = last entry: n = r * n1 + n2,
    # sub-FFT n2 occupies positions [n2 * Q, (n2 + 1) * Q)
    def order(length: int, rs: List[int]) -> List[int]:
        if length == 1:
            return [0]
        r = rs[-1]
        q = length // r
        sub = order(q, rs[:-1])
        out = []
        for n2 in range(r):
            out.extend(r * n1 + n2 for n1 in sub)
        return out

    perm = np.array(order(p, radices), dtype=np.int32)
    sign = 1.0 if inverse else -1.0
    # extended precision, then rounded: symmetric, correctly rounded twiddles
    pi_l = np.longdouble("3.14159265358979323846264338327950288")
    m = np.arange(p)
    ang = 2 * pi_l * m.astype(np.longdouble) / p
    tw = (np.cos(ang).astype(np.float64) + sign * 1j * np.sin(ang).astype(np.float64))
    return perm, tw, radices
