"""Device plan: operator tables on the GPU plus the step's workspaces.

``StepPlan`` owns one ``jsd_plan`` of libjsd.so.  State tensors are torch
complex128 CUDA tensors of shape (3, Nx_modes, Ny, Nz_modes) (the reference's
layout, navier_stokes.py:1778-1813); they are passed to the library as raw
device pointers together with torch's current HIP stream.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib
from .domain_math import Grid, OperatorTables, build_operators, make_grid


def _ptr(a: np.ndarray) -> "ctypes._Pointer":
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t: Optional[torch.Tensor]) -> Optional[ctypes.c_void_p]:
    if t is None:
        return None
    if not t.is_cuda or t.dtype != torch.complex128 or not t.is_contiguous():
        raise _lib.JsdError("expected a contiguous complex128 CUDA tensor")
    return ctypes.c_void_p(t.data_ptr())


class StepPlan:
    """A plan for one (grid, Re, dt, kind) combination."""

    def __init__(
        self,
        grid: Grid,
        re_tau: float,
        dt: float,
        kind: int,
        linearise: bool = False,
        u_max: float = 1.0,
        ops: Optional[OperatorTables] = None,
        explicit_dt: Optional[float] = None,
        plan_re: Optional[float] = None,
        adjoint: bool = False,
        eigen_solve: bool = True,
        coupling_term: bool = True,
        base_profile: Optional[np.ndarray] = None,
    ):
        if not torch.cuda.is_available():
            raise _lib.JsdError("no GPU visible: the channel-flow step runs only through the HIP library")
        lib = _lib.load()
        self.grid = grid
        self.kind = kind
        self.linearise = bool(linearise)
        self.re_tau = re_tau
        self.dt = dt
        # in eigen form the per-|k| inverse tables serve only the plan's
        # cross-check of the factorisation: build them at the checked |k| only
        self.ops = ops if ops is not None else build_operators(grid, re_tau, dt, full=not eigen_solve)
        if not eigen_solve and not self.ops.full:
            raise ValueError("the table-form solve needs full operator tables (build_operators(..., full=True))")
        g, o = grid, self.ops
        self._keep = []

        def arr(x: np.ndarray) -> np.ndarray:
            a = np.ascontiguousarray(x, dtype=np.float64)
            self._keep.append(a)
            return a

        # base-flow profile U(y) of perturbation / dual plans (default: laminar
        # u_max (1 - y^2)) and U'(y) = D U
        base_u = arr(u_max * (1 - g.y**2) if base_profile is None else np.asarray(base_profile, dtype=np.float64))
        if base_u.shape != (g.ny,):
            raise ValueError("base profile needs %d values" % g.ny)
        base_du = arr(g.D @ base_u)
        self.base_profile = base_u.copy()
        desc = _lib.PlanDesc()
        desc.nxm, desc.ny, desc.nzm = g.nxm, g.ny, g.nzm
        desc.px, desc.pz, desc.ppx, desc.ppz = g.px, g.pz, g.ppx, g.ppz
        desc.lx, desc.lz = g.lx, g.lz
        desc.re_tau = plan_re if plan_re is not None else re_tau
        desc.dt = explicit_dt if explicit_dt is not None else dt
        desc.kind = kind
        desc.linearise = int(self.linearise)
        desc.D = _ptr(arr(g.D))
        desc.kx = _ptr(arr(g.kx))
        desc.kz = _ptr(arr(g.kz))
        desc.lhs_inv = _ptr(arr(o.lhs_inv))
        desc.poisson = _ptr(arr(o.poisson))
        desc.v_a = _ptr(arr(o.v_a))
        desc.infl_inv = _ptr(arr(o.infl_inv))
        desc.rhs_base = _ptr(arr(o.rhs_base))
        desc.rhs_ns = _ptr(arr(o.rhs_ns))
        desc.lhs_inv_ns = _ptr(arr(o.lhs_inv_ns))
        desc.base_u = _ptr(base_u)
        desc.base_du = _ptr(base_du)
        desc.y_weights = _ptr(arr(g.y_weights))
        desc.adjoint = int(bool(adjoint))
        # eigen form of the implicit solve (shared E, E^-1 per parity half
        # instead of per-|k| inverse tables); False keeps the tables
        desc.y_eig = _ptr(arr(o.y_eig)) if eigen_solve else None
        # coupling_term = False: perturbation plans add d_z hat(U' v) to h_g
        # (navier_stokes_perturbation.py:214-224), dual plans (d_x^2 + d_z^2) hat(U' v_x)
        # to h_v (navier_stokes_perturbation_dual.py:269-278)
        desc.no_coupling_term = 0 if coupling_term else 1
        self.has_adjoint = bool(adjoint) or (kind == _lib.KIND_PERTURBATION and self.linearise)
        h = ctypes.c_void_p()
        _lib.check(lib.jsd_plan_create(ctypes.byref(desc), ctypes.byref(h)))
        self._h = h
        self._keep = []  # tables were copied to the device

    @property
    def adjoint_needs_state(self) -> bool:
        """The linearised step's transpose is independent of the state."""
        return not (self.kind == _lib.KIND_PERTURBATION and self.linearise)

    def vjp(self, grad: torch.Tensor, nsteps: int = 1) -> torch.Tensor:
        """In place: adjoint after nsteps steps -> adjoint before them
        (linearised perturbation plans)."""
        self._check_state(grad)
        _lib.check(_lib.load().jsd_step_vjp(self._h, _dev(grad), int(nsteps), _stream()))
        return grad

    def record_elems(self) -> int:
        """complex128 elements of one step record (step_record / vjp_recorded)."""
        n = ctypes.c_size_t(0)
        _lib.check(_lib.load().jsd_step_record_bytes(self._h, ctypes.byref(n)))
        return int(n.value) // 16

    def step_record(self, vel_hat: torch.Tensor, record: torch.Tensor, dPdx: float = 0.0) -> torch.Tensor:
        """One step of vel_hat in place, writing the step's record (a complex128
        CUDA tensor of record_elems() elements) for vjp_recorded."""
        self._check_state(vel_hat)
        self._check_record(record)
        _lib.check(_lib.load().jsd_step_record(self._h, _dev(vel_hat), _dev(record), float(dPdx), _stream()))
        return vel_hat

    def vjp_recorded(self, state: torch.Tensor, record: torch.Tensor, grad: torch.Tensor, dPdx: float = 0.0) -> torch.Tensor:
        """vjp_state about the step that step_record recorded from the start
        state `state`: no recomputation.  The record is consumed (its
        transforms are overwritten)."""
        self._check_state(state)
        self._check_state(grad)
        self._check_record(record)
        _lib.check(
            _lib.load().jsd_step_vjp_recorded(self._h, _dev(state), _dev(record), _dev(grad), float(dPdx), _stream())
        )
        return grad

    def _check_record(self, record: torch.Tensor) -> None:
        if record.dtype != torch.complex128 or not record.is_cuda or not record.is_contiguous():
            raise _lib.JsdError("a step record is a contiguous complex128 CUDA tensor")
        if record.numel() < self.record_elems():
            raise _lib.JsdError("step record holds %d elements, needs %d" % (record.numel(), self.record_elems()))

    def vjp_state(self, state: torch.Tensor, grad: torch.Tensor, dPdx: float = 0.0) -> torch.Tensor:
        """In place: adjoint after one step -> adjoint before it, the step
        taken from the forward state `state` (nonlinear plans built with
        adjoint=True; linearised plans ignore `state`)."""
        self._check_state(state)
        self._check_state(grad)
        _lib.check(_lib.load().jsd_step_vjp_state(self._h, _dev(state), _dev(grad), float(dPdx), _stream()))
        return grad

    @property
    def shape(self) -> Tuple[int, int, int, int]:
        return (3, self.grid.nxm, self.grid.ny, self.grid.nzm)

    def set_base_flow(self, base_u, base_du=None) -> None:
        """Replace the base profile U(y) (and U'(y), by default D U) of a
        perturbation plan before the next step.  Host arrays are copied
        before return; float64 tensors on the plan's GPU are copied on the
        current stream without waiting (they must stay alive until the
        stream has run the copies)."""
        if isinstance(base_u, torch.Tensor) and base_u.is_cuda:
            if not (isinstance(base_du, torch.Tensor) and base_du.is_cuda):
                raise ValueError("a device profile needs its derivative on the device too")
            for t in (base_u, base_du):
                if t.dtype != torch.float64 or t.shape != (self.grid.ny,) or not t.is_contiguous():
                    raise ValueError("device base profiles need %d contiguous float64 values" % self.grid.ny)
            _lib.check(
                _lib.load().jsd_set_base_flow_device(
                    self._h, ctypes.c_void_p(base_u.data_ptr()), ctypes.c_void_p(base_du.data_ptr()), _stream()
                )
            )
            return
        u = np.ascontiguousarray(base_u, dtype=np.float64)
        du = np.ascontiguousarray(self.grid.D @ u if base_du is None else base_du, dtype=np.float64)
        if u.shape != (self.grid.ny,) or du.shape != (self.grid.ny,):
            raise ValueError("base profiles need %d values" % self.grid.ny)
        _lib.check(_lib.load().jsd_set_base_flow(self._h, _ptr(u), _ptr(du), _stream()))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().jsd_plan_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass

    def _check_state(self, t: torch.Tensor) -> None:
        if tuple(t.shape) != self.shape:
            raise _lib.JsdError("state shape %s != %s" % (tuple(t.shape), self.shape))

    def step(self, vel_hat: torch.Tensor, dPdx: float = 0.0, nsteps: int = 1) -> torch.Tensor:
        """Advance vel_hat in place by nsteps steps."""
        self._check_state(vel_hat)
        _lib.check(_lib.load().jsd_step(self._h, _dev(vel_hat), float(dPdx), int(nsteps), _stream()))
        return vel_hat

    def step_source(
        self,
        vel_hat: torch.Tensor,
        dPdx: float = 0.0,
        source_hat: Optional[torch.Tensor] = None,
        enforce_base: bool = False,
        fwd_hat: Optional[torch.Tensor] = None,
    ) -> torch.Tensor:
        """One step (a dual step for dual plans, with the forward state
        fwd_hat) with the source term source_hat (3 components) and the
        enforce_base switch; in place."""
        self._check_state(vel_hat)
        if source_hat is not None:
            self._check_state(source_hat)
        if fwd_hat is not None:
            self._check_state(fwd_hat)
        _lib.check(
            _lib.load().jsd_step_source(
                self._h, _dev(vel_hat), _dev(fwd_hat), float(dPdx), _dev(source_hat), int(bool(enforce_base)), _stream()
            )
        )
        return vel_hat

    def substep_terms(
        self,
        vel_hat: torch.Tensor,
        s: int,
        h_v: torch.Tensor,
        h_g: torch.Tensor,
        conv_ns: torch.Tensor,
        dPdx: float = 0.0,
        source_hat: Optional[torch.Tensor] = None,
        enforce_base: bool = False,
    ) -> torch.Tensor:
        """RK substep s with externally supplied nonlinear terms (the
        nonlinear_update_fn plug-in); in place."""
        g = self.grid
        self._check_state(vel_hat)
        self._check_state(conv_ns)
        for t in (h_v, h_g):
            if tuple(t.shape) != (g.nxm, g.ny, g.nzm):
                raise _lib.JsdError("h_v / h_g shape %s != %s" % (tuple(t.shape), (g.nxm, g.ny, g.nzm)))
        if source_hat is not None:
            self._check_state(source_hat)
        h_v, h_g, conv_ns = h_v.contiguous(), h_g.contiguous(), conv_ns.contiguous()
        if source_hat is not None:
            source_hat = source_hat.contiguous()
        _lib.check(
            _lib.load().jsd_substep_terms(
                self._h, _dev(vel_hat), int(s), float(dPdx), _dev(h_v), _dev(h_g), _dev(conv_ns), _dev(source_hat),
                int(bool(enforce_base)), _stream(),
            )
        )
        return vel_hat

    def laplacian(self, f: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
        """scale * laplacian of a spectral field or a stack of them."""
        g = self.grid
        one = (g.nxm, g.ny, g.nzm)
        if tuple(f.shape) == one:
            ncomp = 1
        elif f.dim() == 4 and tuple(f.shape[1:]) == one:
            ncomp = int(f.shape[0])
        else:
            raise _lib.JsdError("laplacian input shape %s" % (tuple(f.shape),))
        f = f.contiguous()
        out = torch.empty_like(f)
        _lib.check(_lib.load().jsd_laplacian(self._h, _dev(f), _dev(out), ncomp, float(scale), _stream()))
        return out

    def dissipation_accumulate(self, vel_hat: torch.Tensor, acc: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
        """acc += scale * sum_ij (d_j u_i)^2 (complex squares of the spectral
        derivatives); in place on acc."""
        g = self.grid
        self._check_state(vel_hat)
        if tuple(acc.shape) != (g.nxm, g.ny, g.nzm):
            raise _lib.JsdError("accumulator shape %s" % (tuple(acc.shape),))
        _lib.check(_lib.load().jsd_dissipation_accumulate(self._h, _dev(vel_hat), _dev(acc), float(scale), _stream()))
        return acc

    def volume_integral(self, f_hat: torch.Tensor) -> float:
        """Volume integral of no_hat(f_hat) for one spectral component."""
        g = self.grid
        if tuple(f_hat.shape) != (g.nxm, g.ny, g.nzm):
            raise _lib.JsdError("field shape %s" % (tuple(f_hat.shape),))
        out = ctypes.c_double(0.0)
        _lib.check(_lib.load().jsd_volume_integral(self._h, _dev(f_hat.contiguous()), ctypes.byref(out), _stream()))
        return out.value

    def dual_step(self, adj_hat: torch.Tensor, fwd_hat: Optional[torch.Tensor]) -> torch.Tensor:
        self._check_state(adj_hat)
        if fwd_hat is not None:
            self._check_state(fwd_hat)
        _lib.check(_lib.load().jsd_dual_step(self._h, _dev(adj_hat), _dev(fwd_hat), _stream()))
        return adj_hat

    def nonlinear(self, vel_hat: torch.Tensor, fwd_hat: Optional[torch.Tensor] = None):
        self._check_state(vel_hat)
        if fwd_hat is not None:
            self._check_state(fwd_hat)
        elif self.kind == _lib.KIND_DUAL and not self.linearise:
            raise _lib.JsdError("a nonlinear dual plan needs the forward state")
        g = self.grid
        h_v = torch.empty((g.nxm, g.ny, g.nzm), dtype=torch.complex128, device=vel_hat.device)
        h_g = torch.empty_like(h_v)
        conv = torch.empty_like(vel_hat)
        _lib.check(
            _lib.load().jsd_nonlinear(self._h, _dev(vel_hat), _dev(fwd_hat), _dev(h_v), _dev(h_g), _dev(conv), _stream())
        )
        return h_v, h_g, conv

    def energy(self, vel_hat: torch.Tensor) -> float:
        self._check_state(vel_hat)
        out = ctypes.c_double(0.0)
        _lib.check(_lib.load().jsd_energy(self._h, _dev(vel_hat), ctypes.byref(out), _stream()))
        return out.value

    def ydiff(self, f: torch.Tensor, order: int = 1) -> torch.Tensor:
        """y-derivative of one field (Nx_modes, Ny, Nz_modes) or of a stack
        (k, Nx_modes, Ny, Nz_modes)."""
        g = self.grid
        one = (g.nxm, g.ny, g.nzm)
        if f.dim() == 3 and tuple(f.shape) == one:
            ncomp = 1
        elif f.dim() == 4 and tuple(f.shape[1:]) == one and f.shape[0] >= 1:
            ncomp = int(f.shape[0])
        else:
            raise _lib.JsdError("ydiff input shape %s: expected %s or (k,) + %s" % (tuple(f.shape), one, one))
        if order not in (1, 2):
            raise _lib.JsdError("ydiff order must be 1 or 2")
        f = f.contiguous()
        out = torch.empty_like(f)
        _lib.check(_lib.load().jsd_ydiff(self._h, _dev(f), _dev(out), ncomp, order, _stream()))
        return out


def make_plan(
    shape: Tuple[int, int, int],
    re: float,
    dt: float,
    kind: int,
    lx: float = 2 * np.pi,
    lz: float = 2 * np.pi,
    linearise: bool = False,
    u_max_over_u_tau: float = 1.0,
) -> StepPlan:
    grid = make_grid(shape, lx, lz)
    return StepPlan(grid, re / u_max_over_u_tau, dt, kind, linearise, u_max_over_u_tau)
