"""Equation base class: field history, hooks and the time loop
(reference equation.py:81-431)."""

from __future__ import annotations

import os
import time
from typing import Any, Callable, Dict, List, Optional

import numpy as np


def write_trajectory(trajectory: Any, filename: str = "trajectory") -> None:
    """<field_dir>/<filename> holding the stacked states as one gzip-9 dataset
    "trajectory" (equation.py:362-372, navier_stokes_perturbation_dual.py:877-889)."""
    from . import hdf5
    from .field import Field

    os.makedirs(Field.field_dir, exist_ok=True)
    data = trajectory.detach().cpu().numpy() if hasattr(trajectory, "detach") else np.asarray(trajectory)
    hdf5.write(Field.field_dir + "/" + filename, {"trajectory": hdf5.Dataset(data)}, compression=9)


class Equation:
    name = "equation"
    write_intermediate_output = False
    write_entire_output = False
    verbosity_level: int = 1

    def __init__(self, domain: Any, *fields: Any, **params: Any):
        self.dt: float = params.get("dt", 1e-2)
        self.domain = domain
        self.fields: Dict[str, List[Any]] = {}
        self.time_step: int = 0
        self.time: float = 0.0
        self.before_time_step_fn: Optional[Callable] = None
        self.after_time_step_fn: Optional[Callable] = None
        self.post_process_fn: Optional[Callable] = None
        self.end_time = params.get("end_time", 0.0)
        self.activate_jit_ = True
        for f in fields:
            self.fields[f.get_name()] = [f]

    @classmethod
    def find_suitable_dt(
        cls,
        domain: Any,
        max_cfl: float = 0.7,
        U_max: tuple = (1.0, 1e-10, 1e-10),
        end_time: Optional[float] = None,
        safety_factor: float = 0.9,
    ) -> float:
        """CFL-limited time step (equation.py:114-162): the smallest
        max_cfl * (extent / cells) / U_max over the directions times the
        safety factor; with end_time, the number of steps is raised until
        its median-factor split into outer x inner steps is balanced
        (sqrt(n) - outer <= outer) and dt = end_time / n."""
        from .navier_stokes import median_factor

        dT = [
            abs(max_cfl * (domain.get_extent(i) / domain.number_of_cells(i)) / U_max[i])
            for i in domain.all_dimensions()
        ]
        dt = safety_factor * min(dT)
        if end_time is None:
            return dt
        n = max(0, int(end_time / dt))

        def bad(n: int) -> bool:
            inner = median_factor(n)
            outer = n // inner
            return abs(np.sqrt(n)) - outer > outer

        while bad(n):
            n += 1
        return end_time / n

    def get_dt(self) -> float:
        return self.dt

    def get_domain(self) -> Any:
        return self.domain

    def activate_jit(self) -> None:
        self.activate_jit_ = True

    def deactivate_jit(self) -> None:
        self.activate_jit_ = False

    # field history (equation.py:173-235)
    def get_field(self, name: str, index: int) -> Any:
        try:
            return self.fields[name][index]
        except KeyError:
            raise KeyError("Expected field named " + name + " in " + self.name + ".")

    def get_fields(self, name: str) -> List[Any]:
        return self.fields[name]

    def get_initial_field(self, name: str) -> Any:
        return self.get_field(name, 0)

    def get_latest_field(self, name: str) -> Any:
        return self.get_field(name, -1)

    def get_number_of_fields(self, name: str) -> int:
        return len(self.fields.get(name, []))

    def set_initial_field(self, name: str, field: Any) -> None:
        self.fields[name] = [field]

    def append_field(self, name: str, field: Any, in_place: bool = True) -> None:
        lst = self.fields.setdefault(name, [])
        if in_place and lst:
            lst[-1] = field
        else:
            lst.append(field)

    def add_field(self, name: str, field: Any = None) -> None:
        assert name not in self.fields, "Field " + name + " already exists!"
        self.fields[name] = [] if field is None else [field]

    def clear_field(self, name: str) -> None:
        self.fields[name] = []

    # time loop (equation.py:298-427)
    def done(self) -> bool:
        return self.time >= self.end_time - 1e-12 * max(1.0, abs(self.end_time))

    def before_time_step(self) -> None:
        if self.before_time_step_fn is not None:
            self.before_time_step_fn(self)

    def after_time_step(self) -> None:
        if self.after_time_step_fn is not None:
            self.after_time_step_fn(self)

    def update_time(self) -> None:
        self.time += self.get_dt()
        self.time_step += 1

    def prepare(self) -> None:
        pass

    def solve(self) -> Any:
        """jit mode: solve_scan; eager mode: step by step with the hooks.
        With JAX_SPECTRAL_DNS_FIELD_DIR set, solve_scan's trajectory is written
        to <field_dir>/trajectory as one gzip-9 dataset "trajectory"
        (equation.py:362-372)."""
        self.prepare()
        if self.activate_jit_:
            t0 = time.time()
            out = self.solve_scan()  # type: ignore[attr-defined]
            self.last_solve_seconds = time.time() - t0
            if os.environ.get("JAX_SPECTRAL_DNS_FIELD_DIR") is not None:
                write_trajectory(out[0])
            return out
        while not self.done():
            self.before_time_step()
            self.perform_time_step(None, None, self.time_step)  # type: ignore[attr-defined]
            self.update_time()
            self.after_time_step()
        return None
