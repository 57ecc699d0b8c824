"""Host-side argument checks that run before any library call: wrongly
shaped tensors and repeated batch members raise instead of reaching a
kernel."""

import pytest
import torch

from jax_spectral_dns_amd import _lib
from jax_spectral_dns_amd import domain_math as DM
from jax_spectral_dns_amd.batch import check_distinct
from jax_spectral_dns_amd.plan import StepPlan


def _plan_shell(shape=(8, 17, 8), kind=_lib.KIND_NS, linearise=False):
    """A StepPlan without a device plan: only the host checks are exercised
    (every check raises before the library is touched)."""
    p = StepPlan.__new__(StepPlan)
    p.grid = DM.make_grid(shape)
    p.kind = kind
    p.linearise = linearise
    p._h = None
    return p


def test_ydiff_rejects_wrong_shapes():
    p = _plan_shell()
    g = p.grid
    for bad in [(g.nxm, g.ny + 1, g.nzm), (g.nxm + 2, g.ny, g.nzm), (2, g.nxm, g.ny, g.nzm + 2), (g.ny, g.nzm),
                (1, 1, g.nxm, g.ny, g.nzm)]:
        with pytest.raises(_lib.JsdError):
            p.ydiff(torch.zeros(bad, dtype=torch.complex128))
    with pytest.raises(_lib.JsdError):
        p.ydiff(torch.zeros((g.nxm, g.ny, g.nzm), dtype=torch.complex128), order=3)


def test_nonlinear_rejects_wrong_shapes():
    p = _plan_shell()
    g = p.grid
    good = torch.zeros((3, g.nxm, g.ny, g.nzm), dtype=torch.complex128)
    with pytest.raises(_lib.JsdError):
        p.nonlinear(torch.zeros((3, g.nxm, g.ny - 1, g.nzm), dtype=torch.complex128))
    with pytest.raises(_lib.JsdError):
        p.nonlinear(good, torch.zeros((3, g.nxm, g.ny, g.nzm + 2), dtype=torch.complex128))
    d = _plan_shell(kind=_lib.KIND_DUAL)
    with pytest.raises(_lib.JsdError):
        d.nonlinear(good, None)


def test_batch_rejects_repeated_plans_and_buffers():
    a, b = object(), object()
    s = [torch.zeros(4), torch.zeros(4)]
    check_distinct([a, b], s)
    check_distinct([a, b], s, [torch.zeros(4), torch.zeros(4)])
    with pytest.raises(_lib.JsdError):
        check_distinct([a, a], s)
    with pytest.raises(_lib.JsdError):
        check_distinct([a, b], [s[0], s[0]])
    with pytest.raises(_lib.JsdError):
        check_distinct([a, b], s, [s[1], torch.zeros(4)])


def test_find_suitable_dt():
    """Equation.find_suitable_dt (equation.py:114-162): CFL bound over the
    directions, then a step count whose median-factor split is balanced."""
    import numpy as np

    from jax_spectral_dns_amd.domain import PhysicalDomain
    from jax_spectral_dns_amd.equation import Equation
    from jax_spectral_dns_amd.navier_stokes import median_factor

    pd = PhysicalDomain.create((24, 65, 4), (True, False, True), scale_factors=(2 * np.pi, 1.0, 2 * np.pi))
    dt = Equation.find_suitable_dt(pd, max_cfl=0.7)
    want = 0.9 * min(0.7 * (2 * np.pi / 38) / 1.0, 0.7 * (2.0 / 65) / 1e-10, 0.7 * (2 * np.pi / 8) / 1e-10)
    assert dt == want
    dt_T = Equation.find_suitable_dt(pd, max_cfl=0.7, end_time=15.0)
    n = round(15.0 / dt_T)
    assert abs(n * dt_T - 15.0) < 1e-12
    assert n >= int(15.0 / dt)
    inner = median_factor(n)
    outer = n // inner
    assert not (np.sqrt(n) - outer > outer)
    # the first count that balances: every smaller count from int(T/dt) is unbalanced
    for m in range(int(15.0 / dt), n):
        i = median_factor(m)
        assert np.sqrt(m) - m // i > m // i
