"""The nonlinear conjugate-gradient method with strong-Wolfe zoom line search
behind OptimiserWrapper (the reference hands its objective to
jaxopt.NonlinearCG, optimiser.py:231-242), checked on the CPU on textbook
problems whose minimisers are known: Rosenbrock's function (minimum at
(1, 1)) and a convex quadratic (minimum A^-1 b).  The dual-gradient
objective itself runs in tests/test_gpu_dual.py."""

import numpy as np
import torch

from jax_spectral_dns_amd.gradient_descent_solver import OptimiserWrapper, _rdot


def _solver(x0, value_and_grad, step_size=1e-3):
    ow = OptimiserWrapper.__new__(OptimiserWrapper)
    ow.value_and_grad = value_and_grad
    ow.normalize_field = lambda v: v
    ow.velocity = lambda v: v
    ow.maxls = 15
    ow.step_size = step_size
    ow.x = x0.clone()
    ow.evaluations = 0
    ow.precondition = False  # the plain method (the preconditioner is the energy metric of the flow problem)
    ow.f, ow.g = value_and_grad(ow.x)
    ow.y = ow.g
    ow.d = -ow.g
    ow.value = -ow.f
    ow.values, ow.steps, ow.wolfe, ow.i = [ow.value], [], [], 0
    return ow


def test_rosenbrock_minimum():
    def vg(x):
        a, b = x.real[0], x.real[1]
        f = 100.0 * (b - a * a) ** 2 + (1.0 - a) ** 2
        ga = -400.0 * a * (b - a * a) - 2.0 * (1.0 - a)
        gb = 200.0 * (b - a * a)
        return float(f), torch.stack([ga, gb]).to(torch.complex128)

    ow = _solver(torch.tensor([-1.2, 1.0], dtype=torch.complex128), vg)
    fs = [ow.f]
    for _ in range(200):
        ow.update()
        fs.append(ow.f)
        if ow.f < 1e-20:
            break
    assert all(b <= a + 1e-15 for a, b in zip(fs, fs[1:]))  # monotone decrease
    assert torch.allclose(ow.x.real, torch.tensor([1.0, 1.0], dtype=torch.float64), atol=1e-6)
    assert sum(ow.wolfe) >= 0.8 * len(ow.wolfe)


def test_convex_quadratic_minimum_and_armijo():
    rng = np.random.default_rng(3)
    n = 30
    Q = rng.standard_normal((n, n))
    A = torch.from_numpy(Q @ Q.T + n * np.eye(n))
    b = torch.from_numpy(rng.standard_normal(n))

    def vg(x):
        xr = x.real
        return float(0.5 * xr @ A @ xr - b @ xr), (A @ xr - b).to(torch.complex128)

    ow = _solver(torch.zeros(n, dtype=torch.complex128), vg, step_size=1e-2)
    for _ in range(60):
        f0, g0, d0 = ow.f, ow.g, ow.d if _rdot(ow.g, ow.d) < 0 else -ow.g
        ow.update()
        if ow.steps:
            # the accepted step satisfies sufficient decrease
            assert ow.f <= f0 + 1e-4 * ow.steps[-1] * _rdot(g0, d0) + 1e-12
    xs = torch.linalg.solve(A, b)
    # f resolves x only to ~sqrt(eps): the accepted steps stop near 2e-8 relative
    assert float((ow.x.real - xs).abs().max()) < 1e-7 * float(xs.abs().max())
