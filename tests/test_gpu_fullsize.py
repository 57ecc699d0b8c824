"""Full-size and long-horizon checks of the HIP path.

* linearised perturbation, 25 x 65 x 5 stored modes, Re 3000,
  dt 0.01: 1500 steps against the committed oracle fixture, <= 1e-8
  relative L2 at T = 15 and <= 1e-11 after one step (the FP64 parity
  tolerances of the step).
* plain nonlinear NS, 129 x 129 x 129 stored modes, Re 5000,
  dt 0.005: ten consecutive steps against the oracle (<= 1e-11 per step),
  200 steps against the oracle's energy history, then size-independent
  properties over several steps: continuity i kx u + D v + i kz w = 0, the
  no-slip / no-penetration wall values, v' = 0 at the walls, Hermitian
  symmetry away from the +-K planes, and the energy budget staying finite.
"""

import os

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import KIND_NS, KIND_PERTURBATION
from jax_spectral_dns_amd import domain_math as DM
from jax_spectral_dns_amd import initial
from jax_spectral_dns_amd.plan import StepPlan
from tests.helpers import rel

pytestmark = pytest.mark.gpu
DEV = "cuda"
HERE = os.path.dirname(os.path.abspath(__file__))


def test_linearised_25x65x5_1500_steps_against_fixture():
    fx = np.load(os.path.join(HERE, "golden", "linearised_25x65x5_1500.npz"))
    re, dt, n = fx["params"][:3]
    g = DM.make_grid((24, 65, 4))
    plan = StepPlan(g, re, dt, KIND_PERTURBATION, linearise=True)
    u0 = torch.from_numpy(fx["u0"])
    assert np.array_equal(fx["u0"], initial.perturbation(g, seed=0, energy=1e-6))
    u = u0.to(DEV)
    plan.step(u, 0.0, 1)
    e1 = rel(u, torch.from_numpy(fx["u1"]))
    plan.step(u, 0.0, 9)
    e10 = rel(u, torch.from_numpy(fx["u10"]))
    energies = [plan.energy(u)]
    for _ in range(int(n) - 10):
        plan.step(u, 0.0, 1)
    eT = rel(u, torch.from_numpy(fx["uT"]))
    gain = plan.energy(u) / plan.energy(u0.to(DEV))
    gain_ref = fx["energy"][-1] / fx["energy"][0]
    print("linearised 25x65x5 rel L2: 1 step %.2e, 10 steps %.2e, 1500 steps %.2e; gain %.12f vs %.12f" % (e1, e10, eT, gain, gain_ref))
    assert e1 < 1e-11
    assert e10 < 1e-10
    assert eT < 1e-8
    assert abs(gain - gain_ref) / gain_ref < 1e-9
    assert abs(energies[0] - fx["energy"][10]) / fx["energy"][10] < 1e-10


def _ns_128x129x128():
    g = DM.make_grid((128, 129, 128), 4 * np.pi, 2 * np.pi)
    base = initial.laminar(g)
    u0 = base + initial.perturbation(g, seed=0, energy=0.05 * initial.spectral_energy(g, base))
    return g, torch.from_numpy(u0)


def test_ns_128x129x128_ten_steps_against_oracle():
    """Ten consecutive steps of the full-size workload, each against the
    oracle advanced from the same start: <= 1e-11 (n + 1) relative L2."""
    from oracle import jsd_oracle as O

    torch.set_num_threads(min(16, len(os.sched_getaffinity(0))))
    g, u0 = _ns_128x129x128()
    re, dt = 5000.0, 0.005
    plan = StepPlan(g, re, dt, KIND_NS)
    u = u0.to(DEV)
    d = O.make_domain((128, 129, 128), 4 * np.pi, 2 * np.pi)
    sp = O.make_ns(d, re, dt, -2.0 / re)
    ref = u0
    for n in range(10):
        plan.step(u, -2.0 / re, 1)
        ref = O.rk_step(sp, ref, n)
        err = rel(u, ref)
        print("ns 128x129x128 step %d rel L2 %.3e" % (n + 1, err), flush=True)
        assert err < 1e-11 * (n + 1)


def test_ns_128x129x128_200_steps_against_oracle_energy_history():
    """The full 128x129x128 horizon: 200 steps of the GPU step against the oracle's
    energy after every step and its (0,0) mean-flow lines of u and w after
    1, 10, 100 and 200 steps (tests/golden/make_golden.py ns)."""
    path = os.path.join(HERE, "golden", "ns_128x129x128_200_energy.npz")
    fx = np.load(path)
    re, dt, n = fx["params"][:3]
    g, u0 = _ns_128x129x128()
    plan = StepPlan(g, re, dt, KIND_NS)
    u = u0.to(DEV)
    e0 = plan.energy(u)
    assert abs(e0 - fx["energy"][0]) <= 1e-13 * fx["energy"][0]
    worst = 0.0
    for t in range(int(n)):
        plan.step(u, -2.0 / re, 1)
        e = plan.energy(u)
        err = abs(e - fx["energy"][t + 1]) / fx["energy"][t + 1]
        worst = max(worst, err)
        assert err < 1e-11, (t + 1, err)
        if t + 1 in (1, 10, 100, 200):
            ref = torch.from_numpy(fx["mean%d" % (t + 1)])
            mean = torch.stack([u[0, 0, :, 0], u[2, 0, :, 0]]).cpu()
            merr = rel(mean, ref)
            print("ns 128x129x128 step %d: energy rel %.2e, mean flow rel %.2e" % (t + 1, err, merr))
            assert merr < 1e-11
    print("ns 128x129x128 worst energy rel over 200 steps %.2e" % worst)


def test_ns_128x129x128_properties_over_steps():
    g, u0 = _ns_128x129x128()
    re, dt = 5000.0, 0.005
    plan = StepPlan(g, re, dt, KIND_NS)
    u = u0.to(DEV)
    e0 = plan.energy(u)
    plan.step(u, -2.0 / re, 20)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(u).all())
    un = u.cpu().numpy()
    D = g.D
    dv = np.einsum("lj,ijk->ilk", D, un[1])
    kx = g.kx[:, None, None]
    kz = g.kz[None, None, :]
    div = 1j * kx * un[0] + dv + 1j * kz * un[2]
    scale = np.abs(un).max() * np.abs(D).max()
    assert np.abs(div).max() < 1e-12 * scale
    assert np.abs(un[:, :, 0]).max() == 0.0 and np.abs(un[:, :, -1]).max() == 0.0
    assert np.abs(dv[:, 0]).max() < 1e-12 * scale and np.abs(dv[:, -1]).max() < 1e-12 * scale
    # Hermitian symmetry of a real field holds away from the +-K planes
    M, Z = g.nxm, g.nzm
    K, KZ = M // 2, Z // 2
    inner_x = [i for i in range(M) if i not in (K, K + 1)]
    inner_z = [j for j in range(Z) if j not in (KZ, KZ + 1)]
    pi = (M - np.arange(M)) % M
    pj = (Z - np.arange(Z)) % Z
    h = un - np.conj(un[:, pi][:, :, :, pj])
    assert np.abs(h[:, inner_x][:, :, :, inner_z]).max() < 1e-12 * np.abs(un).max()
    e1 = plan.energy(u)
    assert 0.9 * e0 < e1 < 1.1 * e0
