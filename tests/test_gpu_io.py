"""Trajectory output of a GPU run in the reference's file format: with
JAX_SPECTRAL_DNS_FIELD_DIR set, Equation.solve writes solve_scan's trajectory
to <field_dir>/trajectory (equation.py:362-372); the file holds exactly the
states the run returned, and a state saved with save_to_file comes back
bit for bit."""

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import hdf5
from jax_spectral_dns_amd.domain import PhysicalDomain
from jax_spectral_dns_amd.field import Field, FourierField, VectorField
from jax_spectral_dns_amd.initial import perturbation
from jax_spectral_dns_amd.navier_stokes_perturbation import NavierStokesVelVortPerturbation

pytestmark = pytest.mark.gpu


def test_solve_writes_trajectory_file(tmp_path, monkeypatch):
    monkeypatch.setenv("JAX_SPECTRAL_DNS_FIELD_DIR", str(tmp_path))
    monkeypatch.setattr(Field, "field_dir", str(tmp_path))
    shape = (24, 33, 16)
    pd = PhysicalDomain.create(shape, (True, False, True), scale_factors=(2 * np.pi, 1.0, np.pi))
    u = torch.from_numpy(perturbation(pd.grid_math, seed=3, energy=1e-3)).to("cuda")
    vh = VectorField.FromData(FourierField, pd, u, "velocity_hat")
    nse = NavierStokesVelVortPerturbation(vh, Re=2000.0, dt=0.01, end_time=0.06)
    nse.write_intermediate_output = True
    traj, _, n = nse.solve()
    t = hdf5.read(str(tmp_path / "trajectory"))
    assert np.array_equal(t["trajectory"].data, traj.cpu().numpy())
    assert t["trajectory"].data.shape[0] == nse.number_of_outer_steps + 1
    final = nse.get_latest_field("velocity_hat")
    VectorField([final[c] for c in range(3)], "velocity_hat").save_to_file("final")
    back = VectorField.FromFile(pd, "final", name="velocity_hat", device="cuda")
    assert torch.equal(back.get_data(), final.get_data())


def test_dlpack_round_trip_keeps_complex128_layout():
    """The torch half of the DLPack binding the integration notes describe:
    a complex128 state exported and re-imported through DLPack is the same
    buffer with the same dtype, shape and strides, and jsd_step through the
    re-imported view equals jsd_step on the original (the JAX half is not
    exercised: JAX is absent)."""
    import torch.utils.dlpack as dl

    from jax_spectral_dns_amd import KIND_NS
    from jax_spectral_dns_amd.plan import StepPlan
    from tests.helpers import ns_field, pair

    g, _ = pair((24, 33, 16), 2 * np.pi, np.pi)
    u = ns_field(g, seed=81).to("cuda")
    v = dl.from_dlpack(dl.to_dlpack(u))
    assert v.data_ptr() == u.data_ptr() and v.dtype == torch.complex128
    assert v.shape == u.shape and v.stride() == u.stride()
    plan = StepPlan(g, 3000.0, 0.005, KIND_NS)
    a = u.clone()
    b = dl.from_dlpack(dl.to_dlpack(u.clone()))
    plan.step(a, -2.0 / 3000.0, 1)
    plan.step(b, -2.0 / 3000.0, 1)
    assert torch.equal(a, b)
