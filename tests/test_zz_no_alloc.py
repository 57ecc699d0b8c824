"""jsd_energy_noalloc (the allocation-free energy: no plan, partial sums in a
static device array) against jsd_energy, through the C ABI, on two calls in
a row, the second with a smaller field; the results are bit-identical (same
kernels and grid)."""

import ctypes

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import KIND_NS, _lib
from jax_spectral_dns_amd import domain_math as DM
from jax_spectral_dns_amd.plan import StepPlan
from tests.helpers import ns_field

pytestmark = pytest.mark.gpu


def _noalloc(g, u, out):
    lib = _lib.load()
    w = torch.from_numpy(np.ascontiguousarray(g.y_weights)).to("cuda")
    vp = ctypes.c_void_p
    rc = lib.jsd_energy_noalloc(
        vp(u.data_ptr()), g.nxm, g.ny, g.nzm, g.px, g.pz, g.lx, g.lz, vp(w.data_ptr()), vp(out.data_ptr()),
        vp(torch.cuda.current_stream().cuda_stream),
    )
    assert rc == 0, lib.jsd_last_error()
    torch.cuda.synchronize()
    return float(out.item())


def test_energy_noalloc_matches_jsd_energy_on_two_calls():
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    for k, (shape, lx, lz) in enumerate([((48, 97, 32), 2 * np.pi, np.pi), ((16, 33, 8), 4 * np.pi, 2 * np.pi)]):
        g = DM.make_grid(shape, lx, lz)
        u = ns_field(g, seed=60 + k).to("cuda")
        want = StepPlan(g, 3000.0, 0.01, KIND_NS).energy(u)
        got = _noalloc(g, u, out)
        print(shape, got, want)
        assert got == want


def test_energy_noalloc_rejects_bad_sizes():
    lib = _lib.load()
    vp = ctypes.c_void_p
    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    st = vp(torch.cuda.current_stream().cuda_stream)
    p = vp(x.data_ptr())
    assert lib.jsd_energy_noalloc(p, 9, 200, 9, 14, 14, 1.0, 1.0, p, p, st) == -1
    assert b"ny" in lib.jsd_last_error()
    assert lib.jsd_energy_noalloc(p, 8, 17, 9, 14, 14, 1.0, 1.0, p, p, st) == -1
    assert lib.jsd_energy_noalloc(None, 9, 17, 9, 14, 14, 1.0, 1.0, p, p, st) == -1
