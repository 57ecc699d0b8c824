"""jsd_step allocates, synchronises and memsets nothing, so a caller can
capture it in a HIP graph (torch.cuda.graph) and replay it: every replay,
with eager work and synchronisation between replays, equals the same steps
run eagerly, bit for bit."""

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import KIND_NS, KIND_PERTURBATION
from jax_spectral_dns_amd.plan import StepPlan
from tests.helpers import ns_field, pair, perturbation_field

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("kind", [KIND_NS, KIND_PERTURBATION])
def test_captured_step_replays_match_eager(kind):
    g, _ = pair((24, 65, 16), 2 * np.pi, np.pi)
    if kind == KIND_NS:
        plan = StepPlan(g, 3000.0, 0.005, KIND_NS)
        u0 = ns_field(g, seed=71).to(DEV)
        dpdx = -2.0 / 3000.0
    else:
        plan = StepPlan(g, 3000.0, 0.005, KIND_PERTURBATION, linearise=False)
        u0 = perturbation_field(g, seed=72, energy=1e-2).to(DEV)
        dpdx = 0.0
    eager = u0.clone()
    ref = []
    for _ in range(4):
        plan.step(eager, dpdx, 2)
        ref.append(eager.clone())
    state = u0.clone()
    plan.step(state, dpdx, 2)  # warm-up on a side buffer before capture
    state.copy_(u0)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.step(state, dpdx, 2)
    torch.cuda.synchronize()
    state.copy_(u0)
    for r in range(4):
        graph.replay()
        torch.cuda.synchronize()
        scratch = torch.randn(1 << 20, device=DEV).sum().item()  # eager work between replays
        assert np.isfinite(scratch)
        assert torch.equal(state, ref[r]), r
