"""The per-step state at the plug-in boundary: the reference's SingleEntryStepper hands SolverDFSPH.step a Newton
State whose particle_q / particle_qd are Warp arrays (alloy/scene/steppers.py:153-184, sim.py:147). The solver views
any such device array through DLPack or __cuda_array_interface__, zero-copy, so a step writes the caller's memory.
The CPU tests check the adapter; the GPU test steps a state seen only through DLPack and compares it bit for bit
with the torch-state path."""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from alloy_amd.sph.solver import as_state_tensor, device_state
from tests.conftest import gpu_available


class DLPackArray:
    """A device array known only by its DLPack export, as a Warp array presents itself."""

    def __init__(self, t: torch.Tensor):
        self._t = t

    def __dlpack__(self, **kwargs):
        return self._t.__dlpack__(**kwargs)

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


class NewtonLikeState:
    def __init__(self, q, qd, f=None):
        self.particle_q, self.particle_qd, self.particle_f = q, qd, f


def test_dlpack_arrays_are_viewed_in_place():
    q = torch.arange(12, dtype=torch.float32)
    v = as_state_tensor(DLPackArray(q), "particle_q", "cpu", 4)
    assert v.shape == (4, 3) and v.data_ptr() == q.data_ptr()
    v[1, 2] = -7.0
    assert q[5] == -7.0  # the caller's memory


def test_torch_states_pass_through_and_others_are_refused():
    st = NewtonLikeState(torch.zeros(2, 3), torch.zeros(2, 3))
    assert device_state(st, torch.device("cpu"), 2) is st
    view = device_state(NewtonLikeState(DLPackArray(torch.zeros(2, 3)), DLPackArray(torch.zeros(2, 3))), "cpu", 2)
    assert view is not st and view.particle_q.shape == (2, 3) and view.particle_f is None
    with pytest.raises(ValueError, match="neither a torch tensor"):
        as_state_tensor([[0.0, 0.0, 0.0]], "particle_q", "cpu")
    with pytest.raises(ValueError, match="lives on"):  # a host array (numpy exports DLPack too) for a device solver
        as_state_tensor(np.zeros((2, 3), np.float32), "particle_q", "cuda")


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
def test_a_dlpack_state_steps_like_the_torch_state_bitwise():
    from alloy_amd.sph import scenes

    sc = scenes.dam_break(10, 0.2, 0, device="cuda")
    print(f"[state] dlpack state: {sc.model.particle_count} particles", flush=True)
    sys.stdout.flush()
    a, b = sc.make_solver(), sc.make_solver()
    st_a = sc.model.state()
    q, qd = st_a.particle_q.clone(), st_a.particle_qd.clone()
    st_b = NewtonLikeState(DLPackArray(q), DLPackArray(qd))
    for _ in range(10):
        a.step(st_a, st_a, None, None, scenes.DT)
        b.step(st_b, st_b, None, None, scenes.DT)
    torch.cuda.synchronize()
    assert torch.equal(st_a.particle_q, q) and torch.equal(st_a.particle_qd, qd)
    assert not torch.equal(q, sc.model.state().particle_q)  # the liquid moved
