"""The block-size autotuner (`DeviceMesh.autotune`): it times every
candidate, restores the state and clock it found, and the block size it
keeps changes no bit of the steps that follow."""

import numpy as np
import pytest

from meshcfd_amd import meshgen
from oracle.euler import primitive_to_conservative


@pytest.mark.gpu
def test_gpu_autotune_restores_the_state_and_keeps_the_bits(cuda_device):
    import torch

    from meshcfd_amd.device import DeviceMesh

    print("case: autotune on a 40x40 Riemann mesh", flush=True)
    case, prim = meshgen.riemann_workload(40, 5)
    u0 = primitive_to_conservative(prim, case.numerics.gamma)
    dm = DeviceMesh(case, cuda_device, fused=True)
    dm.set_state(u0)
    dm.clock.zero_()
    before_u, before_clock = dm.u.clone(), dm.clock.clone()
    best = dm.autotune(candidates=(32, 64, 128), steps=2, repeats=1)
    assert best in (32, 64, 128) and set(dm.timings) == {32, 64, 128}
    assert torch.equal(dm.u, before_u) and torch.equal(dm.clock, before_clock)
    for _ in range(3):
        dm.step()
    tuned = dm.state()
    dm.close()
    ref = DeviceMesh(case, cuda_device, fused=True, block_cells=128 if best != 128 else 32)
    ref.set_state(u0)
    ref.clock.zero_()
    for _ in range(3):
        ref.step()
    np.testing.assert_array_equal(tuned, ref.state())
    ref.close()
