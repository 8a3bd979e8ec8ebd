"""The allocation-free entry point mcfd_sum_squares_noalloc against
mcfd_sum_squares (same bits), called through the C-ABI twice in a row, the
second time on fewer cells, plus its size limit."""

import ctypes as C

import numpy as np
import pytest
import torch

from meshcfd_amd import meshgen, native
from meshcfd_amd.model import NumericalSettings


def _mesh_sums(cuda_device, nx, ny, r):
    from meshcfd_amd.device import DeviceMesh

    case = meshgen.periodic_case(meshgen.rectangle_mesh(nx, ny), NumericalSettings())
    dm = DeviceMesh(case, cuda_device, fused=False)
    assert dm.n == r.shape[0]
    want = dm.sum_squares(r).clone()
    torch.cuda.synchronize()
    dm.close()
    return want


@pytest.mark.gpu
def test_gpu_sum_squares_noalloc_equals_mesh_entry_point(cuda_device):
    lib = native.load()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(11)
    sums = torch.full((4,), -1.0, dtype=torch.float64, device=cuda_device)
    for nx, ny in ((300, 217), (40, 9)):  # 65,100 cells, then 360
        n = nx * ny
        print(f"case: sum_squares_noalloc n={n}", flush=True)
        r = torch.as_tensor(rng.normal(size=(n, 4)) * (1.0 + np.arange(4)), device=cuda_device)
        native.check(lib.mcfd_sum_squares_noalloc(C.c_void_p(r.data_ptr()), n, C.c_void_p(sums.data_ptr()), stream))
        got = sums.clone()
        want = _mesh_sums(cuda_device, nx, ny, r)
        assert torch.equal(got, want), (got, want)
        np.testing.assert_allclose(got.cpu().numpy(), (r.cpu().numpy() ** 2).sum(axis=0), rtol=1e-12)


@pytest.mark.gpu
def test_gpu_sum_squares_noalloc_limits(cuda_device):
    lib = native.load()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    r = torch.zeros((4, 4), dtype=torch.float64, device=cuda_device)
    sums = torch.full((4,), -1.0, dtype=torch.float64, device=cuda_device)
    assert lib.mcfd_sum_squares_noalloc(C.c_void_p(r.data_ptr()), (1 << 24) + 1, C.c_void_p(sums.data_ptr()), stream) != 0
    assert "16777216" in native.last_error()
    native.check(lib.mcfd_sum_squares_noalloc(C.c_void_p(r.data_ptr()), 0, C.c_void_p(sums.data_ptr()), stream))
    torch.cuda.synchronize()
    assert torch.equal(sums.cpu(), torch.zeros(4, dtype=torch.float64))
