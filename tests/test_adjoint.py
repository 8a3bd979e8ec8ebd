"""Adjoint residual: the hand-written VJP against central differences of the
CPU oracle residual (a dot-product test, lambda^T J v), and gradients of a
transient objective against central differences of the forward solver."""

import numpy as np
import pytest

from meshcfd_amd import meshgen
from meshcfd_amd.model import NumericalSettings

pytestmark = pytest.mark.gpu


def _dm(case, cuda_device, fused=True):
    from meshcfd_amd.device import DeviceMesh

    print(f"case: adjoint {case.name} cells={case.mesh.cell_count} flux={case.numerics.flux} "
          f"rec={case.numerics.reconstruction}", flush=True)
    return DeviceMesh(case, cuda_device, fused=fused, block_cells=64)


@pytest.mark.parametrize("flux", ["Rusanov", "HLLC", "Roe"])
@pytest.mark.parametrize("reconstruction", ["First order", "MUSCL linear", "MUSCL parabolic"])
@pytest.mark.parametrize("kind", ["periodic", "mixed", "farfield", "transmissive"])
def test_vjp_matches_central_difference_of_oracle(cuda_device, flux, reconstruction, kind):
    from test_gpu_parity import _case, _perturbed_primitive

    from meshcfd_amd.adjoint import residual_vjp
    from oracle.euler import primitive_to_conservative
    from oracle.solver import Oracle

    case = _case(kind, flux, reconstruction)
    dm = _dm(case, cuda_device)
    u = primitive_to_conservative(_perturbed_primitive(case, seed=5), 1.4)
    rng = np.random.default_rng(11)
    lam = rng.normal(size=u.shape)
    scale = np.max(np.abs(u), axis=0)
    v = rng.normal(size=u.shape) * scale
    o = Oracle(dm.geometry, case.numerics)
    eps = 1e-7
    fd = float(np.sum(lam * (o.residual(u + eps * v) - o.residual(u - eps * v)) / (2 * eps)))
    ubar = dm.to_mesh_order(residual_vjp(dm, dm.to_device_order(u), dm.to_device_order(lam)))
    dm.check_status()
    ad = float(np.sum(ubar * v))
    print("fd", fd, "adjoint", ad)
    assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1.0)
    dm.close()


def test_vjp_rows_match_oracle_jacobian_columns(cuda_device):
    """Every VJP entry of a small first-order case equals a central difference."""
    from test_gpu_parity import _case, _perturbed_primitive

    from meshcfd_amd.adjoint import residual_vjp
    from oracle.euler import primitive_to_conservative
    from oracle.solver import Oracle

    case = _case("mixed", "HLLC", "First order", n=4)
    dm = _dm(case, cuda_device)
    u = primitive_to_conservative(_perturbed_primitive(case, seed=8), 1.4)
    lam = np.random.default_rng(3).normal(size=u.shape)
    o = Oracle(dm.geometry, case.numerics)
    ubar = dm.to_mesh_order(residual_vjp(dm, dm.to_device_order(u), dm.to_device_order(lam)))
    eps = 1e-7
    for i in range(0, u.shape[0], 3):
        for k in range(4):
            e = np.zeros_like(u)
            e[i, k] = eps * max(abs(u[i, k]), 1.0)
            fd = np.sum(lam * (o.residual(u + e) - o.residual(u - e))) / (2 * e[i, k])
            assert abs(ubar[i, k] - fd) <= 1e-6 * max(abs(fd), 1.0), (i, k, ubar[i, k], fd)
    dm.close()


@pytest.mark.parametrize("reconstruction", ["First order", "MUSCL linear"])
def test_kinetic_energy_gradient_wrt_vortex_strength(cuda_device, reconstruction):
    """d KE(U_N) / d beta through 8 checkpointed SSP-RK3 steps vs central
    FD of step 1e-4, to 1e-6 (measured: first order 1.2e-9; MUSCL linear with
    Barth-Jespersen 1.8e-10, its error 1.5e-6 at step 1e-3 and 7e-7 at 1e-7,
    where truncation and rounding take over)."""
    import torch

    from meshcfd_amd.adjoint import kinetic_energy, rk3_rollout, vortex_state
    from meshcfd_amd.geometry import shoelace

    mesh = meshgen.rectangle_mesh(24, 24, (-5.0, 5.0, -5.0, 5.0), jitter=0.2, seed=0)
    numerics = NumericalSettings(flux="HLLC", reconstruction=reconstruction, time_integration="Explicit RK3")
    case = meshgen.periodic_case(mesh, numerics)
    dm = _dm(case, cuda_device)
    centers, areas = shoelace(mesh)
    cdev = dm.to_device_order(centers)
    adev = dm.to_device_order(areas)
    dt, steps = 0.05, 8

    def objective(beta, grad):
        b = torch.tensor(beta, dtype=torch.float64, device=cuda_device, requires_grad=grad)
        u = rk3_rollout(dm, vortex_state(cdev, b), dt, steps, checkpointed=grad)
        ke = kinetic_energy(u, adev)
        if grad:
            ke.backward()
            return float(ke.item()), float(b.grad.item())
        return float(ke.item()), None

    with torch.no_grad():
        kp, _ = objective(5.0 + 1e-4, False)
        km, _ = objective(5.0 - 1e-4, False)
    fd = (kp - km) / 2e-4
    _, ad = objective(5.0, True)
    dm.check_status()
    print(reconstruction, "fd", fd, "adjoint", ad)
    assert abs(ad - fd) <= 1e-6 * abs(fd)
    dm.close()


def test_checkpointed_gradient_equals_stored_gradient(cuda_device):
    import torch

    from meshcfd_amd.adjoint import kinetic_energy, rk3_rollout, vortex_state
    from meshcfd_amd.geometry import shoelace

    mesh = meshgen.rectangle_mesh(16, 16, (-5.0, 5.0, -5.0, 5.0), jitter=0.2, seed=0)
    case = meshgen.periodic_case(mesh, NumericalSettings(flux="HLLC", reconstruction="MUSCL linear"))
    dm = _dm(case, cuda_device)
    centers, areas = shoelace(mesh)
    cdev, adev = dm.to_device_order(centers), dm.to_device_order(areas)
    grads = []
    for ck in (True, False):
        b = torch.tensor(5.0, dtype=torch.float64, device=cuda_device, requires_grad=True)
        kinetic_energy(rk3_rollout(dm, vortex_state(cdev, b), 0.05, 5, checkpointed=ck), adev).backward()
        grads.append(b.grad.item())
    assert grads[0] == grads[1]
    dm.close()
