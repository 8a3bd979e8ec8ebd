"""Shape gradients: the torch restatement of the geometry against the built
geometry and its finite differences (CPU), the device's lambda^T dR/dh
against central differences of the oracle residual on re-built meshes, and
the steady adjoint's d(outlet total-pressure loss)/d(bump height) against
central differences of re-converged solutions (GPU)."""

import math

import numpy as np
import pytest
import torch

from meshcfd_amd import meshgen
from meshcfd_amd.geometry import build_geometry
from meshcfd_amd.geometry_adjoint import GEOMETRY_KEYS, TorchGeometry
from meshcfd_amd.layout import build_layout


def _arrays(case, order=None):
    g = build_geometry(case)
    lay = build_layout(g, order=order)
    out = {k: getattr(lay, k) for k in GEOMETRY_KEYS if k != "area"}
    out["area"] = lay.area
    return g, lay, out


@pytest.mark.parametrize("which", ["bump", "vortex", "split"])
def test_torch_geometry_reproduces_the_layout(which):
    if which == "bump":
        case = meshgen.bump_case(14, 5, 0.07)[0]
    elif which == "vortex":
        case = meshgen.vortex_workload(8, 1)[0]  # periodic pairs
    else:
        case = meshgen.riemann_workload(12, 1)[0]  # quads and triangles
    g, lay, ref = _arrays(case)
    out = TorchGeometry(case, g, lay)(torch.as_tensor(case.mesh.points))
    for k in GEOMETRY_KEYS:
        a = out[k].numpy()
        scale = max(np.max(np.abs(ref[k])), 1e-300)
        assert np.max(np.abs(a - ref[k])) <= 1e-15 * scale, k


def test_bump_nodes_torch_equals_numpy():
    base = meshgen.rectangle_mesh(20, 6, (0.0, 3.0, 0.0, 1.0)).points
    a = meshgen.bump_nodes(base, 0.05)
    b = meshgen.bump_nodes(torch.as_tensor(base), torch.tensor(0.05, dtype=torch.float64), xp=torch).numpy()
    assert np.array_equal(a, b)
    # the bump's top is at (1.5, h), its ends at (1, 0) and (2, 0)
    y = meshgen.bump_surface(np.array([1.0, 1.5, 2.0, 0.5]), 0.05)
    np.testing.assert_allclose(y, [0.0, 0.05, 0.0, 0.0], atol=1e-15)


def test_torch_geometry_derivative_matches_finite_differences():
    """d(geometry arrays)/dh through torch equals central differences of
    build_geometry + build_layout on meshes re-built at h +- eps (same cell
    order)."""
    h, eps = 0.06, 1e-6
    case = meshgen.bump_case(14, 5, h)[0]
    g, lay, _ = _arrays(case)
    tg = TorchGeometry(case, g, lay)
    base = meshgen.rectangle_mesh(14, 5, (0.0, 3.0, 0.0, 1.0)).points
    hp = torch.tensor(h, dtype=torch.float64, requires_grad=True)
    out = tg(meshgen.bump_nodes(torch.as_tensor(base), hp, xp=torch))
    _, _, plus = _arrays(meshgen.bump_case(14, 5, h + eps)[0], order=lay.cell_order)
    _, _, minus = _arrays(meshgen.bump_case(14, 5, h - eps)[0], order=lay.cell_order)
    for k in GEOMETRY_KEYS:
        fd = (plus[k] - minus[k]) / (2 * eps)
        jac = np.array([torch.autograd.grad(out[k][i], hp, retain_graph=True)[0].item() for i in range(len(fd))])
        scale = max(np.max(np.abs(fd)), 1e-12)
        assert np.max(np.abs(jac - fd)) <= 1e-6 * scale, k


# --- GPU ------------------------------------------------------------------------


def _state(case, seed=4):
    g = build_geometry(case)
    x, y = g.centers[:, 0], g.centers[:, 1]
    rng = np.random.default_rng(seed)
    n = len(x)
    s = math.sqrt(1.4) * 0.675
    prim = np.column_stack((1.0 + 0.1 * np.sin(2 * x) * np.cos(3 * y) + 0.02 * rng.random(n),
                            s * (1.0 + 0.1 * np.cos(2 * y)) + 0.02 * rng.random(n),
                            0.05 * np.sin(3 * x) + 0.02 * rng.random(n),
                            1.0 + 0.1 * np.cos(2 * (x - y)) + 0.02 * rng.random(n)))
    return prim


@pytest.mark.gpu
@pytest.mark.parametrize("flux", ["HLLC", "Rusanov"])
@pytest.mark.parametrize("reconstruction", ["First order", "MUSCL linear"])
def test_gpu_shape_derivative_of_residual_matches_fd_of_oracle(cuda_device, flux, reconstruction):
    """lambda^T dR/dh (device geometry adjoint + torch geometry + bump map)
    against central differences of the oracle's lambda^T R on meshes re-built
    at h +- eps, at a fixed non-uniform state."""
    from meshcfd_amd.device import DeviceMesh
    from meshcfd_amd.geometry_adjoint import shape_gradient
    from oracle.euler import primitive_to_conservative
    from oracle.solver import Oracle

    h, eps = 0.06, 1e-7
    nx, ny = 16, 6
    case = meshgen.bump_case(nx, ny, h)[0]
    case.numerics.flux = flux
    case.numerics.reconstruction = reconstruction
    print(f"case: shape gradient {flux} {reconstruction}", flush=True)
    dm = DeviceMesh(case, cuda_device, fused=False)
    u = primitive_to_conservative(_state(case), 1.4)
    lam = np.random.default_rng(7).normal(size=u.shape)
    base = torch.as_tensor(meshgen.rectangle_mesh(nx, ny, (0.0, 3.0, 0.0, 1.0)).points)
    d = shape_gradient(dm, dm.to_device_order(u), dm.to_device_order(lam),
                       lambda p: meshgen.bump_nodes(base, p, xp=torch), h)
    vals = []
    for sgn in (1, -1):
        c = meshgen.bump_case(nx, ny, h + sgn * eps)[0]
        c.numerics = case.numerics
        vals.append(float(np.sum(lam * Oracle(build_geometry(c), c.numerics).residual(u))))
    fd = (vals[0] - vals[1]) / (2 * eps)
    print("adjoint", d, "fd", fd)
    assert abs(d - fd) <= 1e-6 * max(abs(fd), 1.0)
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["periodic", "farfield", "mixed"])
@pytest.mark.parametrize("flux", ["HLLC", "Rusanov"])
def test_gpu_node_gradient_matches_fd(cuda_device, kind, flux):
    """A random displacement of the interior nodes (MUSCL linear; periodic
    pairs, far-field characteristics, walls, inlet and outlet):
    lambda^T dR/dt against central differences of the oracle."""
    from test_gpu_parity import _case, _perturbed_primitive

    from meshcfd_amd.device import DeviceMesh
    from meshcfd_amd.geometry_adjoint import shape_gradient
    from oracle.euler import primitive_to_conservative
    from oracle.solver import Oracle

    case = _case(kind, flux, "MUSCL linear", n=10)
    print(f"case: node gradient {kind} {flux}", flush=True)
    dm = DeviceMesh(case, cuda_device, fused=False)
    pts0 = case.mesh.points.copy()
    rng = np.random.default_rng(3)
    lo, hi = pts0.min(0), pts0.max(0)
    interior = np.all((pts0 > lo + 1e-9) & (pts0 < hi - 1e-9), axis=1)
    field = rng.normal(size=pts0.shape) * interior[:, None] * 0.01
    u = primitive_to_conservative(_perturbed_primitive(case, seed=9), 1.4)
    lam = rng.normal(size=u.shape)
    p0 = torch.as_tensor(pts0)
    v = torch.as_tensor(field)
    d = shape_gradient(dm, dm.to_device_order(u), dm.to_device_order(lam), lambda t: p0 + t * v, 0.0)
    eps = 1e-5  # node displacements of 1e-7: small against the limiter's kinks, large against round-off
    vals = []
    for sgn in (1, -1):
        case.mesh.points = pts0 + sgn * eps * field
        vals.append(float(np.sum(lam * Oracle(build_geometry(case), case.numerics).residual(u))))
    case.mesh.points = pts0
    fd = (vals[0] - vals[1]) / (2 * eps)
    print("adjoint", d, "fd", fd)
    assert abs(d - fd) <= 1e-6 * max(abs(fd), 1.0)
    dm.close()


@pytest.mark.gpu
def test_gpu_outlet_loss_bump_gradient_matches_fd_of_converged_solutions(cuda_device):
    """d(outlet total-pressure loss)/d(bump height) at the converged steady
    state: adjoint (dR/dU)^T lambda = -dJ/dU, then lambda^T dR/dh through the
    geometry, against central differences of re-converged solutions.
    First-order HLLC with explicit local-CFL RK3 pseudo-time steps: the
    Barth-Jespersen MUSCL residual of this channel stalls near 3e-3 (the
    oracle does the same), so its converged states are not smooth in h."""
    from meshcfd_amd.device import DeviceMesh
    from meshcfd_amd.geometry_adjoint import shape_gradient
    from meshcfd_amd.steady_adjoint import OutletTotalPressureLoss, SteadyAdjoint, objective_gradient, residual_rms

    nx, ny, h, dh = 30, 10, 0.06, 1e-4
    p0_ref = (1.0 + 0.2 * 0.675**2) ** 3.5
    base = torch.as_tensor(meshgen.rectangle_mesh(nx, ny, (0.0, 3.0, 0.0, 1.0)).points)

    def converge(hh):
        case, prim = meshgen.bump_case(nx, ny, hh, integrator="Explicit RK3", cfl=0.8)
        case.numerics.reconstruction = "First order"
        dm = DeviceMesh(case, cuda_device, fused=False)
        g = 1.4
        rho, uu, vv, p = prim.T
        dm.set_state(np.column_stack((rho, rho * uu, rho * vv, p / (g - 1) + 0.5 * rho * (uu * uu + vv * vv))))
        rms = residual_rms(dm)
        for chunk in range(40):
            if rms <= 1e-12:
                break
            dm.run(500)
            rms = residual_rms(dm)
            print(f"  h={hh} steps {500 * (chunk + 1)} rms {rms:.3e}", flush=True)
        return dm, rms

    print("case: bump outlet loss", flush=True)
    dm, rms = converge(h)
    assert rms <= 1e-12, rms
    obj = OutletTotalPressureLoss(dm, p0_ref)
    j0, dj = objective_gradient(obj, dm.u)
    adj = SteadyAdjoint(dm)
    lam, hist = adj.solve(dj, tol=1e-11, max_iterations=80)
    print("adjoint residual history", hist[-3:], flush=True)
    assert hist[-1] <= 1e-10, hist[-1]
    d = shape_gradient(dm, dm.u, lam, lambda p: meshgen.bump_nodes(base, p, xp=torch), h)
    vals = []
    for sgn in (1, -1):
        dmp, rmsp = converge(h + sgn * dh)
        assert rmsp <= 1e-12
        vals.append(float(OutletTotalPressureLoss(dmp, p0_ref)(dmp.u).item()))
        dmp.close()
    fd = (vals[0] - vals[1]) / (2 * dh)
    print("J", j0, "adjoint dJ/dh", d, "fd", fd)
    assert abs(d - fd) <= 1e-5 * abs(fd)
    dm.close()


@pytest.mark.gpu
def test_gpu_outlet_loss_gradient_on_an_adapted_bump_matches_fd(cuda_device):
    """The bump channel adapted twice from its own converged density field
    (gradient indicator, threshold 0.05, hanging-node subfaces, bump faces
    split on the arc): d(outlet total-pressure loss)/dh by the adjoint through
    the face normals and areas, for the adapted mesh's nodes as functions of
    h (the same marks replayed on the channel at h +- 1e-6 give dP/dh),
    against central differences of solutions re-converged on the adapted
    meshes at h +- 1e-4 (first-order HLLC, implicit pseudo-time): 1e-6
    (measured 2e-7)."""
    from meshcfd_amd import adaptation as ad
    from meshcfd_amd.device import DeviceMesh
    from meshcfd_amd.geometry_adjoint import shape_gradient
    from meshcfd_amd.solver import conservative_to_primitive, primitive_to_conservative
    from meshcfd_amd.steady_adjoint import OutletTotalPressureLoss, SteadyAdjoint, objective_gradient, solve_steady

    nx, ny, h, dh = 30, 10, 0.06, 1e-4
    p0_ref = (1.0 + 0.2 * 0.675**2) ** 3.5

    def bump(hh, mesh=None):
        case, prim = meshgen.bump_case(nx, ny, hh, integrator="Implicit Euler", cfl=100.0)
        case.numerics.reconstruction = "First order"
        if mesh is not None:
            case.mesh = mesh
            prim = np.tile(prim[0], (mesh.cell_count, 1))
        return case, prim

    def converge(case, prim, warm=None, tol=1e-12):
        dm = DeviceMesh(case, cuda_device, block_cells=64)
        dm.set_state(primitive_to_conservative(prim, 1.4) if warm is None else warm)
        hist = solve_steady(dm, tol=tol, max_steps=600, raise_cfl=2.0)
        assert hist[-1] <= tol, hist[-3:]
        return dm

    print("case: adapted bump outlet loss gradient", flush=True)
    case, prim = bump(h)
    dm = converge(case, prim, tol=1e-8)
    plan = []
    for _ in range(2):
        g = dm.geometry
        u = dm.state()
        marks, coarsen = ad.select_adaptation(case.mesh, case, g.centers, g.areas,
                                              conservative_to_primitive(u, 1.4)[:, 0], 0.05, 2)
        new, groups = ad.adapt_mesh(case.mesh, marks, coarsen)
        u = ad.transfer_cell_values(u, ad.cell_areas(case.mesh), ad.cell_areas(new), groups)
        plan.append((marks, coarsen))
        dm.close()
        case.mesh = new
        dm = converge(case, prim, warm=u, tol=1e-8)
    dm.close()

    def replay(hh):
        m = meshgen.bump_channel(nx, ny, hh)
        for marks, coarsen in plan:
            m, _ = ad.adapt_mesh(m, marks, coarsen)
        return m

    mesh_h = replay(h)
    assert np.array_equal(mesh_h.points, case.mesh.points) and mesh_h.cell_sizes.max() >= 5
    dm = converge(*bump(h, mesh_h))
    u_star = dm.state()
    j0, dj = objective_gradient(OutletTotalPressureLoss(dm, p0_ref), dm.u)
    lam, hist = SteadyAdjoint(dm).solve(dj, tol=1e-11, max_iterations=80)
    assert hist[-1] <= 1e-10, hist[-1]
    eps = 1e-6
    mp, mm = replay(h + eps), replay(h - eps)
    for key in ("faces", "face_owners", "face_neighbours", "cells"):
        assert np.array_equal(getattr(mp, key), getattr(mm, key))
    p_h = torch.as_tensor(mesh_h.points)
    dp_dh = torch.as_tensor((mp.points - mm.points) / (2 * eps))
    d = shape_gradient(dm, dm.u, lam, lambda p: p_h + (p - h) * dp_dh, h)
    dm.close()
    vals = []
    for sgn in (1, -1):
        d2 = converge(*bump(h + sgn * dh, replay(h + sgn * dh)), warm=u_star)
        vals.append(float(OutletTotalPressureLoss(d2, p0_ref)(d2.u).item()))
        d2.close()
    fd = (vals[0] - vals[1]) / (2 * dh)
    print("cells", mesh_h.cell_count, "J", j0, "adjoint dJ/dh", d, "fd", fd)
    assert abs(d - fd) <= 1e-6 * abs(fd)
