"""Steady adjoint: O-grid generator, transposed LU-SGS, adjoint operator,
adjoint GMRES and inflow gradients of wall forces.

The adjoint has no counterpart in the reference. It is pinned by
identities: the transposed LU-SGS against the oracle's dense-factor
transpose and the dot-product test <M^T a, b> = <a, M b>; the adjoint
operator by <A^T v, w> = <v, A w> with A w from central differences of the
device residual; the adjoint GMRES by its true residual; and the inflow
Mach/angle gradients of the wall force against central finite differences
of re-converged steady states.
"""

import math

import numpy as np
import pytest

from meshcfd_amd import meshgen
from meshcfd_amd.geometry import build_geometry
from meshcfd_amd.model import NumericalSettings
from oracle import implicit as oimp
from oracle.euler import primitive_to_conservative


def test_geometric_radii_ratio_and_ends():
    r = meshgen.geometric_radii(64, 0.5, 30.0, 1e-3)
    d = np.diff(r)
    assert r[0] == 0.5 and r[-1] == 30.0
    assert d[0] == pytest.approx(1e-3, rel=1e-12)
    np.testing.assert_allclose(d[1:-1] / d[:-2], d[1] / d[0], rtol=1e-9)
    np.testing.assert_allclose(meshgen.geometric_radii(4, 0.0, 1.0, 0.5), [0, 0.25, 0.5, 0.75, 1.0])


def test_ogrid_covers_the_annulus_polygon():
    nt, nr = 40, 12
    m = meshgen.cylinder_ogrid(nt, nr, 0.5, 10.0, 0.02, 0.05, 3)
    case, _ = meshgen.cylinder_case(nt, nr, reynolds=None, r_outer=10.0, first_spacing=0.02, seed=3)
    g = build_geometry(case)
    assert m.cell_count == nt * nr
    assert np.all(g.areas > 0.0)
    ring = 0.5 * nt * math.sin(2 * math.pi / nt) * (10.0**2 - 0.5**2)  # polygon annulus
    assert g.areas.sum() == pytest.approx(ring, rel=1e-12)
    assert len(g.wall_faces) == nt and len(g.far_field_faces) == nt
    assert len(g.interior_faces) == 2 * nt * nr - nt
    # wall normals point into the cylinder, far-field normals outward
    w = g.wall_faces
    assert np.all(np.sum(g.face_normals[w] * g.face_centers[w], axis=1) < 0)
    f = g.far_field_faces
    assert np.all(np.sum(g.face_normals[f] * g.face_centers[f], axis=1) > 0)


def test_oracle_lusgs_transpose_is_the_adjoint():
    case, prim = meshgen.cylinder_case(16, 6, reynolds=None, r_outer=5.0, first_spacing=0.05, seed=1)
    g = build_geometry(case)
    rng = np.random.default_rng(0)
    prim = prim * (1.0 + 0.05 * rng.random(prim.shape))
    h = 0.02 + 0.01 * rng.random(g.cell_count)
    S, _ = oimp.lusgs_matrix(prim, g, 1.4, h)
    a = rng.normal(size=4 * g.cell_count)
    b = rng.normal(size=4 * g.cell_count)
    lhs = np.dot(oimp.lusgs_apply_transpose(S, a), b)
    rhs = np.dot(a, oimp.lusgs_apply(S, b))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_free_stream_derivatives_by_difference():
    from meshcfd_amd.steady_adjoint import free_stream_derivatives

    dm_, da_ = free_stream_derivatives(0.3, 7.0, 1.4, 1.0, 1.0 / 1.4)
    h = 1e-6
    fd_m = (np.array(meshgen.free_stream(0.3 + h, 7.0)) - np.array(meshgen.free_stream(0.3 - h, 7.0))) / (2 * h)
    hd = math.degrees(h)
    fd_a = (np.array(meshgen.free_stream(0.3, 7.0 + hd)) - np.array(meshgen.free_stream(0.3, 7.0 - hd))) / (2 * h)
    np.testing.assert_allclose(dm_, fd_m, atol=1e-8)
    np.testing.assert_allclose(da_, fd_a, atol=1e-8)


def _inviscid_cylinder(n_theta=48, n_r=24, mach=0.3, alpha=5.0, reconstruction="First order",
                       limiter="Barth-Jespersen"):
    numerics = NumericalSettings(
        flux="HLLC", reconstruction=reconstruction, limiter=limiter, time_integration="Implicit Euler",
        time_step_mode="Local (steady)", cfl=200.0, gmres_tolerance=1e-3, gmres_krylov_dimension=30,
        residual_tolerance=1e-12, max_iterations=400, end_time=1e30,
    )
    return meshgen.cylinder_case(n_theta, n_r, mach=mach, alpha_deg=alpha, reynolds=None, r_outer=10.0,
                                 first_spacing=0.02, seed=5, numerics=numerics)


@pytest.mark.gpu
@pytest.mark.parametrize("per_cell", [False, True])
def test_gpu_lusgs_transpose_matches_oracle(cuda_device, per_cell):
    import torch

    from meshcfd_amd.device import DeviceMesh
    from meshcfd_amd.implicit import ImplicitStepper

    case, prim = _inviscid_cylinder(24, 10)
    print("case: cylinder 24x10 transposed LU-SGS", flush=True)
    dm = DeviceMesh(case, cuda_device, fused=True, block_cells=64)
    st = ImplicitStepper(dm)
    rng = np.random.default_rng(4)
    prim = prim * (1.0 + 0.05 * rng.random(prim.shape))
    u = primitive_to_conservative(prim, 1.4)
    n = dm.n
    h = 0.01 + 0.04 * rng.random(n) if per_cell else 0.03
    S, _ = oimp.lusgs_matrix(prim, dm.geometry, 1.4, h)
    v = rng.normal(size=u.shape)
    want = oimp.lusgs_apply_transpose(S, v.ravel()).reshape(v.shape)
    dt = dm.to_device_order(h) if per_cell else torch.full((1,), h, dtype=torch.float64, device=cuda_device)
    st.assemble(dm.to_device_order(u), dt, per_cell)
    out = torch.empty((n, 4), dtype=torch.float64, device=cuda_device)
    import ctypes as C

    dm.lib.mcfd_lusgs_apply_transpose(dm.handle, C.c_void_p(dm.to_device_order(v).data_ptr()),
                                      C.c_void_p(out.data_ptr()), dm.stream)
    dm.check_status()
    np.testing.assert_allclose(dm.to_mesh_order(out), want, rtol=1e-12, atol=1e-12)
    dm.close()


def _device_state(reconstruction="MUSCL linear", limiter="Venkatakrishnan"):
    from meshcfd_amd.device import DeviceMesh

    case, prim = _inviscid_cylinder(32, 12, reconstruction=reconstruction, limiter=limiter)
    rng = np.random.default_rng(7)
    prim = prim * (1.0 + 0.03 * rng.random(prim.shape))
    dm = DeviceMesh(case, "cuda:0", fused=True, block_cells=64)
    dm.set_state(primitive_to_conservative(prim, 1.4))
    return dm


@pytest.mark.gpu
@pytest.mark.parametrize("reconstruction", ["First order", "MUSCL linear"])
def test_gpu_adjoint_operator_dot_product(cuda_device, reconstruction):
    """<A^T v, w> = <v, A w>, A w = w - D (R(u + e w) - R(u - e w)) / 2e.
    The limited MUSCL residual is only piecewise smooth (min/max and limiter
    branches), so the difference step is small enough (1e-8 relative) not to
    cross a branch on this noisy state."""
    import torch

    from meshcfd_amd.steady_adjoint import SteadyAdjoint

    print(f"case: adjoint operator dot product {reconstruction}", flush=True)
    dm = _device_state(reconstruction)
    sa = SteadyAdjoint(dm, cfl=20.0)
    u = dm.u.clone()
    gen = torch.Generator(device="cpu").manual_seed(3)
    v = torch.randn((dm.n, 4), generator=gen, dtype=torch.float64).to(cuda_device)
    w = torch.randn((dm.n, 4), generator=gen, dtype=torch.float64).to(cuda_device)
    at_v = sa.transpose_apply(u, v)
    dt = sa.dt.clone()
    eps = 1e-8 * float(torch.linalg.norm(u).item()) / float(torch.linalg.norm(w).item())
    jw = (dm.residual(u + eps * w).clone() - dm.residual(u - eps * w).clone()) / (2 * eps)
    a_w = w - dt[:, None] * jw
    lhs = float(torch.sum(at_v * w).item())
    rhs = float(torch.sum(v * a_w).item())
    print("dot product", lhs, rhs)
    assert lhs == pytest.approx(rhs, rel=1e-6)
    dm.close()


@pytest.mark.gpu
def test_gpu_adjoint_gmres_true_residual(cuda_device):
    import torch

    from meshcfd_amd.steady_adjoint import SteadyAdjoint

    print("case: adjoint GMRES true residual", flush=True)
    dm = _device_state()
    sa = SteadyAdjoint(dm, cfl=20.0, rtol=1e-8)
    u = dm.u.clone()
    gen = torch.Generator(device="cpu").manual_seed(5)
    b = torch.randn((dm.n, 4), generator=gen, dtype=torch.float64).to(cuda_device)
    sa._time_steps(u)
    sa.stepper.assemble(u, sa.dt, True)
    import ctypes as C

    from meshcfd_amd import native
    from meshcfd_amd.implicit import McfdGmres

    x = torch.empty_like(b)
    g = McfdGmres()
    g.state, g.rhs, g.x, g.dt = (C.c_void_p(t.data_ptr()) for t in (u, b, x, sa.dt))
    g.dt_per_cell, g.krylov, g.max_iterations, g.tolerance = 1, 30, 600, 1e-8
    native.check(dm.lib.mcfd_adjoint_gmres(dm.handle, C.byref(sa.settings), C.byref(g), dm.stream))
    assert g.info == 0, (g.iterations, g.residual)
    r = sa.transpose_apply(u, x) - b
    rel = float(torch.linalg.norm(r).item() / torch.linalg.norm(b).item())
    print("adjoint gmres", g.iterations, rel)
    assert rel <= 1.5e-8
    dm.close()


def _laminar_cylinder(n_theta=32, n_r=16, mach=0.3, alpha=5.0):
    """Unit free stream (rho = 1, c = 1, gas constant 1) past the unit
    cylinder with constant viscosity mu = 0.3 / 40 (Re 40 at M 0.3), an
    isothermal no-slip wall at the free-stream temperature; HLLC first
    order, implicit Euler, local steps, CFL 200."""
    from meshcfd_amd.model import BoundaryCondition

    numerics = NumericalSettings(
        flux="HLLC", reconstruction="First order", time_integration="Implicit Euler",
        time_step_mode="Local (steady)", cfl=200.0, gmres_tolerance=1e-3, gmres_krylov_dimension=30,
        residual_tolerance=1e-12, max_iterations=400, end_time=1e30,
        viscosity_model="Constant", dynamic_viscosity=0.3 / 40.0, gas_constant=1.0,
    )
    case, prim = meshgen.cylinder_case(n_theta, n_r, mach=mach, alpha_deg=alpha, reynolds=None, r_outer=10.0,
                                       first_spacing=0.02, seed=5, numerics=numerics)
    t_inf = float(prim[0, 3] / prim[0, 0])
    case.boundary_conditions[0] = BoundaryCondition(0, "wall", "Wall", {"temperature": t_inf}, "Temperature",
                                                    "No-slip")
    return case, prim


def _sutherland_cylinder(n_theta=32, n_r=16, mach=0.2, alpha=5.0, reconstruction="First order"):
    """The laminar cylinder's physics in free-stream units (rho = 1, c = 1,
    gas constant 1, T = 1/gamma): Roe flux, Sutherland viscosity with air's
    S / T ratio (110.4 K / 288.15 K) and mu = 0.2 / 40 at the free-stream
    temperature (Re 40 at M 0.2), Pr 0.72, an isothermal no-slip wall at the
    free-stream temperature; implicit Euler, local steps, CFL 200. Mach
    changes at fixed free-stream density and pressure."""
    from meshcfd_amd.model import BoundaryCondition

    numerics = NumericalSettings(
        flux="Roe", reconstruction=reconstruction, time_integration="Implicit Euler",
        time_step_mode="Local (steady)", cfl=200.0, gmres_tolerance=1e-3, gmres_krylov_dimension=30,
        residual_tolerance=1e-12, max_iterations=400, end_time=1e30, viscosity_model="Sutherland",
        prandtl_number=0.72, gas_constant=1.0, sutherland_reference_temperature=1.0 / 1.4,
        sutherland_temperature=110.4 / 288.15 / 1.4, sutherland_reference_viscosity=0.2 / 40.0,
    )
    case, prim = meshgen.cylinder_case(n_theta, n_r, mach=mach, alpha_deg=alpha, reynolds=None, r_outer=10.0,
                                       first_spacing=0.02, seed=5, numerics=numerics)
    case.boundary_conditions[0] = BoundaryCondition(0, "wall", "Wall", {"temperature": 1.0 / 1.4}, "Temperature",
                                                    "No-slip")
    return case, prim


def _steady_force(mach, alpha, warm=None, lift=False, laminar=False, sutherland=False, tighten=False):
    from meshcfd_amd.device import DeviceMesh
    from meshcfd_amd.steady_adjoint import WallForce, solve_steady

    if sutherland:
        case, prim = _sutherland_cylinder(32, 16, mach=mach, alpha=alpha)
    elif laminar:
        case, prim = _laminar_cylinder(32, 16, mach=mach, alpha=alpha)
    else:
        case, prim = _inviscid_cylinder(32, 16, mach=mach, alpha=alpha)
    dm = DeviceMesh(case, "cuda:0", fused=True, block_cells=64)
    dm.set_state(primitive_to_conservative(prim, 1.4) if warm is None else warm)

    def progress(k, rms, cfl):
        if k % 10 == 0:
            print(f"  steady M={mach:.6f} alpha={alpha:.6f}: step {k} rms {rms:.3e} cfl {cfl:g}", flush=True)

    # the reference's Newton test (max|U - U^n - dt R| <= 1e-9) floors the
    # residual near 1e-10 at CFL 200
    hist = solve_steady(dm, tol=2e-10, max_steps=300, progress=progress)
    assert hist[-1] <= 2e-10, hist[-5:]
    if tighten:
        # then raise the CFL number whenever a step stalls, so the last steps
        # are Newton steps on R = 0 (solve_steady's raise_cfl): the laminar
        # case converges to ~1e-14
        hist = hist + solve_steady(dm, tol=1e-14, max_steps=40, raise_cfl=2.0, max_cfl=1e6, progress=progress)
    force = WallForce(dm, alpha, lift=lift)
    return dm, force, float(force(dm.u).item()), hist


@pytest.mark.gpu
@pytest.mark.parametrize("lift", [False, True])
def test_gpu_wall_force_inflow_gradients_match_fd(cuda_device, lift):
    """Inviscid cylinder, HLLC first order, M 0.3, alpha 5 deg, 32x16 O-grid:
    d(force)/dM and d(force)/d(alpha) from one adjoint solve equal central
    differences of re-converged steady states (step 5e-4; steady residual
    2e-10): drag to 1e-6 of the larger derivative (measured 4.3e-7 and
    1.9e-8), lift to 5e-4."""
    from meshcfd_amd.steady_adjoint import SteadyAdjoint, free_stream_derivatives, objective_gradient

    print(f"case: cylinder wall force gradients lift={lift}", flush=True)
    mach, alpha = 0.3, 5.0
    dm, force, j0, hist = _steady_force(mach, alpha, lift=lift)
    print("steady", len(hist), "steps, rms", hist[-1], "force", j0)
    u_star = dm.state()
    sa = SteadyAdjoint(dm, cfl=1e3, rtol=1e-4)
    _, dj_du = objective_gradient(force, dm.u)
    lam, ahist = sa.solve(dj_du, tol=1e-11, max_iterations=60,
                          progress=lambda k, rel: print(f"  adjoint iteration {k}: {rel:.3e}", flush=True))
    assert ahist[-1] <= 1e-11, ahist[-5:]
    sbar = sa.inflow_gradient(lam)
    d_mach, d_alpha = free_stream_derivatives(mach, alpha, 1.4, 1.0, 1.0 / 1.4)
    g_mach = float(sbar @ d_mach)
    g_alpha = float(sbar @ d_alpha) + force.d_alpha(dm.u)
    print("adjoint", g_mach, g_alpha, "iterations", len(ahist), "gmres", sa.gmres_iterations)
    dm.close()
    h = 5e-4
    fd = []
    for dmach, dalpha in ((h, 0.0), (-h, 0.0), (0.0, math.degrees(h)), (0.0, -math.degrees(h))):
        d2, _, j, _ = _steady_force(mach + dmach, alpha + dalpha, warm=u_star, lift=lift)
        d2.close()
        fd.append(j)
    fd_mach = (fd[0] - fd[1]) / (2 * h)
    fd_alpha = (fd[2] - fd[3]) / (2 * h)
    print("fd", fd_mach, fd_alpha)
    # The FD error is absolute (steady residual floor / step, plus the
    # step's truncation), so both derivatives are held to a fraction of the
    # larger one: 1e-6 for drag. Lift measures the circulation, the one
    # steady mode inviscid flow damps only through numerical dissipation; the
    # residual floor leaves it least converged (and Newton steps at a raised
    # CFL do not converge with HLLC's branch switches), so its differences
    # are good to 5e-4 only (1.2e-4 at step 1e-3, 1.3e-3 at 1e-4).
    scale = max(abs(fd_mach), abs(fd_alpha))
    frac = 5e-4 if lift else 1e-6
    assert abs(g_mach - fd_mach) <= frac * scale
    assert abs(g_alpha - fd_alpha) <= frac * scale


@pytest.mark.gpu
def test_gpu_laminar_drag_gradients_match_fd(cuda_device):
    """Laminar cylinder (Re 40 at M 0.3, alpha 5 deg, 32x16 O-grid, HLLC first
    order): the drag includes the viscous wall traction; d(drag)/dM and
    d(drag)/d(alpha) from one adjoint solve (viscous reverse kernels in the
    adjoint operator) equal central differences of re-converged steady
    states to 1e-7 of the larger derivative (measured 5e-9 and 1.6e-9): the
    states are converged to ~1e-14 by Newton steps at a raised CFL, so the
    differences can take a step of 1e-4."""
    from meshcfd_amd.steady_adjoint import SteadyAdjoint, free_stream_derivatives, objective_gradient

    print("case: laminar cylinder drag gradients", flush=True)
    mach, alpha = 0.3, 5.0
    dm, force, j0, hist = _steady_force(mach, alpha, laminar=True, tighten=True)
    fp, fv = force.components(dm.u)
    print("steady", len(hist), "steps, rms", hist[-1], "drag", j0, "pressure", fp.tolist(), "viscous", fv.tolist())
    assert abs(float(fv[0])) > 0.1 * abs(j0)
    u_star = dm.state()
    sa = SteadyAdjoint(dm, cfl=1e3, rtol=1e-4)
    _, dj_du = objective_gradient(force, dm.u)
    lam, ahist = sa.solve(dj_du, tol=1e-12, max_iterations=100,
                          progress=lambda k, rel: print(f"  adjoint iteration {k}: {rel:.3e}", flush=True))
    assert ahist[-1] <= 1e-12, ahist[-5:]
    sbar = sa.inflow_gradient(lam)
    d_mach, d_alpha = free_stream_derivatives(mach, alpha, 1.4, 1.0, 1.0 / 1.4)
    g_mach = float(sbar @ d_mach)
    g_alpha = float(sbar @ d_alpha) + force.d_alpha(dm.u)
    print("adjoint", g_mach, g_alpha, "iterations", len(ahist), "gmres", sa.gmres_iterations)
    dm.close()
    h = 1e-4
    fd = []
    for dmach, dalpha in ((h, 0.0), (-h, 0.0), (0.0, math.degrees(h)), (0.0, -math.degrees(h))):
        d2, _, j, _ = _steady_force(mach + dmach, alpha + dalpha, warm=u_star, laminar=True, tighten=True)
        d2.close()
        fd.append(j)
    fd_mach = (fd[0] - fd[1]) / (2 * h)
    fd_alpha = (fd[2] - fd[3]) / (2 * h)
    scale = max(abs(fd_mach), abs(fd_alpha))
    print("fd", fd_mach, fd_alpha, "rel", abs(g_mach - fd_mach) / scale, abs(g_alpha - fd_alpha) / scale)
    assert abs(g_mach - fd_mach) <= 1e-7 * scale
    assert abs(g_alpha - fd_alpha) <= 1e-7 * scale


@pytest.mark.gpu
def test_gpu_roe_sutherland_drag_gradients_match_fd(cuda_device):
    """The cylinder at the laminar workload's physics (Roe, Sutherland
    viscosity, Pr 0.72, M 0.2, Re 40, alpha 5 deg; 32x16 O-grid, first-order
    reconstruction): d(drag)/dM and d(drag)/d(alpha) from one adjoint solve
    (Roe and Sutherland reverse kernels) equal central differences of
    re-converged steady states (step 5e-4) to 1e-6 of the larger derivative
    (measured: 4.5e-7 for d/dM, 2.9e-8 for d/d(alpha); the Roe flux's
    entropy fix keeps Newton steps at a raised CFL from converging, so the
    states stay at the 2e-10 floor)."""
    from meshcfd_amd.steady_adjoint import SteadyAdjoint, free_stream_derivatives, objective_gradient

    print("case: Roe + Sutherland cylinder drag gradients, M 0.2 Re 40", flush=True)
    mach, alpha = 0.2, 5.0
    dm, force, j0, hist = _steady_force(mach, alpha, sutherland=True)
    fp, fv = force.components(dm.u)
    print("steady", len(hist), "steps, rms", hist[-1], "drag", j0, "pressure", fp.tolist(), "viscous", fv.tolist())
    assert abs(float(fv[0])) > 0.1 * abs(j0)
    u_star = dm.state()
    sa = SteadyAdjoint(dm, cfl=1e3, rtol=1e-4)
    _, dj_du = objective_gradient(force, dm.u)
    lam, ahist = sa.solve(dj_du, tol=1e-11, max_iterations=80,
                          progress=lambda k, rel: print(f"  adjoint iteration {k}: {rel:.3e}", flush=True))
    assert ahist[-1] <= 1e-11, ahist[-5:]
    sbar = sa.inflow_gradient(lam)
    d_mach, d_alpha = free_stream_derivatives(mach, alpha, 1.4, 1.0, 1.0 / 1.4)
    g_mach = float(sbar @ d_mach)
    g_alpha = float(sbar @ d_alpha) + force.d_alpha(dm.u)
    print("adjoint", g_mach, g_alpha, "iterations", len(ahist), "gmres", sa.gmres_iterations)
    dm.close()
    h = 5e-4
    fd = []
    for dmach, dalpha in ((h, 0.0), (-h, 0.0), (0.0, math.degrees(h)), (0.0, -math.degrees(h))):
        d2, _, j, _ = _steady_force(mach + dmach, alpha + dalpha, warm=u_star, sutherland=True)
        d2.close()
        fd.append(j)
    fd_mach = (fd[0] - fd[1]) / (2 * h)
    fd_alpha = (fd[2] - fd[3]) / (2 * h)
    print("fd", fd_mach, fd_alpha, "rel", abs(g_mach - fd_mach) / max(abs(fd_mach), abs(fd_alpha)),
          abs(g_alpha - fd_alpha) / max(abs(fd_mach), abs(fd_alpha)))
    scale = max(abs(fd_mach), abs(fd_alpha))
    assert abs(g_mach - fd_mach) <= 1e-6 * scale
    assert abs(g_alpha - fd_alpha) <= 1e-6 * scale


@pytest.mark.gpu
def test_gpu_laminar_physics_converged_state_is_the_oracles(cuda_device):
    """The laminar workload's physics (Roe, Sutherland, Pr 0.72, M 0.2, Re
    40, first order; 32x16 O-grid, alpha 5 deg) converged on the GPU to the
    reference Newton loop's residual floor (RMS 2e-10) is the oracle's
    steady state: the oracle's residual there is at the same floor, and one
    Newton step with the oracle's own Jacobian (dense, forward differences of
    `Oracle.residual`) lands on R_oracle = 0 (its residual drops below
    1e-14) after moving the state by at most 5e-9 of its largest entry
    (measured: oracle residual 1.236e-10 at the GPU state, the GPU's own to
    four digits; 3.0e-15 after the step, which moves the state 5.7e-10 of
    its largest entry: the floor times the steady Jacobian's inverse, the
    GPU/oracle residual difference being ~1e-15). The MUSCL-parabolic
    reconstruction has no converged state on this grid with the reference's
    first-order Jacobian (the limited residual stalls near 0.1), so the
    comparison uses first order."""
    from oracle.solver import Oracle

    print("case: laminar-physics cylinder converged state vs oracle", flush=True)
    dm, _, _, hist = _steady_force(0.2, 5.0, sutherland=True)
    u = dm.state()
    case, _ = _sutherland_cylinder(32, 16, mach=0.2, alpha=5.0)
    o = Oracle(dm.geometry, case.numerics)
    dm.close()
    r0 = o.residual(u)
    rms0 = math.sqrt(float(np.mean(r0**2)))
    print(f"GPU steady RMS {hist[-1]:.3e}, oracle residual RMS at the GPU state {rms0:.3e}", flush=True)
    assert rms0 <= 2.5e-10
    flat = u.ravel()
    n = flat.size
    jac = np.empty((n, n))
    for j in range(n):
        h = 1e-7 * max(abs(flat[j]), 1.0)
        up = flat.copy()
        up[j] += h
        jac[:, j] = (o.residual(up.reshape(u.shape)).ravel() - r0.ravel()) / h
        if j % 512 == 511:
            print(f"  oracle Jacobian column {j + 1}/{n}", flush=True)
    delta = np.linalg.solve(jac, -r0.ravel())
    rms1 = math.sqrt(float(np.mean(o.residual((flat + delta).reshape(u.shape)) ** 2)))
    move = float(np.abs(delta).max() / np.abs(flat).max())
    print(f"oracle Newton step: residual RMS {rms0:.3e} -> {rms1:.3e}, state moved {move:.3e}", flush=True)
    assert rms1 <= 1e-14
    assert move <= 5e-9
