"""Laminar viscous terms: device lists (CPU) and GPU parity with the oracle.

The CPU tests replay the device algorithm over the lists `build_viscous`
produces (a loop restatement of k_visc_gradient / k_visc_flux / the viscous
gather) and compare it with the oracle's `viscous_residual`, so list order
and sign mistakes fail before a GPU run. The GPU tests compare residuals,
stable time steps and steps with the oracle. Tolerance: per component,
max |gpu - ref| / max |ref| <= 1e-12 for residuals and time steps (the
reference's 2-vector np.dot and pow(T, 1.5) round independently of the
device's), 1e-10 after explicit steps and 1e-8 after an implicit step.
"""

import numpy as np
import pytest

from meshcfd_amd import meshgen
from meshcfd_amd.geometry import build_geometry
from meshcfd_amd.layout import build_layout
from meshcfd_amd.model import BoundaryCondition, Case, NumericalSettings
from meshcfd_amd.viscous import build_viscous
from oracle import euler
from oracle.solver import Oracle
from oracle.viscous import viscous_residual, viscous_time_step_sum

CASES = ("periodic", "channel", "farfield", "sutherland")


def viscous_case(kind, n=10, integrator="Explicit RK3", reconstruction="MUSCL linear", flux="HLLC"):
    """Unit-scale gas (R = 1, rho ~ p ~ T ~ 1, mu = 0.02) so viscous and
    convective terms are of one size; `sutherland` uses air in a 20 um box."""
    common = dict(flux=flux, reconstruction=reconstruction, time_integration=integrator, cfl=0.5, end_time=1e30)
    if kind == "sutherland":
        numerics = NumericalSettings(viscosity_model="Sutherland", gas_constant=287.05, **common)
        bounds = (0.0, 2e-5, 0.0, 2e-5)
    else:
        numerics = NumericalSettings(viscosity_model="Constant", dynamic_viscosity=0.02, gas_constant=1.0,
                                     **common)
        bounds = (0.0, 1.0, 0.0, 1.0)
    mesh = meshgen.rectangle_mesh(n, n, bounds, jitter=0.2, seed=8, split_fraction=0.3)
    if kind == "periodic":
        return meshgen.periodic_case(mesh, numerics)
    # patches: bottom 0, right 1, top 2, left 3
    if kind == "channel":
        bcs = [
            BoundaryCondition(0, "bottom", "Wall", {"temperature": 1.2}, "Temperature", "No-slip"),
            BoundaryCondition(1, "right", "Pressure Outlet", {"pressure": 0.95}),
            BoundaryCondition(2, "top", "Wall", {"heat_flux": 0.3}, "Heat Flux", "Slip"),
            BoundaryCondition(3, "left", "Velocity Inlet",
                              {"density": 1.0, "temperature": 1.0, "velocity_x": 0.4, "velocity_y": 0.05}),
        ]
    elif kind == "farfield":
        ff = {"density": 1.0, "velocity_x": 0.35, "velocity_y": 0.05, "pressure": 1.0}
        bcs = [
            BoundaryCondition(0, "bottom", "Symmetry"),
            BoundaryCondition(1, "right", "Transmissive"),
            BoundaryCondition(2, "top", "Far Field", dict(ff)),
            BoundaryCondition(3, "left", "Far Field", dict(ff)),
        ]
    else:
        bcs = [
            BoundaryCondition(0, "bottom", "Wall", {"heat_flux": -500.0}, "Heat Flux", "No-slip"),
            BoundaryCondition(1, "right", "Pressure Outlet", {"pressure": 1.0e5}),
            BoundaryCondition(2, "top", "Wall", {"temperature": 310.0}, "Temperature", "Slip"),
            BoundaryCondition(3, "left", "Velocity Inlet",
                              {"density": 1.2, "temperature": 290.0, "velocity_x": 30.0, "velocity_y": 0.0}),
        ]
    return Case(mesh=mesh, boundary_conditions=bcs, numerics=numerics, name=kind)


def viscous_primitive(case, seed=4):
    g = build_geometry(case)
    x, y = g.centers[:, 0], g.centers[:, 1]
    rng = np.random.default_rng(seed)
    n = len(x)
    if case.name == "sutherland":
        s = 2 * np.pi / 2e-5
        rho = 1.2 * (1 + 0.05 * np.sin(s * x) * np.cos(s * y)) + 0.01 * rng.random(n)
        u = 30.0 * (1 + 0.2 * np.cos(s * y)) + rng.random(n)
        v = 3.0 * np.sin(s * x)
        p = 1.0e5 * (1 + 0.02 * np.cos(s * (x - y))) + 100 * rng.random(n)
    else:
        t = 2 * np.pi
        rho = 1.0 + 0.15 * np.sin(t * x) * np.cos(t * y) + 0.05 * rng.random(n)
        u = 0.35 + 0.2 * np.cos(t * y) + 0.05 * rng.random(n)
        v = 0.05 + 0.15 * np.sin(t * x)
        p = 1.0 + 0.1 * np.cos(t * (x - y)) + 0.05 * rng.random(n)
    # v alone is a function of x only: the two triangles of a split quad
    # share their centre's x, and equal values put the limiter on its kink
    # (max/min neighbour ties), which a central difference cannot resolve
    v = v + 0.01 * np.max(np.abs(v)) * rng.random(n)
    return np.column_stack((rho, u, v, p))


# --- CPU: replay of the device algorithm over the lists --------------------


def _mu(T, s):
    if s.viscosity_model == "Constant":
        return s.dynamic_viscosity
    tr = s.sutherland_reference_temperature
    mu = s.sutherland_reference_viscosity * (T / tr) ** 1.5
    return mu * ((tr + s.sutherland_temperature) / (T + s.sutherland_temperature))


def _k(mu, s):
    return mu * s.gamma * s.gas_constant / ((s.gamma - 1.0) * s.prandtl_number)


def _inlet(code, w, st, n, s):
    from oracle.solver import far_field_states

    x = np.array(st, dtype=float)
    if code == 5:
        x = far_field_states(w[None], x[None], n[None], s.gamma)[0]
    return np.array((x[1], x[2], x[3] / (x[0] * s.gas_constant)))


def replay_viscous_residual(g, L, v, W, s):
    """Device-order viscous residual from the lists (loop form of the kernels)."""
    n = L.n_cells
    R = s.gas_constant
    T = W[:, 3] / (W[:, 0] * R)
    vals = np.column_stack((W[:, 1], W[:, 2], T))
    VG = np.zeros((n, 8))
    for i in range(n):
        rx = np.zeros(3)
        ry = np.zeros(3)
        for e in range(L.grad_ptr[i], L.grad_ptr[i + 1]):
            d = vals[L.grad_nb[e]] - vals[i]
            rx = rx + v["wl_wx"][e] * d
            ry = ry + v["wl_wy"][e] * d
        mu = _mu(T[i], s)
        kc = _k(mu, s)
        for e in range(v["gh_ptr"][i], v["gh_ptr"][i + 1]):
            f = v["gh_face"][e]
            b = L.face_bc[f]
            code = v["bv_code"][b]
            nrm = np.array((L.face_nx[f], L.face_ny[f]))
            gh = vals[i].copy()
            bg = v["bv_geom"][4 * b:4 * b + 4]
            if code == 1:
                flags = v["bv_flags"][b]
                if flags & 1:
                    gh[:2] = 2.0 * L.bc_wall[2 * b:2 * b + 2] - gh[:2]
                else:
                    gh[:2] = gh[:2] - 2.0 * (gh[0] * nrm[0] + gh[1] * nrm[1]) * nrm
                gh[2] = 2.0 * bg[3] - vals[i, 2] if (flags >> 1) == 1 else gh[2] + 2.0 * bg[0] * bg[3] / kc
            elif code == 2:
                gh[:2] = gh[:2] - 2.0 * (vals[i, 0] * nrm[0] + vals[i, 1] * nrm[1]) * nrm
            elif code in (3, 5):
                gh = 2.0 * _inlet(code, W[i], L.bc_state[4 * b:4 * b + 4], nrm, s) - gh
            d = gh - vals[i]
            rx = rx + v["gh_wx"][e] * d
            ry = ry + v["gh_wy"][e] * d
        xx, xy, yy, inv = v["wl_cell"][4 * i:4 * i + 4]
        if inv != 0.0:
            VG[i, 0:6:2] = (rx * yy - ry * xy) * inv
            VG[i, 1:6:2] = (ry * xx - rx * xy) * inv
        VG[i, 6], VG[i, 7] = mu, kc

    def flux(gr, u, vv, mu, k, nx, ny):
        div = gr[0] + gr[3]
        txx = mu * (2.0 * gr[0] - (2.0 / 3.0) * div)
        tyy = mu * (2.0 * gr[3] - (2.0 / 3.0) * div)
        txy = mu * (gr[1] + gr[2])
        tx = txx * nx + txy * ny
        ty = txy * nx + tyy * ny
        return np.array((0.0, tx, ty, u * tx + vv * ty + k * (gr[4] * nx + gr[5] * ny)))

    VF = np.zeros((L.n_faces, 4))
    for f in range(L.n_faces):
        o = L.face_left[f]
        nx, ny = L.face_nx[f], L.face_ny[f]
        if L.face_kind[f] <= 1:
            r = L.face_right[f]
            gr = 0.5 * (VG[o, :6] + VG[r, :6])
            VF[f] = flux(gr, 0.5 * (W[o, 1] + W[r, 1]), 0.5 * (W[o, 2] + W[r, 2]), 0.5 * (VG[o, 6] + VG[r, 6]),
                         0.5 * (VG[o, 7] + VG[r, 7]), nx, ny)
            continue
        b = L.face_bc[f]
        code, flags = v["bv_code"][b], v["bv_flags"][b]
        dn, tx, ty, val = v["bv_geom"][4 * b:4 * b + 4]
        off = np.array((L.face_lx[f], L.face_ly[f]))
        nrm = np.array((nx, ny))
        tang = np.array((tx, ty))
        gr = VG[o, :6].reshape(3, 2).copy()
        bv = vals[o].copy()
        if code == 1:
            if flags & 1:
                for c in (0, 1):
                    nd = (L.bc_wall[2 * b + c] - bv[c] - gr[c] @ tang) / dn
                    gr[c] += (nd - gr[c] @ nrm) * nrm
                    bv[c] = L.bc_wall[2 * b + c]
            else:
                un = bv[:2] @ nrm
                bv[:2] += gr[:2] @ off
                bv[:2] -= (bv[:2] @ nrm) * nrm
                gnv = nrm[0] * gr[0] + nrm[1] * gr[1]
                nd = -(un + gnv @ tang) / dn
                gr[:2] += np.outer(nd * nrm, nrm) - (gr[:2] @ nrm)[:, None] * nrm
            if (flags >> 1) == 1:
                nd = (val - bv[2] - gr[2] @ tang) / dn
                bv[2] = val
            else:
                nd = val / VG[o, 7]
                bv[2] += gr[2] @ off
            gr[2] += (nd - gr[2] @ nrm) * nrm
        elif code == 2:
            un = bv[:2] @ nrm
            dv = -un / dn * nrm
            for c in (0, 1):
                gr[c] += (dv[c] - gr[c] @ nrm) * nrm
            gr[2] -= (gr[2] @ nrm) * nrm
            bv[:2] -= un * nrm
        elif code in (3, 5):
            bv = _inlet(code, W[o], L.bc_state[4 * b:4 * b + 4], nrm, s)
            for c in range(3):
                nd = (bv[c] - vals[o, c] - gr[c] @ tang) / dn
                gr[c] += (nd - gr[c] @ nrm) * nrm
        else:
            gr -= np.outer(gr @ nrm, nrm)
            bv += gr @ off
        fmu = _mu(bv[2], s)
        VF[f] = flux(gr.reshape(-1), bv[0], bv[1], fmu, _k(fmu, s), nx, ny)
        if code == 1 and (flags >> 1) == 2:
            VF[f, 3] = val

    out = np.zeros((n, 4))
    for i in range(n):
        S = np.zeros(4)
        d = None
        for e in range(v["va_ptr"][i], v["va_ptr"][i + 1]):
            item = v["va_item"][e]
            f, mode = item >> 2, item & 3
            if mode == 0:
                S = S + v["va_coef"][e] * VF[f]
                continue
            if d is None:
                d = 0.0 - S
            x = VF[f] * L.face_len[f]
            d = d + (x / L.area[i] if mode == 1 else (-x) / L.area[i])
        out[i] = 0.0 - S if d is None else d
    return out


@pytest.mark.parametrize("kind", CASES)
def test_list_replay_matches_oracle_viscous_residual(kind):
    case = viscous_case(kind, n=7)
    g = build_geometry(case)
    L = build_layout(g)
    v = build_viscous(g, L)
    prim = viscous_primitive(case)
    ref = viscous_residual(prim, g, case.numerics, Oracle(g, case.numerics).D)
    got = replay_viscous_residual(g, L, v, prim[L.cell_order], case.numerics)
    scale = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
    err = np.max(np.abs(got - ref[L.cell_order]), axis=0) / scale
    assert np.all(err[1:] <= 1e-12), err
    assert np.max(np.abs(ref[:, 0])) == 0.0


@pytest.mark.parametrize("kind", CASES)
def test_viscous_terms_are_significant_in_parity_cases(kind):
    """The parity cases are only meaningful if the viscous residual is of
    the size of the convective one."""
    case = viscous_case(kind, n=7)
    g = build_geometry(case)
    prim = viscous_primitive(case)
    o = Oracle(g, case.numerics)
    total = o.residual(euler.primitive_to_conservative(prim, case.numerics.gamma))
    visc = viscous_residual(prim, g, case.numerics, o.D)
    ratio = np.max(np.abs(visc[:, 1:]), axis=0) / np.max(np.abs(total[:, 1:]), axis=0)
    assert np.all(ratio > 0.02), ratio
    assert np.all(viscous_time_step_sum(prim, g, case.numerics) > 0.0)


def test_lists_follow_reference_sums():
    """wl_cell's xx equals the reference's sequential weighted sum; every
    viscous accumulation list holds interior items first."""
    case = viscous_case("channel", n=6)
    g = build_geometry(case)
    L = build_layout(g)
    v = build_viscous(g, L)
    n = g.cell_count
    new_id = np.empty(n, dtype=np.int64)
    new_id[L.cell_order] = np.arange(n)
    xx = np.zeros(n)
    for s_, dx, dy in zip(g.gradient_sources, g.gradient_dx, g.gradient_dy):
        w = 1.0 / max(dx * dx + dy * dy, 1e-300)
        xx[s_] += (w * dx) * dx
    for f in g.diffusive_boundary_faces:
        o = g.owners[f]
        off = g.face_centers[f] - g.centers[o]
        nrm = g.face_normals[f]
        dn = float(np.dot(off, nrm))
        wall = g.face_kind_code[f] in (1, 2)
        gx = (2.0 * dn * nrm if wall else 2.0 * off)
        w = 1.0 / max(gx[0] * gx[0] + gx[1] * gx[1], 1e-300)
        xx[o] += (w * gx[0]) * gx[0]
    assert np.array_equal(v["wl_cell"].reshape(-1, 4)[new_id, 0], xx)
    for i in range(n):
        modes = v["va_item"][v["va_ptr"][i]:v["va_ptr"][i + 1]] & 3
        nz = np.flatnonzero(modes != 0)
        assert nz.size == 0 or np.all(modes[nz[0]:] != 0)
    assert len(v["gh_face"]) == len(g.diffusive_boundary_faces)


# --- GPU parity --------------------------------------------------------------


def _rel(gpu, ref):
    scale = np.max(np.abs(ref), axis=0)
    scale = np.where(scale > 0, scale, 1.0)
    return np.max(np.abs(gpu - ref), axis=0) / scale


def _device(case, cuda_device, fused=True):
    from meshcfd_amd.device import DeviceMesh

    print(f"viscous case: {case.name} cells={case.mesh.cell_count} model={case.numerics.viscosity_model} "
          f"rec={case.numerics.reconstruction} ti={case.numerics.time_integration}", flush=True)
    return DeviceMesh(case, cuda_device, fused=fused, block_cells=64)


@pytest.mark.gpu
@pytest.mark.parametrize("reconstruction", ["First order", "MUSCL linear"])
@pytest.mark.parametrize("kind", CASES)
def test_gpu_viscous_residual_matches_oracle(cuda_device, kind, reconstruction):
    case = viscous_case(kind, reconstruction=reconstruction)
    dm = _device(case, cuda_device)
    prim = viscous_primitive(case)
    u = euler.primitive_to_conservative(prim, case.numerics.gamma)
    ref = Oracle(dm.geometry, case.numerics).residual(u)
    gpu = dm.to_mesh_order(dm.residual(dm.to_device_order(u)))
    dm.check_status()
    err = _rel(gpu, ref)
    print("viscous residual rel err", err)
    assert np.all(err <= 1e-12), err
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", CASES)
def test_gpu_viscous_time_steps_match_oracle(cuda_device, kind):
    case = viscous_case(kind)
    dm = _device(case, cuda_device)
    prim = viscous_primitive(case)
    u = euler.primitive_to_conservative(prim, case.numerics.gamma)
    ref = Oracle(dm.geometry, case.numerics).stable_time_steps(prim)
    dm.time_steps(dm.to_device_order(u))
    dm.check_status()
    gpu = dm.to_mesh_order(dm.dt_cell)
    assert np.max(np.abs(gpu - ref) / ref) <= 1e-12
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", ["Explicit Euler", "Explicit RK3"])
@pytest.mark.parametrize("kind", ["channel", "farfield"])
def test_gpu_viscous_steps_match_oracle(cuda_device, kind, integrator):
    case = viscous_case(kind, integrator=integrator)
    dm = _device(case, cuda_device)
    prim = viscous_primitive(case)
    dm.set_state(euler.primitive_to_conservative(prim, case.numerics.gamma))
    dm.run(10)
    gpu = dm.state()
    ref, _ = Oracle(dm.geometry, case.numerics).run(prim, 10)
    err = _rel(gpu, ref)
    print(kind, integrator, "rel err", err)
    assert np.all(err <= 1e-10), err
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["channel", "sutherland"])
def test_gpu_viscous_implicit_step_matches_oracle(cuda_device, kind):
    """Implicit Euler (local steps, CFL 20): the Jacobian's first-order
    residual keeps the viscous terms, the LU-SGS preconditioner does not
    (solver.py:1378, 1185-1337)."""
    from oracle import implicit as oimp

    case = viscous_case(kind, integrator="Implicit Euler")
    case.numerics.cfl = 20.0
    case.numerics.time_step_mode = "Local (steady)"
    case.numerics.gmres_tolerance = 1e-6
    dm = _device(case, cuda_device)
    prim = viscous_primitive(case)
    u0 = euler.primitive_to_conservative(prim, case.numerics.gamma)
    dm.set_state(u0)
    dm.step_implicit(0.0)
    dm.check_status()
    gpu = dm.state()
    o = Oracle(dm.geometry, case.numerics)
    ref, _ = oimp.implicit_euler_step(dm.geometry, case.numerics, u0, o.stable_time_steps(prim))
    err = _rel(gpu, ref)
    print(kind, "implicit rel err", err)
    assert np.all(err <= 1e-8), err
    dm.close()


@pytest.mark.gpu
def test_gpu_laminar_cylinder_implicit_step_matches_oracle(cuda_device):
    """Config-3 physics on a small O-grid (24 x 10, r to 5): Roe + MUSCL
    parabolic + Sutherland, no-slip isothermal wall, far field, implicit
    Euler with local steps: one step equals the oracle's to 1e-8."""
    from oracle import implicit as oimp

    case, prim = meshgen.cylinder_case(24, 10, mach=0.2, alpha_deg=0.0, reynolds=40.0, r_outer=5.0,
                                       first_spacing=0.05, seed=2)
    case.numerics.cfl = 20.0
    case.numerics.gmres_tolerance = 1e-6
    rng = np.random.default_rng(3)
    prim = prim * (1.0 + 0.02 * rng.random(prim.shape))
    dm = _device(case, cuda_device)
    u0 = euler.primitive_to_conservative(prim, case.numerics.gamma)
    dm.set_state(u0)
    dm.step_implicit(0.0)
    dm.check_status()
    gpu = dm.state()
    o = Oracle(dm.geometry, case.numerics)
    ref, _ = oimp.implicit_euler_step(dm.geometry, case.numerics, u0, o.stable_time_steps(prim))
    err = _rel(gpu, ref)
    print("laminar cylinder implicit rel err", err)
    assert np.all(err <= 1e-8), err
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("reconstruction", ["First order", "MUSCL linear", "MUSCL parabolic"])
@pytest.mark.parametrize("kind", ["channel", "farfield", "periodic", "sutherland"])
def test_gpu_viscous_vjp_matches_central_difference_of_oracle(cuda_device, kind, reconstruction):
    """lambda^T dR/dU of the viscous residual (hand-written reverse of the
    WLSQ gradients with ghost entries, the boundary corrections, the Stokes
    and Fourier fluxes and the viscosity law) against central differences
    of the oracle's residual: no-slip and slip walls with fixed temperature
    and heat flux, symmetry, velocity inlet, far field, outlet, transmissive
    and periodic faces; constant and Sutherland viscosity."""
    from meshcfd_amd.adjoint import residual_vjp

    case = viscous_case(kind, reconstruction=reconstruction, flux="Roe" if kind == "sutherland" else "HLLC")
    dm = _device(case, cuda_device)
    prim = viscous_primitive(case)
    u = euler.primitive_to_conservative(prim, case.numerics.gamma)
    rng = np.random.default_rng(12)
    lam = rng.normal(size=u.shape)
    scale = np.max(np.abs(u), axis=0)
    v = rng.normal(size=u.shape) * scale
    o = Oracle(dm.geometry, case.numerics)
    eps = 1e-7
    fd = float(np.sum(lam * (o.residual(u + eps * v) - o.residual(u - eps * v)) / (2 * eps)))
    ubar = dm.to_mesh_order(residual_vjp(dm, dm.to_device_order(u), dm.to_device_order(lam)))
    dm.check_status()
    ad = float(np.sum(ubar * v))
    print(kind, reconstruction, "fd", fd, "adjoint", ad)
    assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1.0)
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("reconstruction", ["First order", "MUSCL linear"])
@pytest.mark.parametrize("kind", ["channel", "farfield"])
def test_gpu_viscous_boundary_state_adjoint_matches_fd(cuda_device, kind, reconstruction):
    """lambda^T dR/d(boundary state) of the viscous residual, summed over the
    velocity-inlet (channel) or far-field (farfield) records, against central
    differences of the oracle residual with the boundary condition's values
    moved along a direction d in (rho, u, v, p): the convective faces and the
    viscous ghost entries and boundary values both depend on the state."""
    import torch

    from meshcfd_amd.adjoint import residual_vjp
    from meshcfd_amd.layout import FACE_FAR_FIELD, FACE_VELOCITY_INLET

    case = viscous_case(kind, reconstruction=reconstruction)
    dm = _device(case, cuda_device)
    prim = viscous_primitive(case)
    u = euler.primitive_to_conservative(prim, case.numerics.gamma)
    rng = np.random.default_rng(21)
    lam = rng.normal(size=u.shape)
    n_bc = len(dm.layout.bc_kind)
    sbar = torch.zeros((n_bc, 4), dtype=torch.float64, device=cuda_device)
    residual_vjp(dm, dm.to_device_order(u), dm.to_device_order(lam), state_bar=sbar)
    dm.check_status()
    code = FACE_VELOCITY_INLET if kind == "channel" else FACE_FAR_FIELD
    rec = torch.as_tensor(dm.layout.bc_kind == code, device=cuda_device)
    total = sbar[rec].sum(dim=0).cpu().numpy()
    R = case.numerics.gas_constant
    bcs = [bc for bc in case.boundary_conditions if bc.kind in ("Velocity Inlet", "Far Field")]
    v0 = dict(bcs[0].values)
    if kind == "channel":
        base = np.array((v0["density"], v0["velocity_x"], v0["velocity_y"], v0["density"] * R * v0["temperature"]))
    else:
        base = np.array((v0["density"], v0["velocity_x"], v0["velocity_y"], v0["pressure"]))
    d = rng.normal(size=4) * np.abs(base).clip(0.1)

    def residual_at(state):
        for bc in bcs:
            if kind == "channel":
                bc.values = {"density": state[0], "velocity_x": state[1], "velocity_y": state[2],
                             "temperature": state[3] / (state[0] * R)}
            else:
                bc.values = {"density": state[0], "velocity_x": state[1], "velocity_y": state[2],
                             "pressure": state[3]}
        return Oracle(build_geometry(case), case.numerics).residual(u)

    eps = 1e-7
    fd = float(np.sum(lam * (residual_at(base + eps * d) - residual_at(base - eps * d))) / (2 * eps))
    ad = float(total @ d)
    print(kind, reconstruction, "fd", fd, "adjoint", ad)
    assert abs(ad - fd) <= 1e-6 * max(abs(fd), 1.0)
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("flux", ["Rusanov", "HLLC", "Roe"])
@pytest.mark.parametrize("kind", CASES)
def test_gpu_fused_first_order_viscous_stage_equals_four_kernels(cuda_device, kind, flux):
    """A first-order viscous residual on the fused path (one launch: WLSQ
    gradients of own and ring-1 cells, convective and viscous face fluxes and
    both gathers out of LDS) has the four-kernel path's bits, and so do three
    explicit RK3 steps."""
    import torch

    case = viscous_case(kind, reconstruction="First order", flux=flux)
    a = _device(case, cuda_device, fused=True)
    b = _device(case, cuda_device, fused=False)
    assert a.blocks is not None and a.blocks.vent is not None
    prim = viscous_primitive(case)
    u = euler.primitive_to_conservative(prim, case.numerics.gamma)
    ra = a.residual(a.to_device_order(u)).clone()
    rb = b.residual(b.to_device_order(u)).clone()
    a.check_status()
    b.check_status()
    assert torch.equal(ra, rb), float((ra - rb).abs().max())
    for dm in (a, b):
        dm.set_state(u)
        dm.run(3)
    assert np.array_equal(a.state(), b.state())
    a.close()
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["channel", "farfield", "sutherland"])
def test_gpu_fused_viscous_jacobian_products_give_the_same_implicit_step(cuda_device, kind):
    """Implicit Euler on a viscous case: the Newton loop's Jacobian products
    (first-order viscous residuals) run fused on a mesh with blocks and as
    four kernels without; the implicit step is the same to the last bit."""
    import torch

    case = viscous_case(kind, integrator="Implicit Euler", flux="Roe" if kind == "sutherland" else "HLLC")
    case.numerics.time_step_mode = "Local (steady)"
    case.numerics.cfl = 20.0
    prim = viscous_primitive(case)
    u = euler.primitive_to_conservative(prim, case.numerics.gamma)
    out = []
    for fused in (True, False):
        dm = _device(case, cuda_device, fused=fused)
        dm.set_state(u)
        dm.step_implicit(0.0)
        dm.check_status()
        out.append(dm.u.clone())
        dm.close()
    assert torch.equal(out[0], out[1]), float((out[0] - out[1]).abs().max())
