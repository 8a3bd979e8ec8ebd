"""The reference's seams (`solver.py` module functions) in `meshcfd_amd.solver`.

CPU: the reference's monkeypatch tests re-expressed against the product's
`solve_case` and `_implicit_euler_step` (tests/test_solver.py:786-840 and
:1289-1323): with the seams replaced no GPU call is made.
GPU: each seam against the oracle; `solve_case` through the seams equals the
device-resident run bit for bit; invalid states end the run with the
reference's message and step number through the fused and unfused kernels;
the reported RMS residuals equal the oracle's.
"""

import numpy as np
import pytest

import meshcfd_amd.solver as msolver
from meshcfd_amd import meshgen
from meshcfd_amd.geometry import SolverError, build_geometry
from meshcfd_amd.model import InitialState, NumericalSettings
from oracle import implicit as oimp
from oracle.euler import conservative_to_primitive, primitive_to_conservative
from oracle.solver import Oracle


def periodic_case(nx=4, ny=4, length_x=1.0, **kw):
    mesh = meshgen.rectangle_mesh(nx, ny, (0.0, length_x, 0.0, 1.0))
    return meshgen.periodic_case(mesh, NumericalSettings(**kw))


def test_implicit_euler_halves_time_step_until_nonlinear_solve_converges(monkeypatch):
    """Reference tests/test_solver.py:786-815."""
    case = periodic_case()
    geometry = build_geometry(case)
    state = primitive_to_conservative(np.array(((1.0, 0.2, -0.1, 1.0),)), case.numerics.gamma)
    attempted = []

    def solve_once(initial, time_step, _geometry, _settings):
        attempted.append(time_step)
        if len(attempted) < 3:
            raise SolverError("nonlinear solve did not converge")
        return initial

    monkeypatch.setattr(msolver, "_implicit_euler_solve_once", solve_once)
    solution, accepted = msolver._implicit_euler_step(state, 0.8, geometry, case.numerics)
    assert attempted == pytest.approx((0.8, 0.4, 0.2))
    assert accepted == pytest.approx(0.2)
    np.testing.assert_array_equal(solution, state)


def test_implicit_euler_keeps_a_fixed_time_step_on_nonlinear_failure(monkeypatch):
    """Reference tests/test_solver.py:818-840."""
    case = periodic_case()
    geometry = build_geometry(case)
    state = primitive_to_conservative(np.array(((1.0, 0.2, -0.1, 1.0),)), case.numerics.gamma)
    attempts = []

    def fail_solve_once(_initial, time_step, _geometry, _settings):
        attempts.append(time_step)
        raise SolverError("nonlinear solve did not converge")

    monkeypatch.setattr(msolver, "_implicit_euler_solve_once", fail_solve_once)
    with pytest.raises(SolverError, match="configured fixed time step"):
        msolver._implicit_euler_step(state, 0.8, geometry, case.numerics, allow_time_step_reduction=False)
    assert attempts == [0.8]


def test_implicit_halving_reports_twelve_attempts(monkeypatch):
    case = periodic_case()
    geometry = build_geometry(case)
    attempts = []

    def fail(_initial, time_step, _geometry, _settings):
        attempts.append(time_step)
        raise SolverError("no")

    monkeypatch.setattr(msolver, "_implicit_euler_solve_once", fail)
    with pytest.raises(SolverError, match="after reducing the time step 12 times; last failure: no"):
        msolver._implicit_euler_step(np.ones((1, 4)), np.array([0.8, 0.4]), geometry, case.numerics)
    assert len(attempts) == 12
    np.testing.assert_array_equal(attempts[-1], np.array([0.8, 0.4]) * 0.5**11)


@pytest.mark.parametrize("time_integration", ("Explicit Euler", "Explicit RK2", "Explicit RK3", "Explicit RK4"))
def test_local_time_step_advances_each_cell_with_its_own_cfl_step(monkeypatch, time_integration):
    """Reference tests/test_solver.py:1289-1323 against the product's solve_case."""
    case = periodic_case(nx=2, ny=1, length_x=2.0, time_step_mode="Local (steady)",
                         time_integration=time_integration, max_iterations=1, residual_tolerance=1e-14)
    local_steps = np.array((0.01, 0.02))
    rates = np.array(((0.0, 1.0, 0.0, 0.0), (0.0, 2.0, 0.0, 0.0)))
    monkeypatch.setattr(msolver, "_stable_time_steps", lambda *_args: local_steps.copy())
    monkeypatch.setattr(msolver, "_residual", lambda *_args: rates.copy())
    progress = []
    result = msolver.solve_case(case, (InitialState(density=1.0, pressure=1.0),
                                       InitialState(density=1.0, pressure=1.0)), progress.append)
    assert not result.completed
    assert result.steps == 1
    assert result.time == pytest.approx(np.mean(local_steps))
    assert progress[0].time_step == pytest.approx(np.mean(local_steps))
    assert [state.velocity_x for state in result.states] == pytest.approx((0.01, 0.04))


def test_replaced_residual_failure_is_reported_with_the_step(monkeypatch):
    case = periodic_case(nx=2, ny=1, length_x=2.0, time_step_mode="Global fixed", fixed_time_step=0.1,
                         time_integration="Explicit Euler", max_iterations=3, end_time=10.0)
    calls = []

    def residual(*_args):
        calls.append(1)
        if len(calls) >= 2:
            raise SolverError("Spatial residual contains non-finite values.")
        return np.zeros((2, 4))

    monkeypatch.setattr(msolver, "_residual", residual)
    with pytest.raises(SolverError, match="^Solver failed at step 2: Spatial residual contains non-finite values"):
        msolver.solve_case(case, (InitialState(density=1.0, pressure=1.0),) * 2)


def test_seams_are_detected_as_replaced(monkeypatch):
    assert not msolver.seams_replaced()
    monkeypatch.setattr(msolver, "_stable_time_step", lambda *a: 1.0)
    assert msolver.seams_replaced()


# ----------------------------------------------------------------- GPU -----


def _riemann(n=12, **kw):
    case, prim = meshgen.riemann_workload(n, 10)
    for k, v in kw.items():
        setattr(case.numerics, k, v)
    return case, prim


@pytest.mark.gpu
def test_gpu_seams_match_the_oracle(cuda_device):
    case, prim = _riemann(12)
    print("case: seams vs oracle, Riemann 12x12 HLLC MUSCL", flush=True)
    g = build_geometry(case)
    s = case.numerics
    o = Oracle(g, s)
    u = primitive_to_conservative(prim, s.gamma)
    np.testing.assert_array_equal(msolver._residual(u, g, s), o.residual(u))
    w = conservative_to_primitive(u, s.gamma)
    np.testing.assert_array_equal(msolver._stable_time_steps(w, g, s), o.stable_time_steps(w))
    assert msolver._stable_time_step(w, g, s) == float(np.min(o.stable_time_steps(w)))
    h = 0.5 * float(np.min(o.stable_time_steps(w)))
    np.testing.assert_array_equal(msolver._explicit_rk4_step(u, h, g, s), o.rk4(u, h))
    hc = 0.5 * o.stable_time_steps(w)
    np.testing.assert_array_equal(msolver._explicit_rk4_step(u, hc, g, s), o.rk4(u, hc))


@pytest.mark.gpu
def test_gpu_implicit_seam_matches_the_oracle(cuda_device):
    case, prim = _riemann(8, time_integration="Implicit Euler", flux="Rusanov", reconstruction="First order",
                          cfl=20.0, gmres_tolerance=1e-8)
    print("case: implicit seam vs oracle", flush=True)
    g = build_geometry(case)
    s = case.numerics
    u = primitive_to_conservative(prim, s.gamma)
    dt = float(np.min(Oracle(g, s).stable_time_steps(prim)))
    got, h = msolver._implicit_euler_step(u, dt, g, s)
    want, hr = oimp.implicit_euler_step(g, s, u, dt)
    assert h == hr
    err = np.max(np.abs(got - want), axis=0) / np.max(np.abs(want), axis=0)
    assert np.all(err <= 1e-7), err


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", ["Explicit RK3", "Explicit RK4", "Explicit Euler"])
def test_gpu_solve_case_through_seams_equals_device_run(cuda_device, monkeypatch, integrator):
    """A pass-through replacement of `_residual` sends every step through the
    seams; the result equals the device-resident run bit for bit."""
    case, prim = _riemann(12, time_integration=integrator, max_iterations=6, end_time=1e30)
    print(f"case: solve_case seams vs device {integrator}", flush=True)
    direct = msolver.solve_case(case, prim)
    original = msolver._residual
    calls = []

    def spy(*args):
        calls.append(1)
        return original(*args)

    monkeypatch.setattr(msolver, "_residual", spy)
    seams = msolver.solve_case(case, prim)
    assert calls
    np.testing.assert_array_equal(seams.primitive, direct.primitive)
    assert seams.time == direct.time and seams.steps == direct.steps
    assert seams.residual == pytest.approx(direct.residual, rel=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("mode", ["Global adaptive", "Global fixed", "Local (steady)"])
def test_gpu_invalid_state_fails_at_its_step_with_the_cell(cuda_device, fused, mode):
    """A body force against the flow drives pressures negative in the first
    step: solve_case raises at step 1 with the reference's message and a
    cell whose oracle state is invalid, through the fused and the unfused
    kernels and every time-step mode."""
    from meshcfd_amd.device import DeviceMesh

    case, prim = meshgen.vortex_workload(16, 5)
    s = case.numerics
    s.acceleration_x = -100.0
    s.time_step_mode = mode
    s.fixed_time_step = 0.05
    print(f"case: invalid state fused={fused} mode={mode}", flush=True)
    g = build_geometry(case)
    o = Oracle(g, s)
    u = primitive_to_conservative(prim, s.gamma)
    steps = o.stable_time_steps(prim)
    h = {"Global adaptive": float(np.min(steps)), "Global fixed": 0.05, "Local (steady)": steps}[mode]
    new = u + (np.asarray(h)[:, None] if np.ndim(h) else h) * o.residual(u)
    rho = new[:, 0]
    p = (s.gamma - 1.0) * (new[:, 3] - 0.5 * (new[:, 1] ** 2 + new[:, 2] ** 2) / rho)
    bad = set(np.flatnonzero(~((rho > 0) & (p > 0))).tolist())
    assert bad
    dm = DeviceMesh(case, cuda_device, geometry=g, fused=fused, block_cells=64)
    with pytest.raises(SolverError) as info:
        msolver.solve_case(case, prim, mesh_device=dm)
    message = str(info.value)
    assert message.startswith("Solver failed at step 1: Euler state has non-finite or non-positive density or pressure.")
    cell = int(message.rsplit("First invalid cell: ", 1)[1].rstrip("."))
    assert cell in bad
    dm.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["Global adaptive", "Local (steady)"])
def test_gpu_reported_rms_residuals_match_the_oracle(cuda_device, mode):
    """Per-step component RMS residuals (solver.py:945-957) and the final
    residual (:999) of solve_case against the oracle's NumPy norms."""
    case, prim = meshgen.vortex_workload(24, 8)
    s = case.numerics
    s.time_step_mode = mode
    s.convergence_history_interval = 1
    print(f"case: RMS residuals {mode}", flush=True)
    progress = []
    result = msolver.solve_case(case, prim, progress.append)
    g = build_geometry(case)
    o = Oracle(g, s)
    u = primitive_to_conservative(prim, s.gamma)
    t = 0.0
    for entry in progress:
        start = u
        u, dt, _, d0 = o.step(u, t)
        t += dt
        r = o.residual(u) if mode == "Local (steady)" else d0
        want = np.sqrt(np.mean(r**2, axis=0))
        np.testing.assert_allclose(entry.residuals, want, rtol=1e-12)
        assert entry.time_step == pytest.approx(dt, rel=1e-14)
        del start
    final = float(np.sqrt(np.mean(o.residual(u) ** 2)))
    assert result.residual == pytest.approx(final, rel=1e-12)
    assert len(progress) == 8
