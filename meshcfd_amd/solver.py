"""`solve_case` on the GPU: the reference's driver contract over `DeviceMesh`.

Mirrors `solve_case` (`solver.py:179-1008`): the same arguments, settings
strings and validation messages, the same step order (time step from the
step's start state, the integrator, time += dt, the steady residual after
the step in "Local (steady)" mode), the same progress, solution and cancel
callbacks, `SolverProgress` / `SolverResult` (`solver.py:58-77`) and
`SolverError` with "Solver failed at step {n}: ..." around step failures.
The state stays on the device; the host reads the clock and the status word
once per step and pulls states only for callbacks and the result.

Not covered (raise `SolverError`): SST turbulence, mesh adaptation, the ILU
preconditioner, and evaluating `case.initial_conditions` (the caller passes
the initial states).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from time import thread_time
from typing import Callable

import numpy as np
import torch

from .geometry import SolverError
from .model import Case, FiniteVolumeMesh, InitialState


@dataclass(frozen=True)
class SolverProgress:
    """`SolverProgress` (`solver.py:58-66`)."""

    step: int
    time: float
    time_step: float
    residuals: tuple[float, ...]
    states: tuple[InitialState, ...] | None = None
    mesh: FiniteVolumeMesh | None = None
    cpu_time: float = 0.0
    warning: str | None = None


@dataclass(frozen=True)
class SolverResult:
    """`SolverResult` (`solver.py:69-76`), plus the final primitive state as
    one (C, 4) array in mesh order."""

    states: tuple[InitialState, ...]
    time: float
    steps: int
    residual: float
    completed: bool
    cpu_time: float = 0.0
    primitive: np.ndarray | None = None


def primitive_to_conservative(primitive: np.ndarray, gamma: float) -> np.ndarray:
    """(rho, u, v, p) -> (rho, rho u, rho v, E) (`euler.py:8-17`)."""
    rho, u, v, p = np.asarray(primitive, dtype=np.float64).T
    return np.column_stack((rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u**2 + v**2)))


def conservative_to_primitive(conservative: np.ndarray, gamma: float) -> np.ndarray:
    """Inverse map; raises on non-finite or non-positive density or pressure
    (`euler.py:20-40`)."""
    q = np.asarray(conservative, dtype=np.float64)
    rho = q[:, 0]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        u = q[:, 1] / rho
        v = q[:, 2] / rho
        p = (gamma - 1.0) * (q[:, 3] - 0.5 * (q[:, 1] ** 2 + q[:, 2] ** 2) / rho)
    w = np.column_stack((rho, u, v, p))
    if not np.isfinite(w).all() or np.any(rho <= 0.0) or np.any(p <= 0.0):
        raise SolverError("Euler state has non-finite or non-positive density or pressure.")
    return w


def _states(primitive: np.ndarray) -> tuple[InitialState, ...]:
    return tuple(InitialState(float(r), float(u), float(v), float(p)) for r, u, v, p in primitive)


def _positive_int(value) -> bool:
    return not isinstance(value, bool) and isinstance(value, int) and value > 0


def validate(case: Case, start_time: float, step_offset: int, solution_interval: int,
             cpu_time_offset: float) -> None:
    """The settings checks of `solve_case` (`solver.py:194-374`) for the
    settings this path has, with the reference's messages."""
    if case.mesh is None:
        raise SolverError("Import a mesh before running the solver.")
    s = case.numerics
    if getattr(s, "turbulence_model", "None") != "None":
        raise SolverError("The Menter k-omega SST model is not implemented by the GPU path.")
    if getattr(s, "adaptive_mesh", False):
        raise SolverError("Adaptive mesh refinement is not implemented by the GPU path.")
    if s.reconstruction not in ("First order", "MUSCL linear", "MUSCL parabolic"):
        raise SolverError(f"Unsupported reconstruction '{s.reconstruction}'.")
    if s.limiter not in ("Barth-Jespersen", "Venkatakrishnan", "Minmod", "Van Leer"):
        raise SolverError(f"Unsupported limiter '{s.limiter}'.")
    if s.flux not in ("HLLC", "Roe", "Rusanov"):
        raise SolverError(f"Unsupported convective flux '{s.flux}'.")
    if s.time_integration not in ("Explicit Euler", "Explicit RK2", "Explicit RK3", "Explicit RK4",
                                  "Implicit Euler"):
        raise SolverError(f"Unsupported time integration '{s.time_integration}'.")
    if s.time_step_mode not in ("Global adaptive", "Global fixed", "Local (steady)"):
        raise SolverError(f"Unsupported time-stepping mode '{s.time_step_mode}'.")
    local = s.time_step_mode == "Local (steady)"
    if not math.isfinite(s.fixed_time_step) or s.fixed_time_step <= 0.0:
        raise SolverError("Fixed time step must be positive and finite.")
    if not math.isfinite(s.residual_tolerance) or s.residual_tolerance <= 0.0:
        raise SolverError("Residual tolerance must be positive and finite.")
    maximum_cfl = 1000.0 if s.time_integration == "Implicit Euler" else 1.0
    if not math.isfinite(s.cfl) or not 0.0 < s.cfl <= maximum_cfl:
        raise SolverError(f"{s.time_integration} requires a CFL number in (0, {maximum_cfl:g}].")
    if not _positive_int(s.max_iterations):
        raise SolverError("Maximum iterations must be a positive integer.")
    if not math.isfinite(s.gmres_tolerance) or not 0.0 < s.gmres_tolerance < 1.0:
        raise SolverError("GMRES tolerance must be finite and in (0, 1).")
    for value, label in ((s.gmres_max_iterations, "Maximum GMRES iterations"),
                         (s.gmres_krylov_dimension, "GMRES Krylov dimension")):
        if not _positive_int(value):
            raise SolverError(f"{label} must be a positive integer.")
    if s.implicit_preconditioner not in ("ILU", "LU-SGS"):
        raise SolverError("Unsupported implicit preconditioner.")
    if s.implicit_preconditioner == "ILU" and s.time_integration == "Implicit Euler":
        raise SolverError("The ILU preconditioner is not implemented by the GPU path.")
    if not local and (not math.isfinite(s.end_time) or s.end_time <= 0.0):
        raise SolverError("End time must be positive and finite.")
    if not math.isfinite(start_time) or start_time < 0.0:
        raise SolverError("Current solution time must be finite and non-negative.")
    if not local and start_time >= s.end_time:
        raise SolverError("End time must be greater than the current solution time to continue.")
    if isinstance(step_offset, bool) or not isinstance(step_offset, int) or step_offset < 0:
        raise SolverError("Current solution step must be a non-negative integer.")
    if not _positive_int(solution_interval):
        raise SolverError("Solution output interval must be a positive integer.")
    if not _positive_int(s.convergence_history_interval):
        raise SolverError("Convergence history interval must be a positive integer.")
    if not math.isfinite(cpu_time_offset) or cpu_time_offset < 0.0:
        raise SolverError("CPU time offset must be finite and non-negative.")
    if not math.isfinite(s.gamma) or s.gamma <= 1.0:
        raise SolverError("Specific heat ratio must be finite and greater than 1.")
    if not math.isfinite(s.gas_constant) or s.gas_constant <= 0.0:
        raise SolverError("Gas constant must be positive and finite.")
    if s.viscosity_model not in ("Inviscid", "Constant", "Sutherland"):
        raise SolverError("Unsupported viscosity model.")
    for value, description in ((s.dynamic_viscosity, "Dynamic viscosity"),
                               (s.sutherland_reference_viscosity, "Sutherland reference viscosity"),
                               (s.sutherland_reference_temperature, "Sutherland reference temperature"),
                               (s.sutherland_temperature, "Sutherland temperature"),
                               (s.prandtl_number, "Prandtl number")):
        if not math.isfinite(value) or value <= 0.0:
            raise SolverError(f"{description} must be positive and finite.")
    if not math.isfinite(s.acceleration_x) or not math.isfinite(s.acceleration_y):
        raise SolverError("Body-force acceleration must be finite.")


def solve_case(
    case: Case,
    initial_states: tuple[InitialState, ...] | np.ndarray | None = None,
    progress_callback: Callable[[SolverProgress], None] | None = None,
    cancel_callback: Callable[[], bool] | None = None,
    live_solution_updates: bool = False,
    start_time: float = 0.0,
    step_offset: int = 0,
    pause_callback: Callable[[], None] | None = None,
    solution_callback: Callable[[int, float, tuple[InitialState, ...]], None] | None = None,
    solution_interval: int = 1,
    cpu_time_offset: float = 0.0,
    device: str = "cuda:0",
    mesh_device=None,
    *,
    lusgs_ordering: str = "natural",
) -> SolverResult:
    """Advance a transient case or iterate a steady pseudo-time solution on
    the GPU (`solver.py:179-1008`). ``initial_states`` is a tuple of
    `InitialState` or a (C, 4) primitive array in mesh order. A prepared
    `DeviceMesh` of the same case may be passed as ``mesh_device``.
    ``lusgs_ordering`` "multicolour" sweeps the implicit preconditioner colour
    by colour (a different preconditioner than the reference's natural
    order, with the same converged solutions; much faster on large meshes)."""
    from .device import DeviceMesh

    cpu_start = thread_time()
    validate(case, start_time, step_offset, solution_interval, cpu_time_offset)
    s = case.numerics
    local = s.time_step_mode == "Local (steady)"
    if initial_states is None:
        raise SolverError("Could not evaluate initial conditions: pass the initial states.")
    if isinstance(initial_states, np.ndarray):
        primitive = np.asarray(initial_states, dtype=np.float64)
    else:
        primitive = np.array([(st.density, st.velocity_x, st.velocity_y, st.pressure) for st in initial_states],
                             dtype=float)
    if primitive.shape != (case.mesh.cell_count, 4):
        raise SolverError("Initial state count must match the mesh cell count.")
    try:
        conservative = primitive_to_conservative(primitive, s.gamma)
        conservative_to_primitive(conservative, s.gamma)
    except SolverError as error:
        raise SolverError(f"Invalid initial state: {error}") from error
    if lusgs_ordering not in ("natural", "multicolour"):
        raise SolverError(f"Unsupported LU-SGS ordering '{lusgs_ordering}'.")
    dm = mesh_device if mesh_device is not None else DeviceMesh(case, device)
    if getattr(dm, "_implicit", None) is None:
        dm.lusgs_ordering = lusgs_ordering
    elif dm.lusgs_ordering != lusgs_ordering:
        raise SolverError("The device mesh's implicit structure was built with another LU-SGS ordering.")
    dm.set_state(conservative)
    n = dm.n
    implicit = s.time_integration == "Implicit Euler"
    r = torch.empty((n, 4), dtype=torch.float64, device=dm.device)

    def rms_components(res):
        sums = dm.sum_squares(res).cpu().numpy()
        return np.sqrt(sums / n)

    def host_primitive():
        return conservative_to_primitive(dm.state(), s.gamma)

    time = start_time
    completed = False
    steps = step_offset
    dm.clock[0] = time
    dm.dt_fresh = False
    for local_step in range(1, s.max_iterations + 1):
        step = step_offset + local_step
        if pause_callback is not None:
            pause_callback()
        if cancel_callback is not None and cancel_callback():
            break
        report_due = progress_callback is not None and step % s.convergence_history_interval == 0
        try:
            want_r = progress_callback is not None and not local
            if implicit:
                if want_r:
                    dm.residual(dm.u, out=r)
                new_time = dm.step_implicit(time)
                time_step = new_time - time
            elif local:
                # this step's local steps: their mean is the reported step and the time advance
                if not dm.dt_fresh:
                    dm.time_steps(dm.u)
                    dm.dt_fresh = True
                time_step = float(dm.dt_cell.mean().item())
                dm.step(residual_out=None)
                new_time = time + time_step
            else:
                dm.step(residual_out=r if want_r else None)
                clock = dm.clock.cpu().numpy()
                new_time, time_step = float(clock[0]), float(clock[1])
            dm.check_status()
        except SolverError as error:
            raise SolverError(f"Solver failed at step {step}: {error}") from error
        time = new_time
        steps = step
        steady_converged = False
        if local:
            dm.residual(dm.u, out=r)
            try:
                dm.check_status()
            except SolverError as error:
                raise SolverError(f"Solver failed at step {step}: {error}") from error
            comps = rms_components(r)
            steady_residual = math.sqrt(float(np.mean(comps**2)))
            steady_converged = steady_residual <= s.residual_tolerance
        end_time_reached = not local and time >= s.end_time * (1.0 - 1e-13)
        final_step = (end_time_reached or steady_converged or local_step == s.max_iterations
                      or (cancel_callback is not None and cancel_callback()))
        report_residual = progress_callback is not None and (report_due or final_step)
        progress_states = None
        if progress_callback is not None:
            if live_solution_updates and (local_step == 1 or step % 10 == 0 or end_time_reached or steady_converged
                                          or local_step == s.max_iterations):
                progress_states = _states(host_primitive())
            if report_residual or progress_states is not None:
                component_residuals = ()
                if report_residual:
                    component_residuals = tuple(float(x) for x in (comps if local else rms_components(r)))
                progress_callback(SolverProgress(step, time, time_step, component_residuals, progress_states, None,
                                                 cpu_time_offset + thread_time() - cpu_start, None))
        if solution_callback is not None and step % solution_interval == 0:
            snapshot = progress_states if progress_states is not None else _states(host_primitive())
            solution_callback(step, time, snapshot)
        if end_time_reached:
            time = s.end_time
            completed = True
            break
        if steady_converged:
            completed = True
            break
    final = host_primitive()
    dm.residual(dm.u, out=r)
    try:
        dm.check_status()
    except SolverError as error:
        raise SolverError(f"Could not evaluate final residual: {error}") from error
    residual = math.sqrt(float(np.mean(rms_components(r) ** 2)))
    if mesh_device is None:
        dm.close()
    return SolverResult(_states(final), time, steps, residual, completed,
                        cpu_time_offset + thread_time() - cpu_start, final)
