"""`solve_case` on the GPU: the reference's driver contract over `DeviceMesh`.

Mirrors `solve_case` (`solver.py:179-1008`): the same arguments, settings
strings and validation messages, the same step order (time step from the
step's start state, the integrator, time += dt, the steady residual after
the step in "Local (steady)" mode), the same progress, solution and cancel
callbacks, `SolverProgress` / `SolverResult` (`solver.py:58-77`) and
`SolverError` with "Solver failed at step {n}: ..." around step failures.
The state stays on the device; the host reads the clock and the status word
once per step and pulls states only for callbacks and the result.

With ``adaptive_mesh`` the mesh is adapted every ``adaptation_interval``
steps on the host (`adaptation.py`: indicator, refinement and coarsening, the
conservative transfer of the state) and the device copy is rebuilt
(`solver.py:816-884`).

Not covered (raise `SolverError`): SST turbulence, the ILU preconditioner,
and evaluating `case.initial_conditions` (the caller passes the initial
states).
"""

from __future__ import annotations

import math
import warnings
import weakref
from dataclasses import dataclass
from time import thread_time
from typing import Callable

import numpy as np
import torch

from .adaptation import AdaptationBalanceError
from .geometry import SolverError
from .implicit import halve_until_solved
from .model import Case, FiniteVolumeMesh, InitialState, NumericalSettings


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
    if not isinstance(s.adaptive_mesh, bool):
        raise SolverError("Adaptive mesh setting must be boolean.")
    if isinstance(s.adaptation_interval, bool) or not isinstance(s.adaptation_interval, int) \
            or s.adaptation_interval <= 0:
        raise SolverError("Adaptation interval must be a positive integer.")
    if isinstance(s.adaptation_max_level, bool) or not isinstance(s.adaptation_max_level, int) \
            or not 1 <= s.adaptation_max_level <= 12:
        raise SolverError("Maximum adaptation level must be an integer from 1 to 12.")
    if not math.isfinite(s.adaptation_refine_threshold) or s.adaptation_refine_threshold <= 0.0:
        raise SolverError("Adaptation threshold must be positive and finite.")
    if s.adaptation_variable not in ("Density", "Velocity X", "Velocity Y", "Pressure", "Turbulent k",
                                     "Specific dissipation rate"):
        raise SolverError("Unsupported adaptive mesh indicator variable.")
    if s.adaptive_mesh and s.adaptation_variable in ("Turbulent k", "Specific dissipation rate"):
        raise SolverError("Turbulence-based adaptation indicators require the SST model.")
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


# ---------------------------------------------------------------------------
# Seams: the reference's module functions (`solver.py`), NumPy in and out, in
# mesh order. Each runs its arithmetic on the GPU (a device copy of the
# geometry is kept per geometry object). `solve_case` calls them by module
# name whenever one of them has been replaced (as the reference's tests
# replace them with monkeypatch); otherwise it keeps the state on the device
# and runs the same step sequence with the fused stage kernels.
# ---------------------------------------------------------------------------

SEAM_DEVICE = "cuda:0"
_seam_meshes: dict = {}


def _drop_seam_mesh(key) -> None:
    dm = _seam_meshes.pop(key, None)
    if dm is not None:
        dm.close()


def _seam_mesh(geometry, settings: NumericalSettings):
    """The device mesh of ``geometry`` bound to ``settings`` (kept while the
    geometry object lives)."""
    from .device import DeviceMesh

    key = (id(geometry), settings.reconstruction == "MUSCL parabolic", settings.viscosity_model != "Inviscid")
    dm = _seam_meshes.get(key)
    if dm is None:
        dm = DeviceMesh(Case(mesh=None, boundary_conditions=[], numerics=settings), SEAM_DEVICE, geometry=geometry)
        _seam_meshes[key] = dm
        weakref.finalize(geometry, _drop_seam_mesh, key)
    dm.bind_settings(settings)
    return dm


def _no_turbulence(turbulence) -> None:
    if turbulence is not None:
        raise SolverError("The Menter k-omega SST model is not implemented by the GPU path.")


def _device_steps(dm, time_step):
    """A scalar step or per-cell steps (mesh order) as a device tensor."""
    if np.ndim(time_step) == 0:
        return torch.full((1,), float(time_step), dtype=torch.float64, device=dm.device), False
    return dm.to_device_order(np.asarray(time_step, dtype=np.float64)), True


def _residual(conservative, geometry, settings, turbulence_conservative=None) -> np.ndarray:
    """`_residual` (`solver.py:2437-2718`): the spatial rate dU/dt of a (C,4)
    conservative state."""
    _no_turbulence(turbulence_conservative)
    dm = _seam_mesh(geometry, settings)
    r = dm.residual(dm.to_device_order(conservative))
    dm.check_status()
    return dm.to_mesh_order(r)


def _stable_time_steps(primitive, geometry, settings, turbulence_conservative=None) -> np.ndarray:
    """`_stable_time_steps` (`solver.py:4132-4343`): per-cell CFL steps of a
    (C,4) primitive state."""
    _no_turbulence(turbulence_conservative)
    dm = _seam_mesh(geometry, settings)
    w = dm.to_device_order(primitive)
    dm.time_steps_primitive(w)
    dm.check_status()
    return dm.to_mesh_order(dm.dt_cell)


def _stable_time_step(primitive, geometry, settings, turbulence_conservative=None) -> float:
    """`_stable_time_step` (`solver.py:4117-4129`): the smallest local step."""
    return float(np.min(_stable_time_steps(primitive, geometry, settings, turbulence_conservative)))


def _scale_cell_rate(rate: np.ndarray, time_step) -> np.ndarray:
    """`_scale_cell_rate` (`solver.py:1119-1125`)."""
    if np.ndim(time_step) == 0:
        return time_step * rate
    step_sizes = np.asarray(time_step, dtype=float)
    return step_sizes * rate if rate.ndim == 1 else step_sizes[:, None] * rate


def _time_state_rhs(state, geometry, settings) -> np.ndarray:
    """`_time_state_rhs` (`solver.py:1051-1105`) for the four flow variables."""
    if np.shape(state)[1] != 4:
        _no_turbulence(state)
    return _residual(state, geometry, settings)


def _explicit_rk4_step(state, time_step, geometry, settings) -> np.ndarray:
    """`_explicit_rk4_step` (`solver.py:1128-1182`), SSPRK(5,4): the five
    fused stages on the GPU; through `_time_state_rhs` on the host arrays if
    that seam or `_residual` has been replaced."""
    if _time_state_rhs is not _DEVICE_SEAMS["_time_state_rhs"] or _residual is not _DEVICE_SEAMS["_residual"]:
        return _rk4_through_rhs(np.asarray(state, dtype=np.float64), time_step, geometry, settings)
    dm = _seam_mesh(geometry, settings)
    dm.u.copy_(dm.to_device_order(state))
    dt, per_cell = _device_steps(dm, time_step)
    dm.advance("Explicit RK4", dt, per_cell)
    dm.check_status()
    return dm.to_mesh_order(dm.u)


def _rk4_through_rhs(initial, h, geometry, settings):
    """The SSPRK(5,4) stage sequence of `solver.py:1142-1181` over
    `_time_state_rhs` (its coefficients are the device kernels' RK4_COEFFS)."""
    from .device import RK4_COEFFS as c

    def scaled(rate, k):
        return _scale_cell_rate(rate, k * h) if np.ndim(h) == 0 else _scale_cell_rate(rate, k * np.asarray(h))

    d = _time_state_rhs(initial, geometry, settings)
    s1 = initial + scaled(d, c[0][0])
    d = _time_state_rhs(s1, geometry, settings)
    s2 = c[1][0] * initial + c[1][1] * s1 + scaled(d, c[1][2])
    d = _time_state_rhs(s2, geometry, settings)
    s3 = c[2][0] * initial + c[2][1] * s2 + scaled(d, c[2][2])
    d3 = _time_state_rhs(s3, geometry, settings)
    s4 = c[3][0] * initial + c[3][1] * s3 + scaled(d3, c[3][2])
    d4 = _time_state_rhs(s4, geometry, settings)
    return c[4][0] * s2 + c[4][1] * s3 + scaled(d3, c[4][2]) + c[4][3] * s4 + scaled(d4, c[4][4])


def _implicit_euler_solve_once(initial, time_step, geometry, settings) -> np.ndarray:
    """`_implicit_euler_solve_once` (`solver.py:1371-1486`): one Newton solve
    of U - U^n - dt R(U) = 0 at the given step (matrix-free GMRES with the
    LU-SGS preconditioner on the GPU)."""
    dm = _seam_mesh(geometry, settings)
    stepper = dm.implicit_stepper()
    dt, per_cell = _device_steps(dm, time_step)
    state = stepper.solve_once(dm.to_device_order(initial), dt, per_cell)
    return dm.to_mesh_order(state)


def _implicit_euler_step(initial, time_step, geometry, settings, allow_time_step_reduction: bool = True):
    """`_implicit_euler_step` (`solver.py:1337-1368`): `_implicit_euler_solve_once`
    with the step halved up to 12 times; returns (state, accepted step)."""
    h = float(time_step) if np.ndim(time_step) == 0 else np.asarray(time_step, dtype=float).copy()
    return halve_until_solved(lambda step: _implicit_euler_solve_once(initial, step, geometry, settings), h,
                              allow_time_step_reduction)


_SEAM_NAMES = ("_residual", "_stable_time_steps", "_stable_time_step", "_time_state_rhs", "_explicit_rk4_step",
               "_implicit_euler_solve_once", "_implicit_euler_step")
_DEVICE_SEAMS = {name: globals()[name] for name in _SEAM_NAMES}


def seams_replaced() -> bool:
    """True when a caller has replaced one of this module's seams."""
    g = globals()
    return any(g[name] is not fn for name, fn in _DEVICE_SEAMS.items())


# ---------------------------------------------------------------------------
# solve_case
# ---------------------------------------------------------------------------


class _DeviceRun:
    """The step sequence with the state resident on the GPU (`DeviceMesh`)."""

    def __init__(self, case, conservative, device, mesh_device, lusgs_ordering):
        from .device import DeviceMesh

        self.own = mesh_device is None
        dm = mesh_device if mesh_device is not None else DeviceMesh(case, device)
        if getattr(dm, "_implicit", None) is None:
            dm.lusgs_ordering = lusgs_ordering
        elif dm.lusgs_ordering != lusgs_ordering:
            raise SolverError("The device mesh's implicit structure was built with another LU-SGS ordering.")
        dm.set_state(conservative)
        self.dm = dm
        self.s = case.numerics
        self.r = torch.empty((dm.n, 4), dtype=torch.float64, device=dm.device)
        self.comps = None

    def start(self, time):
        self.dm.clock[0] = time
        self.dm.dt_fresh = False

    def _rms(self):
        return np.sqrt(self.dm.sum_squares(self.r).cpu().numpy() / self.dm.n)

    def step(self, time, want_residual):
        """One step from ``time``; returns (new time, reported step). With
        ``want_residual`` keeps R(start) for the progress report."""
        dm, s = self.dm, self.s
        mode = s.time_step_mode
        if s.time_integration == "Implicit Euler":
            if want_residual:
                dm.residual(dm.u, out=self.r)
            new_time = dm.step_implicit(time)
            time_step = new_time - time
        elif mode == "Local (steady)":
            # this step's local steps: their mean is the reported step and the time advance
            if not dm.dt_fresh:
                dm.time_steps(dm.u)
                dm.dt_fresh = True
            time_step = float(dm.dt_cell.mean().item())
            dm.step(residual_out=None)
            new_time = time + time_step
        else:
            dm.step(residual_out=self.r if want_residual else None)
            clock = dm.clock.cpu().numpy()
            new_time, time_step = float(clock[0]), float(clock[1])
        dm.check_status()
        return new_time, time_step

    def steady_residual(self):
        self.dm.residual(self.dm.u, out=self.r)
        self.dm.check_status()
        self.comps = self._rms()
        return math.sqrt(float(np.mean(self.comps**2)))

    def components(self, local):
        return self.comps if local else self._rms()

    @property
    def geometry(self):
        return self.dm.geometry

    def conservative(self):
        return self.dm.state()

    def remesh(self, case, conservative, device):
        """Continue on an adapted mesh: a new device copy of its geometry,
        the transferred state."""
        from .device import DeviceMesh

        ordering = self.dm.lusgs_ordering
        clock = self.dm.clock.clone()  # [time, dt] of the global modes
        self.close()
        self.own = True
        self.dm = DeviceMesh(case, device)
        self.dm.lusgs_ordering = ordering
        self.dm.set_state(conservative)
        self.dm.clock.copy_(clock)
        self.r = torch.empty((self.dm.n, 4), dtype=torch.float64, device=self.dm.device)

    def primitive(self):
        return conservative_to_primitive(self.dm.state(), self.s.gamma)

    def final_residual(self):
        self.dm.residual(self.dm.u, out=self.r)
        self.dm.check_status()
        return math.sqrt(float(np.mean(self._rms() ** 2)))

    def close(self):
        if self.own:
            self.dm.close()


class _SeamRun:
    """The reference's step sequence (`solver.py:437-812`) through this
    module's seams, on host arrays: used when a seam has been replaced."""

    def __init__(self, case, conservative):
        from .geometry import build_geometry

        self.geometry = build_geometry(case)
        self.s = case.numerics
        self.u = np.asarray(conservative, dtype=np.float64).copy()
        self.derivative = None

    def start(self, time):
        pass

    def step(self, time, want_residual):
        s, g = self.s, self.geometry
        gamma = s.gamma
        mode = s.time_step_mode
        primitive = conservative_to_primitive(self.u, gamma)
        if mode == "Local (steady)":
            sizes = _stable_time_steps(primitive, g, s)
            time_step = float(np.mean(sizes))
        elif mode == "Global fixed":
            time_step = min(s.fixed_time_step, s.end_time - time)
            sizes = time_step
        else:
            time_step = min(_stable_time_step(primitive, g, s), s.end_time - time)
            sizes = time_step
        if not math.isfinite(time_step) or time_step <= 0.0:
            raise SolverError("The stable time step is non-positive or non-finite.")
        start = self.u
        derivative = _residual(start, g, s)
        ti = s.time_integration
        if ti == "Implicit Euler":
            new, accepted = _implicit_euler_step(start, sizes, g, s,
                                                 allow_time_step_reduction=mode != "Global fixed")
            time_step = float(np.mean(np.asarray(accepted)))
        elif ti == "Explicit RK4":
            new = _explicit_rk4_step(start, sizes, g, s)
        elif ti == "Explicit Euler":
            new = start + _scale_cell_rate(derivative, sizes)
        else:
            one = start + _scale_cell_rate(derivative, sizes)
            conservative_to_primitive(one, gamma)
            d1 = _residual(one, g, s)
            if ti == "Explicit RK2":
                new = 0.5 * start + 0.5 * (one + _scale_cell_rate(d1, sizes))
            else:
                two = 0.75 * start + 0.25 * (one + _scale_cell_rate(d1, sizes))
                conservative_to_primitive(two, gamma)
                d2 = _residual(two, g, s)
                new = (1.0 / 3.0) * start + (2.0 / 3.0) * (two + _scale_cell_rate(d2, sizes))
        conservative_to_primitive(new, gamma)
        self.u = new
        self.derivative = derivative
        return time + time_step, time_step

    def steady_residual(self):
        self.derivative = _residual(self.u, self.geometry, self.s)
        return float(np.sqrt(np.mean(self.derivative**2)))

    def components(self, local):
        return np.sqrt(np.mean(self.derivative**2, axis=0))

    def conservative(self):
        return self.u

    def remesh(self, case, conservative, device):
        from .geometry import build_geometry

        self.geometry = build_geometry(case)
        self.u = np.asarray(conservative, dtype=np.float64)

    def primitive(self):
        return conservative_to_primitive(self.u, self.s.gamma)

    def final_residual(self):
        return float(np.sqrt(np.mean(_residual(self.u, self.geometry, self.s) ** 2)))

    def close(self):
        pass


def _adapt(case, run, device, local: bool, want_components: bool):
    """One adaptation of `solve_case` (`solver.py:816-884`): indicator of the
    current state, marks, the adapted mesh, the transferred state; the run
    continues on the new mesh. Returns (adapted mesh or None, the residual
    components of the step's start on the old mesh, for the report)."""
    from . import adaptation as ad

    s = case.numerics
    mesh = case.mesh
    values = ad.indicator_values(run.primitive(), s.adaptation_variable)
    g = run.geometry
    marks, coarsen = ad.select_adaptation(mesh, case, g.centers, g.areas, values, s.adaptation_refine_threshold,
                                          s.adaptation_max_level)
    if not marks and not coarsen:
        return None, None
    old_areas = ad.cell_areas(mesh)
    adapted, groups = ad.adapt_mesh(mesh, marks, coarsen, periodic_pairs=case.periodic_pairs)
    if adapted is mesh:
        return None, None
    comps = run.components(local) if (want_components and not local) else None
    state = ad.transfer_cell_values(run.conservative(), old_areas, ad.cell_areas(adapted), groups)
    case.mesh = adapted
    run.remesh(case, state, device)
    return adapted, comps


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
    order, with the same converged solutions; much faster on large meshes).
    When a caller has replaced one of this module's seams (`_residual`,
    `_stable_time_steps`, ...), every step goes through the seams by name, as
    in the reference."""
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
    if seams_replaced():
        run = _SeamRun(case, conservative)
    else:
        run = _DeviceRun(case, conservative, device, mesh_device, lusgs_ordering)

    time = start_time
    completed = False
    steps = step_offset
    adaptation_disabled = False
    run.start(time)
    for local_step in range(1, s.max_iterations + 1):
        step = step_offset + local_step
        if pause_callback is not None:
            pause_callback()
        if cancel_callback is not None and cancel_callback():
            break
        try:
            time, time_step = run.step(time, progress_callback is not None and not local)
        except (FloatingPointError, ValueError, ZeroDivisionError) as error:
            raise SolverError(f"Solver failed at step {step}: {error}") from error
        steps = step
        adapted_mesh = None
        adaptation_warning = None
        saved_components = None
        if s.adaptive_mesh and not adaptation_disabled and step % s.adaptation_interval == 0:
            try:
                adapted_mesh, saved_components = _adapt(case, run, device, local, progress_callback is not None)
            except AdaptationBalanceError as error:
                adaptation_disabled = True
                adaptation_warning = f"Adaptive mesh disabled at step {step}: {error}"
            except (ValueError, FloatingPointError, np.linalg.LinAlgError) as error:
                raise SolverError(f"Adaptive mesh update failed at step {step}: {error}") from error
        steady_converged = False
        if local:
            try:
                steady_residual = run.steady_residual()
            except SolverError as error:
                raise SolverError(f"Solver failed at step {step}: {error}") from error
            steady_converged = steady_residual <= s.residual_tolerance
        end_time_reached = not local and time >= s.end_time * (1.0 - 1e-13)
        final_step = (end_time_reached or steady_converged or local_step == s.max_iterations
                      or (cancel_callback is not None and cancel_callback()))
        report_residual = progress_callback is not None and (step % s.convergence_history_interval == 0
                                                             or adapted_mesh is not None or final_step)
        progress_states = None
        if progress_callback is not None:
            if (live_solution_updates and (local_step == 1 or step % 10 == 0 or end_time_reached or steady_converged
                                           or local_step == s.max_iterations)) or adapted_mesh is not None:
                progress_states = _states(run.primitive())
            if report_residual or progress_states is not None or adaptation_warning:
                component_residuals = ()
                if report_residual:
                    comps = saved_components if saved_components is not None else run.components(local)
                    component_residuals = tuple(float(x) for x in comps)
                progress_callback(SolverProgress(step, time, time_step, component_residuals, progress_states,
                                                 adapted_mesh, cpu_time_offset + thread_time() - cpu_start,
                                                 adaptation_warning))
        if adaptation_warning and progress_callback is None:
            warnings.warn(adaptation_warning, RuntimeWarning, stacklevel=2)
        if solution_callback is not None and step % solution_interval == 0:
            snapshot = progress_states if progress_states is not None else _states(run.primitive())
            solution_callback(step, time, snapshot)
        if end_time_reached:
            time = s.end_time
            completed = True
            break
        if steady_converged:
            completed = True
            break
    final = run.primitive()
    try:
        residual = run.final_residual()
    except SolverError as error:
        raise SolverError(f"Could not evaluate final residual: {error}") from error
    run.close()
    return SolverResult(_states(final), time, steps, residual, completed,
                        cpu_time_offset + thread_time() - cpu_start, final)
