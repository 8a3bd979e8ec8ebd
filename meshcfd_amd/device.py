"""Device-resident residual / time-step / stage driver over the C-ABI.

`DeviceMesh` uploads a case's geometry once (`mcfd_mesh_create`) and runs the
reference solver's per-step sequence (`solve_case`, `solver.py:437-812`) on
device buffers:

    time step:  mcfd_time_steps -> (global) mcfd_advance_clock
    stages:     mcfd_stage(op)  -- residual + stage update fused per stage

PyTorch (ROCm) is used only to own device memory and to name the stream; the
arithmetic is in libmcfdhip.so.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import inflow, native
from .geometry import SolverError, build_geometry
from .layout import build_layout
from .model import FLUX_CODES, LIMITER_CODES, RECONSTRUCTION_CODES, Case, NumericalSettings

RK4_COEFFS = (
    (0.391752226571890,),
    (0.444370493651235, 0.555629506348765, 0.368410593050371),
    (0.620101851488403, 0.379898148511597, 0.251891774271694),
    (0.178079954393132, 0.821920045606868, 0.544974750228521),
    (0.517231671970585, 0.096059710526147, 0.063692468666290, 0.386708617503269, 0.226007483236906),
)


def native_settings(s: NumericalSettings) -> native.McfdSettings:
    if s.reconstruction not in RECONSTRUCTION_CODES:
        raise SolverError(f"Unsupported reconstruction '{s.reconstruction}'.")
    if s.limiter not in LIMITER_CODES:
        raise SolverError(f"Unsupported limiter '{s.limiter}'.")
    if s.flux not in FLUX_CODES:
        raise SolverError(f"Unsupported convective flux '{s.flux}'.")
    models = {"Inviscid": 0, "Constant": 1, "Sutherland": 2}
    if s.viscosity_model not in models:
        raise SolverError("Unsupported viscosity model.")
    return native.McfdSettings(
        flux=FLUX_CODES[s.flux],
        reconstruction=RECONSTRUCTION_CODES[s.reconstruction],
        limiter=LIMITER_CODES[s.limiter],
        viscosity=models[s.viscosity_model],
        gamma=s.gamma,
        cfl=s.cfl,
        acceleration_x=s.acceleration_x,
        acceleration_y=s.acceleration_y,
        gas_constant=s.gas_constant,
        prandtl_number=s.prandtl_number,
        dynamic_viscosity=s.dynamic_viscosity,
        sutherland_reference_temperature=s.sutherland_reference_temperature,
        sutherland_reference_viscosity=s.sutherland_reference_viscosity,
        sutherland_temperature=s.sutherland_temperature,
    )


def _ptr(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


class DeviceMesh:
    """One mesh uploaded to one GPU, plus the step buffers of its case."""

    def __init__(
        self,
        case: Case,
        device: str | torch.device = "cuda:0",
        geometry=None,
        layout=None,
        fused: bool = True,
        block_cells: int = 128,
        fuse_time_step: bool = True,
    ):
        if not torch.cuda.is_available():
            raise native.NativeError("No ROCm device is visible: the GPU residual cannot run.")
        self.lib = native.load()
        self.case = case
        self.settings = case.numerics
        self.device = torch.device(device)
        self.geometry = geometry if geometry is not None else build_geometry(case)
        self.layout = layout if layout is not None else build_layout(self.geometry)
        self.n = self.layout.n_cells
        with torch.cuda.device(self.device):
            desc, keep = native.mesh_desc(self.layout)
            handle = C.c_void_p()
            native.check(self.lib.mcfd_mesh_create(C.byref(desc), C.byref(handle)))
            del keep
        self.handle = handle
        self.fuse_time_step = fuse_time_step
        self.blocks = None
        # viscous lists first: the fused kernel's first-order viscous path
        # packs their weights and items per block
        self._visc_lists = None
        if self.settings.viscosity_model != "Inviscid":
            from . import viscous

            self._visc_lists = viscous.upload(self.handle, self.geometry, self.layout)
        # pressure-inlet / mass-flow boundaries: exterior states from their own
        # kernels, read by the unfused flux and time-step kernels
        self.inflow = inflow.has_inflow(self.layout)
        if self.inflow:
            inflow.upload(self.lib, self.handle, self.geometry, self.layout)
        elif fused:
            self.set_blocks(block_cells)
        self.native_settings = native_settings(self.settings)
        if self.settings.reconstruction == "MUSCL parabolic":
            self.set_parabolic()
        opts = dict(dtype=torch.float64, device=self.device)
        n = self.n
        self.u = torch.empty((n, 4), **opts)
        self.s1 = torch.empty((n, 4), **opts)
        self.s2 = torch.empty((n, 4), **opts)
        self.s3 = torch.empty((n, 4), **opts)
        self.r3 = torch.empty((n, 4), **opts)
        self.r0 = torch.empty((n, 4), **opts)
        self.dt_cell = torch.empty((n,), **opts)
        self.dt_min = torch.empty((1,), **opts)
        self.clock = torch.zeros((2,), **opts)
        self.sums = torch.zeros((4,), **opts)
        self.case_ptr = None
        self.dt_fresh = False
        self.n_cases = 0
        self.case_clock = None
        self.graph = None
        # LU-SGS ordering of the implicit path: "natural" (the reference's) or
        # "multicolour" (few fully parallel levels); read when the implicit
        # structure is first built
        self.lusgs_ordering = "natural"

    @classmethod
    def from_batch(cls, batch, device="cuda:0", fused: bool = True, block_cells: int = 128):
        """A device mesh for a `batch.Batch`: all instances advance together,
        each with its own CFL step and clock."""
        from .batch import batch_case

        dm = cls(batch_case(batch), device, geometry=batch.geometry, layout=batch.layout, fused=fused,
                 block_cells=block_cells)
        dm.n_cases = batch.n_cases
        dm.case_ptr = torch.as_tensor(batch.case_ptr.astype(np.int32), device=dm.device)
        dm.case_clock = torch.zeros((2 * batch.n_cases,), dtype=torch.float64, device=dm.device)
        return dm

    def capture(self, steps: int = 1):
        """Capture `steps` explicit steps in a HIP graph (replayed by
        `replay()`). Requires a pointer-stable step (RK2/RK3/RK4) and no
        per-launch timing."""
        if self.settings.time_integration == "Explicit Euler" and steps % 2:
            raise ValueError("Euler steps swap buffers: capture an even number of steps")
        self.timing(False)
        saved = [t.clone() for t in (self.u, self.clock)] + ([self.case_clock.clone()] if self.case_clock is not None else [])
        stream = torch.cuda.Stream(self.device)
        stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            for _ in range(2):  # warm up kernels and allocator outside the graph
                self.step()
        torch.cuda.current_stream(self.device).wait_stream(stream)
        torch.cuda.synchronize(self.device)
        self.u.copy_(saved[0])
        self.clock.copy_(saved[1])
        self.dt_fresh = False
        if self.case_clock is not None:
            self.case_clock.copy_(saved[2])
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            for _ in range(steps):
                self.step()
        self.graph = graph
        self.graph_steps = steps
        return graph

    def replay(self):
        self.graph.replay()

    def set_parabolic(self) -> None:
        """Upload the two-ring stencils and precomputed quadratic-fit
        coefficients (MUSCL parabolic)."""
        from .parabolic import build_parabolic

        new_id = np.empty(self.n, dtype=np.int64)
        new_id[self.layout.cell_order] = np.arange(self.n)
        p = build_parabolic(self.geometry, new_id)
        self._parabolic = p
        native.check(
            self.lib.mcfd_mesh_set_parabolic(
                self.handle, len(p["nb"]), C.c_void_p(p["ptr"].ctypes.data), C.c_void_p(p["nb"].ctypes.data),
                C.c_void_p(p["coef"].ctypes.data), C.c_void_p(p["full"].ctypes.data),
            )
        )

    def set_blocks(self, block_cells: int | None) -> None:
        """Switch to the fused one-kernel-per-stage path with blocks of
        ``block_cells`` cells (None: back to the four-kernel path)."""
        from .blocks import build_blocks

        if block_cells is not None and self.inflow:
            raise SolverError("Pressure-inlet and mass-flow boundaries run the unfused kernels.")
        if block_cells is None:
            native.check(self.lib.mcfd_mesh_set_blocks(self.handle, None))
            self.blocks = None
            return
        self.blocks = build_blocks(self.layout, block_cells, viscous=self._visc_lists)
        desc, keep = native.block_desc(self.blocks)
        native.check(self.lib.mcfd_mesh_set_blocks(self.handle, C.byref(desc)))
        del keep
        if self.blocks.vent is not None:
            # first-order viscous residuals (the implicit path's Jacobian
            # products) as one fused launch
            vent = np.ascontiguousarray(self.blocks.vent, dtype=np.float64)
            vloc = np.ascontiguousarray(self.blocks.vacc_local, dtype=np.uint16)
            native.check(self.lib.mcfd_mesh_set_block_viscous(
                self.handle, C.c_void_p(vent.ctypes.data), len(vent), C.c_void_p(vloc.ctypes.data), len(vloc)))

    def autotune(self, candidates=(32, 48, 64, 80, 96, 112, 128), steps: int = 3, repeats: int = 2) -> int:
        """Pick the fused kernel's block size: for each candidate (cells per
        block, each block also limited to 256 faces) time ``steps`` explicit
        steps with HIP events, ``repeats`` times, from the same state; keep
        the fastest and restore the state. Returns the chosen size. Larger
        blocks load fewer halo cells per own cell; smaller ones give more
        workgroups to fill the GPU on small meshes (measured: 128 fastest by
        5-20% from 117k cells up; at 29k cells, where a step is launch-bound,
        every size within 3% and the choice varies between runs)."""
        if self.inflow:
            self.timings = {}
            return None
        saved = (self.u.clone(), self.clock.clone())
        stream = torch.cuda.current_stream(self.device)
        best, best_ms = None, math.inf
        self.timings = {}
        for b in candidates:
            try:
                self.set_blocks(b)
            except (ValueError, native.NativeError):
                continue
            times = []
            for _ in range(repeats):
                self.u.copy_(saved[0])
                self.clock.copy_(saved[1])
                self.dt_fresh = False
                self.step()  # warm-up, first-call set-up
                start = torch.cuda.Event(enable_timing=True)
                stop = torch.cuda.Event(enable_timing=True)
                start.record(stream)
                for _ in range(steps):
                    self.step()
                stop.record(stream)
                stop.synchronize()
                times.append(start.elapsed_time(stop) / steps)
            self.timings[b] = min(times)
            if self.timings[b] < best_ms:
                best, best_ms = b, self.timings[b]
        if best is None:
            raise SolverError("No block size fits this mesh.")
        self.set_blocks(best)
        self.u.copy_(saved[0])
        self.clock.copy_(saved[1])
        # the timed steps overwrote the CFL steps: the next step recomputes them
        self.dt_fresh = False
        self.check_status("autotune: ")
        return best

    def close(self):
        if getattr(self, "handle", None):
            self.lib.mcfd_mesh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- plumbing -----------------------------------------------------------
    @property
    def stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def check_status(self, context: str = "") -> None:
        mask = C.c_int32()
        index = C.c_int32()
        native.check(self.lib.mcfd_status(self.handle, self.stream, C.byref(mask), C.byref(index)))
        if mask.value:
            code = min(k for k in native.ERROR_MESSAGES if mask.value & (1 << k))
            message = native.ERROR_MESSAGES[code]
            if "{name}" in message:  # inflow records report their record index
                message = message.format(name=inflow.record_name(self.geometry, self.layout, index.value))
            elif not mask.value & ~native.CELL_ERROR_MASK and 0 <= index.value < self.n:
                # cell-indexed errors only: the lowest failing cell, in mesh order
                message = f"{message} First invalid cell: {int(self.layout.cell_order[index.value])}."
            raise SolverError(f"{context}{message}")

    # Device cells are stored in the layout's (Hilbert) order; these convert
    # between mesh order (what callers see) and device order.
    def to_device_order(self, array) -> torch.Tensor:
        a = np.asarray(array, dtype=np.float64)[self.layout.cell_order]
        return torch.as_tensor(np.ascontiguousarray(a), device=self.device)

    def to_mesh_order(self, tensor: torch.Tensor) -> np.ndarray:
        a = tensor.detach().cpu().numpy()
        out = np.empty_like(a)
        out[self.layout.cell_order] = a
        return out

    def set_state(self, conservative) -> None:
        self.u.copy_(self.to_device_order(conservative))
        self.dt_fresh = False

    def state(self) -> np.ndarray:
        return self.to_mesh_order(self.u)

    def _stage(self, op, u, out=None, u0=None, u1=None, r1=None, dt=None, per_cell=False, coeffs=(), r_out=None,
               dt_next=None, dt_min=None):
        st = native.McfdStage()
        st.op = op
        st.dt_per_cell = 1 if per_cell else 0
        st.u = _ptr(u)
        st.u0 = _ptr(u0)
        st.u1 = _ptr(u1)
        st.r1 = _ptr(r1)
        st.dt = _ptr(dt)
        st.out = _ptr(out)
        st.r_out = _ptr(r_out)
        st.dt_next = _ptr(dt_next)
        st.dt_min = _ptr(dt_min)
        for k, v in enumerate(coeffs):
            st.c[k] = v
        native.check(self.lib.mcfd_stage(self.handle, C.byref(self.native_settings), C.byref(st), self.stream))

    # -- seams (same meaning as the reference's module functions) ------------
    def residual(self, u: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """`_residual` (`solver.py:2437-2718`) on a device state (C,4)."""
        out = torch.empty_like(u) if out is None else out
        self._stage(native.OP_RESIDUAL, u, r_out=out)
        return out

    def _stage_first_order_residual(self, u, out, settings):
        native.check(self.lib.mcfd_residual(self.handle, C.byref(settings), _ptr(u), _ptr(out), self.stream))
        return out

    def step_implicit(self, time: float) -> float:
        """One implicit-Euler step of `solve_case` (`solver.py:447-513`);
        returns the new time (host-synchronising: Newton needs the norms)."""
        self.implicit_stepper()
        s = self.settings
        mode = s.time_step_mode
        if mode == "Local (steady)":
            self.time_steps(self.u)
            self.check_status()
            dt, per_cell = self.dt_cell.clone(), True
        elif mode == "Global fixed":
            dt = torch.full((1,), min(s.fixed_time_step, s.end_time - time), dtype=torch.float64, device=self.device)
            per_cell = False
        else:
            self.time_steps(self.u)
            self.check_status()
            value = min(float(self.dt_min.item()), s.end_time - time)
            dt = torch.full((1,), value, dtype=torch.float64, device=self.device)
            per_cell = False
        if not bool(torch.all(torch.isfinite(dt))) or bool(torch.any(dt <= 0.0)):
            raise SolverError("The stable time step is non-positive or non-finite.")
        state, accepted = self._implicit.step(self.u, dt, per_cell, allow_reduction=(mode != "Global fixed"))
        self.u.copy_(state)
        return time + float(accepted.mean().item())

    def bind_settings(self, settings: NumericalSettings) -> None:
        """Evaluate with ``settings`` from now on (the reconstruction and
        viscosity data uploaded at creation must cover them)."""
        if settings is self.settings:
            return
        self.settings = settings
        self.case.numerics = settings
        self.native_settings = native_settings(settings)
        if getattr(self, "_implicit", None) is not None:
            self._implicit.bind_settings()

    def implicit_stepper(self):
        """The implicit-Euler machinery of this mesh (built on first use)."""
        from .implicit import ImplicitStepper

        if getattr(self, "_implicit", None) is None:
            self._implicit = ImplicitStepper(self)
        return self._implicit

    def time_steps_primitive(self, w: torch.Tensor) -> torch.Tensor:
        """`_stable_time_steps` of device primitive states (C,4): per-cell
        steps into ``dt_cell`` and their minimum into ``dt_min``."""
        native.check(
            self.lib.mcfd_time_steps_primitive(
                self.handle, C.byref(self.native_settings), _ptr(w), _ptr(self.dt_cell), _ptr(self.dt_min), self.stream
            )
        )
        return self.dt_cell

    def time_steps(self, u: torch.Tensor) -> torch.Tensor:
        """`_stable_time_steps` (`solver.py:4132-4343`): per-cell steps into
        ``dt_cell`` and their minimum into ``dt_min``."""
        native.check(
            self.lib.mcfd_time_steps(
                self.handle, C.byref(self.native_settings), _ptr(u), _ptr(self.dt_cell), _ptr(self.dt_min), self.stream
            )
        )
        return self.dt_cell

    def sum_squares(self, r: torch.Tensor) -> torch.Tensor:
        native.check(self.lib.mcfd_sum_squares(self.handle, _ptr(r), _ptr(self.sums), self.stream))
        return self.sums

    # -- one explicit step --------------------------------------------------
    def step(self, time: float | None = None, residual_out: torch.Tensor | None = None) -> None:
        """Advance ``self.u`` by one explicit step on the stream; no host sync.

        Global modes keep [time, dt] in ``self.clock`` on the device. When
        ``time`` is given the clock's time is set to it first.
        """
        s = self.settings
        mode = s.time_step_mode
        local = mode == "Local (steady)"
        if time is not None:
            self.clock[0] = time
        if self.case_ptr is not None:
            if self.dt_fresh:  # the previous final stage wrote the per-cell steps
                native.check(
                    self.lib.mcfd_batch_case_steps(
                        self.handle, self.n_cases, _ptr(self.case_ptr), s.end_time, _ptr(self.case_clock),
                        _ptr(self.dt_cell), self.stream,
                    )
                )
            else:
                native.check(
                    self.lib.mcfd_batch_time_steps(
                        self.handle, C.byref(self.native_settings), _ptr(self.u), self.n_cases, _ptr(self.case_ptr),
                        s.end_time, _ptr(self.case_clock), _ptr(self.dt_cell), self.stream,
                    )
                )
            dt, per_cell = self.dt_cell, True
        elif local:
            if not self.dt_fresh:
                self.time_steps(self.u)
            dt, per_cell = self.dt_cell, True
        elif mode == "Global fixed":
            native.check(
                self.lib.mcfd_advance_clock(self.handle, None, s.fixed_time_step, s.end_time, _ptr(self.clock), 0,
                                            self.stream)
            )
            dt, per_cell = self.clock[1:], False
        else:
            if not self.dt_fresh:
                self.time_steps(self.u)
            native.check(
                self.lib.mcfd_advance_clock(self.handle, _ptr(self.dt_min), 0.0, s.end_time, _ptr(self.clock), 1,
                                            self.stream)
            )
            dt, per_cell = self.clock[1:], False
        # the fused final stage also computes the next step's CFL steps
        fuse_dt = (
            self.fuse_time_step and self.blocks is not None and mode != "Global fixed"
            and self.settings.viscosity_model == "Inviscid" and self.settings.reconstruction != "MUSCL parabolic"
        )
        nxt = {}
        if fuse_dt:  # the global modes need only the minimum; batches and local steps the per-cell steps
            per_cell_next = local or self.case_ptr is not None
            nxt = dict(dt_next=self.dt_cell, dt_min=self.dt_min) if per_cell_next else dict(dt_min=self.dt_min)
        self.advance(s.time_integration, dt, per_cell, residual_out, nxt)
        if not fuse_dt:
            # the end-of-step state check (solver.py:806-808); the fused final
            # stage makes it while computing the next step's CFL steps
            native.check(self.lib.mcfd_check_state(self.handle, C.byref(self.native_settings), _ptr(self.u),
                                                   self.stream))
        self.dt_fresh = fuse_dt

    def advance(self, ti: str, dt: torch.Tensor, per_cell: bool, residual_out: torch.Tensor | None = None,
                nxt: dict | None = None) -> None:
        """The stages of one explicit step of integrator ``ti`` on ``self.u``
        with the step sizes ``dt`` (device, [1] or [C]); ``nxt`` asks the
        final fused stage for the next step's CFL steps."""
        nxt = nxt or {}
        u = self.u
        if ti == "Explicit Euler":
            self._stage(native.OP_EULER, u, out=self.s1, dt=dt, per_cell=per_cell, coeffs=(1.0,), r_out=residual_out,
                        **nxt)
            self.u, self.s1 = self.s1, u
        elif ti == "Explicit RK2":
            # the last stage reads U^n only at its own cell, so it may write U^n in place
            self._stage(native.OP_EULER, u, out=self.s1, dt=dt, per_cell=per_cell, coeffs=(1.0,), r_out=residual_out)
            self._stage(native.OP_SSP, self.s1, out=u, u0=u, dt=dt, per_cell=per_cell, coeffs=(0.5, 0.5), **nxt)
        elif ti == "Explicit RK3":
            self._stage(native.OP_EULER, u, out=self.s1, dt=dt, per_cell=per_cell, coeffs=(1.0,), r_out=residual_out)
            self._stage(native.OP_SSP, self.s1, out=self.s2, u0=u, dt=dt, per_cell=per_cell, coeffs=(0.75, 0.25))
            self._stage(native.OP_SSP, self.s2, out=u, u0=u, dt=dt, per_cell=per_cell, coeffs=(1.0 / 3.0, 2.0 / 3.0),
                        **nxt)
        elif ti == "Explicit RK4":
            c = RK4_COEFFS
            if residual_out is not None:
                self.residual(u, out=residual_out)
            self._stage(native.OP_EULER, u, out=self.s1, dt=dt, per_cell=per_cell, coeffs=c[0])
            self._stage(native.OP_RK4, self.s1, out=self.s2, u0=u, dt=dt, per_cell=per_cell, coeffs=c[1])
            self._stage(native.OP_RK4, self.s2, out=self.s3, u0=u, dt=dt, per_cell=per_cell, coeffs=c[2], r_out=None)
            # stage four needs R(stage three) kept for the final combination
            self._stage(native.OP_RK4, self.s3, out=self.s1, u0=u, dt=dt, per_cell=per_cell, coeffs=c[3], r_out=self.r3)
            # final: a*S2 + b*S3 + (c*dt)*R3 + d*S4 + (e*dt)*R(S4)
            self._stage(native.OP_RK4_FINAL, self.s1, out=u, u0=self.s2, u1=self.s3, r1=self.r3, dt=dt, per_cell=per_cell,
                        coeffs=c[4], **nxt)
        else:
            raise SolverError(f"Unsupported time integration '{ti}' for the explicit GPU path.")

    def run(self, steps: int, time: float = 0.0, check_every: int = 0) -> float:
        """Advance ``steps`` steps; returns the final time (one host sync)."""
        self.clock[0] = time
        self.dt_fresh = False
        for k in range(steps):
            self.step()
            if check_every and (k + 1) % check_every == 0:
                self.check_status(f"Solver failed at step {k + 1}: ")
        self.check_status("Solver failed: ")
        return float(self.clock[0].item())

    def timing(self, enable: bool):
        total = (C.c_double * 8)()
        counts = (C.c_int32 * 8)()
        native.check(self.lib.mcfd_timing(self.handle, 1 if enable else 0, total, counts))
        return {native.KERNEL_NAMES[k]: (total[k], counts[k]) for k in range(8) if counts[k]}
