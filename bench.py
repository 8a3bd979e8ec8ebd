"""Benchmark of the residual hot path on one MI355X.

Workload: four-quadrant Riemann problem on [0,1]^2, 1000x1000 quads with
0.15h jitter and 30% of quads split into triangles (seed 1): 1,300,000 cells,
HLLC + limited MUSCL-linear (Barth-Jespersen), SSP-RK3, CFL 0.8, FP64,
transmissive boundaries. One "step" = one SSP-RK3 step (CFL time step + three
residual/stage evaluations) over the whole mesh.

Prints one JSON line. ``value`` = cells * steps / seconds / 1e6
(Mcell-updates/s) with the state resident in HBM, from the median of
``--repeats`` timed windows of ``--steps`` steps; ``roofline`` reports the
dominant kernel against the 8 TB/s HBM peak from its algorithmic bytes; the
``stage`` figures use B_alg per cell-update over the whole step;
``cpu_baseline`` times the CPU oracle on one host core on a bounded sample
(median of five runs), and under ``all_cores`` its OpenMP restatement (same
bits) on every core of the process; ``batch`` times 256 cases of uneven size advanced
together (HIP-graph replays); ``host`` names the CPU, GPU and library
versions.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(layout, reconstruction: str, integrator: str):
    """B_alg per cell-update and per-kernel compulsory bytes per launch.

    Per stage: 32 (U in) + 32 (U out) + 32*[stage reads U0] + g_cell +
    g_face*F/C with MUSCL g_cell = 24 (1/A + centroid), g_face = 40
    (owner+neighbour int32, n*L 2xfp64, face midpoint 2xfp64); first order
    g_cell = 8, g_face = 24.
    """
    c = layout.n_cells
    f = layout.n_faces
    muscl = reconstruction != "First order"
    g_cell, g_face = (24.0, 40.0) if muscl else (8.0, 24.0)
    base = 64.0 + g_cell + g_face * f / c
    stages = {"Explicit Euler": [0], "Explicit RK2": [0, 1], "Explicit RK3": [0, 1, 1]}[integrator]
    b_alg = sum(base + 32.0 * u0 for u0 in stages)
    n_grad = len(layout.grad_nb)
    n_lim = len(layout.lim_dx)
    n_acc = len(layout.acc_item)
    n_dts = len(layout.dts_nx)
    kernels = {
        # fused stage kernel: B_alg of one stage (averaged over the step's stages)
        "stage": b_alg * c / len(stages),
        # U in (32), W out (32)
        "primitive": 64.0 * c,
        # W (32) + ptrs (8) + per LSQ entry nb+cx+cy (20) + per limiter entry dx,dy (16) + G,PHI out (96)
        "gradient": (32 + 8 + 96) * c + 20.0 * n_grad + 16.0 * n_lim,
        # per face: ids, kind, bc (16) + n, L (24) + offsets (32) + flux out (32); W,G,PHI once per cell (128)
        "flux": 104.0 * f + (128.0 if muscl else 32.0) * c,
        # per cell: ptr (4) + area (8) + U, U0, out (96); per item 4; per face flux + length once (40)
        "update": (4 + 8 + 96) * c + 4.0 * n_acc + 40.0 * f,
        # per cell: W (32) + ptr (4) + area (8) + dt out (8); per entry n, L, bc (28)
        "time_step": (32 + 4 + 8 + 8) * c + 28.0 * n_dts,
    }
    return b_alg, kernels


def adjoint_rate(dm, torch, repeats: int = 20):
    """Adjoint throughput on the benchmark state: lambda^T dR/dU
    (`mcfd_residual_vjp`, one full reverse residual per gradient) against one
    forward residual (`mcfd_residual`), HIP events on the current stream."""
    from meshcfd_amd import adjoint

    adjoint.enable(dm)
    u = dm.u.clone()
    gen = torch.Generator(device="cpu").manual_seed(0)
    lam = torch.randn(u.shape, generator=gen, dtype=torch.float64).to(u.device)
    out = torch.empty_like(u)
    r = torch.empty_like(u)
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(repeats):
            fn()
        stop.record(stream)
        stop.synchronize()
        return start.elapsed_time(stop) / repeats

    vjp_ms = timed(lambda: adjoint.residual_vjp(dm, u, lam, out=out))
    res_ms = timed(lambda: dm.residual(u, out=r))
    dm.check_status("adjoint timing: ")
    return {
        "grads_per_s": round(1e3 / vjp_ms, 1),
        "vjp_ms": round(vjp_ms, 4),
        "residual_ms": round(res_ms, 4),
        "vjp_over_residual": round(vjp_ms / res_ms, 3),
        "what": "lambda^T dR/dU of the full HLLC + MUSCL-linear residual at the benchmark state",
    }


def host_info() -> dict:
    """CPU model, NumPy/SciPy/torch versions, GPU name and ROCm version."""
    import scipy
    import torch

    model = platform.processor() or platform.machine()
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    rocm = None
    try:
        rocm = Path("/opt/rocm/.info/version").read_text().strip()
    except OSError:
        pass
    return {
        "cpu_model": model,
        "host_cores": os.cpu_count(),
        "affinity_cores": len(os.sched_getaffinity(0)),
        "numpy": np.__version__,
        "scipy": scipy.__version__,
        "torch": torch.__version__,
        "hip_runtime": getattr(torch.version, "hip", None),
        "rocm": rocm,
        "gpu": torch.cuda.get_device_name(0) if torch.cuda.is_available() else None,
        "gpu_arch": getattr(torch.cuda.get_device_properties(0), "gcnArchName", None)
        if torch.cuda.is_available() else None,
        "gpu_cus": torch.cuda.get_device_properties(0).multi_processor_count if torch.cuda.is_available() else None,
    }


def cpu_baseline(case, geometry, primitive, steps: int, samples: int = 5):
    """Time the reference-equivalent CPU path on one core: the NumPy/SciPy
    restatement of the solver with single-threaded C stand-ins for the
    reference's numba kernels (oracle/mcfd_oracle.c), as the reference runs
    them compiled. ``samples`` runs of ``steps`` SSP-RK3 steps each from the
    initial state after one warm-up residual; the rate is the median."""
    from oracle import compiled
    from oracle.euler import primitive_to_conservative
    from oracle.solver import Oracle

    use_c = compiled.available()
    if use_c:
        compiled.enable()

    old = os.sched_getaffinity(0)
    core = min(old)
    os.sched_setaffinity(0, {core})
    try:
        oracle = Oracle(geometry, case.numerics)
        u0 = primitive_to_conservative(primitive, case.numerics.gamma)
        oracle.residual(u0)  # warm-up residual (benchmarks/run_benchmarks.py:266-267)
        times = []
        for _ in range(samples):
            u = u0
            t0 = time.perf_counter()
            t = 0.0
            for _ in range(steps):
                u, dt, _, _ = oracle.step(u, t)
                t += dt
            times.append(time.perf_counter() - t0)
        elapsed = float(np.median(times))
    finally:
        os.sched_setaffinity(0, old)
        if use_c:
            compiled.disable()
    cells = geometry.cell_count
    return {
        "value": cells * steps / elapsed / 1e6,
        "unit": "Mcell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"median of {samples} runs of {steps} SSP-RK3 step(s) of the full {cells}-cell mesh after one "
        "warm-up residual; "
        + ("NumPy/SciPy oracle with C stand-ins for the numba kernels" if use_c else "NumPy/SciPy oracle")
        + ", pinned to 1 core",
        "seconds": elapsed,
        "sample_seconds": [round(x, 3) for x in times],
        "all_cores": cpu_all_cores(case, geometry, primitive, steps, samples),
    }


def cpu_all_cores(case, geometry, primitive, steps: int, samples: int = 5):
    """The multi-core CPU figure: the OpenMP restatement of the same explicit
    step (oracle/cpu_omp.c, the oracle's arithmetic and summation orders,
    equal to it bit for bit) on every core of the process's affinity; same
    protocol as the one-core figure. None when its library is not built."""
    from oracle import omp
    from oracle.euler import primitive_to_conservative

    if not omp.available():
        return None
    t0 = time.perf_counter()
    stepper = omp.OmpStepper(geometry, case.numerics)
    setup = time.perf_counter() - t0
    threads = len(os.sched_getaffinity(0))
    stepper.set_threads(threads)
    u0 = primitive_to_conservative(primitive, case.numerics.gamma)
    stepper.residual(u0)  # warm-up residual
    times = []
    for _ in range(samples):
        t0 = time.perf_counter()
        stepper.run(u0, steps)
        times.append(time.perf_counter() - t0)
    elapsed = float(np.median(times))
    cells = geometry.cell_count
    return {
        "value": cells * steps / elapsed / 1e6,
        "unit": "Mcell-updates/s",
        "cores": stepper.threads,
        "kind": "port",
        "sample": f"median of {samples} runs of {steps} SSP-RK3 step(s) of the full {cells}-cell mesh after one "
        "warm-up residual; OpenMP C restatement of the oracle's step (same bits) on all affinity cores",
        "seconds": elapsed,
        "sample_seconds": [round(x, 3) for x in times],
        "setup_seconds": round(setup, 2),
    }


def cylinder_bench(args) -> int:
    """Implicit-Euler steps of the laminar cylinder (Mach 0.2, Re 40; Roe +
    MUSCL parabolic + Sutherland, local steps at CFL 100, GMRES(30) rtol 1e-3
    + LU-SGS) from the free stream on an n_theta x n_r O-grid: prints one
    JSON line with implicit cell-updates per second (C x steps / t), the
    Newton and GMRES iteration counts and the residual drop."""
    import torch

    from meshcfd_amd import meshgen
    from meshcfd_amd.device import DeviceMesh
    from meshcfd_amd.steady_adjoint import residual_rms

    def note(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    case, prim = meshgen.cylinder_case(args.n_theta, args.n_r)
    note(f"mesh {case.mesh.cell_count} cells in {time.perf_counter() - t0:.1f}s; building device data ...")
    # the Newton residuals (MUSCL parabolic + viscous) run the four-kernel
    # path; the blocks serve the first-order viscous Jacobian products, one
    # fused launch each
    dm = DeviceMesh(case, "cuda:0", fused=True, block_cells=args.block_cells)
    dm.lusgs_ordering = args.lusgs_ordering
    setup_s = time.perf_counter() - t0
    note(f"setup {setup_s:.1f}s")
    gamma = case.numerics.gamma
    rho, u, v, p = prim.T
    dm.set_state(np.column_stack((rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u**2 + v**2))))
    # first step: warm-up (kernel loading, GMRES buffers)
    from meshcfd_amd.implicit import ImplicitStepper

    dm._implicit = ImplicitStepper(dm)
    dm._implicit.on_newton = lambda k, g, r: note(f"  Newton iteration {k}: {g} GMRES iterations, |G| {r:.3e}")
    t = dm.step_implicit(0.0)
    r0 = residual_rms(dm)
    note(f"warm-up step done, rms {r0:.3e}")
    st = dm._implicit
    n0, g0 = st.newton_iterations, st.gmres_iterations
    torch.cuda.synchronize()
    start = time.perf_counter()
    history = []
    if args.converge > 0:
        # steps until the case's residual tolerance or the time limit
        dm._implicit.on_newton = None
        steps = 0
        rms = r0
        from meshcfd_amd.device import native_settings

        nset = case.numerics
        cfl0 = nset.cfl
        while rms > case.numerics.residual_tolerance and time.perf_counter() - start < args.converge:
            t = dm.step_implicit(t)
            steps += 1
            rms = residual_rms(dm)
            if args.ser_max_cfl > 0.0:
                # switched evolution relaxation: CFL grows as the residual
                # falls (same steady state, not the reference's fixed CFL)
                nset.cfl = min(args.ser_max_cfl, max(cfl0, cfl0 * r0 / rms))
                dm.native_settings = native_settings(nset)
            if steps % 20 == 0 or steps <= 5 or rms <= case.numerics.residual_tolerance:
                history.append((steps, round(time.perf_counter() - start, 1), rms,
                                st.newton_iterations - n0, st.gmres_iterations - g0))
                note(f"step {steps}: rms {rms:.4e} cfl {nset.cfl:.3g} at {time.perf_counter() - start:.1f}s")
        nset.cfl = cfl0
        args.steps = max(steps, 1)
    else:
        for k in range(args.steps):
            t = dm.step_implicit(t)
            note(f"step {k + 1}/{args.steps}: {time.perf_counter() - start:.1f}s")
    torch.cuda.synchronize()
    secs = time.perf_counter() - start
    r1 = residual_rms(dm)
    cells = dm.n
    line = {
        "metric": "implicit Mcell-updates/s (laminar cylinder, Roe + MUSCL parabolic + Sutherland, FP64)",
        "value": round(cells * args.steps / secs / 1e6, 4),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "ms_per_step": round(secs / args.steps * 1e3, 2),
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (seeded O-grid, free-stream start)",
        "config": {"workload": f"laminar cylinder O-grid {args.n_theta}x{args.n_r}, r 0.5..30, first spacing 1e-3, "
                               "Mach 0.2, Re 40, implicit Euler local CFL 100, GMRES(30) rtol 1e-3, LU-SGS",
                   "cells": cells, "lusgs_ordering": args.lusgs_ordering},
        "newton_per_step": round((st.newton_iterations - n0) / args.steps, 2),
        "gmres_per_step": round((st.gmres_iterations - g0) / args.steps, 1),
        "residual_rms": [r0, r1],
        "setup_s": round(setup_s, 1),
    }
    if history:
        line["convergence"] = {"limit_s": args.converge, "tolerance": case.numerics.residual_tolerance,
                               "ser_max_cfl": args.ser_max_cfl,
                               "history": [{"step": a, "s": b, "rms": c, "newton": d, "gmres": e}
                                           for a, b, c, d, e in history]}
    print(json.dumps(line), flush=True)
    return 0


def batch_leg(cases: int, steps: int, warmup: int, graph_steps: int, block_cells: int = 128) -> dict:
    """A batch of ``cases`` cases advanced together (cases alternate an
    isentropic vortex and a moving shock tube; cell counts log-uniform in
    [2e3, 2e5] and Mach uniform in [0.1, 2.0] from seed 2; HLLC +
    MUSCL-linear BJ, SSP-RK3, each case its own CFL 0.5 step) on one
    concatenated device mesh: per-case steps by segmented minima, then the
    three fused stages over all cases. Times ``steps`` steps eagerly and as
    replays of a HIP graph of ``graph_steps`` steps; Mcell-updates/s =
    sum(C) x steps / t of the graph replays."""
    import torch

    from meshcfd_amd import batch as mb
    from meshcfd_amd.device import DeviceMesh

    def note(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    specs = mb.batch_specs(cases, seed=2)
    b = mb.build_batch(specs)
    note(f"{b.n_cases} cases, {b.geometry.cell_count} cells built in {time.perf_counter() - t0:.1f}s")
    dm = DeviceMesh.from_batch(b, "cuda:0", block_cells=block_cells)
    setup_s = time.perf_counter() - t0
    gamma = b.numerics.gamma
    prim = np.concatenate(b.primitives)
    rho, u, v, p = prim.T
    dm.set_state(np.column_stack((rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u**2 + v**2))))
    note(f"setup {setup_s:.1f}s")
    for _ in range(warmup):
        dm.step()
    dm.check_status("warm-up: ")
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(steps):
        dm.step()
    torch.cuda.synchronize()
    eager_s = time.perf_counter() - start
    dm.check_status("eager steps: ")
    note(f"eager: {eager_s / steps * 1e3:.3f} ms per step")
    g = max(1, graph_steps)
    dm.capture(g)
    replays = max(1, steps // g)
    dm.replay()
    torch.cuda.synchronize()
    start = time.perf_counter()
    for _ in range(replays):
        dm.replay()
    torch.cuda.synchronize()
    graph_s = time.perf_counter() - start
    dm.check_status("graph steps: ")
    eager_steps = steps
    steps = replays * g
    cells = dm.n
    clocks = dm.case_clock.view(-1, 2)[:, 0].cpu().numpy()
    dm.close()
    return {
        "metric": "batched Mcell-updates/s (HLLC+MUSCL RK3, FP64, per-case CFL)",
        "value": round(cells * steps / graph_s / 1e6, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(graph_s / steps * 1e3, 4),
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (seeded meshes and initial states)",
        "config": {"workload": f"{b.n_cases} cases (vortex / moving shock tube), cells log-uniform in [2e3, 2e5], "
                               "Mach uniform in [0.1, 2.0], seed 2; HLLC + MUSCL-linear (BJ), SSP-RK3, CFL 0.5 per case",
                   "cases": b.n_cases, "cells": cells, "graph_steps": g},
        "eager": {"ms_per_step": round(eager_s / eager_steps * 1e3, 4),
                  "value": round(cells * eager_steps / eager_s / 1e6, 2)},
        "case_time_range": [float(clocks.min()), float(clocks.max())],
        "setup_s": round(setup_s, 1),
    }


def batch_bench(args) -> int:
    """`batch_leg` alone, printed as one JSON line."""
    line = batch_leg(args.cases, args.steps, args.warmup, args.graph_steps, args.block_cells)
    print(json.dumps(line), flush=True)
    return 0


def adaptive_bench(args) -> int:
    """The adaptive bump channel through `solve_case`: Mach 0.675 inflow,
    pressure outlet, bump height 0.06 on a --bump-nx x --bump-ny base grid;
    HLLC + MUSCL linear, SSP-RK3 with local CFL steps (0.8), the density
    indicator refining up to 3 levels (threshold 0.02) every
    --adapt-interval steps, --steps steps from the free stream. Prints one
    JSON line splitting the wall time into the host's adaptations (indicator,
    marks, new mesh, transfer and the device rebuild) and the rest (device
    steps and their progress reads)."""
    import torch

    from meshcfd_amd import meshgen, solver

    case, prim = meshgen.bump_case(args.bump_nx, args.bump_ny, 0.06, integrator="Explicit RK3", cfl=0.8)
    s = case.numerics
    s.adaptive_mesh = True
    s.adaptation_interval = args.adapt_interval
    s.adaptation_max_level = 3
    s.adaptation_refine_threshold = 0.02
    s.adaptation_variable = "Density"
    s.max_iterations = args.steps
    s.residual_tolerance = 1e-9
    s.convergence_history_interval = args.steps + 1
    base = case.mesh.cell_count
    log = []  # (seconds, cells after) per adaptation
    inner = solver._adapt

    def timed(*a, **k):
        t0 = time.perf_counter()
        out = inner(*a, **k)
        torch.cuda.synchronize()
        log.append((time.perf_counter() - t0, case.mesh.cell_count))
        print(f"[bench] adaptation {len(log)}: {case.mesh.cell_count} cells, {log[-1][0]:.2f}s", file=sys.stderr,
              flush=True)
        return out

    solver._adapt = timed
    try:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = solver.solve_case(case, prim, device="cuda:0")
        torch.cuda.synchronize()
        total = time.perf_counter() - t0
    finally:
        solver._adapt = inner
    # cell-updates: each interval runs on the mesh of the adaptation before it
    counts = [base] + [c for _, c in log]
    updates = 0
    for k in range(res.steps):
        updates += counts[min(k // args.adapt_interval, len(counts) - 1)]
    adapt_s = sum(t for t, _ in log)
    line = {
        "metric": "adaptive bump channel: device Mcell-updates/s and host adaptation seconds",
        "value": round(updates / (total - adapt_s) / 1e6, 2),
        "unit": "Mcell-updates/s (wall time outside the adaptations)",
        "n_gpus": 1,
        "steps": res.steps,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (bump channel, free-stream start)",
        "config": {"workload": f"bump channel {args.bump_nx}x{args.bump_ny} base, h 0.06, Mach 0.675, HLLC + MUSCL "
                   f"linear SSP-RK3 local CFL 0.8, density indicator (0.02) up to 3 levels every "
                   f"{args.adapt_interval} steps"},
        "seconds": round(total, 2),
        "adaptation_seconds": round(adapt_s, 2),
        "adaptations": [{"seconds": round(t, 2), "cells": c} for t, c in log],
        "cells": {"base": base, "final": case.mesh.cell_count},
        "residual": res.residual,
    }
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1000, help="quads per side")
    ap.add_argument("--repeats", type=int, default=5, help="timed windows of --steps steps; value from the median")
    ap.add_argument("--cpu-steps", type=int, default=4,
                    help="SSP-RK3 steps per CPU-oracle sample, 5 samples (0 = skip)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--unfused", action="store_true", help="four-kernel path instead of the fused stage kernel")
    ap.add_argument("--block-cells", type=int, default=128, help="cells per workgroup of the fused stage kernel")
    ap.add_argument("--no-adjoint", action="store_true", help="skip timing the adjoint residual (VJP)")
    ap.add_argument("--workload", choices=("riemann", "cylinder", "batch", "adaptive"), default="riemann",
                    help="riemann: the explicit headline benchmark; cylinder: implicit steps of the laminar "
                    "O-grid cylinder (Roe + MUSCL parabolic + Sutherland, LU-SGS + GMRES); batch: 256 cases of "
                    "uneven size advanced together; adaptive: the bump channel adapting through solve_case")
    ap.add_argument("--bump-nx", type=int, default=512, help="adaptive: base cells along the channel")
    ap.add_argument("--bump-ny", type=int, default=128, help="adaptive: base cells across the channel")
    ap.add_argument("--adapt-interval", type=int, default=200, help="adaptive: steps between adaptations")
    ap.add_argument("--cases", type=int, default=256, help="batch: number of cases")
    ap.add_argument("--batch-cases", type=int, default=256,
                    help="riemann: also time a batch of this many cases as the JSON line's 'batch' key (0 = skip)")
    ap.add_argument("--graph-steps", type=int, default=10, help="batch: steps per captured HIP graph")
    ap.add_argument("--n-theta", type=int, default=2048, help="cylinder: cells around")
    ap.add_argument("--n-r", type=int, default=2048, help="cylinder: cells outward")
    ap.add_argument("--converge", type=float, default=0.0, metavar="SECONDS",
                    help="cylinder: step until the case's residual tolerance or this many seconds (0: --steps steps)")
    ap.add_argument("--ser-max-cfl", type=float, default=0.0, metavar="CFL",
                    help="cylinder --converge: raise the CFL number as CFL0 * r0 / r after each step, up to CFL "
                    "(0: the case's fixed CFL)")
    ap.add_argument("--lusgs-ordering", choices=("natural", "multicolour"), default="natural",
                    help="cylinder: LU-SGS ordering (natural: the reference's; multicolour: few parallel levels)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --block-cells as given instead of timing candidate block sizes first")
    ap.add_argument("--separate-time-step", action="store_true",
                    help="compute the CFL steps with their own kernels instead of inside the final fused stage")
    ap.add_argument(
        "--copy-probe",
        type=int,
        default=0,
        metavar="BYTES",
        help="also run the plain copy kernel k_copy over BYTES (HBM counter calibration)",
    )
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(
            json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is design-only", "n_gpus": args.gpus}),
            flush=True,
        )
        return 2
    if args.workload == "cylinder":
        return cylinder_bench(args)
    if args.workload == "batch":
        return batch_bench(args)
    if args.workload == "adaptive":
        return adaptive_bench(args)

    import torch

    from meshcfd_amd import meshgen
    from meshcfd_amd.device import DeviceMesh
    from meshcfd_amd.geometry import build_geometry
    from meshcfd_amd.layout import build_layout

    if args.copy_probe:
        import ctypes

        from meshcfd_amd import native

        lib = native.load()
        count = args.copy_probe // 8
        src = torch.ones(count, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for _ in range(3):
            native.check(lib.mcfd_copy(ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), count, s))
        torch.cuda.synchronize()
        del src, dst

    t0 = time.perf_counter()
    case, primitive = meshgen.riemann_workload(args.n, args.steps)
    geometry = build_geometry(case)
    layout = build_layout(geometry)
    setup_s = time.perf_counter() - t0
    print(f"[bench] setup {setup_s:.1f}s: {layout.n_cells} cells, {layout.n_faces} flux faces", file=sys.stderr, flush=True)

    dm = DeviceMesh(
        case, "cuda:0", geometry=geometry, layout=layout, fused=not args.unfused, block_cells=args.block_cells,
        fuse_time_step=not args.separate_time_step,
    )
    gamma = case.numerics.gamma
    rho, u, v, p = primitive.T
    cons = np.column_stack((rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u**2 + v**2)))
    dm.set_state(cons)
    dm.clock.zero_()
    tuned = None
    if not args.unfused and not args.no_autotune:
        # smaller blocks give small meshes more workgroups
        block = dm.autotune(candidates=tuple(sorted({32, 48, 64, 80, 96, 104, 112, 120, args.block_cells})))
        tuned = {"chosen": block, "ms_per_step": {str(k): round(v, 5) for k, v in dm.timings.items()}}
        args.block_cells = block
        print(f"[bench] autotuned block size {block}: {tuned['ms_per_step']}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        dm.step()
    dm.check_status("warm-up: ")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    windows = []
    for _ in range(max(1, args.repeats)):
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(args.steps):
            dm.step()
        stop.record(stream)
        torch.cuda.synchronize()
        windows.append(start.elapsed_time(stop))
        dm.check_status("timed run: ")
    ms = float(np.median(windows))
    cells = layout.n_cells
    value = cells * args.steps / (ms / 1e3) / 1e6
    print(f"[bench] {len(windows)} windows of {args.steps} steps: {[round(w, 3) for w in windows]} ms",
          file=sys.stderr, flush=True)

    b_alg, kernel_bytes = algorithmic_bytes(layout, case.numerics.reconstruction, case.numerics.time_integration)
    stage_gbs = b_alg * cells * args.steps / (ms / 1e3) / 1e9

    roofline = None
    kernel_times = {}
    if not args.no_kernel_timing:
        dm.timing(True)
        ksteps = max(1, min(args.steps, 50))
        for _ in range(ksteps):
            dm.step()
        torch.cuda.synchronize()
        kernel_times = dm.timing(False)
        dominant = max((k for k in kernel_times if k in kernel_bytes), key=lambda k: kernel_times[k][0])
        total, count = kernel_times[dominant]
        avg_ms = total / count
        achieved = kernel_bytes[dominant] / (avg_ms / 1e3) / 1e9
        traffic = None
        tfile = ROOT / "profiles" / "traffic.json"
        if tfile.exists():
            try:
                tdata = json.loads(tfile.read_text())
                # HIP kernel names: the fused stage kernel is k_stage_fused
                traffic = tdata.get("kernels", {}).get({"stage": "stage_fused"}.get(dominant, dominant))
            except Exception:
                traffic = None
        roofline = {
            "bound": "hbm",
            "kernel": "k_stage_fused" if dominant == "stage" else f"k_{dominant}",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "alg_bytes_per_launch": kernel_bytes[dominant],
            "avg_launch_ms": round(avg_ms, 5),
        }

    adj = None
    if not args.no_adjoint:
        adj = adjoint_rate(dm, torch)
        print(f"[bench] adjoint {adj}", file=sys.stderr, flush=True)

    dm.close()
    batch = None
    if args.batch_cases > 0:
        print("[bench] batch leg ...", file=sys.stderr, flush=True)
        batch = batch_leg(args.batch_cases, 100, 10, 10)

    cpu = None
    if args.cpu_steps > 0:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(case, geometry, primitive, args.cpu_steps)

    line = {
        "metric": "Mcell-updates/s (HLLC+MUSCL RK3, FP64)",
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "repeats": len(windows),
        "window_ms": [round(w, 4) for w in windows],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mesh and initial state)",
        "config": {
            "workload": f"four-quadrant Riemann, {args.n}x{args.n} quads + 30% triangle split, {cells} cells, "
            "HLLC + MUSCL-linear (Barth-Jespersen), SSP-RK3, CFL 0.8, transmissive BCs",
            "cells": cells,
            "flux_faces": layout.n_faces,
            "parallelism": "single GPU",
            "path": "four-kernel" if args.unfused else (
                f"fused stage kernel, blocks of <= {args.block_cells} cells and <= 256 flux faces"
                + ("" if args.separate_time_step else ", next step's CFL steps inside the final stage")
            ),
            "autotune": tuned,
        },
        "roofline": roofline,
        "stage": {
            "b_alg_bytes_per_cell_update": round(b_alg, 2),
            "achieved_GBs": round(stage_gbs, 1),
            "frac": round(stage_gbs / HBM_PEAK_GBS, 4),
        },
        "kernel_ms": {k: round(t / c, 5) for k, (t, c) in kernel_times.items()},
        "cpu_baseline": cpu,
        "adjoint": adj,
        "batch": batch,
        "host": host_info(),
        "setup_s": round(setup_s, 2),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
