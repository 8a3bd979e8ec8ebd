"""Writes riemann_1p3M_200.json: the oracle's state after 200 SSP-RK3 steps
of the four-quadrant Riemann workload at full size (1000 x 1000 quads, 30%
split into triangles: 1.3M cells; HLLC + MUSCL-linear Barth-Jespersen, CFL
0.8, transmissive boundaries, FP64).

The fixture holds the SHA-256 of the final conservative state (mesh order,
float64 little-endian), per-component maxima of |U| and sums, the final time,
and the exact states of every 4001st cell (float.hex), so a GPU run is checked
bit for bit and, should that fail, by the relative-difference metric.

Run from the repository root: python tests/golden/make_riemann_golden.py
(about three minutes on one core; the oracle uses its C stand-ins when
oracle/_lib is built).
"""

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from meshcfd_amd import meshgen  # noqa: E402
from meshcfd_amd.geometry import build_geometry  # noqa: E402
from oracle import compiled  # noqa: E402
from oracle.euler import primitive_to_conservative  # noqa: E402
from oracle.solver import Oracle  # noqa: E402

STEPS = 200
STRIDE = 4001


def main():
    case, prim = meshgen.riemann_workload(1000, STEPS)
    g = build_geometry(case)
    if compiled.available():
        compiled.enable()
    o = Oracle(g, case.numerics)
    u = primitive_to_conservative(prim, case.numerics.gamma)
    t = 0.0
    t0 = time.perf_counter()
    for k in range(STEPS):
        u, dt, _, _ = o.step(u, t)
        t += dt
        if (k + 1) % 20 == 0:
            print(f"step {k + 1}: t = {t:.6g} ({time.perf_counter() - t0:.0f} s)", flush=True)
    u = np.ascontiguousarray(u, dtype="<f8")
    idx = np.arange(0, len(u), STRIDE)
    out = {
        "workload": "four-quadrant Riemann 1000x1000 + 30% split, HLLC MUSCL-linear BJ, SSP-RK3, CFL 0.8",
        "cells": int(len(u)),
        "steps": STEPS,
        "time": float(t).hex(),
        "sha256": hashlib.sha256(u.tobytes()).hexdigest(),
        "max_abs": [float(x).hex() for x in np.max(np.abs(u), axis=0)],
        "sum": [float(x).hex() for x in np.sum(u, axis=0)],
        "sample_stride": STRIDE,
        "samples": [[float(x).hex() for x in row] for row in u[idx]],
    }
    path = Path(__file__).resolve().parent / "riemann_1p3M_200.json"
    path.write_text(json.dumps(out, indent=0))
    print(f"wrote {path} ({len(idx)} sampled cells), sha256 {out['sha256']}")


if __name__ == "__main__":
    main()
