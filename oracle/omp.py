"""ORACLE / CPU BASELINE (test infrastructure only): ctypes driver of
oracle/cpu_omp.c, the OpenMP restatement of the oracle's explicit step for
the four-quadrant Riemann workload's settings (MUSCL linear, HLLC, interior
and transmissive faces, "Global adaptive" CFL, "Explicit RK3").

`OmpStepper` builds the per-cell lists in the orders the oracle sums in
(oracle/solver.py): gradient entries and limiter offsets in entry order, the
divergence matrix's CSR order, transmissive faces and time-step faces in
np.add.at's order, so its steps equal `Oracle.step`'s bit for bit
(tests/test_oracle_c.py). bench.py times it on all host cores as the
multi-core CPU figure.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .solver import LIMITER_CODES, divergence_operator

LIB = Path(__file__).resolve().parent / "_lib" / "libmcfd_omp.so"

_P = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int64)


class _Case(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("nf", C.c_int64), ("gamma", C.c_double), ("cfl", C.c_double), ("limiter", C.c_int),
        ("area", _P),
        ("gptr", _I), ("gnb", _I), ("gcx", _P), ("gcy", _P),
        ("lptr", _I), ("ldx", _P), ("ldy", _P),
        ("fo", _I), ("fn", _I), ("foo", _P), ("fno", _P), ("fnrm", _P),
        ("dptr", _I), ("dcol", _I), ("dval", _P),
        ("tptr", _I), ("toff", _P), ("tnrm", _P), ("tlen", _P),
        ("sptr", _I), ("snrm", _P), ("slen", _P),
        ("prim", _P), ("grads", _P), ("factors", _P), ("flux", _P), ("d0", _P), ("d1", _P), ("stage", _P),
    ]


def available() -> bool:
    return LIB.exists()


def _csr(cells, n):
    """Stable order of list items by cell, and the CSR pointer."""
    order = np.argsort(cells, kind="stable")
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(cells, minlength=n), out=ptr[1:])
    return order, ptr


class OmpStepper:
    """SSP-RK3 steps of the oracle's explicit scheme on all OpenMP threads."""

    def __init__(self, geometry, settings):
        s = settings
        if (s.reconstruction != "MUSCL linear" or s.flux != "HLLC" or s.time_integration != "Explicit RK3"
                or s.time_step_mode != "Global adaptive" or s.viscosity_model != "Inviscid"
                or s.acceleration_x != 0.0 or s.acceleration_y != 0.0):
            raise ValueError("the OpenMP baseline covers MUSCL linear + HLLC + Explicit RK3, Global adaptive, inviscid")
        g = geometry
        other = [k for k in ("wall_faces", "velocity_inlet_faces", "pressure_outlet_faces", "far_field_faces",
                             "pressure_inlet_faces", "mass_flow_inlet_faces", "mass_flow_outlet_faces",
                             "periodic_faces_a") if len(getattr(g, k, ()))]
        if other:
            raise ValueError(f"the OpenMP baseline covers interior and transmissive faces only ({other})")
        self.lib = C.CDLL(str(LIB))
        self.lib.omp_rk3_steps.restype = C.c_int
        self.lib.omp_residual.restype = C.c_int
        self.lib.omp_threads.restype = C.c_int
        n = len(g.areas)
        self.n = n
        self.end_time = s.end_time
        keep = {}

        def arr(name, a, dtype):
            keep[name] = np.ascontiguousarray(a, dtype=dtype)
            return keep[name].ctypes.data_as(_I if dtype == np.int64 else _P)

        c = _Case()
        c.n, c.nf, c.gamma, c.cfl, c.limiter = n, len(g.interior_faces), s.gamma, s.cfl, LIMITER_CODES[s.limiter]
        c.area = arr("area", g.areas, np.float64)
        o, ptr = _csr(g.gradient_sources, n)
        c.gptr, c.gnb = arr("gptr", ptr, np.int64), arr("gnb", g.gradient_neighbours[o], np.int64)
        c.gcx = arr("gcx", g.gradient_coefficients_x[o], np.float64)
        c.gcy = arr("gcy", g.gradient_coefficients_y[o], np.float64)
        o, ptr = _csr(g.limiter_cells, n)
        c.lptr = arr("lptr", ptr, np.int64)
        c.ldx, c.ldy = arr("ldx", g.limiter_dx[o], np.float64), arr("ldy", g.limiter_dy[o], np.float64)
        c.fo, c.fn = arr("fo", g.interior_owners, np.int64), arr("fn", g.interior_neighbours, np.int64)
        c.foo = arr("foo", g.interior_owner_offsets, np.float64)
        c.fno = arr("fno", g.interior_neighbour_offsets, np.float64)
        c.fnrm = arr("fnrm", g.interior_normals, np.float64)
        D = divergence_operator(g)
        c.dptr = arr("dptr", D.indptr, np.int64)
        c.dcol, c.dval = arr("dcol", D.indices, np.int64), arr("dval", D.data, np.float64)
        tf = np.asarray(getattr(g, "transmissive_faces", np.zeros(0, dtype=np.int64)), dtype=np.int64)
        to = g.owners[tf]
        o, ptr = _csr(to, n)
        tf = tf[o]
        c.tptr = arr("tptr", ptr, np.int64)
        c.toff = arr("toff", g.face_centers[tf] - g.centers[g.owners[tf]], np.float64)
        c.tnrm, c.tlen = arr("tnrm", g.face_normals[tf], np.float64), arr("tlen", g.face_lengths[tf], np.float64)
        # time-step sums: every face at its owner (face order), then interior
        # faces at their neighbour (interior order)
        inner = g.interior_faces
        cells = np.concatenate((g.owners, g.neighbours[inner]))
        faces = np.concatenate((np.arange(len(g.owners)), inner))
        o, ptr = _csr(cells, n)
        c.sptr = arr("sptr", ptr, np.int64)
        c.snrm, c.slen = arr("snrm", g.face_normals[faces[o]], np.float64), arr("slen", g.face_lengths[faces[o]], np.float64)
        for name, size in (("prim", 4 * n), ("grads", 8 * n), ("factors", 4 * n), ("flux", 4 * max(len(inner), 1)),
                           ("d0", 4 * n), ("d1", 4 * n), ("stage", 4 * n)):
            setattr(c, name, arr(name, np.zeros(size), np.float64))
        self._keep = keep
        self._case = c

    def set_threads(self, n: int) -> None:
        self.lib.omp_set_threads(C.c_int(int(n)))

    @property
    def threads(self) -> int:
        return int(self.lib.omp_threads())

    def residual(self, conservative):
        u = np.ascontiguousarray(conservative, dtype=np.float64)
        out = np.empty_like(u)
        rc = self.lib.omp_residual(C.byref(self._case), u.ctypes.data_as(_P), out.ctypes.data_as(_P))
        if rc:
            raise ValueError("Euler state has non-finite or non-positive density or pressure.")
        return out

    def run(self, conservative, steps: int, time: float = 0.0):
        """(state after ``steps`` SSP-RK3 steps, time)."""
        u = np.array(conservative, dtype=np.float64, order="C", copy=True)
        t = C.c_double(time)
        rc = self.lib.omp_rk3_steps(C.byref(self._case), u.ctypes.data_as(_P), C.c_int(steps),
                                    C.c_double(self.end_time), C.byref(t))
        if rc == 1:
            raise ValueError("Euler state has non-finite or non-positive density or pressure.")
        if rc == 2:
            raise ValueError("The stable time step is non-positive or non-finite.")
        return u, t.value
