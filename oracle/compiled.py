"""ORACLE (test infrastructure only): ctypes access to oracle/_lib's C
stand-ins of the reference's numba kernels (oracle/mcfd_oracle.c).

`enable()` makes `oracle.kernels` use them; results are identical to the
NumPy restatement (tests/test_oracle_c.py), only faster — the CPU baseline
times the reference-equivalent path (NumPy/SciPy + compiled kernels, one core).
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import kernels

LIB = Path(__file__).resolve().parent / "_lib" / "libmcfd_oracle.so"
_lib = None
_numpy_versions = {}


def available() -> bool:
    return LIB.exists()


def _load():
    global _lib
    if _lib is None:
        _lib = C.CDLL(str(LIB))
        _lib.oracle_linear_gradients_and_factors.restype = None
        _lib.oracle_reconstruct_hllc_flux.restype = None
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def linear_gradients_and_factors(prim, src, nbr, cx, cy, cells, dx, dy, code):
    lib = _load()
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    n = len(prim)
    grads = np.empty((n, 4, 2))
    factors = np.empty((n, 4))
    mins = np.empty((n, 4))
    maxs = np.empty((n, 4))
    src = np.ascontiguousarray(src, dtype=np.int64)
    nbr = np.ascontiguousarray(nbr, dtype=np.int64)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    cx, cy, dx, dy = (np.ascontiguousarray(a, dtype=np.float64) for a in (cx, cy, dx, dy))
    lib.oracle_linear_gradients_and_factors(
        C.c_int64(n), _p(prim), C.c_int64(len(src)), _p(src), _p(nbr), _p(cx), _p(cy),
        C.c_int64(len(cells)), _p(cells), _p(dx), _p(dy), C.c_int(int(code)), _p(grads), _p(factors),
        _p(mins), _p(maxs),
    )
    return grads, factors


def reconstruct_hllc_flux(prim, grads, factors, owners, neighbours, oo, no, normals, gamma):
    lib = _load()
    prim = np.ascontiguousarray(prim, dtype=np.float64)
    grads = np.ascontiguousarray(grads, dtype=np.float64)
    factors = np.ascontiguousarray(factors, dtype=np.float64)
    owners = np.ascontiguousarray(owners, dtype=np.int64)
    neighbours = np.ascontiguousarray(neighbours, dtype=np.int64)
    oo, no, normals = (np.ascontiguousarray(a, dtype=np.float64) for a in (oo, no, normals))
    flux = np.empty((len(owners), 4))
    lib.oracle_reconstruct_hllc_flux(
        C.c_int64(len(owners)), _p(prim), _p(grads), _p(factors), _p(owners), _p(neighbours), _p(oo), _p(no),
        _p(normals), C.c_double(gamma), _p(flux),
    )
    return flux


def enable() -> None:
    """Route oracle.kernels' two compiled-kernel restatements to C."""
    if not _numpy_versions:
        _numpy_versions["lgf"] = kernels.linear_gradients_and_factors
        _numpy_versions["rhf"] = kernels.reconstruct_hllc_flux
    kernels.linear_gradients_and_factors = linear_gradients_and_factors
    kernels.reconstruct_hllc_flux = reconstruct_hllc_flux


def disable() -> None:
    if _numpy_versions:
        kernels.linear_gradients_and_factors = _numpy_versions["lgf"]
        kernels.reconstruct_hllc_flux = _numpy_versions["rhf"]
