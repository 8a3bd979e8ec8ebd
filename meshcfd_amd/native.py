"""ctypes binding of libmcfdhip.so (the C-ABI declared in include/mcfd.h).

The library is built in-tree (meshcfd_amd/csrc/libmcfdhip.so); loading fails
loudly when it is missing — there is no CPU fallback on the product path.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "csrc" / "libmcfdhip.so"

ERROR_MESSAGES = {
    1: "Euler state has non-finite or non-positive density or pressure.",
    2: "HLLC flux produced a non-finite value.",
    3: "Roe average has a non-positive sound speed.",
    4: "Roe flux produced a non-finite value.",
    5: "Far-field characteristics produced invalid sound speed.",
    6: "Far-field characteristics produced an invalid state.",
    7: "Spatial residual contains non-finite values.",
    8: "Local stable time steps must be positive and finite.",
    9: "The stable time step is non-positive or non-finite.",
    10: "Viscous flow requires finite positive temperature.",
    11: "Sutherland law produced an invalid viscosity.",
}

# codes whose status index is a device cell (state, residual, local step, temperature)
CELL_ERROR_MASK = (1 << 1) | (1 << 7) | (1 << 8) | (1 << 10)

OP_EULER, OP_SSP, OP_RK4, OP_RK4_FINAL, OP_RESIDUAL = 0, 1, 2, 3, 4
KERNEL_NAMES = ("primitive", "gradient", "flux", "update", "time_step", "reduce", "stage", "viscous")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class McfdMeshDesc(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int32),
        ("n_faces", C.c_int32),
        ("n_grad", C.c_int32),
        ("n_lim", C.c_int32),
        ("n_acc", C.c_int32),
        ("n_dts", C.c_int32),
        ("n_bc", C.c_int32),
        ("reserved", C.c_int32),
        ("area", _dp),
        ("grad_ptr", _ip),
        ("grad_nb", _ip),
        ("grad_cx", _dp),
        ("grad_cy", _dp),
        ("lim_ptr", _ip),
        ("lim_dx", _dp),
        ("lim_dy", _dp),
        ("face_left", _ip),
        ("face_right", _ip),
        ("face_kind", _ip),
        ("face_bc", _ip),
        ("face_nx", _dp),
        ("face_ny", _dp),
        ("face_len", _dp),
        ("face_lx", _dp),
        ("face_ly", _dp),
        ("face_rx", _dp),
        ("face_ry", _dp),
        ("acc_ptr", _ip),
        ("acc_item", _ip),
        ("dts_ptr", _ip),
        ("dts_nx", _dp),
        ("dts_ny", _dp),
        ("dts_len", _dp),
        ("dts_bc", _ip),
        ("bc_kind", _ip),
        ("bc_state", _dp),
        ("bc_wall", _dp),
        ("acc_coef", _dp),
    ]


_hp = C.POINTER(C.c_uint16)


class McfdBlockDesc(C.Structure):
    _fields_ = [
        ("n_blocks", C.c_int32),
        ("block_cells", C.c_int32),
        ("max_local", C.c_int32),
        ("max_grad", C.c_int32),
        ("max_faces", C.c_int32),
        ("n_local", C.c_int32),
        ("n_gptr", C.c_int32),
        ("n_gnb", C.c_int32),
        ("n_lent", C.c_int32),
        ("n_block_faces", C.c_int32),
        ("info", _ip),
        ("cells", _ip),
        ("gptr", _ip),
        ("gnb", _hp),
        ("gent", _dp),
        ("lptr", _ip),
        ("lent", _dp),
        ("faces", _ip),
        ("frec", C.c_void_p),
        ("acc_local", _hp),
        ("dts_local", _hp),
    ]


def block_desc(bl) -> tuple[McfdBlockDesc, list]:
    keep = []

    def arr(a, dtype, ctype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return as_ptr(a, ctype)

    frec = np.ascontiguousarray(bl.frec)
    keep.append(frec)
    desc = McfdBlockDesc(
        n_blocks=bl.n_blocks,
        block_cells=bl.block_cells,
        max_local=bl.max_local,
        max_grad=bl.max_grad_cells,
        max_faces=bl.max_faces,
        n_local=len(bl.cells),
        n_gptr=len(bl.gnb_ptr),
        n_gnb=len(bl.gnb),
        n_lent=len(bl.lent),
        n_block_faces=len(bl.faces),
        info=arr(bl.info.reshape(-1), np.int32, C.c_int32),
        cells=arr(bl.cells, np.int32, C.c_int32),
        gptr=arr(bl.gnb_ptr, np.int32, C.c_int32),
        gnb=arr(bl.gnb, np.uint16, C.c_uint16),
        gent=arr(bl.gent.reshape(-1), np.float64, C.c_double),
        lptr=arr(bl.lptr, np.int32, C.c_int32),
        lent=arr(bl.lent.reshape(-1), np.float64, C.c_double),
        faces=arr(bl.faces, np.int32, C.c_int32),
        frec=frec.ctypes.data,
        acc_local=arr(bl.acc_local, np.uint16, C.c_uint16),
        dts_local=arr(bl.dts_local, np.uint16, C.c_uint16),
    )
    return desc, keep


class McfdSettings(C.Structure):
    _fields_ = [
        ("flux", C.c_int32),
        ("reconstruction", C.c_int32),
        ("limiter", C.c_int32),
        ("viscosity", C.c_int32),
        ("gamma", C.c_double),
        ("cfl", C.c_double),
        ("acceleration_x", C.c_double),
        ("acceleration_y", C.c_double),
        ("gas_constant", C.c_double),
        ("prandtl_number", C.c_double),
        ("dynamic_viscosity", C.c_double),
        ("sutherland_reference_temperature", C.c_double),
        ("sutherland_reference_viscosity", C.c_double),
        ("sutherland_temperature", C.c_double),
    ]


class McfdStage(C.Structure):
    _fields_ = [
        ("op", C.c_int32),
        ("dt_per_cell", C.c_int32),
        ("u", C.c_void_p),
        ("u0", C.c_void_p),
        ("u1", C.c_void_p),
        ("r1", C.c_void_p),
        ("dt", C.c_void_p),
        ("out", C.c_void_p),
        ("r_out", C.c_void_p),
        ("c", C.c_double * 5),
        ("dt_next", C.c_void_p),
        ("dt_min", C.c_void_p),
    ]


EXPORTS = {
    "mcfd_version": (C.c_int, []),
    "mcfd_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "mcfd_mesh_create": (C.c_int, [C.POINTER(McfdMeshDesc), C.POINTER(C.c_void_p)]),
    "mcfd_mesh_destroy": (C.c_int, [C.c_void_p]),
    "mcfd_mesh_set_blocks": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mcfd_mesh_set_block_viscous": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]),
    "mcfd_residual": (C.c_int, [C.c_void_p, C.POINTER(McfdSettings), C.c_void_p, C.c_void_p, C.c_void_p]),
    "mcfd_time_steps": (
        C.c_int,
        [C.c_void_p, C.POINTER(McfdSettings), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    ),
    "mcfd_time_steps_primitive": (
        C.c_int,
        [C.c_void_p, C.POINTER(McfdSettings), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    ),
    "mcfd_check_state": (C.c_int, [C.c_void_p, C.POINTER(McfdSettings), C.c_void_p, C.c_void_p]),
    "mcfd_advance_clock": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_int32, C.c_void_p],
    ),
    "mcfd_stage": (C.c_int, [C.c_void_p, C.POINTER(McfdSettings), C.POINTER(McfdStage), C.c_void_p]),
    "mcfd_sum_squares": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mcfd_mesh_set_parabolic": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mcfd_batch_time_steps": (
        C.c_int,
        [C.c_void_p, C.POINTER(McfdSettings), C.c_void_p, C.c_int32, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p,
         C.c_void_p],
    ),
    "mcfd_batch_case_steps": (
        C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p],
    ),
    "mcfd_sum_squares_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "mcfd_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "mcfd_status": (C.c_int, [C.c_void_p, C.c_void_p, _ip, _ip]),
    "mcfd_timing": (C.c_int, [C.c_void_p, C.c_int32, _dp, _ip]),
}

_lib = None


class NativeError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load libmcfdhip.so, raising if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("MCFD_LIB", LIB_PATH))
    if not path.exists():
        raise NativeError(
            f"{path} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')."
        )
    lib = C.CDLL(str(path))
    for name, (restype, argtypes) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    load().mcfd_last_error(buf, len(buf))
    return buf.value.decode()


def check(rc: int) -> None:
    if rc != 0:
        raise NativeError(last_error())


def as_ptr(array: np.ndarray, ctype):
    return array.ctypes.data_as(C.POINTER(ctype))


def mesh_desc(layout) -> tuple[McfdMeshDesc, list]:
    """Fill a McfdMeshDesc from a DeviceLayout; returns (desc, keepalive)."""
    keep = []

    def d(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return as_ptr(a, C.c_double)

    def i(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return as_ptr(a, C.c_int32)

    L = layout
    desc = McfdMeshDesc(
        n_cells=L.n_cells,
        n_faces=L.n_faces,
        n_grad=len(L.grad_nb),
        n_lim=len(L.lim_dx),
        n_acc=len(L.acc_item),
        n_dts=len(L.dts_nx),
        n_bc=len(L.bc_kind),
        reserved=0,
        area=d(L.area),
        grad_ptr=i(L.grad_ptr),
        grad_nb=i(L.grad_nb),
        grad_cx=d(L.grad_cx),
        grad_cy=d(L.grad_cy),
        lim_ptr=i(L.lim_ptr),
        lim_dx=d(L.lim_dx),
        lim_dy=d(L.lim_dy),
        face_left=i(L.face_left),
        face_right=i(L.face_right),
        face_kind=i(L.face_kind),
        face_bc=i(L.face_bc),
        face_nx=d(L.face_nx),
        face_ny=d(L.face_ny),
        face_len=d(L.face_len),
        face_lx=d(L.face_lx),
        face_ly=d(L.face_ly),
        face_rx=d(L.face_rx),
        face_ry=d(L.face_ry),
        acc_ptr=i(L.acc_ptr),
        acc_item=i(L.acc_item),
        dts_ptr=i(L.dts_ptr),
        dts_nx=d(L.dts_nx),
        dts_ny=d(L.dts_ny),
        dts_len=d(L.dts_len),
        dts_bc=i(L.dts_bc),
        bc_kind=i(L.bc_kind),
        bc_state=d(L.bc_state),
        bc_wall=d(L.bc_wall),
        acc_coef=d(L.acc_coef),
    )
    return desc, keep
