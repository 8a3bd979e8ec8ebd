"""Laminar viscous terms: device lists and their upload.

The reference evaluates the viscous residual in Python on every call
(`src/meshcfd/core/solver.py`): weighted least-squares gradients of
(u, v, T) with one ghost entry per non-periodic boundary face
(`_viscous_gradients`, `:3615-3714`, over the numba kernel
`weighted_least_squares_gradients`, `_kernels.py:163-207`), face-averaged
Stokes/Fourier fluxes with boundary gradient corrections
(`_viscous_flux`, `_viscous_residual`, `:3717-4114`) and a diffusive term in
the stable time step (`_stable_time_steps`, `:4246-4332`).

Everything that depends on geometry only is computed here once, with the
same numpy operations the reference applies per call, so the device sees
identical values: LSQ weights w*dx, w*dy of every entry, the per-cell 2x2
normal matrix {xx, xy, yy} and its inverse determinant (0 where the
reference's determinant test leaves the gradient at zero), ghost offsets and
weights, the normal distance dn and tangential offset of boundary faces, and
the clamped distances of the diffusive time step. Per-cell lists keep the
reference's summation order.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import native
from .geometry import BOUNDARY_CODES, Geometry, SolverError
from .layout import DeviceLayout

VIS_WALL, VIS_SYMMETRY, VIS_INLET, VIS_INFLOW, VIS_FAR_FIELD = 1, 2, 3, 4, 5
VIS_MODELS = {"Inviscid": 0, "Constant": 1, "Sutherland": 2}


class McfdViscousDesc(C.Structure):
    _fields_ = [
        ("n_ghost", C.c_int32),
        ("n_vacc", C.c_int32),
        ("n_vdt", C.c_int32),
        ("reserved", C.c_int32),
        ("wl_wx", C.c_void_p),
        ("wl_wy", C.c_void_p),
        ("wl_cell", C.c_void_p),
        ("gh_ptr", C.c_void_p),
        ("gh_face", C.c_void_p),
        ("gh_wx", C.c_void_p),
        ("gh_wy", C.c_void_p),
        ("bv_code", C.c_void_p),
        ("bv_flags", C.c_void_p),
        ("bv_geom", C.c_void_p),
        ("va_ptr", C.c_void_p),
        ("va_item", C.c_void_p),
        ("va_coef", C.c_void_p),
        ("vd_ptr", C.c_void_p),
        ("vd_nb", C.c_void_p),
        ("vd_len", C.c_void_p),
        ("vd_dist", C.c_void_p),
    ]


native.EXPORTS["mcfd_mesh_set_viscous"] = (C.c_int, [C.c_void_p, C.c_void_p])


def _csr(keys: np.ndarray, n: int):
    order = np.argsort(keys, kind="stable")
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(keys, minlength=n), out=ptr[1:])
    return ptr.astype(np.int32), order


def _weights(dx: np.ndarray, dy: np.ndarray):
    """w*dx, w*dy with w = 1/max(dx^2 + dy^2, 1e-300) (_kernels.py:177-181)."""
    w = 1.0 / np.maximum(dx * dx + dy * dy, 1e-300)
    return w * dx, w * dy


def build_viscous(g: Geometry, L: DeviceLayout) -> dict:
    """Device-order viscous lists of `g` for the layout `L`."""
    if g.diffusive_boundary_faces is None:
        raise SolverError("Geometry was built without viscous boundary data.")
    n = g.cell_count
    new_id = np.empty(n, dtype=np.int64)
    new_id[L.cell_order] = np.arange(n)
    flux = L.mesh_face_flux
    c, fc, normals = g.centers, g.face_centers, g.face_normals

    # --- boundary faces: ghost offsets and per-record geometry --------------
    dbf = np.asarray(g.diffusive_boundary_faces, dtype=np.int64)
    nb_ = len(dbf)
    owners_b = g.owners[dbf]
    dn = np.empty(nb_)
    for k, f in enumerate(dbf):
        dn[k] = float(np.dot(fc[f] - c[owners_b[k]], normals[f]))
    if nb_ and np.any(dn <= 1e-14):
        raise SolverError("Boundary face has invalid cell-center distance.")
    kinds = g.face_kind_code[dbf]
    mirror = (kinds == BOUNDARY_CODES["Wall"]) | (kinds == BOUNDARY_CODES["Symmetry"])
    offset = fc[dbf] - c[owners_b]
    goff = np.where(mirror[:, None], (2.0 * dn)[:, None] * normals[dbf], 2.0 * offset)
    tang = offset - dn[:, None] * normals[dbf]

    n_bc = len(L.bc_kind)
    face_bc = L.face_bc[flux[dbf]] if nb_ else np.zeros(0, dtype=np.int64)
    code_map = {BOUNDARY_CODES["Wall"]: VIS_WALL, BOUNDARY_CODES["Symmetry"]: VIS_SYMMETRY,
                BOUNDARY_CODES["Velocity Inlet"]: VIS_INLET, BOUNDARY_CODES["Far Field"]: VIS_FAR_FIELD,
                # exterior states of the inflow records, from the owner cells' states
                BOUNDARY_CODES["Pressure Inlet"]: VIS_INFLOW, BOUNDARY_CODES["Mass Flow Inlet"]: VIS_INFLOW}
    bv_code = np.zeros(n_bc, dtype=np.int32)
    bv_flags = np.zeros(n_bc, dtype=np.int32)
    bv_geom = np.zeros((n_bc, 4))
    if nb_:
        if np.any(face_bc < 0):
            raise SolverError("Boundary face without a boundary record.")
        bv_code[face_bc] = [code_map.get(int(k), 0) for k in kinds]
        bv_flags[face_bc] = g.wall_no_slip[dbf].astype(np.int32) | (g.wall_thermal[dbf].astype(np.int32) << 1)
        bv_geom[face_bc, 0] = dn
        bv_geom[face_bc, 1:3] = tang
        bv_geom[face_bc, 3] = g.wall_thermal_value[dbf]

    # --- WLSQ: face entries in LSQ order, then ghost entries ----------------
    fwx, fwy = _weights(g.gradient_dx, g.gradient_dy)
    gwx, gwy = _weights(goff[:, 0], goff[:, 1]) if nb_ else (np.zeros(0), np.zeros(0))
    src_all = np.concatenate((g.gradient_sources, owners_b)).astype(np.int64)
    dx_all = np.concatenate((g.gradient_dx, goff[:, 0]))
    dy_all = np.concatenate((g.gradient_dy, goff[:, 1]))
    wx_all = np.concatenate((fwx, gwx))
    wy_all = np.concatenate((fwy, gwy))
    xx = np.zeros(n)
    xy = np.zeros(n)
    yy = np.zeros(n)
    np.add.at(xx, src_all, wx_all * dx_all)
    np.add.at(xy, src_all, wx_all * dy_all)
    np.add.at(yy, src_all, wy_all * dy_all)
    det = xx * yy - xy * xy
    ok = det > 1e-14 * np.maximum(xx * yy, 1e-300)
    inv = np.zeros(n)
    inv[ok] = 1.0 / det[ok]
    wl_cell = np.empty((n, 4))
    wl_cell[new_id] = np.column_stack((xx, xy, yy, inv))
    _, order = _csr(new_id[g.gradient_sources], n)
    wl_wx = fwx[order]
    wl_wy = fwy[order]
    gh_ptr, gorder = _csr(new_id[owners_b], n)
    gh_face = flux[dbf][gorder].astype(np.int32)
    gh_wx = gwx[gorder]
    gh_wy = gwy[gorder]

    # --- viscous accumulation items -----------------------------------------
    inter = g.interior_faces
    io, ine = g.interior_owners, g.interior_neighbours
    n_int = len(inter)
    k_int = np.arange(n_int, dtype=np.int64)
    int_cells = np.concatenate((io, ine))
    int_k = np.concatenate((k_int, k_int))
    lint = g.face_lengths[inter]
    int_coef = np.concatenate((-lint / g.areas[io], lint / g.areas[ine]))
    o = np.lexsort((int_k, int_cells))
    int_cells, int_k, int_coef = int_cells[o], int_k[o], int_coef[o]
    int_items = flux[inter][int_k] * 4
    fa, fb = g.periodic_faces_a, g.periodic_faces_b
    oa, ob = g.owners[fa], g.owners[fb]
    pf = flux[fa]
    seq_cells = np.concatenate((oa, ob, owners_b)).astype(np.int64)
    seq_items = np.concatenate((pf * 4 + 1, pf * 4 + 2, flux[dbf] * 4 + 1))
    all_cells = np.concatenate((int_cells, seq_cells))
    all_items = np.concatenate((int_items, seq_items))
    all_coef = np.concatenate((int_coef, np.zeros(len(seq_cells))))
    va_ptr, order = _csr(new_id[all_cells], n)
    va_item = all_items[order]
    va_coef = all_coef[order]
    if va_item.size and va_item.max() >= 2**31:
        raise ValueError("Too many faces for 32-bit accumulation items.")

    # --- diffusive time-step items -------------------------------------------
    d_int = np.abs(np.sum((c[ine] - c[io]) * normals[inter], axis=1))
    d_per = np.abs(np.sum((fc[fa] - c[oa]) * normals[fa], axis=1)) + np.abs(
        np.sum((fc[fb] - c[ob]) * normals[fb], axis=1)
    )
    d_bnd = np.empty(nb_)
    for k, f in enumerate(dbf):
        d_bnd[k] = abs(float(np.dot(fc[f] - c[owners_b[k]], normals[f])))
    cells = np.concatenate((io, ine, oa, ob, owners_b)).astype(np.int64)
    partner = np.concatenate((ine, io, ob, oa, np.full(nb_, -1))).astype(np.int64)
    lens = np.concatenate((lint, lint, g.face_lengths[fa], g.face_lengths[fa], g.face_lengths[dbf]))
    dist = np.concatenate((d_int, d_int, d_per, d_per, d_bnd))
    vd_ptr, order = _csr(new_id[cells], n)
    part = partner[order]
    vd_nb = np.where(part >= 0, new_id[np.maximum(part, 0)], -1).astype(np.int32)

    return {
        "wl_wx": np.ascontiguousarray(wl_wx),
        "wl_wy": np.ascontiguousarray(wl_wy),
        "wl_cell": np.ascontiguousarray(wl_cell.reshape(-1)),
        "gh_ptr": gh_ptr,
        "gh_face": gh_face,
        "gh_wx": np.ascontiguousarray(gh_wx),
        "gh_wy": np.ascontiguousarray(gh_wy),
        "bv_code": bv_code,
        "bv_flags": bv_flags,
        "bv_geom": np.ascontiguousarray(bv_geom.reshape(-1)),
        "va_ptr": va_ptr,
        "va_item": va_item.astype(np.int32),
        "va_coef": np.ascontiguousarray(va_coef),
        "vd_ptr": vd_ptr,
        "vd_nb": vd_nb,
        "vd_len": np.ascontiguousarray(lens[order]),
        "vd_dist": np.ascontiguousarray(np.maximum(dist, 1e-14)[order]),
    }


_INT_FIELDS = ("gh_ptr", "gh_face", "bv_code", "bv_flags", "va_ptr", "va_item", "vd_ptr", "vd_nb")


def viscous_desc(v: dict) -> tuple[McfdViscousDesc, list]:
    keep = []
    kw = {}
    for name, _ in McfdViscousDesc._fields_[4:]:
        a = np.ascontiguousarray(v[name], dtype=np.int32 if name in _INT_FIELDS else np.float64)
        keep.append(a)
        kw[name] = a.ctypes.data
    desc = McfdViscousDesc(n_ghost=len(v["gh_face"]), n_vacc=len(v["va_item"]), n_vdt=len(v["vd_nb"]),
                           reserved=0, **kw)
    return desc, keep


def upload(handle, g: Geometry, L: DeviceLayout) -> dict:
    """Build the viscous lists of `g`, upload them to the mesh `handle` and
    return what the fused kernel's block packing needs (the WLSQ weights of
    the LSQ entries and the accumulation lists)."""
    lib = native.load()
    lists = build_viscous(g, L)
    desc, keep = viscous_desc(lists)
    native.check(lib.mcfd_mesh_set_viscous(handle, C.byref(desc)))
    del keep
    return {k: lists[k] for k in ("wl_wx", "wl_wy", "va_ptr", "va_item")}
