"""Device layout of a `Geometry`: flux faces and per-cell CSR lists.

Each per-cell list keeps the reference's summation order, so the device
gather reproduces the reference residual's sums term for term:

* LSQ entries: the order of `gradient_sources` (`solver.py:2004-2019`);
* accumulation: interior faces ascending (the CSR divergence operator's row
  order, `solver.py:2157-2174,2501`), then periodic a-sides, periodic
  b-sides, wall/symmetry, velocity inlet, pressure outlet, far field,
  pressure inlet, mass-flow inlet, mass-flow outlet and transmissive faces,
  in that order (`solver.py:2502-2705`);
* spectral radius: owners of all faces ascending, then neighbours of interior
  faces ascending (`solver.py:4226-4245`).

Flux faces are numbered interior faces first, then periodic pairs, then the
boundary groups in the order above.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .geometry import Geometry

FACE_INTERIOR, FACE_PERIODIC, FACE_WALL, FACE_VELOCITY_INLET = 0, 1, 2, 3
FACE_PRESSURE_OUTLET, FACE_FAR_FIELD, FACE_TRANSMISSIVE = 4, 5, 6
FACE_PRESSURE_INLET, FACE_MASS_FLOW_INLET, FACE_MASS_FLOW_OUTLET = 7, 8, 9
INFLOW_KINDS = (FACE_PRESSURE_INLET, FACE_MASS_FLOW_INLET, FACE_MASS_FLOW_OUTLET)
ACC_OWNER, ACC_NEIGHBOUR, ACC_BOUNDARY, ACC_PERIODIC_B = 0, 1, 2, 3


def hilbert_order(points: np.ndarray, bits: int = 16) -> np.ndarray:
    """Permutation sorting points along a Hilbert curve (2^bits grid).

    Cells that are close in this order are close in space, so a contiguous
    block of cells (one workgroup) touches a compact neighbourhood and its
    neighbour gathers hit the same L2 lines.
    """
    lo = points.min(axis=0)
    span = np.maximum(points.max(axis=0) - lo, 1e-300)
    side = (1 << bits) - 1
    q = np.clip(((points - lo) / span.max() * side).astype(np.int64), 0, side)
    x, y = q[:, 0].copy(), q[:, 1].copy()
    d = np.zeros(len(points), dtype=np.int64)
    s = 1 << (bits - 1)
    while s > 0:
        rx = ((x & s) > 0).astype(np.int64)
        ry = ((y & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        # rotate the quadrant
        flip = ry == 0
        swap_x = flip & (rx == 1)
        x = np.where(swap_x, side - x, x)
        y = np.where(swap_x, side - y, y)
        x2 = np.where(flip, y, x)
        y2 = np.where(flip, x, y)
        x, y = x2, y2
        s >>= 1
    return np.argsort(d, kind="stable")


@dataclass
class DeviceLayout:
    n_cells: int
    cell_order: np.ndarray  # device cell j holds mesh cell cell_order[j]
    area: np.ndarray
    grad_ptr: np.ndarray
    grad_nb: np.ndarray
    grad_cx: np.ndarray
    grad_cy: np.ndarray
    grad_rev: np.ndarray  # entry of the neighbour's list that points back
    lim_ptr: np.ndarray
    lim_dx: np.ndarray
    lim_dy: np.ndarray
    face_left: np.ndarray
    face_right: np.ndarray
    face_kind: np.ndarray
    face_bc: np.ndarray
    face_nx: np.ndarray
    face_ny: np.ndarray
    face_len: np.ndarray
    face_lx: np.ndarray
    face_ly: np.ndarray
    face_rx: np.ndarray
    face_ry: np.ndarray
    acc_ptr: np.ndarray
    acc_item: np.ndarray
    acc_coef: np.ndarray  # -L/A_owner, L/A_neighbour (solver.py:2157-2174); 0 on boundary items
    dts_ptr: np.ndarray
    dts_nx: np.ndarray
    dts_ny: np.ndarray
    dts_len: np.ndarray
    dts_bc: np.ndarray
    bc_kind: np.ndarray
    bc_state: np.ndarray
    bc_wall: np.ndarray
    mesh_face_flux: np.ndarray = None  # mesh face -> flux face (periodic b faces: their pair's)
    dts_flux: np.ndarray = None  # flux face of each spectral-radius item
    dts_own: np.ndarray = None  # item uses its flux face's normal and length (not a periodic b side, no exterior state)
    flux_mesh_face: np.ndarray = None  # flux face -> mesh face of its normal, length and left offset
    flux_right_face: np.ndarray = None  # flux face -> mesh face of its right offset (-1: boundary)
    grad_src_order: np.ndarray = None  # device LSQ entry -> geometry LSQ entry
    lim_src_order: np.ndarray = None  # device limiter entry -> geometry limiter entry

    @property
    def n_faces(self) -> int:
        return int(self.face_left.shape[0])

    def nbytes(self) -> int:
        return int(sum(v.nbytes for v in vars(self).values() if isinstance(v, np.ndarray)))


def _csr(keys: np.ndarray, n: int):
    """Stable grouping by key: returns (ptr, order)."""
    order = np.argsort(keys, kind="stable")
    counts = np.bincount(keys, minlength=n)
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    return ptr.astype(np.int32), order


def build_layout(g: Geometry, reorder: bool = True, order: np.ndarray | None = None) -> DeviceLayout:
    """Device layout; with ``reorder`` cells are renumbered along a Hilbert
    curve and flux faces sorted by their (renumbered) left cell. Per-cell
    lists keep the reference summation order, which depends on original face
    ids only, so reordering changes no result bit."""
    n = g.cell_count
    if n >= 2**31 - 1:
        raise ValueError("Mesh too large for 32-bit cell indices.")
    if order is not None:
        order_cells = np.asarray(order, dtype=np.int64)
    else:
        order_cells = hilbert_order(g.centers) if reorder else np.arange(n)
    new_id = np.empty(n, dtype=np.int64)
    new_id[order_cells] = np.arange(n)
    c = g.centers
    fc = g.face_centers

    # --- flux faces --------------------------------------------------------
    groups = []  # (kind, left, right, nx, ny, len, lx, ly, rx, ry, bc_state, bc_wall)
    inter = g.interior_faces
    io, ine = g.interior_owners, g.interior_neighbours
    groups.append(
        (
            FACE_INTERIOR, io, ine, g.face_normals[inter], g.face_lengths[inter],
            g.interior_owner_offsets, g.interior_neighbour_offsets, None, None,
        )
    )
    fa, fb = g.periodic_faces_a, g.periodic_faces_b
    oa, ob = g.owners[fa], g.owners[fb]
    groups.append(
        (FACE_PERIODIC, oa, ob, g.face_normals[fa], g.face_lengths[fa], fc[fa] - c[oa], fc[fb] - c[ob], None, None)
    )

    def bgroup(kind, faces, states=None, pressures=None):
        faces = np.asarray(faces, dtype=np.int64)
        owners = g.owners[faces]
        st = np.zeros((len(faces), 4))
        if states is not None:
            st[:] = states
        if pressures is not None:
            st[:, 3] = pressures
        wall = g.wall_velocities[faces] if kind == FACE_WALL else np.zeros((len(faces), 2))
        groups.append(
            (kind, owners, np.full(len(faces), -1, np.int64), g.face_normals[faces], g.face_lengths[faces],
             fc[faces] - c[owners], np.zeros((len(faces), 2)), st, wall)
        )
        return faces

    bc_faces = [
        bgroup(FACE_WALL, g.wall_faces),
        bgroup(FACE_VELOCITY_INLET, g.velocity_inlet_faces, states=g.velocity_inlet_states),
        bgroup(FACE_PRESSURE_OUTLET, g.pressure_outlet_faces, pressures=g.pressure_outlet_pressures),
        bgroup(FACE_FAR_FIELD, g.far_field_faces, states=g.far_field_states),
        bgroup(FACE_PRESSURE_INLET, g.pressure_inlet_faces),
        bgroup(FACE_MASS_FLOW_INLET, g.mass_flow_inlet_faces),
        bgroup(FACE_MASS_FLOW_OUTLET, g.mass_flow_outlet_faces),
        bgroup(FACE_TRANSMISSIVE, g.transmissive_faces),
    ]
    bc_kinds_order = (FACE_WALL, FACE_VELOCITY_INLET, FACE_PRESSURE_OUTLET, FACE_FAR_FIELD) + INFLOW_KINDS + (
        FACE_TRANSMISSIVE,)
    kinds = np.concatenate([np.full(len(gr[1]), gr[0], np.int32) for gr in groups])
    left = np.concatenate([gr[1] for gr in groups]).astype(np.int32)
    right = np.concatenate([gr[2] for gr in groups]).astype(np.int32)
    normals = np.concatenate([gr[3] for gr in groups])
    lengths = np.concatenate([gr[4] for gr in groups])
    loff = np.concatenate([gr[5] for gr in groups])
    roff = np.concatenate([gr[6] for gr in groups])
    nf = len(kinds)
    n_int, n_per = len(inter), len(fa)
    bstart = n_int + n_per
    n_bc = nf - bstart
    face_bc = np.full(nf, -1, np.int32)
    face_bc[bstart:] = np.arange(n_bc, dtype=np.int32)
    bc_state = np.concatenate([gr[7] for gr in groups[2:]]) if n_bc else np.zeros((0, 4))
    bc_wall = np.concatenate([gr[8] for gr in groups[2:]]) if n_bc else np.zeros((0, 2))
    bc_kind = kinds[bstart:].copy()

    # --- renumbering: faces sorted by their renumbered left cell ----------------
    fperm = np.argsort(new_id[left], kind="stable")  # new face -> old face
    fnew = np.empty(nf, dtype=np.int64)
    fnew[fperm] = np.arange(nf)

    mesh_face_flux = np.full(len(g.owners), -1, dtype=np.int64)
    old_mesh = np.concatenate([inter, fa] + bc_faces).astype(np.int64)
    old_right = np.concatenate([inter, fb, np.full(n_bc, -1, np.int64)]).astype(np.int64)
    mesh_face_flux[old_mesh] = fnew[np.arange(nf)]
    mesh_face_flux[fb] = fnew[n_int + np.arange(n_per)]

    # --- accumulation items ------------------------------------------------
    k_int = np.arange(n_int, dtype=np.int64)
    # interior: row order ascending (original) face index over owner and neighbour entries
    int_cells = np.concatenate((io, ine))
    int_k = np.concatenate((k_int, k_int))
    int_mode = np.concatenate((np.full(n_int, ACC_OWNER), np.full(n_int, ACC_NEIGHBOUR)))
    o = np.lexsort((int_k, int_cells))
    int_cells, int_k, int_mode = int_cells[o], int_k[o], int_mode[o]
    k_per = n_int + np.arange(n_per, dtype=np.int64)
    k_bnd = bstart + np.arange(n_bc, dtype=np.int64)
    b_owners = left[bstart:].astype(np.int64)
    seq_cells = np.concatenate((oa, ob, b_owners))
    seq_k = np.concatenate((k_per, k_per, k_bnd))
    seq_mode = np.concatenate(
        (np.full(n_per, ACC_BOUNDARY), np.full(n_per, ACC_PERIODIC_B), np.full(n_bc, ACC_BOUNDARY))
    )
    # interior entries (already cell-grouped, ascending face) precede the sequence entries
    all_cells = np.concatenate((int_cells, seq_cells))
    all_items = fnew[np.concatenate((int_k, seq_k))] * 4 + np.concatenate((int_mode, seq_mode))
    acc_ptr, order = _csr(new_id[all_cells], n)
    acc_item = all_items[order]
    # CSR data of the divergence operator: (-L)/A for the owner, L/A for the neighbour
    lint = g.face_lengths[inter]
    int_coef = np.where(int_mode == ACC_OWNER, -lint[int_k] / g.areas[int_cells], lint[int_k] / g.areas[int_cells])
    all_coef = np.concatenate((int_coef, np.zeros(len(seq_cells))))
    acc_coef = all_coef[order]
    if acc_item.size and acc_item.max() >= 2**31:
        raise ValueError("Too many faces for 32-bit accumulation items.")

    # --- LSQ and limiter entries -------------------------------------------
    grad_ptr, order = _csr(new_id[g.gradient_sources], n)
    grad_src_order = order
    grad_nb = new_id[g.gradient_neighbours[order]].astype(np.int32)
    # LSQ entries come in (a -> b, b -> a) pairs: 2k and 2k+1 in the reference list
    pos = np.empty(len(order), dtype=np.int64)
    pos[order] = np.arange(len(order))
    grad_rev = pos[order ^ 1].astype(np.int32)
    grad_cx = g.gradient_coefficients_x[order]
    grad_cy = g.gradient_coefficients_y[order]
    lim_ptr, order = _csr(new_id[g.limiter_cells], n)
    lim_src_order = order
    lim_dx = g.limiter_dx[order]
    lim_dy = g.limiter_dy[order]

    # --- spectral-radius entries -------------------------------------------
    F = len(g.owners)
    mesh_bc = np.full(F, -1, np.int64)
    # boundary records (numbered from the first boundary flux face) of faces
    # whose exterior state enters the time step
    rec = 0
    for kind, faces in zip(bc_kinds_order, bc_faces):
        if kind in (FACE_VELOCITY_INLET, FACE_PRESSURE_OUTLET, FACE_FAR_FIELD) + INFLOW_KINDS:
            mesh_bc[faces] = rec + np.arange(len(faces))
        rec += len(faces)
    d_cells = np.concatenate((g.owners, ine))
    d_faces = np.concatenate((np.arange(F), inter))
    d_bc = np.concatenate((mesh_bc, np.full(len(inter), -1, np.int64)))
    dts_ptr, order = _csr(new_id[d_cells], n)
    d_faces = d_faces[order]
    d_bc_sorted = d_bc[order]
    right_new = np.where(right >= 0, new_id[np.maximum(right, 0)], -1)
    left = new_id[left][fperm].astype(np.int32)
    right = right_new[fperm].astype(np.int32)
    kinds = kinds[fperm]
    face_bc = face_bc[fperm]
    normals = normals[fperm]
    lengths = lengths[fperm]
    loff = loff[fperm]
    roff = roff[fperm]
    return DeviceLayout(
        n_cells=n,
        cell_order=order_cells,
        area=np.ascontiguousarray(g.areas[order_cells]),
        grad_ptr=grad_ptr,
        grad_nb=grad_nb,
        grad_cx=np.ascontiguousarray(grad_cx),
        grad_cy=np.ascontiguousarray(grad_cy),
        grad_rev=grad_rev,
        lim_ptr=lim_ptr,
        lim_dx=np.ascontiguousarray(lim_dx),
        lim_dy=np.ascontiguousarray(lim_dy),
        face_left=left,
        face_right=right,
        face_kind=kinds,
        face_bc=face_bc,
        face_nx=np.ascontiguousarray(normals[:, 0]),
        face_ny=np.ascontiguousarray(normals[:, 1]),
        face_len=np.ascontiguousarray(lengths),
        face_lx=np.ascontiguousarray(loff[:, 0]),
        face_ly=np.ascontiguousarray(loff[:, 1]),
        face_rx=np.ascontiguousarray(roff[:, 0]),
        face_ry=np.ascontiguousarray(roff[:, 1]),
        acc_ptr=acc_ptr,
        acc_item=acc_item.astype(np.int32),
        acc_coef=np.ascontiguousarray(acc_coef),
        dts_ptr=dts_ptr,
        dts_nx=np.ascontiguousarray(g.face_normals[d_faces, 0]),
        dts_ny=np.ascontiguousarray(g.face_normals[d_faces, 1]),
        dts_len=np.ascontiguousarray(g.face_lengths[d_faces]),
        dts_bc=d_bc[order].astype(np.int32),
        bc_kind=bc_kind.astype(np.int32),
        bc_state=np.ascontiguousarray(bc_state.reshape(-1)),
        bc_wall=np.ascontiguousarray(bc_wall.reshape(-1)),
        mesh_face_flux=mesh_face_flux,
        dts_flux=mesh_face_flux[d_faces].astype(np.int32),
        dts_own=(~np.isin(d_faces, fb)) & (d_bc_sorted < 0),
        flux_mesh_face=old_mesh[fperm],
        flux_right_face=old_right[fperm],
        grad_src_order=grad_src_order,
        lim_src_order=lim_src_order,
    )
