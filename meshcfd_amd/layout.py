"""Device layout of a `Geometry`: flux faces and per-cell CSR lists.

Each per-cell list keeps the reference's summation order, so the device
gather reproduces the reference residual's sums term for term:

* LSQ entries: the order of `gradient_sources` (`solver.py:2004-2019`);
* accumulation: interior faces ascending (the CSR divergence operator's row
  order, `solver.py:2157-2174,2501`), then periodic a-sides, periodic
  b-sides, wall/symmetry, velocity inlet, pressure outlet, far field and
  transmissive faces, in that order (`solver.py:2502-2625`);
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
ACC_OWNER, ACC_NEIGHBOUR, ACC_BOUNDARY, ACC_PERIODIC_B = 0, 1, 2, 3


@dataclass
class DeviceLayout:
    n_cells: int
    area: np.ndarray
    grad_ptr: np.ndarray
    grad_nb: np.ndarray
    grad_cx: np.ndarray
    grad_cy: np.ndarray
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
    dts_ptr: np.ndarray
    dts_nx: np.ndarray
    dts_ny: np.ndarray
    dts_len: np.ndarray
    dts_bc: np.ndarray
    bc_kind: np.ndarray
    bc_state: np.ndarray
    bc_wall: np.ndarray

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


def build_layout(g: Geometry) -> DeviceLayout:
    n = g.cell_count
    if n >= 2**31 - 1:
        raise ValueError("Mesh too large for 32-bit cell indices.")
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
        bgroup(FACE_TRANSMISSIVE, g.transmissive_faces),
    ]
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

    # --- accumulation items ------------------------------------------------
    k_int = np.arange(n_int, dtype=np.int64)
    # interior: row order ascending face index over owner and neighbour entries
    int_cells = np.concatenate((io, ine))
    int_items = np.concatenate((k_int * 4 + ACC_OWNER, k_int * 4 + ACC_NEIGHBOUR))
    int_face = np.concatenate((k_int, k_int))
    o = np.lexsort((int_face, int_cells))
    int_cells, int_items = int_cells[o], int_items[o]
    k_per = n_int + np.arange(n_per, dtype=np.int64)
    k_bnd = bstart + np.arange(n_bc, dtype=np.int64)
    b_owners = left[bstart:].astype(np.int64)
    seq_cells = np.concatenate((oa, ob, b_owners))
    seq_items = np.concatenate((k_per * 4 + ACC_BOUNDARY, k_per * 4 + ACC_PERIODIC_B, k_bnd * 4 + ACC_BOUNDARY))
    # interior entries (already cell-grouped, ascending face) precede the sequence entries
    all_cells = np.concatenate((int_cells, seq_cells))
    all_items = np.concatenate((int_items, seq_items))
    acc_ptr, order = _csr(all_cells, n)
    acc_item = all_items[order]
    if acc_item.size and acc_item.max() >= 2**31:
        raise ValueError("Too many faces for 32-bit accumulation items.")

    # --- LSQ and limiter entries -------------------------------------------
    grad_ptr, order = _csr(g.gradient_sources, n)
    grad_nb = g.gradient_neighbours[order].astype(np.int32)
    grad_cx = g.gradient_coefficients_x[order]
    grad_cy = g.gradient_coefficients_y[order]
    lim_ptr, order = _csr(g.limiter_cells, n)
    lim_dx = g.limiter_dx[order]
    lim_dy = g.limiter_dy[order]

    # --- spectral-radius entries -------------------------------------------
    F = len(g.owners)
    mesh_bc = np.full(F, -1, np.int64)
    # boundary records for faces whose exterior state enters the time step
    off = bstart - bstart  # records are numbered from the first boundary flux face
    rec = 0
    for kind, faces in zip(
        (FACE_WALL, FACE_VELOCITY_INLET, FACE_PRESSURE_OUTLET, FACE_FAR_FIELD, FACE_TRANSMISSIVE), bc_faces
    ):
        if kind in (FACE_VELOCITY_INLET, FACE_PRESSURE_OUTLET, FACE_FAR_FIELD):
            mesh_bc[faces] = off + rec + np.arange(len(faces))
        rec += len(faces)
    d_cells = np.concatenate((g.owners, ine))
    d_faces = np.concatenate((np.arange(F), inter))
    d_bc = np.concatenate((mesh_bc, np.full(len(inter), -1, np.int64)))
    dts_ptr, order = _csr(d_cells, n)
    d_faces = d_faces[order]
    return DeviceLayout(
        n_cells=n,
        area=np.ascontiguousarray(g.areas),
        grad_ptr=grad_ptr,
        grad_nb=grad_nb,
        grad_cx=np.ascontiguousarray(grad_cx),
        grad_cy=np.ascontiguousarray(grad_cy),
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
        dts_ptr=dts_ptr,
        dts_nx=np.ascontiguousarray(g.face_normals[d_faces, 0]),
        dts_ny=np.ascontiguousarray(g.face_normals[d_faces, 1]),
        dts_len=np.ascontiguousarray(g.face_lengths[d_faces]),
        dts_bc=d_bc[order].astype(np.int32),
        bc_kind=bc_kind.astype(np.int32),
        bc_state=np.ascontiguousarray(bc_state.reshape(-1)),
        bc_wall=np.ascontiguousarray(bc_wall.reshape(-1)),
    )
