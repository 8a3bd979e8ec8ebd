"""Implicit Euler on the device: LU-SGS structure, Newton driver, dt halving.

Mirrors the reference's implicit chain (`src/meshcfd/core/solver.py`):

* `implicit_desc`     -- structure of `_implicit_preconditioner`'s scalar
  matrix (`:1185-1301`) in device cell ids: face-pair and boundary records,
  per-cell diagonal items in add.at order, off-diagonal items split into the
  natural-order lower / upper triangles;
* `ImplicitStepper.solve_once` -- `_implicit_euler_solve_once` (`:1371-1486`):
  Newton on G(U) = U - U^n - dt R(U) with variable scaling, matrix-free GMRES
  (`mcfd_gmres`) preconditioned by LU-SGS (`mcfd_lusgs_*`), backtracking;
* `ImplicitStepper.step` -- `_implicit_euler_step` (`:1337-1368`): up to 12
  dt halvings in adaptive modes, 1 attempt in fixed mode.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import native
from .geometry import Geometry, SolverError
from .layout import DeviceLayout


LUSGS_ORDERINGS = ("natural", "multicolour")


def halve_until_solved(solve_once, time_step, allow_reduction: bool = True):
    """The retry loop of `_implicit_euler_step` (`solver.py:1337-1368`):
    ``solve_once(step)``, halving the step after each `SolverError`, at most
    12 attempts (1 without reduction). Returns (solution, accepted step);
    ``time_step`` is a float, a NumPy array or a device tensor."""
    h = time_step
    last = None
    for _ in range(12 if allow_reduction else 1):
        try:
            return solve_once(h), h
        except SolverError as error:
            last = error
            h = h * 0.5
    if not allow_reduction:
        raise SolverError(f"Implicit Euler could not converge at the configured fixed time step: {last}") from last
    raise SolverError(
        f"Implicit Euler could not converge after reducing the time step 12 times; last failure: {last}"
    ) from last


def colour_cells(ptr: np.ndarray, nbr: np.ndarray, seed: int = 0) -> np.ndarray:
    """Greedy colouring of the cell graph (CSR ``ptr``, ``nbr``) in
    Jones-Plassmann rounds: each round the uncoloured cells whose seeded
    random priority exceeds that of every uncoloured neighbour take the
    smallest colour no coloured neighbour has (two such cells are never
    neighbours). At most max degree + 1 colours."""
    n = len(ptr) - 1
    prio = np.random.default_rng(seed).permutation(n).astype(np.int64)
    colour = np.full(n, -1, dtype=np.int64)
    deg = np.diff(ptr)
    has = deg > 0
    starts = ptr[:-1][has]
    if len(nbr) and deg.max() > 60:
        raise ValueError("a cell has more than 60 neighbours")
    owner = np.repeat(np.arange(n), deg)
    if np.any(nbr == owner):
        raise ValueError("the cell graph lists a cell as its own neighbour")
    while True:
        unc = colour < 0
        if not unc.any():
            return colour
        m = np.full(n, -1, dtype=np.int64)
        if len(nbr):
            m[has] = np.maximum.reduceat(np.where(unc[nbr], prio[nbr], -1), starts)
        sel = unc & (prio > m)
        if not sel.any():  # cannot happen without self-edges: the highest uncoloured priority wins
            raise ValueError("colouring made no progress")
        used = np.zeros(n, dtype=np.int64)
        if len(nbr):
            bits = np.where(colour[nbr] >= 0, np.left_shift(np.int64(1), np.maximum(colour[nbr], 0)), 0)
            used[has] = np.bitwise_or.reduceat(bits, starts)
        free = ~used & (used + 1)  # lowest colour not used
        colour[sel] = np.log2(free[sel].astype(np.float64)).astype(np.int64)


def two_colouring(ptr: np.ndarray, nbr: np.ndarray):
    """A 2-colouring of the cell graph (CSR ``ptr``, ``nbr``), or None when
    the graph has an odd cycle. On the bipartite double cover (cells (v, 0)
    and (v, 1), each edge u-v joining (u, 0)-(v, 1) and (u, 1)-(v, 0)) a
    component of the graph splits in two exactly when it is bipartite; a
    cell takes colour 0 when (v, 0) lies in the lower-numbered half."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    n = len(ptr) - 1
    src = np.repeat(np.arange(n), np.diff(ptr))
    rows = np.concatenate((src, src + n))
    cols = np.concatenate((nbr + n, nbr))
    graph = coo_matrix((np.ones(len(rows), dtype=np.int8), (rows, cols)), shape=(2 * n, 2 * n)).tocsr()
    _, lab = connected_components(graph, directed=False)
    if np.any(lab[:n] == lab[n:]):
        return None
    return (lab[:n] > lab[n:]).astype(np.int64)


def implicit_desc(g: Geometry, L: DeviceLayout, ordering: str = "natural"):
    """The LU-SGS structure on the device cells. ``ordering`` "natural" is
    the reference's (`solver.py:1309-1323`: triangular solves in mesh
    order); "multicolour" (not in the reference; a different, equally valid
    SGS preconditioner, so GMRES takes other iterations to the same
    tolerance) orders the cells colour by colour, so each sweep has as many
    levels as colours, each of them fully parallel."""
    if ordering not in LUSGS_ORDERINGS:
        raise ValueError(f"Unsupported LU-SGS ordering '{ordering}'.")
    n = g.cell_count
    new_id = np.empty(n, dtype=np.int64)
    new_id[L.cell_order] = np.arange(n)
    inner = g.interior_faces
    fa, fb = g.periodic_faces_a, g.periodic_faces_b
    bnd = np.flatnonzero(g.neighbours < 0)
    ni, npair = len(inner), len(fa)
    nrm = g.face_normals
    rec_o = np.concatenate((g.owners[inner], g.owners[fa], g.owners[bnd]))
    rec_n = np.concatenate((g.neighbours[inner], g.owners[fb], np.full(len(bnd), -1)))
    n_o = np.concatenate((nrm[inner], nrm[fa], nrm[bnd]))
    n_n = np.concatenate((nrm[inner], nrm[fb], nrm[bnd]))
    rlen = np.concatenate((g.face_lengths[inner], g.face_lengths[fa], g.face_lengths[bnd]))
    r_int = np.arange(ni)
    r_per = ni + np.arange(npair)
    r_bnd = ni + npair + np.arange(len(bnd))
    # A cell that owns both faces of a periodic pair (a mesh one cell wide)
    # has the pair's two off-diagonal entries on its own row and column: the
    # reference's COO sum adds them to the diagonal after the diagonal value,
    # a->b then b->a. They are listed as diagonal items -(item) - 1.
    self_pair = g.owners[fa] == g.owners[fb]
    sp_cells = np.repeat(g.owners[fa][self_pair], 2)
    sp_items = -np.column_stack((2 * r_per[self_pair], 2 * r_per[self_pair] + 1)).ravel() - 1
    # diagonal: interior owners, interior neighbours, periodic a, periodic b,
    # boundary, then the self-pair entries
    d_cells = np.concatenate((g.owners[inner], g.neighbours[inner], g.owners[fa], g.owners[fb], g.owners[bnd], sp_cells))
    d_items = np.concatenate((2 * r_int, 2 * r_int + 1, 2 * r_per, 2 * r_per + 1, 2 * r_bnd, sp_items))
    o = np.argsort(new_id[d_cells], kind="stable")
    diag_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(new_id[d_cells], minlength=n), out=diag_ptr[1:])
    diag_item = d_items[o]
    # off-diagonal, COO order: owner->neighbour, neighbour->owner, a->b, b->a
    o_cells = np.concatenate((g.owners[inner], g.neighbours[inner], g.owners[fa], g.owners[fb]))
    o_cols = np.concatenate((g.neighbours[inner], g.owners[inner], g.owners[fb], g.owners[fa]))
    o_items = np.concatenate((2 * r_int, 2 * r_int + 1, 2 * r_per, 2 * r_per + 1))
    off_diag = o_cells != o_cols
    o_cells, o_cols, o_items = o_cells[off_diag], o_cols[off_diag], o_items[off_diag]
    seq = np.arange(len(o_cells))
    if ordering == "natural":
        rank = np.arange(n)  # position of each mesh cell in the sweep order
    else:
        # colour the mesh cells (neighbours through the off-diagonal items)
        co = np.argsort(o_cells, kind="stable")
        cptr = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(np.bincount(o_cells, minlength=n), out=cptr[1:])
        colour = two_colouring(cptr, o_cols[co])
        if colour is None:
            colour = colour_cells(cptr, o_cols[co])
        rank = np.empty(n, dtype=np.int64)
        rank[np.lexsort((np.arange(n), colour))] = np.arange(n)
    # a cell's items: lower (earlier in the sweep order) first
    o = np.lexsort((seq, rank[o_cols], new_id[o_cells]))
    off_cells, off_cols, off_items = o_cells[o], o_cols[o], o_items[o]
    off_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(new_id[off_cells], minlength=n), out=off_ptr[1:])
    lower = rank[off_cols] < rank[off_cells]
    n_lower = np.bincount(new_id[off_cells[lower]], minlength=n)
    off_split = off_ptr[:-1] + n_lower
    # item of (record, side) -> position; the transposed entry is the other side's item
    pos = np.full(2 * (ni + npair), -1, dtype=np.int64)
    pos[off_items] = np.arange(len(off_items))
    off_t = pos[off_items ^ 1]
    return dict(
        n_records=len(rec_o),
        rec_o=new_id[rec_o],
        rec_n=np.where(rec_n >= 0, new_id[np.maximum(rec_n, 0)], -1),
        rec_nox=n_o[:, 0].copy(),
        rec_noy=n_o[:, 1].copy(),
        rec_nnx=n_n[:, 0].copy(),
        rec_nny=n_n[:, 1].copy(),
        rec_len=rlen,
        diag_ptr=diag_ptr,
        diag_item=diag_item,
        off_ptr=off_ptr,
        off_split=off_split,
        off_item=off_items,
        off_nb=new_id[off_cols],
        natural_order=new_id[np.argsort(rank)],
        off_t=off_t,
    )


LAUNCH_PATHS = {"auto": 0, "workgroup": 1, "levels": 2}


class McfdImplicitDesc(C.Structure):
    _fields_ = [
        ("n_records", C.c_int32),
        ("launch_path", C.c_int32),
        ("rec_o", C.c_void_p),
        ("rec_n", C.c_void_p),
        ("rec_nox", C.c_void_p),
        ("rec_noy", C.c_void_p),
        ("rec_nnx", C.c_void_p),
        ("rec_nny", C.c_void_p),
        ("rec_len", C.c_void_p),
        ("diag_ptr", C.c_void_p),
        ("diag_item", C.c_void_p),
        ("off_ptr", C.c_void_p),
        ("off_split", C.c_void_p),
        ("off_item", C.c_void_p),
        ("off_nb", C.c_void_p),
        ("natural_order", C.c_void_p),
        ("off_t", C.c_void_p),
    ]


class McfdGmres(C.Structure):
    _fields_ = [
        ("state", C.c_void_p),
        ("rate", C.c_void_p),
        ("rhs", C.c_void_p),
        ("x", C.c_void_p),
        ("dt", C.c_void_p),
        ("dt_per_cell", C.c_int32),
        ("krylov", C.c_int32),
        ("max_iterations", C.c_int32),
        ("iterations", C.c_int32),
        ("info", C.c_int32),
        ("reserved", C.c_int32),
        ("sigma", C.c_double * 4),
        ("state_norm", C.c_double),
        ("tolerance", C.c_double),
        ("residual", C.c_double),
    ]


native.EXPORTS.update(
    {
        "mcfd_mesh_set_implicit": (C.c_int, [C.c_void_p, C.c_void_p]),
        "mcfd_lusgs_assemble": (
            C.c_int,
            [C.c_void_p, C.POINTER(native.McfdSettings), C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p],
        ),
        "mcfd_lusgs_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "mcfd_jacobian_apply": (
            C.c_int,
            [C.c_void_p, C.POINTER(native.McfdSettings), C.POINTER(McfdGmres), C.c_void_p, C.c_void_p, C.c_void_p],
        ),
        "mcfd_gmres": (C.c_int, [C.c_void_p, C.POINTER(native.McfdSettings), C.POINTER(McfdGmres), C.c_void_p]),
        "mcfd_newton_equation": (
            C.c_int,
            [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_double),
             C.c_void_p, C.c_void_p, C.c_void_p],
        ),
        "mcfd_newton_candidate": (
            C.c_int,
            [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_void_p, C.c_void_p],
        ),
        "mcfd_norm2": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "mcfd_lusgs_apply_transpose": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "mcfd_adjoint_apply": (
            C.c_int,
            [C.c_void_p, C.POINTER(native.McfdSettings), C.POINTER(McfdGmres), C.c_void_p, C.c_void_p, C.c_void_p],
        ),
        "mcfd_adjoint_gmres": (
            C.c_int, [C.c_void_p, C.POINTER(native.McfdSettings), C.POINTER(McfdGmres), C.c_void_p]
        ),
    }
)


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class ImplicitStepper:
    """Implicit-Euler steps of one DeviceMesh. ``launch_path`` fixes how the
    LU-SGS sweeps launch ("workgroup": one workgroup per sweep, "levels": a
    launch per level); "auto" lets the library choose per mesh. All paths
    give the same bits."""

    def __init__(self, dm, launch_path: str = "auto"):
        if launch_path not in LAUNCH_PATHS:
            raise ValueError(f"Unsupported LU-SGS launch path '{launch_path}'.")
        self.dm = dm
        self.lib = native.load()
        for name in ("mcfd_mesh_set_implicit", "mcfd_lusgs_assemble", "mcfd_lusgs_apply", "mcfd_jacobian_apply",
                     "mcfd_gmres", "mcfd_newton_equation", "mcfd_newton_candidate", "mcfd_norm2",
                     "mcfd_lusgs_apply_transpose", "mcfd_adjoint_apply", "mcfd_adjoint_gmres"):
            restype, argtypes = native.EXPORTS[name]
            fn = getattr(self.lib, name)
            fn.restype, fn.argtypes = restype, argtypes
        d = implicit_desc(dm.geometry, dm.layout, getattr(dm, "lusgs_ordering", "natural"))
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return C.c_void_p(a.ctypes.data)

        desc = McfdImplicitDesc(
            n_records=d["n_records"],
            launch_path=LAUNCH_PATHS[launch_path],
            rec_o=arr(d["rec_o"], np.int32),
            rec_n=arr(d["rec_n"], np.int32),
            rec_nox=arr(d["rec_nox"], np.float64),
            rec_noy=arr(d["rec_noy"], np.float64),
            rec_nnx=arr(d["rec_nnx"], np.float64),
            rec_nny=arr(d["rec_nny"], np.float64),
            rec_len=arr(d["rec_len"], np.float64),
            diag_ptr=arr(d["diag_ptr"], np.int32),
            diag_item=arr(d["diag_item"], np.int32),
            off_ptr=arr(d["off_ptr"], np.int32),
            off_split=arr(d["off_split"], np.int32),
            off_item=arr(d["off_item"], np.int32),
            off_nb=arr(d["off_nb"], np.int32),
            natural_order=arr(d["natural_order"], np.int32),
            off_t=arr(d["off_t"], np.int32),
        )
        native.check(self.lib.mcfd_mesh_set_implicit(dm.handle, C.byref(desc)))
        del keep
        s = dm.settings
        self.bind_settings()
        opts = dict(dtype=torch.float64, device=dm.device)
        n = dm.n
        self.rate1 = torch.empty((n, 4), **opts)
        self.R = torch.empty((n, 4), **opts)
        self.scaled = torch.empty((n, 4), **opts)
        self.rhs = torch.empty((n, 4), **opts)
        self.x = torch.empty((n, 4), **opts)
        self.cand = torch.empty((n, 4), **opts)
        self.norms = torch.zeros((2,), **opts)
        self.gmres_iterations = 0
        self.newton_iterations = 0
        # called after each Newton iteration's GMRES solve with (Newton
        # iterations so far, GMRES iterations of this solve, residual norm)
        self.on_newton = None
        del s

    def bind_settings(self) -> None:
        """The Jacobian's settings: the mesh's, with first-order
        reconstruction (`solver.py:1378`)."""
        self.first_settings = native.McfdSettings.from_buffer_copy(self.dm.native_settings)
        self.first_settings.reconstruction = 0

    # -- pieces -----------------------------------------------------------
    def assemble(self, u, dt, per_cell):
        native.check(
            self.lib.mcfd_lusgs_assemble(self.dm.handle, C.byref(self.first_settings), _p(u), _p(dt),
                                         1 if per_cell else 0, self.dm.stream)
        )

    def lusgs_apply(self, v, out):
        native.check(self.lib.mcfd_lusgs_apply(self.dm.handle, _p(v), _p(out), self.dm.stream))
        return out

    def gmres_args(self, state, rate, dt, per_cell, sigma, rhs=None, x=None):
        g = McfdGmres()
        g.state, g.rate = _p(state), _p(rate)
        g.rhs, g.x = _p(rhs), _p(x)
        g.dt = _p(dt)
        g.dt_per_cell = 1 if per_cell else 0
        s = self.dm.settings
        g.krylov = s.gmres_krylov_dimension
        g.max_iterations = s.gmres_max_iterations
        for k in range(4):
            g.sigma[k] = float(sigma[k])
        g.state_norm = -1.0  # ||state / sigma||_2, reduced on the device by the call
        g.tolerance = s.gmres_tolerance
        return g

    def jacobian_apply(self, args, v, out):
        native.check(
            self.lib.mcfd_jacobian_apply(self.dm.handle, C.byref(self.first_settings), C.byref(args), _p(v), _p(out),
                                         self.dm.stream)
        )
        return out

    def equation(self, u, u0, dt, per_cell, sigma):
        sig = (C.c_double * 4)(*sigma)
        native.check(
            self.lib.mcfd_newton_equation(self.dm.handle, _p(u), _p(u0), _p(self.R), _p(dt), 1 if per_cell else 0, sig,
                                          _p(self.scaled), _p(self.norms), self.dm.stream)
        )
        values = self.norms.cpu().numpy()
        return float(values[0]), float(values[1])

    def _residual_ok(self, u) -> bool:
        self.dm.residual(u, out=self.R)
        try:
            self.dm.check_status()
        except SolverError:
            return False
        return True

    # -- Newton -----------------------------------------------------------
    def solve_once(self, initial: torch.Tensor, dt: torch.Tensor, per_cell: bool) -> torch.Tensor:
        s = self.dm.settings
        gamma = s.gamma
        state = initial.clone()
        host = state.cpu().numpy()
        rho = host[:, 0]
        vel = host[:, 1:3] / rho[:, None]
        p = (gamma - 1.0) * (host[:, 3] - 0.5 * (host[:, 1] ** 2 + host[:, 2] ** 2) / rho)
        rho_scale = max(float(np.max(rho)), 1e-12)
        speed = max(float(np.max(np.abs(vel))), float(np.max(np.sqrt(gamma * p / rho))), 1e-12)
        sc = np.maximum(np.max(np.abs(host), axis=0), 1e-12)
        sc[0] = max(sc[0], rho_scale)
        sc[1:3] = np.maximum(sc[1:3], rho_scale * speed)
        sc[3] = max(sc[3], rho_scale * speed**2, float(np.max(p)) / (gamma - 1.0))
        sigma = [float(v) for v in sc]
        sig = (C.c_double * 4)(*sigma)
        for _ in range(12):
            if not self._residual_ok(state):
                raise SolverError("Spatial residual contains non-finite values.")
            norm, upd = self.equation(state, initial, dt, per_cell, sigma)
            if norm <= 1e-9 * max(1.0, upd):
                return state
            self.newton_iterations += 1
            self.assemble(state, dt, per_cell)
            self.dm.check_status()
            self.dm._stage_first_order_residual(state, self.rate1, self.first_settings)
            torch.neg(self.scaled, out=self.rhs)
            args = self.gmres_args(state, self.rate1, dt, per_cell, sigma, rhs=self.rhs, x=self.x)
            native.check(self.lib.mcfd_gmres(self.dm.handle, C.byref(self.first_settings), C.byref(args), self.dm.stream))
            self.gmres_iterations += args.iterations
            if self.on_newton is not None:
                self.on_newton(self.newton_iterations, args.iterations, norm)
            if args.info != 0:
                raise SolverError(
                    f"Implicit Euler GMRES did not converge within {s.gmres_max_iterations} iterations "
                    f"(relative tolerance {s.gmres_tolerance:g})."
                )
            damping = 1.0
            accepted = False
            for _ in range(8):
                native.check(
                    self.lib.mcfd_newton_candidate(self.dm.handle, _p(state), _p(self.x), damping, sig, _p(self.cand),
                                                   self.dm.stream)
                )
                if self._residual_ok(self.cand):
                    cn, _ = self.equation(self.cand, initial, dt, per_cell, sigma)
                else:
                    cn = math.inf
                if cn < norm:
                    state.copy_(self.cand)
                    accepted = True
                    break
                damping *= 0.5
            if not accepted:
                raise SolverError("Implicit Euler Newton iteration could not reduce its residual.")
        raise SolverError("Implicit Euler Newton iteration did not converge.")

    def step(self, initial: torch.Tensor, dt, per_cell: bool, allow_reduction: bool = True):
        """`_implicit_euler_step`: returns (state, accepted dt tensor)."""
        return halve_until_solved(lambda h: self.solve_once(initial, h, per_cell), dt.clone(), allow_reduction)
