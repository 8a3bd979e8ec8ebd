"""Shape gradients: lambda^T dR/d(mesh nodes) and gradients with respect to
shape parameters.

No counterpart in the reference, which has no gradients. The residual reads
the mesh only through the geometry arrays `build_geometry` derives from the
nodes (`solver.py:1654-2238`): face normals and lengths, cell areas, the
reconstruction offsets (face centre - centroid), the LSQ coefficients and
the limiter offsets. The device computes lambda^T dR/d(each array)
(`mcfd_residual_geometry_vjp`, hand-written reverse of the face fluxes,
reconstruction, limiter and LSQ gradient; the area term from the forward
residual). `TorchGeometry` restates the geometry formulas as torch
functions of the node coordinates, with the discrete choices (normal
orientation, periodic shifts, the LSQ determinant test) taken from the
built geometry, so autograd carries those adjoints to the nodes and, through
a node map such as `meshgen.bump_nodes`, to shape parameters.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import adjoint, native
from .geometry import SolverError

native.EXPORTS["mcfd_residual_geometry_vjp"] = (
    C.c_int,
    [C.c_void_p, C.POINTER(native.McfdSettings), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
     C.c_void_p],
)

GEOMETRY_KEYS = ("face_nx", "face_ny", "face_len", "face_lx", "face_ly", "face_rx", "face_ry", "area",
                 "grad_cx", "grad_cy", "lim_dx", "lim_dy")


def _p(t):
    return C.c_void_p(t.data_ptr())


class TorchGeometry:
    """The geometry arrays of a DeviceLayout as torch functions of the mesh
    nodes, in device order (the formulas of `geometry.build_geometry`)."""

    def __init__(self, case, geometry, layout, device="cpu"):
        mesh = case.mesh
        g = geometry
        opts = dict(device=device)
        self.device = torch.device(device)
        lt = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int64), **opts)  # noqa: E731
        self.cells = lt(np.where(mesh.cells >= 0, mesh.cells, 0))
        self.sizes = lt(mesh.cell_sizes)
        self.fa, self.fb = lt(mesh.faces[:, 0]), lt(mesh.faces[:, 1])
        owners = np.asarray(g.owners, dtype=np.int64)
        self.owners = lt(owners)
        # orientation of each face normal, as build_geometry chose it
        pts = mesh.points
        ex = pts[mesh.faces[:, 1], 0] - pts[mesh.faces[:, 0], 0]
        ey = pts[mesh.faces[:, 1], 1] - pts[mesh.faces[:, 0], 1]
        sign = np.where(ey * g.face_normals[:, 0] - ex * g.face_normals[:, 1] < 0.0, -1.0, 1.0)
        self.sign = torch.as_tensor(sign, dtype=torch.float64, **opts)
        inter = np.asarray(g.interior_faces, dtype=np.int64)
        ine = np.asarray(g.interior_neighbours, dtype=np.int64)
        # flux faces in device order: mesh face, left cell, right mesh face and cell
        mf = layout.flux_mesh_face
        rf = layout.flux_right_face
        self.flux_face = lt(mf)
        self.flux_left = lt(owners[mf])
        has_right = rf >= 0
        neighbours = np.asarray(g.neighbours, dtype=np.int64)
        # interior faces: the face's own neighbour; periodic pairs: the b face's owner
        rcell = np.where(has_right & (neighbours[np.maximum(rf, 0)] >= 0), neighbours[np.maximum(rf, 0)],
                         owners[np.maximum(rf, 0)])
        self.flux_has_right = torch.as_tensor(has_right, **opts)
        self.flux_rface = lt(np.maximum(rf, 0))
        self.flux_rcell = lt(np.where(has_right, rcell, 0))
        self.cell_order = lt(layout.cell_order)
        # LSQ entries (geometry order): interior pairs, then periodic pairs with their shifts
        io = np.asarray(g.interior_owners, dtype=np.int64)
        pa, pb = np.asarray(g.periodic_faces_a, dtype=np.int64), np.asarray(g.periodic_faces_b, dtype=np.int64)
        oa, ob = owners[pa], owners[pb]
        self.io, self.ine = lt(io), lt(ine)
        self.oa, self.ob = lt(oa), lt(ob)
        self.shift = torch.as_tensor(np.asarray(g.periodic_shifts, dtype=np.float64).reshape(-1, 2), **opts)
        src = np.asarray(g.gradient_sources, dtype=np.int64)
        self.src = lt(src)
        self.n_cells = g.cell_count
        # LSQ determinant test of build_geometry (kept as decided there)
        gdx, gdy = g.gradient_dx, g.gradient_dy
        w = 1.0 / np.maximum(gdx * gdx + gdy * gdy, 1e-300)
        mxx, mxy, myy = (np.zeros(self.n_cells) for _ in range(3))
        np.add.at(mxx, src, w * gdx * gdx)
        np.add.at(mxy, src, w * gdx * gdy)
        np.add.at(myy, src, w * gdy * gdy)
        det = mxx * myy - mxy * mxy
        self.valid = torch.as_tensor(det > 1e-14 * np.maximum(mxx * myy, 1e-300), **opts)
        self.grad_order = lt(layout.grad_src_order)
        self.lim_order = lt(layout.lim_src_order)
        self.inter = lt(inter)

    def shoelace(self, nodes):
        """Centroids and areas (`geometry.shoelace`)."""
        n = self.cells.shape[0]
        twice = nodes.new_zeros(n)
        xs = nodes.new_zeros(n)
        ys = nodes.new_zeros(n)
        rows = torch.arange(n, device=self.device)
        for k in range(self.cells.shape[1]):
            active = self.sizes > k
            a = self.cells[:, k]
            nxt = torch.where(self.sizes > k + 1, torch.full_like(a, k + 1), torch.zeros_like(a))
            b = self.cells[rows, nxt]
            x0, y0 = nodes[a, 0], nodes[a, 1]
            x1, y1 = nodes[b, 0], nodes[b, 1]
            cross = torch.where(active, x0 * y1 - x1 * y0, torch.zeros_like(x0))
            twice = twice + cross
            xs = xs + (x0 + x1) * cross
            ys = ys + (y0 + y1) * cross
        centers = torch.stack((xs / (3.0 * twice), ys / (3.0 * twice)), 1)
        return centers, 0.5 * torch.abs(twice)

    def __call__(self, nodes) -> dict:
        c, areas = self.shoelace(nodes)
        a, b = nodes[self.fa], nodes[self.fb]
        ex = b[:, 0] - a[:, 0]
        ey = b[:, 1] - a[:, 1]
        length = torch.sqrt(ex * ex + ey * ey)
        nx = self.sign * ey / length
        ny = -self.sign * ex / length
        fc = 0.5 * (a + b)
        out = {}
        f = self.flux_face
        out["face_nx"] = nx[f]
        out["face_ny"] = ny[f]
        out["face_len"] = length[f]
        loff = fc[f] - c[self.flux_left]
        out["face_lx"], out["face_ly"] = loff[:, 0], loff[:, 1]
        roff = torch.where(self.flux_has_right[:, None], fc[self.flux_rface] - c[self.flux_rcell],
                           torch.zeros_like(loff))
        out["face_rx"], out["face_ry"] = roff[:, 0], roff[:, 1]
        out["area"] = areas[self.cell_order]
        # LSQ coefficients (build_geometry's formulas)
        delta = c[self.ine] - c[self.io]
        parts_x = [torch.stack((delta[:, 0], -delta[:, 0]), 1).reshape(-1)]
        parts_y = [torch.stack((delta[:, 1], -delta[:, 1]), 1).reshape(-1)]
        if self.oa.numel():
            pd = (c[self.ob] - self.shift) - c[self.oa]
            parts_x.append(torch.stack((pd[:, 0], -pd[:, 0]), 1).reshape(-1))
            parts_y.append(torch.stack((pd[:, 1], -pd[:, 1]), 1).reshape(-1))
        gdx, gdy = torch.cat(parts_x), torch.cat(parts_y)
        w = 1.0 / torch.clamp(gdx * gdx + gdy * gdy, min=1e-300)
        wdx, wdy = w * gdx, w * gdy
        z = nodes.new_zeros(self.n_cells)
        mxx = z.index_add(0, self.src, wdx * gdx)
        mxy = z.index_add(0, self.src, wdx * gdy)
        myy = z.index_add(0, self.src, wdy * gdy)
        det = torch.where(self.valid, mxx * myy - mxy * mxy, torch.ones_like(mxx))
        ixx = torch.where(self.valid, myy / det, z)
        ixy = torch.where(self.valid, -mxy / det, z)
        iyy = torch.where(self.valid, mxx / det, z)
        cx = ixx[self.src] * wdx + ixy[self.src] * wdy
        cy = ixy[self.src] * wdx + iyy[self.src] * wdy
        out["grad_cx"], out["grad_cy"] = cx[self.grad_order], cy[self.grad_order]
        owners = self.owners
        ldx = torch.cat((fc[:, 0] - c[owners, 0], fc[self.inter, 0] - c[self.ine, 0]))
        ldy = torch.cat((fc[:, 1] - c[owners, 1], fc[self.inter, 1] - c[self.ine, 1]))
        out["lim_dx"], out["lim_dy"] = ldx[self.lim_order], ldy[self.lim_order]
        return out


def geometry_vjp(dm, u: torch.Tensor, lam: torch.Tensor) -> dict:
    """lambda^T dR/d(geometry array) at device state u for every array in
    GEOMETRY_KEYS (device order)."""
    adjoint.enable(dm)
    L = dm.layout
    opts = dict(dtype=torch.float64, device=dm.device)
    face = torch.empty((L.n_faces, 7), **opts)
    grad = torch.empty((max(len(L.grad_nb), 1), 2), **opts)
    lim = torch.empty((max(len(L.lim_dx), 1), 2), **opts)
    u = u.contiguous()
    lam = lam.contiguous()
    native.check(dm.lib.mcfd_residual_geometry_vjp(dm.handle, C.byref(dm.native_settings), _p(u), _p(lam), _p(face),
                                                   _p(grad), _p(lim), dm.stream))
    r = dm.residual(u).clone()
    dm.check_status("geometry adjoint: ")
    s = dm.settings
    if s.acceleration_x != 0.0 or s.acceleration_y != 0.0:
        rho = u[:, 0]
        vx, vy = u[:, 1] / rho, u[:, 2] / rho
        r[:, 1] -= rho * s.acceleration_x
        r[:, 2] -= rho * s.acceleration_y
        r[:, 3] -= rho * (s.acceleration_x * vx + s.acceleration_y * vy)
    area = torch.as_tensor(L.area, **opts)
    out = {k: face[:, j] for j, k in enumerate(GEOMETRY_KEYS[:7])}
    out["area"] = -(lam * r).sum(1) / area
    out["grad_cx"], out["grad_cy"] = grad[: len(L.grad_nb), 0], grad[: len(L.grad_nb), 1]
    out["lim_dx"], out["lim_dy"] = lim[: len(L.lim_dx), 0], lim[: len(L.lim_dx), 1]
    return out


def node_gradient(tg: TorchGeometry, nodes: np.ndarray, bars: dict) -> np.ndarray:
    """sum over the geometry arrays of bar . d(array)/d(nodes)."""
    x = torch.tensor(nodes, dtype=torch.float64, device=tg.device, requires_grad=True)
    geo = tg(x)
    total = sum(torch.sum(geo[k] * bars[k].to(tg.device)) for k in GEOMETRY_KEYS)
    (gx,) = torch.autograd.grad(total, x)
    return gx.cpu().numpy()


def shape_gradient(dm, u, lam, node_map, parameter: float, tg: TorchGeometry | None = None) -> float:
    """lambda^T dR/d(parameter) for nodes = node_map(parameter) (a torch
    function of a 0-d float64 tensor), at device state u."""
    if tg is None:
        tg = TorchGeometry(dm.case, dm.geometry, dm.layout)
    bars = geometry_vjp(dm, u, lam)
    p = torch.tensor(parameter, dtype=torch.float64, requires_grad=True)
    nodes = node_map(p)
    xbar = node_gradient(tg, nodes.detach().numpy(), bars)
    (d,) = torch.autograd.grad(nodes, p, grad_outputs=torch.as_tensor(xbar))
    return float(d.item())


def check_supported(dm) -> None:
    s = dm.settings
    if s.flux not in ("Rusanov", "HLLC") or s.reconstruction == "MUSCL parabolic" or s.viscosity_model != "Inviscid":
        raise SolverError("Geometry gradients cover the inviscid Rusanov and HLLC residual (first order or MUSCL "
                          "linear).")
