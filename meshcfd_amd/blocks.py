"""Cell blocks with two-ring halos for the fused stage kernel.

The device mesh is cut into runs of consecutive cells in its
space-filling-curve order, each as long as possible up to ``B`` cells while
it touches at most 256 flux faces (one per thread of the workgroup). One
workgroup owns one block and evaluates a whole stage for it out of LDS:

    local cells = [own cells (B)] + [ring 1] + [ring 2]

* ring 1: cells outside the block that share a face (or a periodic pair) with
  an own cell — their gradients and limiters are needed for the own cells'
  face reconstructions;
* ring 2: cells outside block + ring 1 that neighbour a ring-1 cell — only
  their primitive states are needed (ring-1 gradients and min/max).

Per block the layout lists the local cells' global ids, every LSQ entry of an
own or ring-1 cell as a local neighbour id (in the reference's entry order),
the flux faces touching an own cell (interior faces between two blocks are
evaluated by both, with identical inputs, so both get the same bits), and the
accumulation items re-encoded as local face slots.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .layout import DeviceLayout

# cells_start, n_own, n_ring1, n_ring2, faces_start, n_faces, gnb_start,
# gptr_start, own_start, lent_start, (2 spare)
INFO_FIELDS = 12
FACE_RECORD = np.dtype([("geom", np.float64, 7), ("kind_bc", np.int32), ("lr", np.uint32)])  # 64 bytes
MAX_BLOCK_FACES = 256  # one flux face per thread of the 256-thread workgroup
MAX_BLOCK_CELLS = 128  # two threads per own cell in the gather/update phase


def partition(L: DeviceLayout, max_cells: int, max_faces: int = MAX_BLOCK_FACES) -> np.ndarray:
    """Block start offsets over the device cell order: each block takes as
    many consecutive cells as it can (at most ``max_cells``) while the flux
    faces touching its cells stay within ``max_faces``, so the stage kernel
    evaluates every face of a block in one pass of its threads.

    A face is counted for block [s, e) when one of its cells lies in [s, e);
    with ``key`` = its smallest cell index >= s, that is key < e, so the
    largest admissible e is the (max_faces+1)-th smallest key."""
    n = L.n_cells
    fl = L.face_left.astype(np.int64)
    fr = L.face_right.astype(np.int64)
    nf = len(fl)
    # incident faces per cell (CSR)
    has_r = fr >= 0
    ends = np.concatenate((fl, fr[has_r]))
    fids = np.concatenate((np.arange(nf), np.flatnonzero(has_r)))
    order = np.argsort(ends, kind="stable")
    inc_face = fids[order]
    inc_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(ends, minlength=n), out=inc_ptr[1:])
    other = np.where(fl[inc_face] == ends[order], fr[inc_face], fl[inc_face])  # the face's other cell
    starts = [0]
    s = 0
    while s < n:
        w = min(n, s + max_cells)
        a, b = inc_ptr[s], inc_ptr[w]
        cell = np.repeat(np.arange(s, w), np.diff(inc_ptr[s:w + 1]))
        oth = other[a:b]
        # a face whose other cell lies in [s, cell) was already counted at that cell
        first_visit = ~((oth >= s) & (oth < cell))
        keys = cell[first_visit]
        if len(keys) > max_faces:
            e = int(np.sort(keys)[max_faces])
        else:
            e = w
        if e <= s:
            raise ValueError(f"cell {s} alone touches more than {max_faces} flux faces")
        starts.append(e)
        s = e
    return np.asarray(starts, dtype=np.int64)


@dataclass
class BlockLayout:
    block_cells: int
    n_blocks: int
    info: np.ndarray  # (n_blocks, 8) int32
    cells: np.ndarray  # int32 global ids of local cells, per block
    gnb_ptr: np.ndarray  # int32, per block n_own + n_ring1 + 1 offsets into the block's LSQ entries
    gnb: np.ndarray  # uint16 local neighbour ids of the LSQ entries
    gent: np.ndarray  # float64 (n, 2) LSQ coefficients (cx, cy) of the entries, same order as gnb
    lptr: np.ndarray  # int32, per block n_own + n_ring1 + 1 offsets into the block's limiter entries
    lent: np.ndarray  # float64 (n, 2) limiter offsets (dx, dy) of own + ring-1 cells, per block
    faces: np.ndarray  # int32 global flux-face ids, per block
    frec: np.ndarray  # FACE_RECORD per block face: nx, ny, lx, ly, rx, ry, length; kind | bc << 4; l | r << 16
    acc_local: np.ndarray  # uint16, per global acc item: local face slot * 4 + mode
    dts_local: np.ndarray  # uint16, per global spectral-radius item: local face slot | 0x8000 (global item)
    max_local: int
    max_faces: int
    max_grad_cells: int
    # first-order viscous fused stages (build_blocks with viscous lists)
    vent: np.ndarray | None = None  # float64 (n, 2) WLSQ weights (wx, wy), same order as gnb
    vacc_local: np.ndarray | None = None  # uint16, per global viscous acc item: local face slot * 4 + mode

    def nbytes(self) -> int:
        return int(sum(v.nbytes for v in vars(self).values() if isinstance(v, np.ndarray)))


def _lookup(keys_sorted: np.ndarray, values: np.ndarray, query: np.ndarray) -> np.ndarray:
    pos = np.searchsorted(keys_sorted, query)
    if np.any(pos >= len(keys_sorted)) or np.any(keys_sorted[np.minimum(pos, len(keys_sorted) - 1)] != query):
        raise ValueError("block halo lookup failed")
    return values[pos]


def build_blocks(L: DeviceLayout, block_cells: int = 128, viscous: dict | None = None) -> BlockLayout:
    """Blocks of at most ``block_cells`` cells. With ``viscous`` (the device
    lists of `viscous.build_viscous`: wl_wx, wl_wy, va_ptr, va_item) also the
    first-order viscous fused stage's arrays: the WLSQ weights of every block
    LSQ entry (gent's order) and every viscous accumulation item as a local
    face slot * 4 + mode of its cell's block."""
    n = L.n_cells
    B = block_cells
    if B < 1:
        raise ValueError("block_cells must be positive")
    bstart = partition(L, min(B, MAX_BLOCK_CELLS))
    nb = len(bstart) - 1
    blk = np.repeat(np.arange(nb, dtype=np.int64), np.diff(bstart))
    gptr = L.grad_ptr.astype(np.int64)
    deg = np.diff(gptr)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    nbr = L.grad_nb.astype(np.int64)

    # ring 1: (block, cell) pairs for neighbours outside the block
    cross = blk[src] != blk[nbr]
    r1 = np.unique(blk[src[cross]] * n + nbr[cross])
    r1_b, r1_c = r1 // n, r1 % n
    # ring 2: neighbours of ring-1 cells outside the block and outside ring 1
    cnt = deg[r1_c]
    exp_b = np.repeat(r1_b, cnt)
    starts = gptr[r1_c]
    offs = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    exp_c = nbr[np.repeat(starts, cnt) + offs]
    cand = exp_b * n + exp_c
    keep = (blk[exp_c] != exp_b) & ~np.isin(cand, r1)
    r2 = np.unique(cand[keep])
    r2_b, r2_c = r2 // n, r2 % n

    n_own = np.bincount(blk, minlength=nb)
    n_r1 = np.bincount(r1_b, minlength=nb)
    n_r2 = np.bincount(r2_b, minlength=nb)
    n_loc = n_own + n_r1 + n_r2
    cstart = np.concatenate(([0], np.cumsum(n_loc)[:-1]))
    # local ids: own cells 0..n_own-1, ring 1 next, ring 2 last
    own_local = np.arange(n) - bstart[blk]
    r1_rank = np.arange(len(r1)) - np.concatenate(([0], np.cumsum(n_r1)[:-1]))[r1_b]
    r2_rank = np.arange(len(r2)) - np.concatenate(([0], np.cumsum(n_r2)[:-1]))[r2_b]
    r1_local = n_own[r1_b] + r1_rank
    r2_local = n_own[r2_b] + n_r1[r2_b] + r2_rank
    cells = np.empty(int(n_loc.sum()), dtype=np.int64)
    cells[cstart[blk] + own_local] = np.arange(n)
    cells[cstart[r1_b] + r1_local] = r1_c
    cells[cstart[r2_b] + r2_local] = r2_c
    # (block, global cell) -> local id lookup
    keys = np.concatenate((blk * n + np.arange(n), r1, r2))
    vals = np.concatenate((own_local, r1_local, r2_local))
    o = np.argsort(keys)
    keys, vals = keys[o], vals[o]

    # LSQ entries of own + ring-1 cells as local neighbour ids
    gc_b = np.concatenate((blk, r1_b))  # block of each gradient cell
    gc_c = np.concatenate((np.arange(n), r1_c))
    gc_local = np.concatenate((own_local, r1_local))
    o = np.lexsort((gc_local, gc_b))
    gc_b, gc_c, gc_local = gc_b[o], gc_c[o], gc_local[o]
    gcnt = deg[gc_c]
    g_b = np.repeat(gc_b, gcnt)
    g_off = np.arange(gcnt.sum()) - np.repeat(np.cumsum(gcnt) - gcnt, gcnt)
    g_nb = nbr[np.repeat(gptr[gc_c], gcnt) + g_off]
    gnb = _lookup(keys, vals, g_b * n + g_nb)
    if gnb.max(initial=0) >= 0xFFFF:
        raise ValueError("block too large for 16-bit local ids")
    g_idx = np.repeat(gptr[gc_c], gcnt) + g_off  # global LSQ entry of each block entry
    gent = np.column_stack((L.grad_cx[g_idx], L.grad_cy[g_idx]))
    n_gcells = n_own + n_r1
    # per block n_gcells + 1 offsets (block relative) into the entry lists
    gstart_cells = np.concatenate(([0], np.cumsum(n_gcells + 1)[:-1]))

    def block_offsets(counts):
        """Per gradient cell (sorted by block, local id) entry counts ->
        per block [0, c0, c0+c1, ...] and each block's first entry."""
        total = np.bincount(gc_b, weights=counts, minlength=nb).astype(np.int64)
        start = np.concatenate(([0], np.cumsum(total)[:-1]))
        cum = np.cumsum(counts) - counts  # global exclusive prefix
        rel = cum - start[gc_b]
        ptr = np.empty(int((n_gcells + 1).sum()), dtype=np.int64)
        slot = gstart_cells[gc_b] + (np.arange(len(gc_b)) - np.concatenate(([0], np.cumsum(n_gcells)[:-1]))[gc_b])
        ptr[slot] = rel
        ptr[gstart_cells + n_gcells] = total
        return ptr, start

    gnb_ptr, gnb_start = block_offsets(gcnt)
    lp = L.lim_ptr.astype(np.int64)
    lcnt = lp[gc_c + 1] - lp[gc_c]
    l_off = np.arange(lcnt.sum()) - np.repeat(np.cumsum(lcnt) - lcnt, lcnt)
    l_idx = np.repeat(lp[gc_c], lcnt) + l_off
    lent = np.column_stack((L.lim_dx[l_idx], L.lim_dy[l_idx]))
    lptr, lent_start = block_offsets(lcnt)

    # flux faces per block: the left cell's block, plus the right cell's block if different
    fl = L.face_left.astype(np.int64)
    fr = L.face_right.astype(np.int64)
    nf = len(fl)
    fb_l = blk[fl]
    has_r = fr >= 0
    fb_r = np.where(has_r, blk[np.maximum(fr, 0)], -1)
    second = has_r & (fb_r != fb_l)
    pf_b = np.concatenate((fb_l, fb_r[second]))
    pf_k = np.concatenate((np.arange(nf), np.flatnonzero(second)))
    o = np.lexsort((pf_k, pf_b))
    pf_b, pf_k = pf_b[o], pf_k[o]
    nfb = np.bincount(pf_b, minlength=nb)
    fstart = np.concatenate(([0], np.cumsum(nfb)[:-1]))
    slot = np.arange(len(pf_k)) - fstart[pf_b]
    lr = np.empty((len(pf_k), 2), dtype=np.int64)
    lr[:, 0] = _lookup(keys, vals, pf_b * n + fl[pf_k])
    rr = fr[pf_k]
    lr[:, 1] = 0xFFFF
    hr = rr >= 0
    lr[hr, 1] = _lookup(keys, vals, pf_b[hr] * n + rr[hr])
    # accumulation items -> local slots of the owning cell's block
    fkeys = pf_b * nf + pf_k
    fo = np.argsort(fkeys)
    fkeys_s, slots_s = fkeys[fo], slot[fo]
    aptr = L.acc_ptr.astype(np.int64)
    acell = np.repeat(np.arange(n), np.diff(aptr))
    item = L.acc_item.astype(np.int64)
    aslot = _lookup(fkeys_s, slots_s, blk[acell] * nf + (item >> 2))
    if aslot.max(initial=0) >= (1 << 14):
        raise ValueError("too many faces per block for 14-bit slots")
    acc_local = (aslot << 2) | (item & 3)
    # spectral-radius items -> local slots (bit 15: the item's own normal,
    # length or exterior state is read from the global item arrays)
    dptr = L.dts_ptr.astype(np.int64)
    dcell = np.repeat(np.arange(n), np.diff(dptr))
    dslot = _lookup(fkeys_s, slots_s, blk[dcell] * nf + L.dts_flux.astype(np.int64))
    if dslot.max(initial=0) >= (1 << 15):
        raise ValueError("too many faces per block for 15-bit slots")
    dts_local = dslot | np.where(L.dts_own, 0, 1 << 15)
    vent = vacc_local = None
    if viscous is not None:
        vent = np.column_stack((np.asarray(viscous["wl_wx"])[g_idx], np.asarray(viscous["wl_wy"])[g_idx]))
        vptr = np.asarray(viscous["va_ptr"], dtype=np.int64)
        vitem = np.asarray(viscous["va_item"], dtype=np.int64)
        vcell = np.repeat(np.arange(n), np.diff(vptr))
        vslot = _lookup(fkeys_s, slots_s, blk[vcell] * nf + (vitem >> 2))
        vacc_local = ((vslot << 2) | (vitem & 3)).astype(np.uint16)

    info = np.zeros((nb, INFO_FIELDS), dtype=np.int64)
    info[:, 0] = cstart
    info[:, 1] = n_own
    info[:, 2] = n_r1
    info[:, 3] = n_r2
    info[:, 4] = fstart
    info[:, 5] = nfb
    info[:, 6] = gnb_start
    info[:, 7] = gstart_cells
    info[:, 8] = bstart[:-1]
    info[:, 9] = lent_start
    frec = np.zeros(len(pf_k), dtype=FACE_RECORD)
    frec["geom"] = np.column_stack((L.face_nx[pf_k], L.face_ny[pf_k], L.face_lx[pf_k], L.face_ly[pf_k],
                                    L.face_rx[pf_k], L.face_ry[pf_k], L.face_len[pf_k]))
    frec["kind_bc"] = (L.face_kind[pf_k].astype(np.int64) + 16 * L.face_bc[pf_k].astype(np.int64)).astype(np.int32)
    frec["lr"] = (lr[:, 0] | (lr[:, 1] << 16)).astype(np.uint32)
    if nfb.max(initial=0) > MAX_BLOCK_FACES:
        raise ValueError("block partition left a block with too many faces")
    if n_gcells.max(initial=0) > MAX_BLOCK_FACES:
        raise ValueError("a block has more than 256 own + ring-1 cells; use smaller blocks")
    return BlockLayout(
        block_cells=B,
        n_blocks=nb,
        info=info.astype(np.int32),
        cells=cells.astype(np.int32),
        gnb_ptr=gnb_ptr.astype(np.int32),
        gnb=gnb.astype(np.uint16),
        gent=np.ascontiguousarray(gent),
        lptr=lptr.astype(np.int32),
        lent=np.ascontiguousarray(lent),
        faces=pf_k.astype(np.int32),
        frec=frec,
        acc_local=acc_local.astype(np.uint16),
        dts_local=dts_local.astype(np.uint16),
        vent=None if vent is None else np.ascontiguousarray(vent),
        vacc_local=vacc_local,
        max_local=int(n_loc.max()),
        max_faces=int(nfb.max()),
        max_grad_cells=int(n_gcells.max()),
    )
