"""The C-ABI library loads on a CPU-only host and exports every entry point
declared in include/mcfd.h; host-side layout invariants (no GPU calls)."""

import re
from pathlib import Path

import numpy as np
import pytest

from meshcfd_amd import meshgen, native
from meshcfd_amd.geometry import build_geometry
from meshcfd_amd.layout import ACC_BOUNDARY, ACC_NEIGHBOUR, ACC_OWNER, ACC_PERIODIC_B, build_layout

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "mcfd.h").read_text()
    return sorted(set(re.findall(r"^\s*int\s+(mcfd_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "mcfd_stage" in names and "mcfd_residual" in names and len(names) >= 10


def test_library_exports_every_declared_symbol():
    import meshcfd_amd.adjoint  # noqa: F401  (registers its entry points)
    import meshcfd_amd.geometry_adjoint  # noqa: F401
    import meshcfd_amd.implicit  # noqa: F401
    import meshcfd_amd.inflow  # noqa: F401
    import meshcfd_amd.viscous  # noqa: F401
    import meshcfd_amd.steady_adjoint  # noqa: F401

    lib = native.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert set(declared_symbols()) <= set(native.EXPORTS)
    assert lib.mcfd_version() == 1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(native, "_lib", None)
    monkeypatch.setenv("MCFD_LIB", str(tmp_path / "nope.so"))
    with pytest.raises(native.NativeError):
        native.load()
    monkeypatch.setattr(native, "_lib", None)


def test_ctypes_structs_match_header_layout():
    import ctypes as C

    # 8 int32 counts + 30 pointers
    assert C.sizeof(native.McfdMeshDesc) == 8 * 4 + 30 * 8
    assert C.sizeof(native.McfdSettings) == 4 * 4 + 10 * 8
    from meshcfd_amd.viscous import McfdViscousDesc

    assert C.sizeof(McfdViscousDesc) == 4 * 4 + 17 * 8
    from meshcfd_amd.inflow import McfdInflowDesc

    assert C.sizeof(McfdInflowDesc) == 6 * 4 + 8 * 8
    assert C.sizeof(native.McfdStage) == 8 + 7 * 8 + 5 * 8 + 2 * 8


@pytest.mark.parametrize("reorder", [False, True])
def test_layout_lists_follow_reference_order(reorder):
    case, _ = meshgen.vortex_workload(8, 1)
    g = build_geometry(case)
    L = build_layout(g, reorder=reorder)
    n = g.cell_count
    new_id = np.empty(n, dtype=np.int64)
    new_id[L.cell_order] = np.arange(n)
    # every interior face appears once as owner and once as neighbour item
    items = L.acc_item
    modes = items & 3
    assert (modes == ACC_OWNER).sum() == len(g.interior_faces)
    assert (modes == ACC_NEIGHBOUR).sum() == len(g.interior_faces)
    assert (modes == ACC_PERIODIC_B).sum() == len(g.periodic_faces_a)
    assert (modes == ACC_BOUNDARY).sum() == len(g.periodic_faces_a)
    # per cell: interior items precede boundary items
    for c in range(n):
        m = modes[L.acc_ptr[c]:L.acc_ptr[c + 1]]
        k = np.flatnonzero(m >= 2)
        assert k.size == 0 or np.all(m[k[0]:] >= 2)
    # LSQ rows hold the reference's entries of that cell, in entry order
    for c_old in range(0, n, 7):
        c = new_id[c_old]
        want = g.gradient_neighbours[g.gradient_sources == c_old]
        got = L.grad_nb[L.grad_ptr[c]:L.grad_ptr[c + 1]]
        assert np.array_equal(new_id[want], got)
    assert L.grad_ptr[-1] == len(g.gradient_sources)
    assert L.lim_ptr[-1] == len(g.limiter_cells)
    assert L.dts_ptr[-1] == len(g.owners) + len(g.interior_faces)
