"""Pins of the oracle's time integrators and LU-SGS matrix (runs on the CPU).

Each check fixes what the mathematics fixes, so that a slip in the oracle
(a mistyped coefficient digit, a dropped factor of one half, a missing or
extra boundary term) fails it:

* explicit integrators: on the linear problem dU/dt = lam U one step
  multiplies U by the method's amplification polynomial P(z), z = lam dt.
  SSP-RK2 and SSP-RK3 have P(z) = sum_{k<=s} z^k/k!. "Explicit RK4" is
  Spiteri and Ruuth's SSPRK(5,4) (SIAM J. Numer. Anal. 40 (2002) 469-491):
  its published Shu-Osher table is checked here against the fourth-order
  conditions and the published SSP coefficient 1.508, and the oracle's step
  must equal that table's P(z), evaluated by a generic Shu-Osher loop in
  extended precision, to 2e-17: a one-unit change in the 15th decimal of any
  coefficient moves P(z) by about 1e-15;
* the constant-rate local time step known answer of the reference
  (tests/test_solver.py:1289-1323): each cell advances by its own step, so
  the velocities become (0.01, 0.04) for every integrator;
* LU-SGS: on an axis-aligned grid with a uniform state the spectral radius
  of every face is |u.n| + c in closed form, so the scalar matrix S of
  `_implicit_preconditioner` (solver.py:1185-1301) is known entry by entry:
  off-diagonals -dt_i (|u.n| + c) L / (2 A_i) towards each face neighbour,
  diagonal 1 + dt_i sum_faces (|u.n| + c) L / (2 A_i) + dt_i sum_{faces
  without a neighbour} (|u.n| + c) L / A_i (periodic faces have no
  neighbour in the face arrays, so they count twice).
"""

from fractions import Fraction

import numpy as np
import pytest

from meshcfd_amd import meshgen
from meshcfd_amd.device import RK4_COEFFS
from meshcfd_amd.geometry import build_geometry
from meshcfd_amd.model import BoundaryCondition, Case, NumericalSettings
from oracle import implicit as oimp
from oracle.euler import primitive_to_conservative, sound_speed
from oracle.solver import Oracle

LD = np.longdouble

# Spiteri-Ruuth SSPRK(5,4), Shu-Osher form, as published:
# u_i = sum_j (alpha_ij u_j + beta_ij dt L(u_j)), u_0 = U^n, U^{n+1} = u_5.
SSPRK54_ALPHA = {
    (1, 0): "1",
    (2, 0): "0.444370493651235", (2, 1): "0.555629506348765",
    (3, 0): "0.620101851488403", (3, 2): "0.379898148511597",
    (4, 0): "0.178079954393132", (4, 3): "0.821920045606868",
    (5, 2): "0.517231671970585", (5, 3): "0.096059710526147", (5, 4): "0.386708617503269",
}
SSPRK54_BETA = {
    (1, 0): "0.391752226571890",
    (2, 1): "0.368410593050371",
    (3, 2): "0.251891774271694",
    (4, 3): "0.544974750228521",
    (5, 3): "0.063692468666290", (5, 4): "0.226007483236906",
}


def shu_osher_amplification(alpha, beta, z, stages=5, one=1):
    """u_i = sum_j (alpha_ij + beta_ij z) u_j over a Shu-Osher table; returns
    u_stages (the amplification factor). Works on Fractions, polynomials
    and extended floats alike."""
    u = [one]
    for i in range(1, stages + 1):
        acc = 0 * one
        for j in range(i):
            a = alpha.get((i, j))
            b = beta.get((i, j))
            if a is not None:
                acc = acc + a * u[j]
            if b is not None:
                acc = acc + b * z * u[j]
        u.append(acc)
    return u[stages]


def _poly_coefficients(alpha, beta):
    """Exact amplification polynomial coefficients of a Shu-Osher table with
    decimal-string entries (rational arithmetic)."""
    a = {k: Fraction(v) for k, v in alpha.items()}
    b = {k: Fraction(v) for k, v in beta.items()}

    class Poly:
        def __init__(self, c):
            self.c = list(c)

        def __add__(self, o):
            n = max(len(self.c), len(o.c))
            return Poly([(self.c[k] if k < len(self.c) else 0) + (o.c[k] if k < len(o.c) else 0) for k in range(n)])

        def __rmul__(self, s):
            if isinstance(s, Poly):
                out = [Fraction(0)] * (len(self.c) + len(s.c) - 1)
                for i, x in enumerate(s.c):
                    for j, y in enumerate(self.c):
                        out[i + j] += x * y
                return Poly(out)
            return Poly([s * x for x in self.c])

        __mul__ = __rmul__

    z = Poly([Fraction(0), Fraction(1)])
    p = shu_osher_amplification(a, b, z, one=Poly([Fraction(1)]))
    return p.c


def test_ssprk54_table_is_fourth_order_with_published_ssp_coefficient():
    c = _poly_coefficients(SSPRK54_ALPHA, SSPRK54_BETA)
    want = [1, 1, Fraction(1, 2), Fraction(1, 6), Fraction(1, 24)]
    for k in range(5):
        # the table is published to 15 decimals
        assert abs(float(c[k] - want[k])) < 5e-15, (k, float(c[k]))
    # fifth-order condition deliberately not met (a 5-stage SSP method is at most 4th order)
    assert abs(float(c[5]) - 1.0 / 120.0) > 1e-4
    ratios = [Fraction(SSPRK54_ALPHA[k]) / Fraction(v) for k, v in SSPRK54_BETA.items() if k in SSPRK54_ALPHA]
    assert abs(float(min(ratios)) - 1.508) < 1e-3
    # every stage's state weights are non-negative and (to the published digits) sum to one
    for i in range(1, 6):
        w = [Fraction(v) for (r, _), v in SSPRK54_ALPHA.items() if r == i]
        assert all(x >= 0 for x in w) and abs(float(sum(w)) - 1.0) < 2e-15


def test_device_ssprk54_coefficients_are_the_published_table():
    c = RK4_COEFFS
    a, b = SSPRK54_ALPHA, SSPRK54_BETA
    f = float
    assert c[0] == (f(b[1, 0]),)
    assert c[1] == (f(a[2, 0]), f(a[2, 1]), f(b[2, 1]))
    assert c[2] == (f(a[3, 0]), f(a[3, 2]), f(b[3, 2]))
    assert c[3] == (f(a[4, 0]), f(a[4, 3]), f(b[4, 3]))
    assert c[4] == (f(a[5, 2]), f(a[5, 3]), f(b[5, 3]), f(a[5, 4]), f(b[5, 4]))


def _linear_oracle(integrator, lam, dt):
    mesh = meshgen.rectangle_mesh(3, 3)
    numerics = NumericalSettings(time_integration=integrator, time_step_mode="Global fixed", fixed_time_step=dt,
                                 end_time=1e30, reconstruction="First order", flux="Rusanov")
    g = build_geometry(meshgen.periodic_case(mesh, numerics))
    o = Oracle(g, numerics)
    o.residual = lambda u: lam * u
    return o


def _linear_state():
    prim = np.array([[1.0, 0.3, -0.2, 1.0]] * 9)
    return primitive_to_conservative(prim, 1.4).astype(LD)


@pytest.mark.parametrize("z", [-0.05, -0.3, -0.7, 0.2])
def test_oracle_ssprk54_step_is_the_published_table(z):
    dt = 0.25
    lam = LD(z) / LD(dt)
    o = _linear_oracle("Explicit RK4", lam, dt)
    u = _linear_state()
    new, _, _, _ = o.step(u, 0.0)
    assert new.dtype == LD
    alpha = {k: LD(float(v)) for k, v in SSPRK54_ALPHA.items()}
    beta = {k: LD(float(v)) for k, v in SSPRK54_BETA.items()}
    want = shu_osher_amplification(alpha, beta, lam * LD(dt), one=LD(1))
    got = new / u
    assert np.max(np.abs(got - want)) < 2e-17, (np.max(np.abs(got - want)), want)
    # and P(z) agrees with exp(z) through z^4
    assert abs(float(want) - np.exp(z)) < abs(z) ** 5 / 60


@pytest.mark.parametrize("integrator,order", [("Explicit Euler", 1), ("Explicit RK2", 2), ("Explicit RK3", 3)])
@pytest.mark.parametrize("z", [-0.05, -0.3, -0.7, 0.2])
def test_oracle_ssp_steps_have_taylor_amplification(integrator, order, z):
    dt = 0.125
    lam = LD(z) / LD(dt)
    o = _linear_oracle(integrator, lam, dt)
    u = _linear_state()
    new, _, _, _ = o.step(u, 0.0)
    zz = lam * LD(dt)
    want = sum(zz**k / LD(np.prod(np.arange(1, k + 1), dtype=float)) for k in range(order + 1))
    got = new / u
    # SSP-RK3's weights 1/3 and 2/3 are rounded doubles (6e-17 short of one)
    assert np.max(np.abs(got - want)) < 1e-16, (integrator, np.max(np.abs(got - want)))


@pytest.mark.parametrize("integrator", ["Explicit Euler", "Explicit RK2", "Explicit RK3", "Explicit RK4"])
def test_oracle_local_steps_advance_each_cell_by_its_own_step(integrator):
    """Reference tests/test_solver.py:1289-1323 on the oracle's step."""
    mesh = meshgen.rectangle_mesh(2, 1, (0.0, 2.0, 0.0, 1.0))
    numerics = NumericalSettings(time_integration=integrator, time_step_mode="Local (steady)", max_iterations=1,
                                 residual_tolerance=1e-14)
    g = build_geometry(meshgen.periodic_case(mesh, numerics))
    o = Oracle(g, numerics)
    local = np.array((0.01, 0.02))
    rates = np.array(((0.0, 1.0, 0.0, 0.0), (0.0, 2.0, 0.0, 0.0)))
    o.stable_time_steps = lambda *_: local.copy()
    o.residual = lambda *_: rates.copy()
    u = primitive_to_conservative(np.array(((1.0, 0.0, 0.0, 1.0), (1.0, 0.0, 0.0, 1.0))), 1.4)
    new, dt, sizes, _ = o.step(u, 0.0)
    assert dt == pytest.approx(np.mean(local))
    np.testing.assert_allclose(new[:, 1] / new[:, 0], (0.01, 0.04), rtol=1e-14)


def _grid_case(nx, ny, bounds, kinds):
    mesh = meshgen.rectangle_mesh(nx, ny, bounds)
    numerics = NumericalSettings(reconstruction="First order", flux="Rusanov", time_integration="Implicit Euler")
    if kinds == "periodic":
        return meshgen.periodic_case(mesh, numerics), mesh
    from meshcfd_amd.model import PeriodicBoundaryPair

    x = mesh.points[:, 0]
    bcs = [BoundaryCondition(0, "bottom", kind="Wall"), BoundaryCondition(1, "right", kind="Periodic"),
           BoundaryCondition(2, "top", kind="Symmetry"), BoundaryCondition(3, "left", kind="Periodic")]
    pairs = [PeriodicBoundaryPair(1, 3, "x", float(x.max() - x.min()))]
    return Case(mesh=mesh, boundary_conditions=bcs, periodic_pairs=pairs, numerics=numerics), mesh


@pytest.mark.parametrize("kinds", ["periodic", "walls"])
def test_lusgs_matrix_entries_closed_form(kinds):
    nx, ny = 4, 3
    x0, x1, y0, y1 = 0.0, 2.0, 0.0, 0.9
    case, mesh = _grid_case(nx, ny, (x0, x1, y0, y1), kinds)
    g = build_geometry(case)
    gamma = 1.4
    rho, u, v, p = 1.3, 0.7, -0.25, 0.9
    prim = np.tile([rho, u, v, p], (g.cell_count, 1))
    dt = np.random.default_rng(4).uniform(0.02, 0.05, g.cell_count)
    S, diag = oimp.lusgs_matrix(prim, g, gamma, dt)

    c = float(sound_speed(prim[:1], gamma)[0])
    hx, hy = (x1 - x0) / nx, (y1 - y0) / ny
    area = hx * hy
    lam_x = abs(u) + c  # vertical faces (normal +-x), length hy
    lam_y = abs(v) + c  # horizontal faces (normal +-y), length hx
    ij = np.floor(g.centers / (hx, hy)).astype(int)
    cell_of = {(int(i), int(j)): k for k, (i, j) in enumerate(ij)}
    want = np.zeros((g.cell_count, g.cell_count))
    for k, (i, j) in enumerate(ij):
        want[k, k] = 1.0
        for di, dj, lam, length in ((1, 0, lam_x, hy), (-1, 0, lam_x, hy), (0, 1, lam_y, hx), (0, -1, lam_y, hx)):
            ni, nj = i + di, j + dj
            across_x = not 0 <= ni < nx
            across_y = not 0 <= nj < ny
            half = dt[k] * 0.5 * lam * length / area
            if across_y and kinds == "walls":  # wall / symmetry: boundary term only
                want[k, k] += dt[k] * lam * length / area
                continue
            other = cell_of[(ni % nx, nj % ny)]
            want[k, other] -= half
            want[k, k] += half
            if across_x or across_y:  # a periodic face: also counted as a face without a neighbour
                want[k, k] += dt[k] * lam * length / area
    got = S.toarray()
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(diag, np.diag(want), rtol=1e-13)
