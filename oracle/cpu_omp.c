/*
 * ORACLE / CPU BASELINE (test infrastructure only): an OpenMP restatement of
 * the oracle's explicit step (oracle/solver.py: Oracle.residual,
 * stable_time_steps, step) for the four-quadrant Riemann workload's settings:
 * MUSCL linear with any of the four limiters, HLLC, interior and
 * transmissive faces, "Global adaptive" CFL steps, "Explicit RK3". It gives
 * the all-core CPU figure reported beside the one-core reference-equivalent
 * path; the product never calls it.
 *
 * Every sum runs per cell over a CSR list built in the order the oracle
 * accumulates (entry order of the gradient loop, scipy's CSR order of the
 * divergence matrix, np.add.at's face order), and the per-element
 * arithmetic is the oracle's, so with -ffp-contract=off its steps equal the
 * oracle's bit for bit (tests/test_oracle_c.py) for any thread count.
 *
 *   gradients, limiter factors  <- src/meshcfd/core/_kernels.py:239-273, :9-67
 *   interior HLLC (fused order) <- src/meshcfd/core/_kernels.py:70-160
 *   boundary HLLC (vectorised)  <- src/meshcfd/core/euler.py:96-166
 *   residual assembly           <- src/meshcfd/core/solver.py:2437-2560
 *   CFL steps                   <- src/meshcfd/core/solver.py:4117-4137
 *   SSP-RK3                     <- src/meshcfd/core/solver.py:1095-1126
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

typedef struct {
  int64_t n, nf;  /* cells, interior faces */
  double gamma, cfl;
  int limiter;    /* LIMITER_CODES */
  const double *area;
  const int64_t *gptr, *gnb; /* gradient entries per cell, entry order */
  const double *gcx, *gcy;
  const int64_t *lptr;       /* limiter offsets per cell */
  const double *ldx, *ldy;
  const int64_t *fo, *fn;    /* interior faces */
  const double *foo, *fno, *fnrm;
  const int64_t *dptr, *dcol; /* divergence matrix, CSR (cells x interior faces) */
  const double *dval;
  const int64_t *tptr;       /* transmissive faces per owner cell, face order */
  const double *toff, *tnrm, *tlen;
  const int64_t *sptr;       /* time-step faces per cell: owned faces, then interior faces as neighbour */
  const double *snrm, *slen;
  double *prim, *grads, *factors, *flux, *d0, *d1, *stage;
} OmpCase;

static int to_primitive(const OmpCase *c, const double *u) {
  int bad = 0;
  const double g = c->gamma;
#pragma omp parallel for reduction(| : bad) schedule(static)
  for (int64_t i = 0; i < c->n; ++i) {
    const double *q = u + 4 * i;
    double rho = q[0];
    double uu = q[1] / rho, vv = q[2] / rho;
    double p = (g - 1.0) * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / rho);
    double *w = c->prim + 4 * i;
    w[0] = rho;
    w[1] = uu;
    w[2] = vv;
    w[3] = p;
    if (!(isfinite(rho) && isfinite(uu) && isfinite(vv) && isfinite(p)) || rho <= 0.0 || p <= 0.0) bad = 1;
  }
  return bad;
}

static double limiter_value(double r, int code) {
  if (code <= 1) return (1.0 < r) ? 1.0 : r;
  if (code == 2) {
    if (r > 0.0) {
      double t = 2.0 * r / (1.0 + r);
      return (t < 1.0) ? t : 1.0;
    }
    return 0.0;
  }
  double eps = 1e-12;
  double f = (r * r + 2.0 * r + eps) / (r * r + r + 2.0 + eps);
  return (1.0 < f) ? 1.0 : f;
}

static void gradients_and_factors(const OmpCase *c) {
  const double *W = c->prim;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c->n; ++i) {
    double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double lo[4], hi[4], f[4] = {1.0, 1.0, 1.0, 1.0};
    for (int v = 0; v < 4; ++v) lo[v] = hi[v] = W[4 * i + v];
    for (int64_t e = c->gptr[i]; e < c->gptr[i + 1]; ++e) {
      const int64_t m = c->gnb[e];
      for (int v = 0; v < 4; ++v) {
        double d = W[4 * m + v] - W[4 * i + v];
        g[2 * v] += c->gcx[e] * d;
        g[2 * v + 1] += c->gcy[e] * d;
        double val = W[4 * m + v];
        if (val < lo[v]) lo[v] = val;
        if (val > hi[v]) hi[v] = val;
      }
    }
    for (int64_t e = c->lptr[i]; e < c->lptr[i + 1]; ++e) {
      for (int v = 0; v < 4; ++v) {
        double w = W[4 * i + v];
        double inc = g[2 * v] * c->ldx[e] + g[2 * v + 1] * c->ldy[e];
        double aw = fabs(w);
        double tol = 1e-14 * (aw >= 1.0 ? aw : 1.0);
        double r = 1.0;
        if (inc > tol)
          r = (hi[v] - w) / inc;
        else if (inc < -tol)
          r = (lo[v] - w) / inc;
        if (0.0 > r) r = 0.0;
        double phi = limiter_value(r, c->limiter);
        if (phi < f[v]) f[v] = phi;
      }
    }
    memcpy(c->grads + 8 * i, g, sizeof g);
    memcpy(c->factors + 4 * i, f, sizeof f);
  }
}

static void reconstruct(const OmpCase *c, int64_t cell, double dx, double dy, double out[4]) {
  for (int v = 0; v < 4; ++v)
    out[v] = c->prim[4 * cell + v] +
             c->factors[4 * cell + v] * (c->grads[8 * cell + 2 * v] * dx + c->grads[8 * cell + 2 * v + 1] * dy);
}

/* HLLC with Davis speeds; fused: branch order S_L >= 0, S_R <= 0, S* >= 0 (the
 * compiled interior kernel), else S_L >= 0, S* >= 0, S_R > 0 (vectorised). */
static void hllc(const double L[4], const double R[4], double nx, double ny, double gamma, int fused,
                 double out[4]) {
  double rl = L[0], ul = L[1], vl = L[2], pl = L[3];
  double rr = R[0], ur = R[1], vr = R[2], pr = R[3];
  double unl = ul * nx + vl * ny, unr = ur * nx + vr * ny;
  double cl = sqrt(gamma * pl / rl), cr = sqrt(gamma * pr / rr);
  double a = unl - cl, b = unr - cr;
  double sl = (b < a) ? b : a;
  a = unl + cl;
  b = unr + cr;
  double sr = (b > a) ? b : a;
  double el = pl / (gamma - 1.0) + 0.5 * rl * (ul * ul + vl * vl);
  double er = pr / (gamma - 1.0) + 0.5 * rr * (ur * ur + vr * vr);
  double ul4[4] = {rl, rl * ul, rl * vl, el};
  double ur4[4] = {rr, rr * ur, rr * vr, er};
  double fl[4] = {rl * unl, rl * ul * unl + pl * nx, rl * vl * unl + pl * ny, (el + pl) * unl};
  double fr[4] = {rr * unr, rr * ur * unr + pr * nx, rr * vr * unr + pr * ny, (er + pr) * unr};
  double den = rl * (sl - unl) - rr * (sr - unr);
  if (fabs(den) <= 1e-14) den = copysign(1e-14, den);
  double star = (pr - pl + rl * unl * (sl - unl) - rr * unr * (sr - unr)) / den;
  int side; /* 0: F_L, 1: F_R, 2: left star, 3: right star */
  if (sl >= 0.0)
    side = 0;
  else if (fused)
    side = (sr <= 0.0) ? 1 : (star >= 0.0 ? 2 : 3);
  else
    side = (star >= 0.0) ? 2 : (sr > 0.0 ? 3 : 1);
  if (side < 2) {
    memcpy(out, side == 0 ? fl : fr, sizeof fl);
    return;
  }
  double rho, u, v, p, un, w, e;
  const double *st, *ph;
  if (side == 2) {
    rho = rl; u = ul; v = vl; p = pl; un = unl; w = sl; e = el; st = ul4; ph = fl;
  } else {
    rho = rr; u = ur; v = vr; p = pr; un = unr; w = sr; e = er; st = ur4; ph = fr;
  }
  double factor = rho * (w - un) / (w - star);
  double sx = u + (star - un) * nx;
  double sy = v + (star - un) * ny;
  double se = factor * (e / rho + (star - un) * (star + p / (rho * (w - un))));
  double q[4] = {factor, factor * sx, factor * sy, se};
  for (int k = 0; k < 4; ++k) out[k] = ph[k] + w * (q[k] - st[k]);
}

static void residual(const OmpCase *c, double *deriv) {
  gradients_and_factors(c);
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < c->nf; ++f) {
    double L[4], R[4];
    reconstruct(c, c->fo[f], c->foo[2 * f], c->foo[2 * f + 1], L);
    reconstruct(c, c->fn[f], c->fno[2 * f], c->fno[2 * f + 1], R);
    hllc(L, R, c->fnrm[2 * f], c->fnrm[2 * f + 1], c->gamma, 1, c->flux + 4 * f);
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < c->n; ++i) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t j = c->dptr[i]; j < c->dptr[i + 1]; ++j) {
      const double a = c->dval[j];
      const double *x = c->flux + 4 * c->dcol[j];
      for (int k = 0; k < 4; ++k) acc[k] += a * x[k];
    }
    for (int64_t t = c->tptr[i]; t < c->tptr[i + 1]; ++t) {
      double W[4], F[4];
      reconstruct(c, i, c->toff[2 * t], c->toff[2 * t + 1], W);
      hllc(W, W, c->tnrm[2 * t], c->tnrm[2 * t + 1], c->gamma, 0, F);
      for (int k = 0; k < 4; ++k) acc[k] += (-(F[k] * c->tlen[t])) / c->area[i];
    }
    memcpy(deriv + 4 * i, acc, sizeof acc);
  }
}

static double min_time_step(const OmpCase *c) {
  double best = INFINITY;
  const double g = c->gamma;
#pragma omp parallel for reduction(min : best) schedule(static)
  for (int64_t i = 0; i < c->n; ++i) {
    const double *w = c->prim + 4 * i;
    const double cs = sqrt(g * w[3] / w[0]);
    double total = 0.0;
    for (int64_t e = c->sptr[i]; e < c->sptr[i + 1]; ++e) {
      double un = w[1] * c->snrm[2 * e] + w[2] * c->snrm[2 * e + 1];
      total += (fabs(un) + cs) * c->slen[e];
    }
    double h = c->cfl * c->area[i] / total;
    if (h < best) best = h;
  }
  return best;
}

/* ``steps`` SSP-RK3 steps in place on u (n x 4 conservative); *time advances.
 * Returns 0, or 1 on a non-physical state, 2 on a non-positive step. */
int omp_rk3_steps(OmpCase *c, double *u, int steps, double end_time, double *time) {
  const int64_t N = 4 * c->n;
  for (int s = 0; s < steps; ++s) {
    if (to_primitive(c, u)) return 1;
    double dt = min_time_step(c);
    if (end_time - *time < dt) dt = end_time - *time;
    if (!(dt > 0.0) || !isfinite(dt)) return 2;
    residual(c, c->d0);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < N; ++k) c->stage[k] = u[k] + dt * c->d0[k];
    if (to_primitive(c, c->stage)) return 1;
    residual(c, c->d1);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < N; ++k) c->stage[k] = 0.75 * u[k] + 0.25 * (c->stage[k] + dt * c->d1[k]);
    if (to_primitive(c, c->stage)) return 1;
    residual(c, c->d1);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < N; ++k) u[k] = (1.0 / 3.0) * u[k] + (2.0 / 3.0) * (c->stage[k] + dt * c->d1[k]);
    *time += dt;
  }
  return to_primitive(c, u);
}

/* Threads of the parallel loops (the process's cores). */
void omp_set_threads(int n) { omp_set_num_threads(n); }

int omp_threads(void) { return omp_get_max_threads(); }

/* The residual alone (for tests): deriv = R(u). */
int omp_residual(OmpCase *c, const double *u, double *deriv) {
  if (to_primitive(c, u)) return 1;
  residual(c, deriv);
  return 0;
}
