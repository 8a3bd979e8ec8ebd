/*
 * ORACLE (test infrastructure only): single-threaded C stand-ins for the
 * reference's numba kernels, used to time the reference-equivalent CPU path
 * (the reference runs these loops compiled by numba, without parallel=True).
 *
 *   linear_gradients_and_factors  <- src/meshcfd/core/_kernels.py:239-273
 *                                     (with limited_gradient_factors, :9-67)
 *   reconstruct_hllc_flux         <- src/meshcfd/core/_kernels.py:70-160
 *
 * Same loops, same per-element arithmetic order; built with
 * -ffp-contract=off so results equal the NumPy restatement bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

void oracle_linear_gradients_and_factors(
    int64_t cells, const double *prim, int64_t entries, const int64_t *src, const int64_t *nbr,
    const double *cx, const double *cy, int64_t lim_entries, const int64_t *lim_cells,
    const double *lim_dx, const double *lim_dy, int code, double *grads, double *factors,
    double *mins, double *maxs) {
  memset(grads, 0, sizeof(double) * (size_t)cells * 8);
  for (int64_t e = 0; e < entries; ++e) {
    int64_t s = src[e], n = nbr[e];
    for (int v = 0; v < 4; ++v) {
      double d = prim[4 * n + v] - prim[4 * s + v];
      grads[8 * s + 2 * v] += cx[e] * d;
      grads[8 * s + 2 * v + 1] += cy[e] * d;
    }
  }
  /* limited_gradient_factors: local min/max over face neighbours, then phi */
  for (int64_t i = 0; i < cells * 4; ++i) factors[i] = 1.0;
  memcpy(mins, prim, sizeof(double) * (size_t)cells * 4);
  memcpy(maxs, prim, sizeof(double) * (size_t)cells * 4);
  for (int64_t e = 0; e < entries; ++e) {
    int64_t s = src[e], n = nbr[e];
    for (int v = 0; v < 4; ++v) {
      double val = prim[4 * n + v];
      if (val < mins[4 * s + v]) mins[4 * s + v] = val;
      if (val > maxs[4 * s + v]) maxs[4 * s + v] = val;
    }
  }
  for (int64_t e = 0; e < lim_entries; ++e) {
    int64_t c = lim_cells[e];
    double dx = lim_dx[e], dy = lim_dy[e];
    for (int v = 0; v < 4; ++v) {
      double w = prim[4 * c + v];
      double inc = grads[8 * c + 2 * v] * dx + grads[8 * c + 2 * v + 1] * dy;
      double aw = fabs(w);
      double tol = 1e-14 * (aw >= 1.0 ? aw : 1.0);
      double r = 1.0;
      if (inc > tol)
        r = (maxs[4 * c + v] - w) / inc;
      else if (inc < -tol)
        r = (mins[4 * c + v] - w) / inc;
      if (0.0 > r) r = 0.0;
      double f;
      if (code <= 1) {
        f = (1.0 < r) ? 1.0 : r;
      } else if (code == 2) {
        if (r > 0.0) {
          double t = 2.0 * r / (1.0 + r);
          f = (t < 1.0) ? t : 1.0;
        } else {
          f = 0.0;
        }
      } else {
        double eps = 1e-12;
        f = (r * r + 2.0 * r + eps) / (r * r + r + 2.0 + eps);
        if (1.0 < f) f = 1.0;
      }
      if (f < factors[4 * c + v]) factors[4 * c + v] = f;
    }
  }
}

void oracle_reconstruct_hllc_flux(int64_t faces, const double *prim, const double *grads,
                                  const double *factors, const int64_t *owners,
                                  const int64_t *neighbours, const double *oo, const double *no,
                                  const double *normals, double gamma, double *flux) {
  for (int64_t f = 0; f < faces; ++f) {
    int64_t o = owners[f], n = neighbours[f];
    double L[4], R[4];
    for (int v = 0; v < 4; ++v) {
      L[v] = prim[4 * o + v] +
             factors[4 * o + v] * (grads[8 * o + 2 * v] * oo[2 * f] + grads[8 * o + 2 * v + 1] * oo[2 * f + 1]);
      R[v] = prim[4 * n + v] +
             factors[4 * n + v] * (grads[8 * n + 2 * v] * no[2 * f] + grads[8 * n + 2 * v + 1] * no[2 * f + 1]);
    }
    double nx = normals[2 * f], ny = normals[2 * f + 1];
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
    double *out = flux + 4 * f;
    if (sl >= 0.0) {
      memcpy(out, fl, sizeof fl);
      continue;
    }
    if (sr <= 0.0) {
      memcpy(out, fr, sizeof fr);
      continue;
    }
    double rho, u, v, p, un, w, e;
    const double *st, *ph;
    if (star >= 0.0) {
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
}
