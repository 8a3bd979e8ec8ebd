// mcfd.hip — gfx950 kernels and C-ABI of the residual hot path.
//
// Everything is FP64. Arithmetic inside each kernel follows the operation
// order of the reference solver (src/meshcfd/core/euler.py, _kernels.py,
// solver.py) and the library is built with -ffp-contract=off, so a residual
// reproduces the CPU oracle bit for bit. Face contributions are summed by a
// per-cell CSR gather in the reference's summation order: no atomics on
// data, so reruns are bit-identical.
//
// Per residual evaluation:
//   k_primitive   cell   U -> W (+ state validity)
//   k_gradient    cell   LSQ gradient + limiter factor (MUSCL linear only)
//   k_flux        face   reconstruction + ghost state + numerical flux
//   k_update      cell   CSR gather of face fluxes, body force, stage update
// plus k_time_steps (cell) + k_min_reduce (one block) for the CFL step.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mcfd.h"

namespace {

thread_local std::string g_last_error;

int fail(const std::string& msg) {
  g_last_error = msg;
  return -1;
}

#define HIP_CHECK(expr)                                                        \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess)                                                      \
      return fail(std::string(#expr) + ": " + hipGetErrorString(_e));        \
  } while (0)

constexpr int kBlock = 256;

struct P4 {
  double r, u, v, p;
};
struct F4 {
  double a, b, c, d;
};

__device__ __forceinline__ P4 load4(const double* __restrict__ base, long i) {
  const double2* q = reinterpret_cast<const double2*>(base + 4 * i);
  double2 x = q[0];
  double2 y = q[1];
  return P4{x.x, x.y, y.x, y.y};
}

__device__ __forceinline__ void store4(double* __restrict__ base, long i, double a, double b,
                                       double c, double d) {
  double2* q = reinterpret_cast<double2*>(base + 4 * i);
  q[0] = make_double2(a, b);
  q[1] = make_double2(c, d);
}

__device__ __forceinline__ void report(int* status, int code, int index) {
  atomicOr(status, 1 << code);
  atomicMin(status + 1, index);
}

__device__ __forceinline__ bool finite(double x) { return isfinite(x); }

// numpy.minimum / maximum for non-NaN operands.
__device__ __forceinline__ double vmin(double a, double b) { return (a <= b) ? a : b; }
__device__ __forceinline__ double vmax(double a, double b) { return (a >= b) ? a : b; }

// euler.py:8-17 energy, euler.py:43-57 physical flux.
struct Side {
  double rho, u, v, p, un, E;
  double U[4];
  double F[4];
};

__device__ __forceinline__ void side_state(const P4& w, double nx, double ny, double g, Side& s) {
  s.rho = w.r;
  s.u = w.u;
  s.v = w.v;
  s.p = w.p;
  s.un = w.u * nx + w.v * ny;
  s.E = w.p / (g - 1.0) + 0.5 * w.r * (w.u * w.u + w.v * w.v);
  s.U[0] = w.r;
  s.U[1] = w.r * w.u;
  s.U[2] = w.r * w.v;
  s.U[3] = s.E;
  s.F[0] = w.r * s.un;
  s.F[1] = w.r * w.u * s.un + w.p * nx;
  s.F[2] = w.r * w.v * s.un + w.p * ny;
  s.F[3] = (s.E + w.p) * s.un;
}

__device__ __forceinline__ double sound(const P4& w, double g) { return sqrt(g * w.p / w.r); }

// euler.py:80-93
__device__ F4 rusanov(const P4& L, const P4& R, double nx, double ny, double g) {
  Side a, b;
  side_state(L, nx, ny, g, a);
  side_state(R, nx, ny, g, b);
  double s = vmax(fabs(a.un) + sound(L, g), fabs(b.un) + sound(R, g));
  F4 f;
  f.a = 0.5 * (a.F[0] + b.F[0] - s * (b.U[0] - a.U[0]));
  f.b = 0.5 * (a.F[1] + b.F[1] - s * (b.U[1] - a.U[1]));
  f.c = 0.5 * (a.F[2] + b.F[2] - s * (b.U[2] - a.U[2]));
  f.d = 0.5 * (a.F[3] + b.F[3] - s * (b.U[3] - a.U[3]));
  return f;
}

// HLLC star-region flux for one side (euler.py:122-154, _kernels.py:151-159).
__device__ __forceinline__ F4 hllc_star(const Side& s, double wave, double star, double nx,
                                        double ny) {
  double factor = s.rho * (wave - s.un) / (wave - star);
  double vx = s.u + (star - s.un) * nx;
  double vy = s.v + (star - s.un) * ny;
  double energy = factor * (s.E / s.rho + (star - s.un) * (star + s.p / (s.rho * (wave - s.un))));
  F4 f;
  f.a = s.F[0] + wave * (factor - s.U[0]);
  f.b = s.F[1] + wave * (factor * vx - s.U[1]);
  f.c = s.F[2] + wave * (factor * vy - s.U[2]);
  f.d = s.F[3] + wave * (energy - s.U[3]);
  return f;
}

// HLLC with Davis wave speeds (euler.py:96-166). fused_order: branch order of
// the compiled interior-face kernel (_kernels.py:136-159).
__device__ F4 hllc(const P4& L, const P4& R, double nx, double ny, double g, bool fused_order) {
  Side a, b;
  side_state(L, nx, ny, g, a);
  side_state(R, nx, ny, g, b);
  double cl = sound(L, g), cr = sound(R, g);
  double sl = vmin(a.un - cl, b.un - cr);
  double sr = vmax(a.un + cl, b.un + cr);
  double den = a.rho * (sl - a.un) - b.rho * (sr - b.un);
  if (!(fabs(den) > 1e-14)) den = copysign(1e-14, den);
  double star = (b.p - a.p + a.rho * a.un * (sl - a.un) - b.rho * b.un * (sr - b.un)) / den;
  F4 f;
  if (sl >= 0.0) {
    f = F4{a.F[0], a.F[1], a.F[2], a.F[3]};
  } else if (fused_order) {
    if (sr <= 0.0)
      f = F4{b.F[0], b.F[1], b.F[2], b.F[3]};
    else if (star >= 0.0)
      f = hllc_star(a, sl, star, nx, ny);
    else
      f = hllc_star(b, sr, star, nx, ny);
  } else {
    if (star >= 0.0)
      f = hllc_star(a, sl, star, nx, ny);
    else if (sr > 0.0)
      f = hllc_star(b, sr, star, nx, ny);
    else
      f = F4{b.F[0], b.F[1], b.F[2], b.F[3]};
  }
  return f;
}

// Roe flux with Harten entropy fix (euler.py:169-262). The "total enthalpy"
// average uses p/rho + |u|^2/2 per side, as the reference does (euler.py:181-185).
__device__ F4 roe(const P4& L, const P4& R, double nx, double ny, double g, int* err) {
  Side a, b;
  side_state(L, nx, ny, g, a);
  side_state(R, nx, ny, g, b);
  double sl = sqrt(L.r), sr = sqrt(R.r);
  double den = sl + sr;
  double u = (sl * L.u + sr * R.u) / den;
  double v = (sl * L.v + sr * R.v) / den;
  double h = (sl * (L.p / L.r + 0.5 * (L.u * L.u + L.v * L.v)) +
              sr * (R.p / R.r + 0.5 * (R.u * R.u + R.v * R.v))) /
             den;
  double un = u * nx + v * ny;
  double ut = u * -ny + v * nx;
  double a2 = (g - 1.0) * (h - 0.5 * (u * u + v * v));
  if (!(a2 > 0.0) || !finite(a2)) *err = MCFD_ERR_ROE_SOUND;
  double c = sqrt(a2);
  double rho_avg = sl * sr;
  double unl = a.un, unr = b.un;
  double utl = L.u * -ny + L.v * nx;
  double utr = R.u * -ny + R.v * nx;
  double drho = R.r - L.r;
  double dun = unr - unl;
  double dut = utr - utl;
  double dp = R.p - L.p;
  double am = (dp - rho_avg * c * dun) / (2.0 * a2);
  double ap = (dp + rho_avg * c * dun) / (2.0 * a2);
  double ac = drho - dp / a2;
  double as = rho_avg * dut;
  double fix = 0.1 * c;
  double l0 = fabs(un - c), l1 = fabs(un), l3 = fabs(un + c);
  if (l0 < fix) l0 = 0.5 * (l0 * l0 / fix + fix);
  if (l1 < fix) l1 = 0.5 * (l1 * l1 / fix + fix);
  if (l3 < fix) l3 = 0.5 * (l3 * l3 / fix + fix);
  double w0 = l0 * am, w1 = l1 * ac, w2 = l1 * as, w3 = l3 * ap;
  double d0 = w0 + w1 + 0.0 * w2 + w3;
  double d1 = (un - c) * w0 + un * w1 + 0.0 * w2 + (un + c) * w3;
  double d2 = ut * w0 + ut * w1 + w2 + ut * w3;
  double d3 = (h - un * c) * w0 + 0.5 * (un * un + ut * ut) * w1 + ut * w2 + (h + un * c) * w3;
  F4 f;
  f.a = 0.5 * (a.F[0] + b.F[0] - d0);
  f.b = 0.5 * (a.F[1] + b.F[1] - (d1 * nx - d2 * ny));
  f.c = 0.5 * (a.F[2] + b.F[2] - (d1 * ny + d2 * nx));
  f.d = 0.5 * (a.F[3] + b.F[3] - d3);
  return f;
}

// Characteristic far-field ghost state (solver.py:4350-4409).
__device__ P4 far_field(const P4& in, const P4& out, double nx, double ny, double g, int* err) {
  double un_i = in.u * nx + in.v * ny;
  double un_f = out.u * nx + out.v * ny;
  double c_i = sound(in, g);
  double c_f = sound(out, g);
  if (un_i <= -c_i) return out;
  if (un_i >= c_i) return in;
  double r_out = un_i + 2.0 * c_i / (g - 1.0);
  double r_in = un_f - 2.0 * c_f / (g - 1.0);
  double vb = 0.5 * (r_out + r_in);
  double cb = 0.25 * (g - 1.0) * (r_out - r_in);
  if (!finite(cb) || !(cb > 0.0)) {
    *err = MCFD_ERR_FAR_SOUND;
    return in;
  }
  bool inflow = vb < 0.0;
  const P4& src = inflow ? out : in;
  double entropy = src.p / pow(src.r, g);
  double rho = pow(cb * cb / (g * entropy), 1.0 / (g - 1.0));
  double p = entropy * pow(rho, g);
  double tx = -ny, ty = nx;
  double vt = src.u * tx + src.v * ty;
  P4 e{rho, vb * nx + vt * tx, vb * ny + vt * ty, p};
  if (!finite(e.r) || !finite(e.u) || !finite(e.v) || !finite(e.p) || !(e.r > 0.0) ||
      !(e.p > 0.0))
    *err = MCFD_ERR_FAR_STATE;
  return e;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// euler.py:20-40
__global__ void __launch_bounds__(kBlock)
    k_primitive(int n, const double* __restrict__ U, double* __restrict__ W, double g,
                int* __restrict__ status) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  P4 q = load4(U, i);
  double rho = q.r;
  double u = q.u / rho;
  double v = q.v / rho;
  double p = (g - 1.0) * (q.p - 0.5 * (q.u * q.u + q.v * q.v) / rho);
  if (!finite(rho) || !finite(u) || !finite(v) || !finite(p) || !(rho > 0.0) || !(p > 0.0))
    report(status, MCFD_ERR_STATE, i);
  store4(W, i, rho, u, v, p);
}

// LSQ gradient (_kernels.py:252-262) + limiter factor (_kernels.py:9-67).
__global__ void __launch_bounds__(kBlock)
    k_gradient(int n, const double* __restrict__ W, const int* __restrict__ gptr,
               const int* __restrict__ gnb, const double* __restrict__ gcx,
               const double* __restrict__ gcy, const int* __restrict__ lptr,
               const double* __restrict__ ldx, const double* __restrict__ ldy, int code,
               double* __restrict__ G, double* __restrict__ PHI) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  P4 w = load4(W, i);
  double wv[4] = {w.r, w.u, w.v, w.p};
  double gx[4] = {0.0, 0.0, 0.0, 0.0};
  double gy[4] = {0.0, 0.0, 0.0, 0.0};
  double mn[4] = {wv[0], wv[1], wv[2], wv[3]};
  double mx[4] = {wv[0], wv[1], wv[2], wv[3]};
  int e0 = gptr[i], e1 = gptr[i + 1];
  for (int e = e0; e < e1; ++e) {
    P4 o = load4(W, gnb[e]);
    double ov[4] = {o.r, o.u, o.v, o.p};
    double cx = gcx[e], cy = gcy[e];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double d = ov[k] - wv[k];
      gx[k] = gx[k] + cx * d;
      gy[k] = gy[k] + cy * d;
      if (ov[k] < mn[k]) mn[k] = ov[k];
      if (ov[k] > mx[k]) mx[k] = ov[k];
    }
  }
  double phi[4] = {1.0, 1.0, 1.0, 1.0};
  int l0 = lptr[i], l1 = lptr[i + 1];
  for (int e = l0; e < l1; ++e) {
    double dx = ldx[e], dy = ldy[e];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double inc = gx[k] * dx + gy[k] * dy;
      double tol = 1e-14 * vmax(fabs(wv[k]), 1.0);
      double ratio = 1.0;
      if (inc > tol)
        ratio = (mx[k] - wv[k]) / inc;
      else if (inc < -tol)
        ratio = (mn[k] - wv[k]) / inc;
      ratio = (0.0 > ratio) ? 0.0 : ratio;
      double f;
      if (code <= 1) {
        f = (ratio < 1.0) ? ratio : 1.0;
      } else if (code == 2) {
        if (ratio > 0.0) {
          double t = 2.0 * ratio / (1.0 + ratio);
          f = (t < 1.0) ? t : 1.0;
        } else {
          f = 0.0;
        }
      } else {
        const double eps = 1e-12;
        f = (ratio * ratio + 2.0 * ratio + eps) / (ratio * ratio + ratio + 2.0 + eps);
        f = (f < 1.0) ? f : 1.0;
      }
      if (f < phi[k]) phi[k] = f;
    }
  }
  double2* gq = reinterpret_cast<double2*>(G + 8 * (long)i);
  gq[0] = make_double2(gx[0], gy[0]);
  gq[1] = make_double2(gx[1], gy[1]);
  gq[2] = make_double2(gx[2], gy[2]);
  gq[3] = make_double2(gx[3], gy[3]);
  store4(PHI, i, phi[0], phi[1], phi[2], phi[3]);
}

// W_f = W + phi*(g_x*dx + g_y*dy) (_kernels.py:90-98, solver.py:2418-2434).
__device__ __forceinline__ P4 reconstruct(const double* __restrict__ W,
                                          const double* __restrict__ G,
                                          const double* __restrict__ PHI, int c, double dx,
                                          double dy, bool muscl) {
  P4 w = load4(W, c);
  if (!muscl) return w;
  const double2* gq = reinterpret_cast<const double2*>(G + 8 * (long)c);
  double2 g0 = gq[0], g1 = gq[1], g2 = gq[2], g3 = gq[3];
  P4 f = load4(PHI, c);
  P4 o;
  o.r = w.r + f.r * (g0.x * dx + g0.y * dy);
  o.u = w.u + f.u * (g1.x * dx + g1.y * dy);
  o.v = w.v + f.v * (g2.x * dx + g2.y * dy);
  o.p = w.p + f.p * (g3.x * dx + g3.y * dy);
  return o;
}

struct FaceArrays {
  const int* left;
  const int* right;
  const int* kind;
  const int* bc;
  const double* nx;
  const double* ny;
  const double* lx;
  const double* ly;
  const double* rx;
  const double* ry;
  const double* bc_state;
  const double* bc_wall;
};

__global__ void __launch_bounds__(kBlock)
    k_flux(int nf, FaceArrays fa, const double* __restrict__ W, const double* __restrict__ G,
           const double* __restrict__ PHI, int scheme, int muscl, double g,
           double* __restrict__ FLUX, int* __restrict__ status) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  int kind = fa.kind[k];
  double nx = fa.nx[k], ny = fa.ny[k];
  P4 L = reconstruct(W, G, PHI, fa.left[k], fa.lx[k], fa.ly[k], muscl);
  P4 R;
  int err = 0;
  if (kind <= MCFD_FACE_PERIODIC) {
    R = reconstruct(W, G, PHI, fa.right[k], fa.rx[k], fa.ry[k], muscl);
  } else {
    int b = fa.bc[k];
    const double* st = fa.bc_state + 4 * (long)b;
    if (kind == MCFD_FACE_WALL) {
      double rx = L.u - fa.bc_wall[2 * (long)b];
      double ry = L.v - fa.bc_wall[2 * (long)b + 1];
      double un = rx * nx + ry * ny;
      R = L;
      R.u = L.u - 2.0 * un * nx;
      R.v = L.v - 2.0 * un * ny;
    } else if (kind == MCFD_FACE_VELOCITY_INLET) {
      R = P4{st[0], st[1], st[2], st[3]};
    } else if (kind == MCFD_FACE_PRESSURE_OUTLET) {
      R = L;
      R.p = st[3];
    } else if (kind == MCFD_FACE_FAR_FIELD) {
      R = far_field(L, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, &err);
    } else {
      R = L;
    }
  }
  F4 f;
  if (scheme == 1) {
    f = hllc(L, R, nx, ny, g, kind == MCFD_FACE_INTERIOR);
    if (kind != MCFD_FACE_INTERIOR &&
        !(finite(f.a) && finite(f.b) && finite(f.c) && finite(f.d)))
      err = MCFD_ERR_HLLC;
  } else if (scheme == 2) {
    f = roe(L, R, nx, ny, g, &err);
    if (!err && !(finite(f.a) && finite(f.b) && finite(f.c) && finite(f.d))) err = MCFD_ERR_ROE;
  } else {
    f = rusanov(L, R, nx, ny, g);
  }
  if (err) report(status, err, k);
  store4(FLUX, k, f.a, f.b, f.c, f.d);
}

struct StageArgs {
  int op;
  int dt_per_cell;
  const double* u;
  const double* u0;
  const double* u1;
  const double* r1;
  const double* dt;
  double* out;
  double* r_out;
  double c0, c1, c2, c3, c4;
};

// Face -> cell gather in the reference order (solver.py:2501-2704), body force
// (solver.py:2709-2715), finiteness check (:2716-2717), then the stage update.
__global__ void __launch_bounds__(kBlock)
    k_update(int n, const int* __restrict__ aptr, const int* __restrict__ aitem,
             const double* __restrict__ FLUX, const double* __restrict__ flen,
             const double* __restrict__ area, const double* __restrict__ W, double ax, double ay,
             StageArgs s, int* __restrict__ status) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double A = area[i];
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
  int e0 = aptr[i], e1 = aptr[i + 1];
  for (int e = e0; e < e1; ++e) {
    int item = aitem[e];
    int f = item >> 2;
    int mode = item & 3;
    double L = flen[f];
    P4 F = load4(FLUX, f);
    if (mode == MCFD_ACC_OWNER) {
      double c = -L / A;
      r0 = r0 + c * F.r;
      r1 = r1 + c * F.u;
      r2 = r2 + c * F.v;
      r3 = r3 + c * F.p;
    } else if (mode == MCFD_ACC_NEIGHBOUR) {
      double c = L / A;
      r0 = r0 + c * F.r;
      r1 = r1 + c * F.u;
      r2 = r2 + c * F.v;
      r3 = r3 + c * F.p;
    } else if (mode == MCFD_ACC_BOUNDARY) {
      r0 = r0 + (-(F.r * L)) / A;
      r1 = r1 + (-(F.u * L)) / A;
      r2 = r2 + (-(F.v * L)) / A;
      r3 = r3 + (-(F.p * L)) / A;
    } else {
      r0 = r0 + (F.r * L) / A;
      r1 = r1 + (F.u * L) / A;
      r2 = r2 + (F.v * L) / A;
      r3 = r3 + (F.p * L) / A;
    }
  }
  if (ax != 0.0 || ay != 0.0) {
    P4 w = load4(W, i);
    r1 = r1 + w.r * ax;
    r2 = r2 + w.r * ay;
    r3 = r3 + w.r * (ax * w.u + ay * w.v);
  }
  if (!(finite(r0) && finite(r1) && finite(r2) && finite(r3))) report(status, MCFD_ERR_RESIDUAL, i);
  if (s.r_out) store4(s.r_out, i, r0, r1, r2, r3);
  if (s.op == MCFD_OP_RESIDUAL) return;
  double h = s.dt_per_cell ? s.dt[i] : s.dt[0];
  P4 u = load4(s.u, i);
  double o0, o1, o2, o3;
  if (s.op == MCFD_OP_EULER) {
    double k = s.c0 * h;
    o0 = u.r + k * r0;
    o1 = u.u + k * r1;
    o2 = u.v + k * r2;
    o3 = u.p + k * r3;
  } else if (s.op == MCFD_OP_SSP) {
    P4 z = load4(s.u0, i);
    o0 = s.c0 * z.r + s.c1 * (u.r + h * r0);
    o1 = s.c0 * z.u + s.c1 * (u.u + h * r1);
    o2 = s.c0 * z.v + s.c1 * (u.v + h * r2);
    o3 = s.c0 * z.p + s.c1 * (u.p + h * r3);
  } else if (s.op == MCFD_OP_RK4) {
    P4 z = load4(s.u0, i);
    double k = s.c2 * h;
    o0 = s.c0 * z.r + s.c1 * u.r + k * r0;
    o1 = s.c0 * z.u + s.c1 * u.u + k * r1;
    o2 = s.c0 * z.v + s.c1 * u.v + k * r2;
    o3 = s.c0 * z.p + s.c1 * u.p + k * r3;
  } else {
    P4 z = load4(s.u0, i);
    P4 y = load4(s.u1, i);
    P4 q = load4(s.r1, i);
    double k2 = s.c2 * h, k4 = s.c4 * h;
    o0 = s.c0 * z.r + s.c1 * y.r + k2 * q.r + s.c3 * u.r + k4 * r0;
    o1 = s.c0 * z.u + s.c1 * y.u + k2 * q.u + s.c3 * u.u + k4 * r1;
    o2 = s.c0 * z.v + s.c1 * y.v + k2 * q.v + s.c3 * u.v + k4 * r2;
    o3 = s.c0 * z.p + s.c1 * y.p + k2 * q.p + s.c3 * u.p + k4 * r3;
  }
  store4(s.out, i, o0, o1, o2, o3);
}

// Spectral-radius sum and local step (solver.py:4132-4343, inviscid part).
__global__ void __launch_bounds__(kBlock)
    k_time_steps(int n, const double* __restrict__ W, const int* __restrict__ dptr,
                 const double* __restrict__ dnx, const double* __restrict__ dny,
                 const double* __restrict__ dlen, const int* __restrict__ dbc,
                 const int* __restrict__ bc_kind, const double* __restrict__ bc_state,
                 const double* __restrict__ area, double g, double cfl,
                 double* __restrict__ dt_cell, double* __restrict__ partial,
                 int* __restrict__ status) {
  __shared__ double red[kBlock];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double dt = INFINITY;
  if (i < n) {
    P4 w = load4(W, i);
    double c = sound(w, g);
    double sum = 0.0;
    int e0 = dptr[i], e1 = dptr[i + 1];
    for (int e = e0; e < e1; ++e) {
      double nx = dnx[e], ny = dny[e];
      double rad = fabs(w.u * nx + w.v * ny) + c;
      int b = dbc[e];
      if (b >= 0) {
        int kind = bc_kind[b];
        const double* st = bc_state + 4 * (long)b;
        P4 x = w;
        int err = 0;
        if (kind == MCFD_FACE_VELOCITY_INLET)
          x = P4{st[0], st[1], st[2], st[3]};
        else if (kind == MCFD_FACE_PRESSURE_OUTLET)
          x.p = st[3];
        else if (kind == MCFD_FACE_FAR_FIELD)
          x = far_field(w, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, &err);
        if (err) report(status, err, i);
        double xr = fabs(x.u * nx + x.v * ny) + sound(x, g);
        rad = vmax(rad, xr);
      }
      sum = sum + rad * dlen[e];
    }
    dt = cfl * area[i] / sum;
    if (!finite(dt) || !(dt > 0.0)) report(status, MCFD_ERR_LOCAL_DT, i);
    dt_cell[i] = dt;
  }
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(1024)
    k_min_reduce(int n, const double* __restrict__ partial, double* __restrict__ out) {
  __shared__ double red[1024];
  double m = INFINITY;
  for (int j = threadIdx.x; j < n; j += blockDim.x) m = fmin(m, partial[j]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_clock(const double* __restrict__ src, double fixed_dt, double end_time,
                        double* __restrict__ clock, int* __restrict__ status) {
  if (threadIdx.x != 0) return;
  double t = clock[0];
  double v = src ? src[0] : fixed_dt;
  double rest = end_time - t;
  double dt = (rest < v) ? rest : v;
  if (!finite(dt) || !(dt > 0.0)) report(status, MCFD_ERR_GLOBAL_DT, 0);
  clock[1] = dt;
  clock[0] = t + dt;
}

// Per-component sum of squares: block partials in a fixed order, then one
// block combines them in a fixed order (bit-reproducible).
__global__ void __launch_bounds__(kBlock)
    k_sumsq(int n, const double* __restrict__ R, double* __restrict__ partial) {
  __shared__ double red[4][kBlock];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  P4 r = (i < n) ? load4(R, i) : P4{0.0, 0.0, 0.0, 0.0};
  red[0][threadIdx.x] = r.r * r.r;
  red[1][threadIdx.x] = r.u * r.u;
  red[2][threadIdx.x] = r.v * r.v;
  red[3][threadIdx.x] = r.p * r.p;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[4 * (long)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256)
    k_sumsq_final(int nb, const double* __restrict__ partial, double* __restrict__ out) {
  __shared__ double red[4][256];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = threadIdx.x; j < nb; j += blockDim.x)
    for (int k = 0; k < 4; ++k) acc[k] += partial[4 * (long)j + k];
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = red[threadIdx.x][0];
}

// Plain streaming copy of n doubles: the reference point for HBM counters.
__global__ void __launch_bounds__(kBlock)
    k_copy(long n2, const double2* __restrict__ src, double2* __restrict__ dst) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long stride = (long)gridDim.x * blockDim.x;
  for (; i < n2; i += stride) dst[i] = src[i];
}

__global__ void k_status_init(int* status) {
  if (threadIdx.x == 0) {
    status[0] = 0;
    status[1] = 0x7fffffff;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Mesh handle
// ---------------------------------------------------------------------------

struct McfdMesh {
  int n_cells = 0, n_faces = 0, n_grad = 0, n_lim = 0, n_acc = 0, n_dts = 0, n_bc = 0;
  int device = 0;
  std::vector<void*> allocations;
  double* area = nullptr;
  int* grad_ptr = nullptr;
  int* grad_nb = nullptr;
  double* grad_cx = nullptr;
  double* grad_cy = nullptr;
  int* lim_ptr = nullptr;
  double* lim_dx = nullptr;
  double* lim_dy = nullptr;
  FaceArrays faces{};
  double* face_len = nullptr;
  int* acc_ptr = nullptr;
  int* acc_item = nullptr;
  int* dts_ptr = nullptr;
  double* dts_nx = nullptr;
  double* dts_ny = nullptr;
  double* dts_len = nullptr;
  int* dts_bc = nullptr;
  int* bc_kind = nullptr;
  // scratch
  double* W = nullptr;
  double* G = nullptr;
  double* PHI = nullptr;
  double* FLUX = nullptr;
  double* partial = nullptr;
  int* status = nullptr;
  // timing
  bool timing = false;
  std::vector<hipEvent_t> event_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> timed;
  size_t next_event = 0;
};

namespace {

template <typename T>
int upload(McfdMesh* m, T** dst, const T* src, size_t count) {
  size_t bytes = sizeof(T) * (count ? count : 1);
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, bytes));
  m->allocations.push_back(p);
  if (count) {
    if (!src) return fail("mcfd_mesh_create: missing array");
    HIP_CHECK(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
  }
  *dst = static_cast<T*>(p);
  return 0;
}

template <typename T>
int scratch(McfdMesh* m, T** dst, size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
  m->allocations.push_back(p);
  *dst = static_cast<T*>(p);
  return 0;
}

void free_mesh(McfdMesh* m) {
  for (void* p : m->allocations) (void)hipFree(p);
  for (hipEvent_t e : m->event_pool) (void)hipEventDestroy(e);
  delete m;
}

int blocks(int n) { return (n + kBlock - 1) / kBlock; }

struct TimedLaunch {
  McfdMesh* m;
  hipStream_t s;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  TimedLaunch(McfdMesh* mesh, hipStream_t stream, int k) : m(mesh), s(stream), kind(k) {
    if (m->timing && m->next_event + 2 <= m->event_pool.size()) {
      a = m->event_pool[m->next_event++];
      b = m->event_pool[m->next_event++];
      (void)hipEventRecord(a, s);
    }
  }
  ~TimedLaunch() {
    if (a) {
      (void)hipEventRecord(b, s);
      m->timed.push_back({kind, {a, b}});
    }
  }
};

int check_settings(const McfdSettings* s) {
  if (!s) return fail("settings is NULL");
  if (s->flux < 0 || s->flux > 2) return fail("Unsupported convective flux code.");
  if (s->reconstruction < 0 || s->reconstruction > 1)
    return fail("Unsupported reconstruction code for the GPU residual.");
  if (s->limiter < 0 || s->limiter > 3) return fail("Unsupported limiter code.");
  return 0;
}

int residual_front(McfdMesh* m, const McfdSettings* st, const double* u, hipStream_t s) {
  int n = m->n_cells;
  {
    TimedLaunch t(m, s, MCFD_KERNEL_PRIMITIVE);
    hipLaunchKernelGGL(k_primitive, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, m->W, st->gamma,
                       m->status);
  }
  bool muscl = st->reconstruction == 1 && m->n_grad > 0;
  if (muscl) {
    TimedLaunch t(m, s, MCFD_KERNEL_GRADIENT);
    hipLaunchKernelGGL(k_gradient, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->grad_ptr,
                       m->grad_nb, m->grad_cx, m->grad_cy, m->lim_ptr, m->lim_dx, m->lim_dy,
                       st->limiter, m->G, m->PHI);
  }
  if (m->n_faces) {
    TimedLaunch t(m, s, MCFD_KERNEL_FLUX);
    hipLaunchKernelGGL(k_flux, dim3(blocks(m->n_faces)), dim3(kBlock), 0, s, m->n_faces,
                       m->faces, m->W, m->G, m->PHI, st->flux, muscl ? 1 : 0, st->gamma,
                       m->FLUX, m->status);
  }
  HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int mcfd_version(void) { return MCFD_ABI_VERSION; }

int mcfd_last_error(char* buffer, size_t size) {
  if (!buffer || size == 0) return -1;
  std::snprintf(buffer, size, "%s", g_last_error.c_str());
  return 0;
}

int mcfd_mesh_create(const McfdMeshDesc* d, McfdMesh** out) {
  if (!d || !out) return fail("mcfd_mesh_create: NULL argument");
  if (d->n_cells <= 0) return fail("mcfd_mesh_create: mesh has no cells");
  McfdMesh* m = new McfdMesh();
  m->n_cells = d->n_cells;
  m->n_faces = d->n_faces;
  m->n_grad = d->n_grad;
  m->n_lim = d->n_lim;
  m->n_acc = d->n_acc;
  m->n_dts = d->n_dts;
  m->n_bc = d->n_bc;
  (void)hipGetDevice(&m->device);
  size_t C = d->n_cells, F = d->n_faces;
  int rc = 0;
  FaceArrays& fa = m->faces;
  rc |= upload(m, &m->area, d->area, C);
  rc |= upload(m, &m->grad_ptr, d->grad_ptr, C + 1);
  rc |= upload(m, &m->grad_nb, d->grad_nb, d->n_grad);
  rc |= upload(m, &m->grad_cx, d->grad_cx, d->n_grad);
  rc |= upload(m, &m->grad_cy, d->grad_cy, d->n_grad);
  rc |= upload(m, &m->lim_ptr, d->lim_ptr, C + 1);
  rc |= upload(m, &m->lim_dx, d->lim_dx, d->n_lim);
  rc |= upload(m, &m->lim_dy, d->lim_dy, d->n_lim);
  rc |= upload(m, const_cast<int**>(&fa.left), d->face_left, F);
  rc |= upload(m, const_cast<int**>(&fa.right), d->face_right, F);
  rc |= upload(m, const_cast<int**>(&fa.kind), d->face_kind, F);
  rc |= upload(m, const_cast<int**>(&fa.bc), d->face_bc, F);
  rc |= upload(m, const_cast<double**>(&fa.nx), d->face_nx, F);
  rc |= upload(m, const_cast<double**>(&fa.ny), d->face_ny, F);
  rc |= upload(m, &m->face_len, d->face_len, F);
  rc |= upload(m, const_cast<double**>(&fa.lx), d->face_lx, F);
  rc |= upload(m, const_cast<double**>(&fa.ly), d->face_ly, F);
  rc |= upload(m, const_cast<double**>(&fa.rx), d->face_rx, F);
  rc |= upload(m, const_cast<double**>(&fa.ry), d->face_ry, F);
  rc |= upload(m, &m->acc_ptr, d->acc_ptr, C + 1);
  rc |= upload(m, &m->acc_item, d->acc_item, d->n_acc);
  rc |= upload(m, &m->dts_ptr, d->dts_ptr, C + 1);
  rc |= upload(m, &m->dts_nx, d->dts_nx, d->n_dts);
  rc |= upload(m, &m->dts_ny, d->dts_ny, d->n_dts);
  rc |= upload(m, &m->dts_len, d->dts_len, d->n_dts);
  rc |= upload(m, &m->dts_bc, d->dts_bc, d->n_dts);
  rc |= upload(m, &m->bc_kind, d->bc_kind, d->n_bc);
  rc |= upload(m, const_cast<double**>(&fa.bc_state), d->bc_state, 4 * (size_t)d->n_bc);
  rc |= upload(m, const_cast<double**>(&fa.bc_wall), d->bc_wall, 2 * (size_t)d->n_bc);
  rc |= scratch(m, &m->W, 4 * C);
  rc |= scratch(m, &m->G, 8 * C);
  rc |= scratch(m, &m->PHI, 4 * C);
  rc |= scratch(m, &m->FLUX, 4 * F);
  rc |= scratch(m, &m->partial, 4 * (size_t)blocks(d->n_cells) + 4);
  rc |= scratch(m, &m->status, 2);
  if (rc) {
    std::string msg = g_last_error;
    free_mesh(m);
    return fail(msg);
  }
  hipLaunchKernelGGL(k_status_init, dim3(1), dim3(64), 0, 0, m->status);
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    free_mesh(m);
    return fail(std::string("mcfd_mesh_create: ") + hipGetErrorString(e));
  }
  *out = m;
  return 0;
}

int mcfd_mesh_destroy(McfdMesh* m) {
  if (!m) return 0;
  (void)hipDeviceSynchronize();
  free_mesh(m);
  return 0;
}

int mcfd_residual(McfdMesh* m, const McfdSettings* st, const double* u, double* r,
                  void* stream) {
  McfdStage stage{};
  stage.op = MCFD_OP_RESIDUAL;
  stage.u = u;
  stage.r_out = r;
  return mcfd_stage(m, st, &stage, stream);
}

int mcfd_stage(McfdMesh* m, const McfdSettings* st, const McfdStage* sg, void* stream) {
  if (!m || !sg) return fail("mcfd_stage: NULL argument");
  if (check_settings(st)) return -1;
  if (!sg->u) return fail("mcfd_stage: state is NULL");
  if (sg->op < 0 || sg->op > MCFD_OP_RESIDUAL) return fail("mcfd_stage: unknown op");
  if (sg->op != MCFD_OP_RESIDUAL && (!sg->out || !sg->dt))
    return fail("mcfd_stage: output state and dt are required");
  if ((sg->op == MCFD_OP_SSP || sg->op == MCFD_OP_RK4 || sg->op == MCFD_OP_RK4_FINAL) && !sg->u0)
    return fail("mcfd_stage: u0 is required for this op");
  if (sg->op == MCFD_OP_RK4_FINAL && (!sg->u1 || !sg->r1))
    return fail("mcfd_stage: u1 and r1 are required for the final SSPRK(5,4) combination");
  if (sg->op == MCFD_OP_RESIDUAL && !sg->r_out) return fail("mcfd_residual: output is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (residual_front(m, st, sg->u, s)) return -1;
  StageArgs a{sg->op, sg->dt_per_cell, sg->u, sg->u0, sg->u1, sg->r1, sg->dt, sg->out,
              sg->r_out, sg->c[0], sg->c[1], sg->c[2], sg->c[3], sg->c[4]};
  {
    TimedLaunch t(m, s, MCFD_KERNEL_UPDATE);
    hipLaunchKernelGGL(k_update, dim3(blocks(m->n_cells)), dim3(kBlock), 0, s, m->n_cells,
                       m->acc_ptr, m->acc_item, m->FLUX, m->face_len, m->area, m->W,
                       st->acceleration_x, st->acceleration_y, a, m->status);
  }
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_time_steps(McfdMesh* m, const McfdSettings* st, const double* u, double* dt_cell,
                    double* dt_min, void* stream) {
  if (!m || !u || !dt_cell) return fail("mcfd_time_steps: NULL argument");
  if (check_settings(st)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int n = m->n_cells;
  {
    TimedLaunch t(m, s, MCFD_KERNEL_PRIMITIVE);
    hipLaunchKernelGGL(k_primitive, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, m->W, st->gamma,
                       m->status);
  }
  {
    TimedLaunch t(m, s, MCFD_KERNEL_TIME_STEP);
    hipLaunchKernelGGL(k_time_steps, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->dts_ptr,
                       m->dts_nx, m->dts_ny, m->dts_len, m->dts_bc, m->bc_kind,
                       m->faces.bc_state, m->area, st->gamma, st->cfl, dt_cell, m->partial,
                       m->status);
  }
  if (dt_min) {
    TimedLaunch t(m, s, MCFD_KERNEL_REDUCE);
    hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(1024), 0, s, blocks(n), m->partial, dt_min);
  }
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_advance_clock(McfdMesh* m, const double* dt_source, double fixed_dt, double end_time,
                       double* clock, void* stream) {
  if (!m || !clock) return fail("mcfd_advance_clock: NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, s, dt_source, fixed_dt, end_time, clock,
                     m->status);
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_sum_squares(McfdMesh* m, const double* r, double* sums, void* stream) {
  if (!m || !r || !sums) return fail("mcfd_sum_squares: NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int nb = blocks(m->n_cells);
  hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(kBlock), 0, s, m->n_cells, r, m->partial);
  hipLaunchKernelGGL(k_sumsq_final, dim3(1), dim3(256), 0, s, nb, m->partial, sums);
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_copy(const double* src, double* dst, size_t count, void* stream) {
  if (!src || !dst || (count & 1)) return fail("mcfd_copy: NULL pointer or odd count");
  hipStream_t s = static_cast<hipStream_t>(stream);
  long n2 = (long)(count / 2);
  int grid = 256 * 8 * 4;
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, s, n2,
                     reinterpret_cast<const double2*>(src), reinterpret_cast<double2*>(dst));
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_status(McfdMesh* m, void* stream, int32_t* mask, int32_t* index) {
  if (!m) return fail("mcfd_status: NULL mesh");
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_CHECK(hipStreamSynchronize(s));
  int host[2] = {0, 0};
  HIP_CHECK(hipMemcpy(host, m->status, sizeof(host), hipMemcpyDeviceToHost));
  if (host[0]) {
    hipLaunchKernelGGL(k_status_init, dim3(1), dim3(64), 0, s, m->status);
    HIP_CHECK(hipStreamSynchronize(s));
  }
  if (mask) *mask = host[0];
  if (index) *index = host[0] ? host[1] : -1;
  return 0;
}

int mcfd_timing(McfdMesh* m, int32_t enable, double* total_ms, int32_t* counts) {
  if (!m) return fail("mcfd_timing: NULL mesh");
  if (total_ms)
    for (int k = 0; k < MCFD_KERNEL_KINDS; ++k) total_ms[k] = 0.0;
  if (counts)
    for (int k = 0; k < MCFD_KERNEL_KINDS; ++k) counts[k] = 0;
  for (auto& t : m->timed) {
    HIP_CHECK(hipEventSynchronize(t.second.second));
    float ms = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&ms, t.second.first, t.second.second));
    if (t.first >= 0 && t.first < MCFD_KERNEL_KINDS) {
      if (total_ms) total_ms[t.first] += ms;
      if (counts) counts[t.first] += 1;
    }
  }
  if (enable) {
    if (m->event_pool.empty()) {
      m->event_pool.resize(16384);
      for (auto& e : m->event_pool) HIP_CHECK(hipEventCreate(&e));
    }
    m->timed.clear();
    m->next_event = 0;
    m->timing = true;
  } else {
    m->timing = false;
  }
  return 0;
}

}  // extern "C"
