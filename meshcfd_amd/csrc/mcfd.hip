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
#include <algorithm>
#include <mutex>
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
constexpr int kBlockInfo = 12;  // ints per fused-kernel block record (McfdBlockDesc.info)
// ints per block record on the device: the 12 of McfdBlockDesc.info (field 0
// rewritten to the block's fixed-stride offset into the local-cell list), then
// the block's LSQ and limiter entry counts and the first item and count of
// its own cells' accumulation and spectral-radius lists, so that no load of
// the fused kernel waits for another load to learn its address
constexpr int kBlockInfoDev = 16;

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

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8). Remap so that each XCD walks one contiguous range of blocks:
// with cells in space-filling-curve order, an XCD's L2 then holds one compact
// region of the mesh and the neighbour gathers hit in it. Bijective for any
// grid size.
__device__ __forceinline__ int xcd_block() {
  int nb = gridDim.x;
  int b = blockIdx.x;
  int q = nb >> 3, r = nb & 7;
  int x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

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
  // Branch choice first (both orders of euler.py:96-166 / _kernels.py:136-159),
  // then one evaluation on the chosen side: lanes of a wave that take the
  // left and the right star state share one pass.
  bool left, use_star;
  if (sl >= 0.0) {
    left = true;
    use_star = false;
  } else if (fused_order) {
    left = !(sr <= 0.0) && star >= 0.0;
    use_star = !(sr <= 0.0);
  } else {
    left = star >= 0.0;
    use_star = left || sr > 0.0;
  }
  const Side& s = left ? a : b;
  F4 f;
  if (use_star)
    f = hllc_star(s, left ? sl : sr, star, nx, ny);
  else
    f = F4{s.F[0], s.F[1], s.F[2], s.F[3]};
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

// The validity test of euler.py:20-40 alone (no primitive states written).
__global__ void __launch_bounds__(kBlock)
    k_check_state(int n, const double* __restrict__ U, double g, int* __restrict__ status) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  P4 q = load4(U, i);
  double rho = q.r;
  double u = q.u / rho;
  double v = q.v / rho;
  double p = (g - 1.0) * (q.p - 0.5 * (q.u * q.u + q.v * q.v) / rho);
  if (!finite(rho) || !finite(u) || !finite(v) || !finite(p) || !(rho > 0.0) || !(p > 0.0))
    report(status, MCFD_ERR_STATE, i);
}

// euler.py:20-40
__global__ void __launch_bounds__(kBlock)
    k_primitive(int n, const double* __restrict__ U, double* __restrict__ W, double g,
                int* __restrict__ status) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
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


// One limiter entry (_kernels.py:36-65): phi_k = min(phi_k, f(r)). For
// Barth-Jespersen / Minmod (f = min(r, 1)) an entry whose ratio provably is
// >= phi_k cannot lower phi_k and its division is skipped:
// |num| > fl(phi*|inc|)*(1+1e-14) implies fl(num/inc) > phi. The phi values
// are identical to evaluating every entry.
__device__ __forceinline__ void limiter_entry(double inc, double wk, double mxk, double mnk, int code,
                                              double& phik) {
  double tol = 1e-14 * vmax(fabs(wk), 1.0);
  double num;
  if (inc > tol) {
    num = mxk - wk;
  } else if (inc < -tol) {
    num = mnk - wk;
  } else {
    if (code <= 2) return;  // ratio = 1: f = 1 for these limiters
    const double eps = 1e-12;
    double f = (1.0 * 1.0 + 2.0 * 1.0 + eps) / (1.0 * 1.0 + 1.0 + 2.0 + eps);
    f = (f < 1.0) ? f : 1.0;
    if (f < phik) phik = f;
    return;
  }
  if (code <= 1) {
    bool same = (num > 0.0 && inc > 0.0) || (num < 0.0 && inc < 0.0);
    if (same && fabs(num) > phik * fabs(inc) * (1.0 + 1e-14)) return;
  }
  double ratio = num / inc;
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
  if (f < phik) phik = f;
}

// limiter_entry with the cell's per-variable constants hoisted (dmx = max - w,
// dmn = min - w, tol = 1e-14 max(|w|, 1): the same operations, so the same
// bits) and its branches turned into selects; only the division stays behind
// a branch, which a wave skips when none of its lanes needs it.
template <int CODE>
__device__ __forceinline__ void limiter_entry_sel(double inc, double dmx, double dmn, double tol,
                                                  double& phik) {
  const bool up = inc > tol, down = inc < -tol;
  const double num = up ? dmx : dmn;
  bool need = up || down;
  if constexpr (CODE == 3) {
    // |inc| <= tol: ratio 1
    const double eps = 1e-12;
    double f1 = (1.0 * 1.0 + 2.0 * 1.0 + eps) / (1.0 * 1.0 + 1.0 + 2.0 + eps);
    f1 = (f1 < 1.0) ? f1 : 1.0;
    if (!need && f1 < phik) phik = f1;
  }
  if constexpr (CODE <= 1) {
    const bool same = (num > 0.0 && inc > 0.0) || (num < 0.0 && inc < 0.0);
    need = need && !(same && fabs(num) > phik * fabs(inc) * (1.0 + 1e-14));
  }
  if (need) {
    double ratio = num / inc;
    ratio = (0.0 > ratio) ? 0.0 : ratio;
    double f;
    if constexpr (CODE <= 1) {
      f = (ratio < 1.0) ? ratio : 1.0;
    } else if constexpr (CODE == 2) {
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
    if (f < phik) phik = f;
  }
}

// LSQ gradient (_kernels.py:252-262) + limiter factor (_kernels.py:9-67).
// With ARGS (for the adjoint) also per variable k: ARG[4i+k] = limiting
// limiter entry | max entry << 8 | min entry << 16, entries counted from the
// cell's first, 0xFF = none (the limiter stays 1) / the cell itself.
template <bool ARGS>
__global__ void __launch_bounds__(kBlock)
    k_gradient_t(int n, const double* __restrict__ W, const int* __restrict__ gptr,
                 const int* __restrict__ gnb, const double* __restrict__ gcx,
                 const double* __restrict__ gcy, const int* __restrict__ lptr,
                 const double* __restrict__ ldx, const double* __restrict__ ldy, int code,
                 double* __restrict__ G, double* __restrict__ PHI, unsigned* __restrict__ ARG) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned arg[4] = {0xFFFFFFu, 0xFFFFFFu, 0xFFFFFFu, 0xFFFFFFu};
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
      if (ov[k] < mn[k]) {
        mn[k] = ov[k];
        if (ARGS) arg[k] = (arg[k] & 0x00FFFFu) | ((unsigned)(e - e0) << 16);
      }
      if (ov[k] > mx[k]) {
        mx[k] = ov[k];
        if (ARGS) arg[k] = (arg[k] & 0xFF00FFu) | ((unsigned)(e - e0) << 8);
      }
    }
  }
  double phi[4] = {1.0, 1.0, 1.0, 1.0};
  int l0 = lptr[i], l1 = lptr[i + 1];
  for (int e = l0; e < l1; ++e) {
    double dx = ldx[e], dy = ldy[e];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double before = phi[k];
      limiter_entry(gx[k] * dx + gy[k] * dy, wv[k], mx[k], mn[k], code, phi[k]);
      if (ARGS && phi[k] < before) arg[k] = (arg[k] & 0xFFFF00u) | (unsigned)(e - l0);
    }
  }
  if (ARGS) {
    uint4* aq = reinterpret_cast<uint4*>(ARG + 4 * (long)i);
    *aq = make_uint4(arg[0], arg[1], arg[2], arg[3]);
  }
  double2* gq = reinterpret_cast<double2*>(G + 8 * (long)i);
  gq[0] = make_double2(gx[0], gy[0]);
  gq[1] = make_double2(gx[1], gy[1]);
  gq[2] = make_double2(gx[2], gy[2]);
  gq[3] = make_double2(gx[3], gy[3]);
  store4(PHI, i, phi[0], phi[1], phi[2], phi[3]);
}


// MUSCL parabolic (solver.py:2322-2415): linear LSQ gradient and limiter as
// k_gradient; for full-rank cells the quadratic fit replaces them:
// gradient/curvature = sum_e P_e (W_e - W_i) with P_e precomputed per stencil
// entry, min/max over the two-ring stencil, quadratic face increments.
// With ARGS (for the adjoint) the limiting and extremum entries as in
// k_gradient_t; extremum entries count in the two-ring list for full-rank
// cells and in the LSQ list otherwise.
template <bool ARGS>
__global__ void __launch_bounds__(kBlock)
    k_gradient_parabolic_t(int n, const double* __restrict__ W, const int* __restrict__ gptr,
                           const int* __restrict__ gnb, const double* __restrict__ gcx,
                           const double* __restrict__ gcy, const int* __restrict__ lptr,
                           const double* __restrict__ ldx, const double* __restrict__ ldy, int code,
                           const int* __restrict__ pptr, const int* __restrict__ pnb,
                           const double* __restrict__ pcoef, const int* __restrict__ pfull,
                           double* __restrict__ G, double* __restrict__ K, double* __restrict__ PHI,
                           unsigned* __restrict__ ARG) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned arg[4] = {0xFFFFFFu, 0xFFFFFFu, 0xFFFFFFu, 0xFFFFFFu};
  P4 w = load4(W, i);
  double wv[4] = {w.r, w.u, w.v, w.p};
  double gx[4] = {0.0, 0.0, 0.0, 0.0};
  double gy[4] = {0.0, 0.0, 0.0, 0.0};
  double kk[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // [k][xx, xy, yy]
  double mn[4] = {wv[0], wv[1], wv[2], wv[3]};
  double mx[4] = {wv[0], wv[1], wv[2], wv[3]};
  double phi[4] = {1.0, 1.0, 1.0, 1.0};
  if (pfull[i]) {
    double c[5][4] = {{0}};
    const int p0 = pptr[i], p1 = pptr[i + 1];
    // entries in pairs: both entries' loads are issued before either is
    // summed (sums in entry order)
    auto entry = [&](int e, const P4& o, const double* pe) {
      double ov[4] = {o.r, o.u, o.v, o.p};
      for (int k = 0; k < 4; ++k) {
        double d = ov[k] - wv[k];
        for (int a = 0; a < 5; ++a) c[a][k] = c[a][k] + pe[a] * d;
        if (ov[k] < mn[k]) {
          mn[k] = ov[k];
          if (ARGS) arg[k] = (arg[k] & 0x00FFFFu) | ((unsigned)(e - p0) << 16);
        }
        if (ov[k] > mx[k]) {
          mx[k] = ov[k];
          if (ARGS) arg[k] = (arg[k] & 0xFF00FFu) | ((unsigned)(e - p0) << 8);
        }
      }
    };
    int e = p0;
    for (; e + 1 < p1; e += 2) {
      const int n0 = pnb[e], n1 = pnb[e + 1];
      const P4 o0 = load4(W, n0), o1 = load4(W, n1);
      double pe0[5], pe1[5];
      for (int a = 0; a < 5; ++a) {
        pe0[a] = pcoef[5 * (long)e + a];
        pe1[a] = pcoef[5 * (long)(e + 1) + a];
      }
      entry(e, o0, pe0);
      entry(e + 1, o1, pe1);
    }
    for (; e < p1; ++e) entry(e, load4(W, pnb[e]), pcoef + 5 * (long)e);
    for (int k = 0; k < 4; ++k) {
      gx[k] = c[0][k];
      gy[k] = c[1][k];
      kk[3 * k] = c[2][k];
      kk[3 * k + 1] = c[3][k];
      kk[3 * k + 2] = c[4][k];
    }
    for (int e = lptr[i]; e < lptr[i + 1]; ++e) {
      double dx = ldx[e], dy = ldy[e];
      for (int k = 0; k < 4; ++k) {
        double inc = gx[k] * dx + gy[k] * dy + 0.5 * kk[3 * k] * (dx * dx) +
                     kk[3 * k + 1] * (dx * dy) + 0.5 * kk[3 * k + 2] * (dy * dy);
        double before = phi[k];
        limiter_entry(inc, wv[k], mx[k], mn[k], code, phi[k]);
        if (ARGS && phi[k] < before) arg[k] = (arg[k] & 0xFFFF00u) | (unsigned)(e - lptr[i]);
      }
    }
  } else {
    const int g0 = gptr[i];
    for (int e = g0; e < gptr[i + 1]; ++e) {
      P4 o = load4(W, gnb[e]);
      double ov[4] = {o.r, o.u, o.v, o.p};
      double cx = gcx[e], cy = gcy[e];
      for (int k = 0; k < 4; ++k) {
        double d = ov[k] - wv[k];
        gx[k] = gx[k] + cx * d;
        gy[k] = gy[k] + cy * d;
        if (ov[k] < mn[k]) {
          mn[k] = ov[k];
          if (ARGS) arg[k] = (arg[k] & 0x00FFFFu) | ((unsigned)(e - g0) << 16);
        }
        if (ov[k] > mx[k]) {
          mx[k] = ov[k];
          if (ARGS) arg[k] = (arg[k] & 0xFF00FFu) | ((unsigned)(e - g0) << 8);
        }
      }
    }
    for (int e = lptr[i]; e < lptr[i + 1]; ++e) {
      double dx = ldx[e], dy = ldy[e];
      for (int k = 0; k < 4; ++k) {
        double before = phi[k];
        limiter_entry(gx[k] * dx + gy[k] * dy, wv[k], mx[k], mn[k], code, phi[k]);
        if (ARGS && phi[k] < before) arg[k] = (arg[k] & 0xFFFF00u) | (unsigned)(e - lptr[i]);
      }
    }
  }
  if (ARGS) {
    uint4* aq = reinterpret_cast<uint4*>(ARG + 4 * (long)i);
    *aq = make_uint4(arg[0], arg[1], arg[2], arg[3]);
  }
  double2* gq = reinterpret_cast<double2*>(G + 8 * (long)i);
  for (int k = 0; k < 4; ++k) gq[k] = make_double2(gx[k], gy[k]);
  double* kq = K + 12 * (long)i;
  for (int a = 0; a < 12; ++a) kq[a] = kk[a];
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

// With curvature (solver.py:2424-2434): the quadratic increment is added to
// the linear one as a group.
__device__ __forceinline__ P4 reconstruct_q(const double* __restrict__ W,
                                            const double* __restrict__ G,
                                            const double* __restrict__ K,
                                            const double* __restrict__ PHI, int c, double dx,
                                            double dy) {
  P4 w = load4(W, c);
  const double* g = G + 8 * (long)c;
  const double* k = K + 12 * (long)c;
  P4 f = load4(PHI, c);
  double q[4];
  for (int v = 0; v < 4; ++v) {
    double lin = g[2 * v] * dx + g[2 * v + 1] * dy;
    double quad = 0.5 * k[3 * v] * (dx * dx) + k[3 * v + 1] * (dx * dy) + 0.5 * k[3 * v + 2] * (dy * dy);
    q[v] = lin + quad;
  }
  return P4{w.r + f.r * q[0], w.u + f.u * q[1], w.v + f.v * q[2], w.p + f.p * q[3]};
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
  const double* inflow;  // exterior states of the inflow records (kinds 7-9), or NULL
  int inflow_first;      // boundary record of inflow[0]
};

__global__ void __launch_bounds__(kBlock)
    k_flux(int nf, FaceArrays fa, const double* __restrict__ W, const double* __restrict__ G,
           const double* __restrict__ PHI, const double* __restrict__ K, int scheme, int muscl,
           double g, double* __restrict__ FLUX, int* __restrict__ status) {
  int k = xcd_block() * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  int kind = fa.kind[k];
  double nx = fa.nx[k], ny = fa.ny[k];
  const bool quad = K != nullptr;
  P4 L = quad ? reconstruct_q(W, G, K, PHI, fa.left[k], fa.lx[k], fa.ly[k])
              : reconstruct(W, G, PHI, fa.left[k], fa.lx[k], fa.ly[k], muscl);
  P4 R;
  int err = 0;
  if (kind <= MCFD_FACE_PERIODIC) {
    R = quad ? reconstruct_q(W, G, K, PHI, fa.right[k], fa.rx[k], fa.ry[k])
             : reconstruct(W, G, PHI, fa.right[k], fa.rx[k], fa.ry[k], muscl);
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
    } else if (kind >= MCFD_FACE_PRESSURE_INLET) {
      R = load4(fa.inflow, b - fa.inflow_first);
    } else {
      R = L;
    }
  }
  F4 f;
  if (scheme == 1) {
    // interior faces take the compiled kernel's branch order unless the
    // reconstruction is parabolic (solver.py:2479-2496)
    bool fused_order = kind == MCFD_FACE_INTERIOR && !quad;
    f = hllc(L, R, nx, ny, g, fused_order);
    if (!fused_order &&
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
  double* dt_next;                 // per-cell stable step of the updated state, or NULL
  double* dt_partial;              // per-block minimum of the next step's dt (fused final stage), or NULL
  double cfl;
  // kOpJacobian (fused stage only): u = state, u1 = direction v, u0 = rate
  // R1(state), out = v - dt (R1(state + eps sigma v) - rate) / (eps sigma)
  // with eps = jv_num / ||v|| (||v|| = jv_norm[0] on the device; 0: out = v)
  const double* jv_norm;
  double jv_num;
  double jv_sigma[4];
};

// Internal stage op of the fused kernel (not an McfdStage op): the
// implicit path's Jacobian product in one launch (k_perturb, the first-order
// residual and k_jv_finish together, same operations).
constexpr int kOpJacobian = 100;

// Arrays of the spectral-radius sum (solver.py:4132-4343, inviscid part).
struct DtArrays {
  const int* ptr;
  const double* nx;
  const double* ny;
  const double* len;
  const int* bc;
  const int* bc_kind;
  const double* bc_state;
  const double* area;
  const double* inflow;  // exterior states of the inflow records (from cell states), or NULL
  int inflow_first;
};

// Local stable step of one cell from its primitive state: sum over its
// faces of (|u.n| + c) L with the exterior state's radius at inlet, outlet and
// far-field faces, then cfl * A / sum. Shared by the time-step kernel and the
// fused final stage so both give the same bits.
__device__ double cell_spectral_sum(int i, const P4& w, const DtArrays& d, double g,
                                    int* status) {
  double c = sound(w, g);
  double sum = 0.0;
  int e0 = d.ptr[i], e1 = d.ptr[i + 1];
  for (int e = e0; e < e1; ++e) {
    double nx = d.nx[e], ny = d.ny[e];
    double rad = fabs(w.u * nx + w.v * ny) + c;
    int b = d.bc[e];
    if (b >= 0) {
      int kind = d.bc_kind[b];
      const double* st = d.bc_state + 4 * (long)b;
      P4 x = w;
      int err = 0;
      if (kind == MCFD_FACE_VELOCITY_INLET)
        x = P4{st[0], st[1], st[2], st[3]};
      else if (kind == MCFD_FACE_PRESSURE_OUTLET)
        x.p = st[3];
      else if (kind == MCFD_FACE_FAR_FIELD)
        x = far_field(w, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, &err);
      else if (kind >= MCFD_FACE_PRESSURE_INLET)
        x = load4(d.inflow, b - d.inflow_first);
      if (err) report(status, err, i);
      double xr = fabs(x.u * nx + x.v * ny) + sound(x, g);
      rad = vmax(rad, xr);
    }
    sum = sum + rad * d.len[e];
  }
  return sum;
}

__device__ double cell_time_step(int i, const P4& w, const DtArrays& d, double g, double cfl,
                                 int* status) {
  double sum = cell_spectral_sum(i, w, d, g, status);
  double dt = cfl * d.area[i] / sum;
  if (!finite(dt) || !(dt > 0.0)) report(status, MCFD_ERR_LOCAL_DT, i);
  return dt;
}

// ---------------------------------------------------------------------------
// Pressure-inlet, mass-flow-inlet and mass-flow-outlet states
// (solver.py:4412-4577). One pass over the inflow records computes the
// pressure-inlet and mass-flow-outlet states and the mass-flow inlets'
// interior pressures from the reconstructed face states (residual) or the
// owner cells' states (time steps); one workgroup per mass-flow-inlet patch
// then bisects the patch's inflow speed. The flux and time-step kernels read
// the states from `ext`.

struct InflowArrays {
  int first;              // boundary record of record 0
  int n;                  // records
  const int* face;        // [n] flux face
  const double* params;   // [4n]
  const double* proj;     // [n] n.d (pressure inlet), -(n.d) (mass-flow inlet)
  const int* patch_rec;   // [patches+1]
  const int* leaf_ptr;    // [patches+1]
  const int* leaf;        // [2 leaves] {start, size} within the patch
  const int* prog_ptr;    // [patches+1]
  const int* program;     // summation tokens
  double* ext;            // [4n] exterior states
};

// _pressure_inlet_boundary_states for one face (solver.py:4431-4496): the
// inflow Mach number M in [0, 1] whose incoming invariant
// c0/sqrt(1 + (g-1)/2 M^2) (d.n M + 2/(g-1)) equals the interior's outgoing
// one, by 64 bisection steps; then the isentropic static state along d.
__device__ P4 pressure_inlet_state(const P4& in, double nx, double ny, const double* par, double dn,
                                   double g, double R, int* err) {
  double c_in = sqrt(g * in.p / in.r);
  double un = in.u * nx + in.v * ny;
  double m_in = un / c_in;
  if (m_in >= 1.0) {
    *err = MCFD_ERR_INLET_OUTFLOW;
    return in;
  }
  if (m_in <= -1.0) {
    *err = MCFD_ERR_INLET_SUPERSONIC;
    return in;
  }
  const double p0 = par[0], t0 = par[1];
  const double c0 = sqrt(g * R * t0);
  const double half_gm1 = 0.5 * (g - 1.0);
  const double two_gm1 = 2.0 / (g - 1.0);
  const double outgoing = un + 2.0 * c_in / (g - 1.0);
  auto invariant = [&](double m) {
    double ratio = 1.0 + half_gm1 * (m * m);
    double c = c0 / sqrt(ratio);
    return c * (dn * m + two_gm1);
  };
  if (!(invariant(1.0) <= outgoing && outgoing <= invariant(0.0))) {
    *err = MCFD_ERR_INLET_TOTALS;
    return in;
  }
  double low = 0.0, high = 1.0;
  for (int it = 0; it < 64; ++it) {
    double mid = 0.5 * (low + high);
    if (invariant(mid) > outgoing)
      low = mid;
    else
      high = mid;
  }
  double mach = 0.5 * (low + high);
  double T = t0 / (1.0 + half_gm1 * (mach * mach));
  double p = p0 * pow(T / t0, g / (g - 1.0));
  double rho = p / (R * T);
  double speed = mach * sqrt(g * R * T);
  P4 e{rho, speed * par[2], speed * par[3], p};
  if (!finite(e.r) || !finite(e.u) || !finite(e.v) || !finite(e.p) || !(e.r > 0.0) || !(e.p > 0.0))
    *err = MCFD_ERR_INLET_STATE;
  return e;
}

// _mass_flow_outlet_boundary_states (solver.py:4557-4577): normal velocity
// mdot / (sum L * rho), tangential velocity, density and pressure kept.
__device__ P4 mass_flow_outlet_state(const P4& in, double nx, double ny, const double* par, int* err) {
  if (!finite(in.r) || !(in.r > 0.0)) *err = MCFD_ERR_OUTLET_DENSITY;
  double un = par[0] / (par[1] * in.r);
  double tx = -ny, ty = nx;
  double ut = in.u * tx + in.v * ty;
  return P4{in.r, un * nx + ut * tx, un * ny + ut * ty, in.p};
}

// Reverse of one inflow record's exterior state e = ext(in), the adjoint eb
// of e added into the interior primitive adjoint inb.
// Pressure inlet: e depends on in through the outgoing invariant
// J = u.n + 2 c / (g-1) alone; the Mach number M solves inv(M) = J, so
// dM/dJ = 1 / inv'(M) with inv(M) = c0 th^(-1/2) (dn M + 2/(g-1)),
// th = 1 + (g-1)/2 M^2, and along M: d rho/dM = -rho M / th,
// d speed/dM = a / th (a = sqrt(g R T)), dp/dM = -g p M / th (M, a and th
// from e itself). Mass-flow outlet: rho and p copied, the normal velocity
// mdot / (L rho) and the tangential velocity u.t. Mass-flow inlet: the face's
// own pressure (rho = p / (R T) and p); the patch speed's dependence on every
// face's pressure is added per patch by k_mass_flow_inlet_adjoint.
__device__ void inflow_state_b(int kind, const P4& in, const P4& e, double nx, double ny, const double* par,
                               double dn, double g, double R, const P4& eb, P4& inb) {
  if (kind == MCFD_FACE_PRESSURE_INLET) {
    const double h = 0.5 * (g - 1.0);
    const double speed = sqrt(e.u * e.u + e.v * e.v);
    const double a = sqrt(g * e.p / e.r);
    const double M = speed / a;
    const double th = 1.0 + h * (M * M);
    const double c0 = sqrt(g * R * par[1]);
    const double dinv = c0 / sqrt(th) * (dn - h * M * (dn * M + 2.0 / (g - 1.0)) / th);
    const double Mb = eb.r * (-e.r * M / th) + (eb.u * par[2] + eb.v * par[3]) * (a / th) +
                      eb.p * (-g * e.p * M / th);
    const double Jb = Mb / dinv;
    const double c_in = sqrt(g * in.p / in.r);
    inb.u += Jb * nx;
    inb.v += Jb * ny;
    inb.p += Jb * g / ((g - 1.0) * in.r * c_in);
    inb.r -= Jb * c_in / ((g - 1.0) * in.r);
  } else if (kind == MCFD_FACE_MASS_FLOW_INLET) {
    inb.p += eb.r * (e.r / e.p) + eb.p;
  } else {
    const double un = par[0] / (par[1] * in.r);
    const double tx = -ny, ty = nx;
    const double unb = eb.u * nx + eb.v * ny;
    const double utb = eb.u * tx + eb.v * ty;
    inb.r += eb.r - unb * un / in.r;
    inb.u += utb * tx;
    inb.v += utb * ty;
    inb.p += eb.p;
  }
}

__global__ void __launch_bounds__(kBlock)
    k_inflow_records(InflowArrays ia, FaceArrays fa, const double* __restrict__ W,
                     const double* __restrict__ G, const double* __restrict__ PHI,
                     const double* __restrict__ K, int muscl, int reconstructed, double g, double R,
                     int* __restrict__ status) {
  int r = xcd_block() * blockDim.x + threadIdx.x;
  if (r >= ia.n) return;
  int k = ia.face[r];
  int kind = fa.kind[k];
  double nx = fa.nx[k], ny = fa.ny[k];
  P4 in;
  if (!reconstructed)
    in = load4(W, fa.left[k]);
  else if (K)
    in = reconstruct_q(W, G, K, PHI, fa.left[k], fa.lx[k], fa.ly[k]);
  else
    in = reconstruct(W, G, PHI, fa.left[k], fa.lx[k], fa.ly[k], muscl);
  const double* par = ia.params + 4 * (long)r;
  int err = 0;
  P4 e;
  if (kind == MCFD_FACE_PRESSURE_INLET) {
    e = pressure_inlet_state(in, nx, ny, par, ia.proj[r], g, R, &err);
  } else if (kind == MCFD_FACE_MASS_FLOW_INLET) {
    if (!finite(in.p) || !(in.p > 0.0)) err = MCFD_ERR_MASS_PRESSURE;
    e = P4{0.0, 0.0, 0.0, in.p};  // the patch kernel completes the state
  } else {
    e = mass_flow_outlet_state(in, nx, ny, par, &err);
  }
  if (err) report(status, err, r);
  store4(ia.ext, r, e.r, e.u, e.v, e.p);
}

// Reverse of the convective faces' inflow states: per record, the exterior
// adjoint in the right slot of the face's FADJ record goes through
// inflow_state_b into the left slot (the reconstructed left state's adjoint).
__global__ void __launch_bounds__(kBlock)
    k_inflow_adjoint(InflowArrays ia, FaceArrays fa, const double* __restrict__ W,
                     const double* __restrict__ G, const double* __restrict__ PHI,
                     const double* __restrict__ K, int muscl, double g, double R, double* __restrict__ FADJ) {
  int r = xcd_block() * blockDim.x + threadIdx.x;
  if (r >= ia.n) return;
  const int k = ia.face[r];
  const int kind = fa.kind[k];
  const double nx = fa.nx[k], ny = fa.ny[k];
  const P4 in = K ? reconstruct_q(W, G, K, PHI, fa.left[k], fa.lx[k], fa.ly[k])
                  : reconstruct(W, G, PHI, fa.left[k], fa.lx[k], fa.ly[k], muscl);
  const P4 e = load4(ia.ext, r);
  double* q = FADJ + 8 * (long)k;
  const P4 eb{q[4], q[5], q[6], q[7]};
  P4 inb{0.0, 0.0, 0.0, 0.0};
  inflow_state_b(kind, in, e, nx, ny, ia.params + 4 * (long)r, ia.proj[r], g, R, eb, inb);
  q[0] += inb.r;
  q[1] += inb.u;
  q[2] += inb.v;
  q[3] += inb.p;
}

// Sum over a mass-flow-inlet patch of (p_i / (R T)) * speed * proj_i * L_i
// (the reference's flow_at, solver.py:4521-4535) in NumPy's pairwise order:
// leaves of < 8 values are summed in turn from 0; larger leaves with eight
// strided accumulators combined as ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) plus
// the tail in turn; the leaves combined by the patch's token program. Every
// thread of the workgroup calls it and gets the sum.
__device__ double patch_flow(const InflowArrays& ia, const double* __restrict__ face_len, int p, int a,
                             int n, int la, int nl, double speed, double t0, double cp, double R,
                             double* terms, double* rr, double* leafsum, double* stack, double* bcast) {
  double RT = R * (t0 - speed * speed / (2.0 * cp));
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int r = a + i;
    terms[i] = ia.ext[4 * (long)r + 3] / RT * speed * ia.proj[r] * face_len[ia.face[r]];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 8 * nl; t += blockDim.x) {
    int l = t >> 3, j = t & 7;
    int s = ia.leaf[2 * (la + l)], m = ia.leaf[2 * (la + l) + 1];
    if (m < 8) {
      if (j == 0) {
        double res = 0.0;
        for (int i = 0; i < m; ++i) res += terms[s + i];
        leafsum[l] = res;
      }
    } else {
      double acc = terms[s + j];
      int end = m - (m & 7);
      for (int i = 8; i < end; i += 8) acc += terms[s + i + j];
      rr[t] = acc;
    }
  }
  __syncthreads();
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    int s = ia.leaf[2 * (la + l)], m = ia.leaf[2 * (la + l) + 1];
    if (m >= 8) {
      const double* q = rr + 8 * l;
      double res = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
      for (int i = m - (m & 7); i < m; ++i) res += terms[s + i];
      leafsum[l] = res;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int sp = 0;
    for (int k = ia.prog_ptr[p]; k < ia.prog_ptr[p + 1]; ++k) {
      int tok = ia.program[k];
      if (tok >= 0) {
        stack[sp++] = leafsum[tok];
      } else {
        double rhs = stack[--sp];
        double lhs = stack[--sp];
        stack[sp++] = lhs + rhs;
      }
    }
    *bcast = stack[0];
  }
  __syncthreads();
  return *bcast;
}

// _mass_flow_inlet_boundary_states (solver.py:4499-4554): one workgroup per
// patch; 80 bisection steps on the inflow speed in [0, vmax] against the
// target flow, then rho = p / (R T), velocity = speed * d, pressure kept.
// LDS: terms [max_n], accumulators [8 max_leaves], leaf sums [max_leaves],
// token stack [64], broadcast [1]; the host checks the patch sizes.
__global__ void __launch_bounds__(256)
    k_mass_flow_inlet(InflowArrays ia, const double* __restrict__ face_len, int max_n, int max_leaves,
                      double g, double R, int* __restrict__ status) {
  extern __shared__ double sh[];
  double* terms = sh;
  double* rr = terms + max_n;
  double* leafsum = rr + 8 * max_leaves;
  double* stack = leafsum + max_leaves;
  double* bcast = stack + 64;
  const int p = blockIdx.x;
  const int a = ia.patch_rec[p], n = ia.patch_rec[p + 1] - a;
  const int la = ia.leaf_ptr[p], nl = ia.leaf_ptr[p + 1] - la;
  const double* par = ia.params + 4 * (long)a;
  const double mdot = par[0], t0 = par[1];
  const double cp = g * R / (g - 1.0);
  const double vmax = sqrt(2.0 * cp * t0) * (1.0 - 1e-12);
  double low = 0.0, high = vmax;
  double f = patch_flow(ia, face_len, p, a, n, la, nl, high, t0, cp, R, terms, rr, leafsum, stack, bcast);
  if (f < mdot && threadIdx.x == 0) report(status, MCFD_ERR_MASS_FLOW, a);
  for (int it = 0; it < 80; ++it) {
    double mid = 0.5 * (low + high);
    f = patch_flow(ia, face_len, p, a, n, la, nl, mid, t0, cp, R, terms, rr, leafsum, stack, bcast);
    if (f < mdot)
      low = mid;
    else
      high = mid;
  }
  const double speed = 0.5 * (low + high);
  const double RT = R * (t0 - speed * speed / (2.0 * cp));
  const double u = speed * par[2], v = speed * par[3];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int r = a + i;
    double pr = ia.ext[4 * (long)r + 3];
    store4(ia.ext, r, pr / RT, u, v, pr);
  }
}

// The patch speed s of a mass-flow inlet solves
// G(s, p) = sum_g p_g / (R T(s)) s w_g - mdot = 0, T = t0 - s^2 / (2 cp),
// w_g = (-n.d)_g L_g. With S = sum_g eb_g . d e_g/ds (d rho_g/ds =
// p_g s / (R cp T^2), d(u, v)/ds = d), each face's pressure adjoint gains
// -S (dG/dp_f) / (dG/ds), dG/dp_f = s w_f / (R T),
// dG/ds = sum_g p_g w_g / R (1/T + s^2 / (cp T^2)). Sums over the patch by
// a fixed tree (deterministic). Returns S / (dG/ds) * (s / (R T)) to every
// thread; the caller multiplies by -w_f.
__device__ double mass_flow_patch_factor(const InflowArrays& ia, const double* __restrict__ ext,
                                         const double* __restrict__ face_len, const double* __restrict__ ebar,
                                         int ebar_stride, int ebar_off, const int* __restrict__ ebar_face,
                                         int a, int n, double g, double R, double* red) {
  const double* par = ia.params + 4 * (long)a;
  const double t0 = par[1], d0 = par[2], d1 = par[3];
  const double cp = g * R / (g - 1.0);
  const double* e0 = ext + 4 * (long)a;
  const double s = (fabs(d0) >= fabs(d1)) ? e0[1] / d0 : e0[2] / d1;
  const double T = t0 - s * s / (2.0 * cp);
  double dg = 0.0, sb = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int r = a + i;
    const double p = ext[4 * (long)r + 3];
    const double w = ia.proj[r] * face_len[ia.face[r]];
    const long row = ebar_face ? (long)ebar_face[r] : (long)r;
    const double* eb = ebar + ebar_stride * row + ebar_off;
    dg += p * w / R * (1.0 / T + s * s / (cp * T * T));
    sb += eb[0] * (p * s / (R * cp * T * T)) + eb[1] * d0 + eb[2] * d1;
  }
  red[threadIdx.x] = dg;
  red[blockDim.x + threadIdx.x] = sb;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      red[threadIdx.x] += red[threadIdx.x + k];
      red[blockDim.x + threadIdx.x] += red[blockDim.x + threadIdx.x + k];
    }
    __syncthreads();
  }
  const double f = red[blockDim.x] / red[0] * (s / (R * T));
  __syncthreads();
  return f;
}

// Convective faces: one workgroup per mass-flow-inlet patch; the exterior
// state adjoints are the right slots of the faces' FADJ records, the
// pressure adjoint goes to the left slot (reconstructed left state).
__global__ void __launch_bounds__(256)
    k_mass_flow_inlet_adjoint(InflowArrays ia, const double* __restrict__ face_len, double g, double R,
                              double* __restrict__ FADJ) {
  __shared__ double red[512];
  const int p = blockIdx.x;
  const int a = ia.patch_rec[p], n = ia.patch_rec[p + 1] - a;
  const double f = mass_flow_patch_factor(ia, ia.ext, face_len, FADJ, 8, 4, ia.face, a, n, g, R, red);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int r = a + i;
    const int k = ia.face[r];
    FADJ[8 * (long)k + 3] += -f * ia.proj[r] * face_len[k];
  }
}

// Viscous boundary values (exterior states from the owner cells): the
// exterior adjoints were gathered per record into vbar; the pressure adjoint
// goes to the owner cells' primitive adjoints. One workgroup walks the
// patches and then their records in turn, so cells owning several records
// are updated in a fixed order.
__global__ void __launch_bounds__(256)
    k_mass_flow_inlet_visc_adjoint(InflowArrays ia, int patches, const double* __restrict__ ext_v,
                                   const double* __restrict__ face_len, const int* __restrict__ owner,
                                   const double* __restrict__ vbar, double g, double R,
                                   double* __restrict__ wbar) {
  __shared__ double red[512];
  for (int p = 0; p < patches; ++p) {
    const int a = ia.patch_rec[p], n = ia.patch_rec[p + 1] - a;
    const double f = mass_flow_patch_factor(ia, ext_v, face_len, vbar, 4, 0, nullptr, a, n, g, R, red);
    if (threadIdx.x == 0)
      for (int i = 0; i < n; ++i) {
        const int r = a + i;
        const int k = ia.face[r];
        wbar[4 * (long)owner[k] + 3] += -f * ia.proj[r] * face_len[k];
      }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Laminar viscous terms (solver.py:2721-2737, 3615-3747, 3750-4114, 4246-4332)

struct ViscParams {
  int model;  // MCFD_VISCOSITY_*
  double gamma, R, Pr, mu, tref, muref, S;
};

// _dynamic_viscosity (solver.py:2721-2737)
__device__ double visc_mu(double T, const ViscParams& v, int* status, int index) {
  if (v.model == MCFD_VISCOSITY_CONSTANT) return v.mu;
  double mu = v.muref * pow(T / v.tref, 1.5);
  mu = mu * ((v.tref + v.S) / (T + v.S));
  if (!finite(mu) || !(mu > 0.0)) report(status, MCFD_ERR_VISCOSITY, index);
  return mu;
}

// conductivity = mu * gamma * R / ((gamma - 1) * Pr)
__device__ __forceinline__ double visc_k(double mu, const ViscParams& v) {
  return mu * v.gamma * v.R / ((v.gamma - 1.0) * v.Pr);
}

struct ViscArrays {
  const double* inf_ext;  // [4 * inflow records] exterior states from the owner cells (kVisInflow)
  int inf_first;          // boundary record of inflow record 0
  const double* wl_wx;
  const double* wl_wy;
  const double* wl_cell;
  const int* gh_ptr;
  const int* gh_face;
  const double* gh_wx;
  const double* gh_wy;
  const int* bv_code;
  const int* bv_flags;
  const double* bv_geom;
};

enum { kVisWall = 1, kVisSymmetry = 2, kVisInlet = 3, kVisInflow = 4, kVisFarField = 5 };

// Exterior state read by visc_inlet: the boundary record's own state, or for
// pressure / mass-flow inlets the inflow record's state from the owner cell.
__device__ __forceinline__ const double* visc_inlet_state(int code, int b, const ViscArrays& va,
                                                          const FaceArrays& fa) {
  return code == kVisInflow ? va.inf_ext + 4 * (long)(b - va.inf_first) : fa.bc_state + 4 * (long)b;
}

// Exterior (u, v, T) of a velocity inlet or far-field face from the owner's
// cell state (_viscous_inlet_states, solver.py:2740-2785).
__device__ __forceinline__ void visc_inlet(int code, const P4& w, const double* st, double nx,
                                           double ny, const ViscParams& v, double out[3],
                                           int* status, int index) {
  P4 x{st[0], st[1], st[2], st[3]};
  if (code == kVisFarField) {
    int err = 0;
    x = far_field(w, x, nx, ny, v.gamma, &err);
    if (err) report(status, err, index);
  }
  out[0] = x.u;
  out[1] = x.v;
  out[2] = x.p / (x.r * v.R);
}

// Second half of one cell's viscous gradient (shared by k_visc_gradient and
// the fused stage kernel): the viscosity and conductivity at the cell, the
// ghost entries of its non-periodic boundary faces in list order, then the
// 2x2 solve. out = {du/dx, du/dy, dv/dx, dv/dy, dT/dx, dT/dy, mu, k}.
__device__ __forceinline__ void visc_gradient_finish(int i, const P4& w, const double val[3], double rx[3],
                                                     double ry[3], const ViscArrays& va, const FaceArrays& fa,
                                                     const ViscParams& vp, int* status, double out[8]) {
  const double T = val[2];
  double mu = visc_mu(T, vp, status, i);
  double kc = visc_k(mu, vp);
  for (int e = va.gh_ptr[i]; e < va.gh_ptr[i + 1]; ++e) {
    int f = va.gh_face[e];
    int b = fa.bc[f];
    int code = va.bv_code[b];
    double nx = fa.nx[f], ny = fa.ny[f];
    double gh[3] = {val[0], val[1], val[2]};
    if (code == kVisWall) {
      int flags = va.bv_flags[b];
      const double* bg = va.bv_geom + 4 * (long)b;
      if (flags & 1) {
        const double* wv = fa.bc_wall + 2 * (long)b;
        gh[0] = 2.0 * wv[0] - gh[0];
        gh[1] = 2.0 * wv[1] - gh[1];
      } else {
        double t = 2.0 * (gh[0] * nx + gh[1] * ny);
        gh[0] = gh[0] - t * nx;
        gh[1] = gh[1] - t * ny;
      }
      if (((flags >> 1) & 3) == 1)
        gh[2] = 2.0 * bg[3] - val[2];
      else
        gh[2] = gh[2] + 2.0 * bg[0] * bg[3] / kc;
    } else if (code == kVisSymmetry) {
      double t = 2.0 * (val[0] * nx + val[1] * ny);
      gh[0] = gh[0] - t * nx;
      gh[1] = gh[1] - t * ny;
    } else if (code == kVisInlet || code == kVisInflow || code == kVisFarField) {
      double x[3];
      visc_inlet(code, w, visc_inlet_state(code, b, va, fa), nx, ny, vp, x, status, i);
#pragma unroll
      for (int k = 0; k < 3; ++k) gh[k] = 2.0 * x[k] - gh[k];
    }
    double wx = va.gh_wx[e], wy = va.gh_wy[e];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double d = gh[k] - val[k];
      rx[k] = rx[k] + wx * d;
      ry[k] = ry[k] + wy * d;
    }
  }
  const double* m = va.wl_cell + 4 * (long)i;
  double xx = m[0], xy = m[1], yy = m[2], inv = m[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    out[2 * k] = (inv != 0.0) ? (rx[k] * yy - ry[k] * xy) * inv : 0.0;
    out[2 * k + 1] = (inv != 0.0) ? (ry[k] * xx - rx[k] * xy) * inv : 0.0;
  }
  out[6] = mu;
  out[7] = kc;
}

// Weighted least-squares gradients of (u, v, T) with one ghost entry per
// non-periodic boundary face (_viscous_gradients + the numba kernel
// weighted_least_squares_gradients, _kernels.py:163-207), then the cell's
// viscosity and conductivity. VG[i] = {du/dx, du/dy, dv/dx, dv/dy, dT/dx,
// dT/dy, mu, k}.
__global__ void __launch_bounds__(kBlock)
    k_visc_gradient(int n, const double* __restrict__ W, const int* __restrict__ grad_ptr,
                    const int* __restrict__ grad_nb, ViscArrays va, FaceArrays fa, ViscParams vp,
                    double* __restrict__ VG, int* __restrict__ status) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  P4 w = load4(W, i);
  double T = w.p / (w.r * vp.R);
  if (!finite(T) || !(T > 0.0)) report(status, MCFD_ERR_TEMPERATURE, i);
  double val[3] = {w.u, w.v, T};
  double rx[3] = {0.0, 0.0, 0.0}, ry[3] = {0.0, 0.0, 0.0};
  // entries in batches of 4: the batch's loads are issued together, the sums
  // still run in entry order
  const int g0 = grad_ptr[i], g1 = grad_ptr[i + 1];
  for (int base = g0; base < g1; base += 4) {
    int nb[4];
    double wxb[4], wyb[4];
    P4 ob[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool ok = base + q < g1;
      nb[q] = ok ? grad_nb[base + q] : i;
      wxb[q] = ok ? va.wl_wx[base + q] : 0.0;
      wyb[q] = ok ? va.wl_wy[base + q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) ob[q] = load4(W, nb[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (base + q >= g1) break;
      const P4& o = ob[q];
      double d[3] = {o.u - w.u, o.v - w.v, o.p / (o.r * vp.R) - T};
      double wx = wxb[q], wy = wyb[q];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        rx[k] = rx[k] + wx * d[k];
        ry[k] = ry[k] + wy * d[k];
      }
    }
  }
  double out[8];
  visc_gradient_finish(i, w, val, rx, ry, va, fa, vp, status, out);
  double2* q = reinterpret_cast<double2*>(VG + 8 * (long)i);
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = make_double2(out[2 * k], out[2 * k + 1]);
}

// _viscous_flux (solver.py:3717-3747): Stokes stress and Fourier heat flux
// through a face of unit normal n; returns {0, tx, ty, u tx + v ty + q}.
__device__ __forceinline__ void visc_flux(const double g[6], double u, double v, double mu,
                                          double k, double nx, double ny, double F[4]) {
  double div = g[0] + g[3];
  double txx = mu * (2.0 * g[0] - (2.0 / 3.0) * div);
  double tyy = mu * (2.0 * g[3] - (2.0 / 3.0) * div);
  double txy = mu * (g[1] + g[2]);
  double tx = txx * nx + txy * ny;
  double ty = txy * nx + tyy * ny;
  double q = k * (g[4] * nx + g[5] * ny);
  F[0] = 0.0;
  F[1] = tx;
  F[2] = ty;
  F[3] = u * tx + v * ty + q;
}

// Viscous flux through a non-periodic boundary face of record b (shared by
// k_visc_flux and the fused stage kernel): the owner's gradients go with the
// reference's boundary corrections (solver.py:3822-4112), boundary values
// from the owner state w, face offset (offx, offy) from the owner centroid.
__device__ __forceinline__ void visc_boundary_flux(int b, double nx, double ny, double offx, double offy,
                                                   const P4& w, const double* go, const ViscArrays& va,
                                                   const FaceArrays& fa, const ViscParams& vp, int* status,
                                                   int o, double F[4]) {
  int code = va.bv_code[b];
  const double* bg = va.bv_geom + 4 * (long)b;
  double dn = bg[0], tx = bg[1], ty = bg[2];
  double T = w.p / (w.r * vp.R);
  double g[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) g[k] = go[k];
  double bv[3] = {w.u, w.v, T};
  int flags = va.bv_flags[b];
  int thermal = (flags >> 1) & 3;
  if (code == kVisWall) {
    if (flags & 1) {
      const double* wv = fa.bc_wall + 2 * (long)b;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        double nd = (wv[c] - bv[c] - (g[2 * c] * tx + g[2 * c + 1] * ty)) / dn;
        double s = nd - (g[2 * c] * nx + g[2 * c + 1] * ny);
        g[2 * c] = g[2 * c] + s * nx;
        g[2 * c + 1] = g[2 * c + 1] + s * ny;
        bv[c] = wv[c];
      }
    } else {
      double un = bv[0] * nx + bv[1] * ny;
      // boundary_values[:2] += face_gradient[:2] @ offset
      bv[0] = bv[0] + (g[0] * offx + g[1] * offy);
      bv[1] = bv[1] + (g[2] * offx + g[3] * offy);
      double un2 = bv[0] * nx + bv[1] * ny;
      bv[0] = bv[0] - un2 * nx;
      bv[1] = bv[1] - un2 * ny;
      // gradient of the normal velocity, n0*grad(u) + n1*grad(v)
      double gnx = nx * g[0] + ny * g[2];
      double gny = nx * g[1] + ny * g[3];
      double nd = -(un + (gnx * tx + gny * ty)) / dn;
      double vn0 = g[0] * nx + g[1] * ny;
      double vn1 = g[2] * nx + g[3] * ny;
      g[0] = g[0] + (nd * nx * nx - vn0 * nx);
      g[1] = g[1] + (nd * nx * ny - vn0 * ny);
      g[2] = g[2] + (nd * ny * nx - vn1 * nx);
      g[3] = g[3] + (nd * ny * ny - vn1 * ny);
    }
    double nd;
    if (thermal == 1) {
      nd = (bg[3] - bv[2] - (g[4] * tx + g[5] * ty)) / dn;
      bv[2] = bg[3];
    } else {
      nd = bg[3] / go[7];
      bv[2] = bv[2] + (g[4] * offx + g[5] * offy);
    }
    double s = nd - (g[4] * nx + g[5] * ny);
    g[4] = g[4] + s * nx;
    g[5] = g[5] + s * ny;
  } else if (code == kVisSymmetry) {
    double un = bv[0] * nx + bv[1] * ny;
    double dvx = -un / dn * nx, dvy = -un / dn * ny;
    double dv[2] = {dvx, dvy};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      double s = dv[c] - (g[2 * c] * nx + g[2 * c + 1] * ny);
      g[2 * c] = g[2 * c] + s * nx;
      g[2 * c + 1] = g[2 * c + 1] + s * ny;
    }
    double gt = g[4] * nx + g[5] * ny;
    g[4] = g[4] - gt * nx;
    g[5] = g[5] - gt * ny;
    bv[0] = bv[0] - un * nx;
    bv[1] = bv[1] - un * ny;
  } else if (code == kVisInlet || code == kVisInflow || code == kVisFarField) {
    double x[3];
    visc_inlet(code, w, visc_inlet_state(code, b, va, fa), nx, ny, vp, x, status, o);
    double cv[3] = {w.u, w.v, T};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bv[c] = x[c];
      double nd = (bv[c] - cv[c] - (g[2 * c] * tx + g[2 * c + 1] * ty)) / dn;
      double s = nd - (g[2 * c] * nx + g[2 * c + 1] * ny);
      g[2 * c] = g[2 * c] + s * nx;
      g[2 * c + 1] = g[2 * c + 1] + s * ny;
    }
  } else {
    // zero normal gradient: g -= outer(g @ n, n); values += g @ offset
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double gn = g[2 * c] * nx + g[2 * c + 1] * ny;
      g[2 * c] = g[2 * c] - gn * nx;
      g[2 * c + 1] = g[2 * c + 1] - gn * ny;
      bv[c] = bv[c] + (g[2 * c] * offx + g[2 * c + 1] * offy);
    }
  }
  double fmu = visc_mu(bv[2], vp, status, o);
  visc_flux(g, bv[0], bv[1], fmu, visc_k(fmu, vp), nx, ny, F);
  if (code == kVisWall && thermal == 2) F[3] = bg[3];
}

// One viscous flux per flux face: face averages on interior faces and
// periodic pairs, the reference's boundary gradient corrections on boundary
// faces (solver.py:3822-4112).
__global__ void __launch_bounds__(kBlock)
    k_visc_flux(int nf, FaceArrays fa, const double* __restrict__ W,
                const double* __restrict__ VG, ViscArrays va, ViscParams vp,
                double* __restrict__ VF, int* __restrict__ status) {
  int f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int kind = fa.kind[f];
  double nx = fa.nx[f], ny = fa.ny[f];
  int o = fa.left[f];
  const double* go = VG + 8 * (long)o;
  double F[4];
  if (kind == MCFD_FACE_INTERIOR || kind == MCFD_FACE_PERIODIC) {
    int r = fa.right[f];
    const double* gr = VG + 8 * (long)r;
    double g[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) g[k] = 0.5 * (go[k] + gr[k]);
    P4 a = load4(W, o), c = load4(W, r);
    visc_flux(g, 0.5 * (a.u + c.u), 0.5 * (a.v + c.v), 0.5 * (go[6] + gr[6]),
              0.5 * (go[7] + gr[7]), nx, ny, F);
    store4(VF, f, F[0], F[1], F[2], F[3]);
    return;
  }
  visc_boundary_flux(fa.bc[f], nx, ny, fa.lx[f], fa.ly[f], load4(W, o), go, va, fa, vp, status, o, F);
  store4(VF, f, F[0], F[1], F[2], F[3]);
}

// Effective diffusivity max(4/3 mu/rho, mu/(rho min(Pr, 0.9))) per cell
// (solver.py:4256-4259).
__global__ void __launch_bounds__(kBlock)
    k_visc_diffusivity(int n, const double* __restrict__ W, ViscParams vp,
                       double* __restrict__ EFF, int* __restrict__ status) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  P4 w = load4(W, i);
  double T = w.p / (w.r * vp.R);
  double mu = visc_mu(T, vp, status, i);
  double pr = vp.Pr < 0.9 ? vp.Pr : 0.9;
  double a = (4.0 / 3.0) * mu / w.r;
  double b = mu / (w.r * pr);
  EFF[i] = vmax(a, b);
}

struct ViscDt {
  const int* ptr;
  const int* nb;
  const double* len;
  const double* dist;
  const double* eff;
};

// Diffusion sum of one cell in the reference's order (interior owners,
// interior neighbours, periodic a, periodic b, boundary faces).
__device__ __forceinline__ double visc_dt_sum(int i, const ViscDt& vd) {
  double ei = vd.eff[i];
  double sum = 0.0;
  for (int e = vd.ptr[i]; e < vd.ptr[i + 1]; ++e) {
    int j = vd.nb[e];
    double fe = (j >= 0) ? vmax(ei, vd.eff[j]) : ei;
    sum = sum + fe * vd.len[e] / vd.dist[e];
  }
  return sum;
}

// Viscous accumulation of one cell: 0 - S over the interior items, then the
// periodic and boundary items (derivative -= D @ flux, np.add.at, +=).
struct ViscAcc {
  const int* ptr;
  const int* item;
  const double* coef;
  const double* VF;
};

// Items in batches of 4: the batch's item, coefficient and flux loads are
// issued together; the sums still run in item order.
__device__ __forceinline__ void visc_gather(int i, const ViscAcc& v, const double* flen, double A,
                                            double r[4]) {
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  double d[4] = {0.0, 0.0, 0.0, 0.0};
  bool interior = true;
  const int e0 = v.ptr[i], e1 = v.ptr[i + 1];
  for (int base = e0; base < e1; base += 4) {
    int it[4];
    double cb[4], lb[4];
    P4 Fb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool ok = base + q < e1;
      it[q] = ok ? v.item[base + q] : 0;
      cb[q] = ok ? v.coef[base + q] : 0.0;
    }
    // lengths of the items from the first periodic or boundary item on
    bool after = !interior;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool ok = base + q < e1;
      after = after || (ok && (it[q] & 3) != 0);
      Fb[q] = ok ? load4(v.VF, it[q] >> 2) : P4{0.0, 0.0, 0.0, 0.0};
      lb[q] = (ok && after) ? flen[it[q] >> 2] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (base + q >= e1) break;
      const P4& F = Fb[q];
      // interior items (0 - S) first, then the periodic and boundary items
      if (interior && (it[q] & 3) == 0) {
        double c = cb[q];
        s[0] = s[0] + c * F.r;
        s[1] = s[1] + c * F.u;
        s[2] = s[2] + c * F.v;
        s[3] = s[3] + c * F.p;
        continue;
      }
      if (interior) {
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = 0.0 - s[k];
        interior = false;
      }
      double L = lb[q];
      double x[4] = {F.r * L, F.u * L, F.v * L, F.p * L};
      if ((it[q] & 3) == 1) {
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = d[k] + x[k] / A;
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) d[k] = d[k] + (-x[k]) / A;
      }
    }
  }
  if (interior) {
#pragma unroll
    for (int k = 0; k < 4; ++k) d[k] = 0.0 - s[k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = r[k] + d[k];
}

// cell_time_step for the fused final stage: the normals and lengths of the
// cell's faces come from the block's LDS slots; items flagged 0x8000 (periodic
// b sides and faces with an exterior state) use the global item as above.
// FARFIELD false: the mesh has no far-field record, so that branch is left out.
template <bool FARFIELD>
__device__ double block_time_step(int i, double area, const P4& w, int e0, int e1,
                                  const unsigned short* dloc, int dbase,
                                  const double* sN, const double* sL, const DtArrays& d, double g,
                                  double cfl, int* status) {
  double c = sound(w, g);
  double sum = 0.0;
  for (int e = e0; e < e1; ++e) {
    int item = dloc[e - dbase];
    double nx, ny, len;
    int b = -1;
    if (item & 0x8000) {
      nx = d.nx[e];
      ny = d.ny[e];
      len = d.len[e];
      b = d.bc[e];
    } else {
      nx = sN[2 * item];
      ny = sN[2 * item + 1];
      len = sL[item];
    }
    double rad = fabs(w.u * nx + w.v * ny) + c;
    if (b >= 0) {
      int kind = d.bc_kind[b];
      const double* st = d.bc_state + 4 * (long)b;
      P4 x = w;
      int err = 0;
      if (kind == MCFD_FACE_VELOCITY_INLET)
        x = P4{st[0], st[1], st[2], st[3]};
      else if (kind == MCFD_FACE_PRESSURE_OUTLET)
        x.p = st[3];
      else if (FARFIELD && kind == MCFD_FACE_FAR_FIELD)
        x = far_field(w, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, &err);
      if (err) report(status, err, i);
      double xr = fabs(x.u * nx + x.v * ny) + sound(x, g);
      rad = vmax(rad, xr);
    }
    sum = sum + rad * len;
  }
  double dt = cfl * area / sum;
  if (!finite(dt) || !(dt > 0.0)) report(status, MCFD_ERR_LOCAL_DT, i);
  return dt;
}

// Face -> cell gather in the reference order (solver.py:2501-2704), body force
// (solver.py:2709-2715), finiteness check (:2716-2717), then the stage update.
__global__ void __launch_bounds__(kBlock)
    k_update(int n, const int* __restrict__ aptr, const int* __restrict__ aitem,
             const double* __restrict__ acoef, const double* __restrict__ FLUX,
             const double* __restrict__ flen,
             const double* __restrict__ area, const double* __restrict__ W, double ax, double ay,
             StageArgs s, ViscAcc vacc, int* __restrict__ status) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double A = area[i];
  double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
  int e0 = aptr[i], e1 = aptr[i + 1];
  // items in batches of 4: the batch's loads are issued together, the sums
  // still run in item order
  for (int base = e0; base < e1; base += 4) {
    int it[4];
    double cb[4];
    P4 Fb[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool ok = base + q < e1;
      it[q] = ok ? aitem[base + q] : 0;
      cb[q] = ok ? acoef[base + q] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) Fb[q] = (base + q < e1) ? load4(FLUX, it[q] >> 2) : P4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (base + q >= e1) break;
      const int mode = it[q] & 3;
      const P4& F = Fb[q];
      if (mode <= MCFD_ACC_NEIGHBOUR) {
        double c = cb[q];  // (-L)/A or L/A, precomputed like the reference's CSR data
        r0 = r0 + c * F.r;
        r1 = r1 + c * F.u;
        r2 = r2 + c * F.v;
        r3 = r3 + c * F.p;
        continue;
      }
      double L = flen[it[q] >> 2];
      if (mode == MCFD_ACC_BOUNDARY) {
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
  }
  if (vacc.ptr) {  // derivative += _viscous_residual(...) (solver.py:2705-2708)
    double r[4] = {r0, r1, r2, r3};
    visc_gather(i, vacc, flen, A, r);
    r0 = r[0];
    r1 = r[1];
    r2 = r[2];
    r3 = r[3];
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


// ---------------------------------------------------------------------------
// Fused stage: one workgroup evaluates a whole residual + stage update for a
// block of consecutive cells out of LDS.
//   P0  primitive states of own + ring-1 + ring-2 cells  -> LDS
//   P1  LSQ gradient + limiter of own + ring-1 cells     -> LDS
//   P2  flux of every face touching an own cell          -> LDS
//   P3  per own cell: gather in reference order, stage update -> HBM
// Arithmetic is the same as the unfused kernels (same helpers, same order).
// ---------------------------------------------------------------------------

// One block face: geometry, kind | bc << 4 (bc -1: none), left | right << 16
// (local cell ids, 0xFFFF: none). 64 bytes, read with four 16-byte loads.
struct alignas(16) FaceRec {
  double nx, ny, lx, ly, rx, ry, len;
  int kind_bc;
  unsigned lr;
};
static_assert(sizeof(FaceRec) == 64, "face record is 64 bytes");

struct BlockArrays {
  const int* info;          // [nb][kBlockInfoDev]
  const int* cells;         // local -> global cell, block b at b * cstride
  int cstride;              // entries per block in cells (>= max local cells)
  const int* gptr;          // per block, n_grad + 1 offsets into gnb / gent (block relative)
  const unsigned short* gnb;
  const double2* gent;      // LSQ coefficients (cx, cy)
  const int* lptr;          // per block, n_grad + 1 offsets into lent (block relative)
  const double2* lent;      // limiter offsets (dx, dy)
  const int* faces;         // local face -> global flux face
  const FaceRec* frec;
  const unsigned short* acc_local;
  const unsigned short* dts_local;
  // first-order viscous stages: WLSQ weights (wx, wy) of the block's LSQ
  // entries (same order as gent) and, per viscous accumulation item of the
  // global list, its face as a local slot*4 + mode of the owning cell's block
  const double2* vent;
  const unsigned short* vacc_local;
};

struct GeomArrays {
  const double* area;
  const int* grad_ptr;
  const double* grad_cx;
  const double* grad_cy;
  const int* lim_ptr;
  const double* lim_dx;
  const double* lim_dy;
  const int* acc_ptr;
  const double* face_len;
  const double* acc_coef;
};

__device__ __forceinline__ P4 lds4(const double* s, int j) {
  return P4{s[4 * j], s[4 * j + 1], s[4 * j + 2], s[4 * j + 3]};
}

__device__ __forceinline__ P4 rec_lds(const double* sW, const double* sG, const double* sP, int c,
                                      double dx, double dy, bool muscl) {
  P4 w = lds4(sW, c);
  if (!muscl) return w;
  const double* g = sG + 8 * c;
  const double* f = sP + 4 * c;
  P4 o;
  o.r = w.r + f[0] * (g[0] * dx + g[1] * dy);
  o.u = w.u + f[1] * (g[2] * dx + g[3] * dy);
  o.v = w.v + f[2] * (g[4] * dx + g[5] * dy);
  o.p = w.p + f[3] * (g[6] * dx + g[7] * dy);
  return o;
}

typedef __attribute__((address_space(3))) void lds_void_t;

// Copy n items of SIZE bytes from global memory to LDS with direct-to-LDS
// loads (global_load_lds): wave w of the workgroup moves 64-item chunks w,
// w + 4, ...; lane l of chunk c moves item 64c + l (the data never passes
// through VGPRs). Completion is awaited by the s_waitcnt before the next
// barrier.
template <int SIZE>  // 4 or 16
__device__ __forceinline__ void lds_copy(const void* src, void* dst, int n) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int c = w; c * 64 < n; c += kBlock / 64) {
    const int e = c * 64 + lane;
    if (e < n) {
      const void* g = static_cast<const char*>(src) + (size_t)e * SIZE;
      lds_void_t* l = (lds_void_t*)(static_cast<char*>(dst) + (size_t)c * 64 * SIZE);
      if constexpr (SIZE == 16)
        __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
      else
        __builtin_amdgcn_global_load_lds(g, l, 4, 0, 0);
    }
  }
}

// Instantiated per flux scheme, reconstruction, far-field presence and
// next-step dt so that each launch carries only the code it runs (the full
// kernel is ~48 KB, most of a CU pair's instruction cache).
// With VISC (first-order reconstruction only) P1 computes the viscous
// gradients of own and ring-1 cells (k_visc_gradient's arithmetic, from LDS),
// P2 adds each face's viscous flux (k_visc_flux's) and P3 the viscous gather
// (k_update's), so a first-order viscous residual is one launch.
// Waves per SIMD: the MUSCL stages target 6 (78 VGPRs, 80 with the next
// dt; 26 KB of LDS at 128 cells), the viscous stages 5 (96 VGPRs).
template <int SCHEME, bool MUSCL, int LIM, bool FARFIELD, bool DTNEXT, bool VISC = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MUSCL ? 6 : 5, 8)))
    k_stage_fused(int max_acc, int stage_off, int max_local, int max_grad, int max_faces, int max_lsq,
                  BlockArrays ba,
                  GeomArrays ga, FaceArrays fa, DtArrays da, double g, double ax, double ay,
                  StageArgs s, int* __restrict__ status, ViscArrays va, ViscParams vp, ViscAcc vacc) {
  static_assert(!(VISC && (MUSCL || DTNEXT)), "viscous fused stages are first order without the next dt");
  extern __shared__ double smem[];
  constexpr bool muscl = MUSCL;
  const int b = xcd_block();
  const int t = threadIdx.x;
  // The block's local-cell list sits at a fixed stride (>= kBlock, padded with
  // a valid cell id), so local cell t's id is loaded from the block id alone,
  // in the same round trip as the block record, and P0's state loads follow
  // after one trip instead of two.
  const int* __restrict__ cl = ba.cells + (long)ba.cstride * b;
  const int c_first = cl[t];
  const int* inf = ba.info + kBlockInfoDev * b;
  const int n_own = inf[1], n_r1 = inf[2], n_r2 = inf[3];
  const int fs = inf[4], n_faces = inf[5], gs = inf[6], gps = inf[7];
  const int first = inf[8], ls = inf[9];
  const int n_loc = n_own + n_r1 + n_r2;
  const int n_grad = n_own + n_r1;
  // the id is needed now: the wait overlaps the block record's
  asm volatile("" ::"v"(c_first));
  // LDS: W of every local cell, then one region that holds in turn the
  // block's LSQ and limiter entries (MUSCL: copied in at kernel start), the
  // gradients and limiter factors of own + ring-1 cells until the face
  // fluxes are computed, and the face fluxes and lengths afterwards (and the
  // block reduction of the next time step at the end).
  double* sW = smem;
  double* sG = sW + 4 * max_local;
  double* sP = sG + 8 * max_grad;
  double* sF = sG;
  double* sL = sF + 4 * max_faces;
  double* sN = sL + max_faces;  // face normals (final stage with the next step's dt)
  double* sV = sL + max_faces;  // viscous face fluxes (VISC, which has no next-step dt)
  // own cells' accumulation coefficients and items, and spectral-radius
  // items, for P3: written after P2's last read of W and the gradients, over
  // W when they fit there (stage_off 0), else after the shared region
  double* sAC = smem + stage_off;
  unsigned short* sAI = reinterpret_cast<unsigned short*>(sAC + max_acc);
  unsigned short* sDI = sAI + max_acc;
  const double* __restrict__ U = s.u;

  // Index loads of the later phases are issued now, so that their latency
  // overlaps P0 instead of adding a round trip to each phase. Thread t works
  // on gradient cell t in P1, face t in P2 and own cell t in P3 (blocks have
  // at most kBlock of each).
  int e0 = 0, e1 = 0, l0 = 0, l1 = 0;
  if ((MUSCL || VISC) && t < n_grad) {
    e0 = ba.gptr[gps + t];
    e1 = ba.gptr[gps + t + 1];
  }
  if (MUSCL && t < n_grad) {
    l0 = ba.lptr[gps + t];
    l1 = ba.lptr[gps + t + 1];
  }
  // The block's LSQ coefficients and LSQ neighbour ids go straight to LDS
  // while P0 runs, so the gradient loop reads only LDS. The limiter offsets
  // are read from HBM by the limiter loop: with them in LDS the block needs
  // 31 KB at 128 cells and a CU holds 5 workgroups, without them 26 KB and 6.
  const double2* sGE = reinterpret_cast<const double2*>(sG);  // [max_lsq]
  const unsigned short* sGN = reinterpret_cast<const unsigned short*>(sGE + max_lsq) + (gs & 1);
  if constexpr (MUSCL) {
    const int n_lsq = inf[10];
    lds_copy<16>(ba.gent + gs, const_cast<double2*>(sGE), n_lsq);
    lds_copy<4>(ba.gnb + (gs & ~1), const_cast<double2*>(sGE + max_lsq), ((gs & 1) + n_lsq + 1) >> 1);
  }

  // P0: primitive states (euler.py:20-40)
  for (int j = t; j < n_loc; j += kBlock) {
    int c = (j == t) ? c_first : cl[j];
    P4 q = load4(U, c);
    if (s.op == kOpJacobian) {
      // perturbed = s + (eps*sigma_k) * v   (k_perturb, solver.py:1553-1556)
      const P4 dv = load4(s.u1, c);
      const double dn = s.jv_norm[0];
      const double eps = (dn == 0.0) ? 0.0 : s.jv_num / dn;
      q = P4{q.r + eps * s.jv_sigma[0] * dv.r, q.u + eps * s.jv_sigma[1] * dv.u, q.v + eps * s.jv_sigma[2] * dv.v,
             q.p + eps * s.jv_sigma[3] * dv.p};
    }
    double rho = q.r;
    double u = q.u / rho;
    double v = q.v / rho;
    double p = (g - 1.0) * (q.p - 0.5 * (q.u * q.u + q.v * q.v) / rho);
    if (j < n_own) {
      if (!finite(rho) || !finite(u) || !finite(v) || !finite(p) || !(rho > 0.0) || !(p > 0.0))
        report(status, MCFD_ERR_STATE, c);
    }
    sW[4 * j] = rho;
    sW[4 * j + 1] = u;
    sW[4 * j + 2] = v;
    sW[4 * j + 3] = p;
  }
  __syncthreads();

  // P1: LSQ gradient (_kernels.py:252-262) + limiter (_kernels.py:9-67)
  // P2's face record: read from HBM while the block waits at P1's barriers.
  const int f = t;
  FaceRec fr{};
  if constexpr (MUSCL) {
    double gx[4] = {0.0, 0.0, 0.0, 0.0};
    double gy[4] = {0.0, 0.0, 0.0, 0.0};
    double phi[4] = {1.0, 1.0, 1.0, 1.0};
    if (t < n_grad) {
      const int j = t;
      double wv[4] = {sW[4 * j], sW[4 * j + 1], sW[4 * j + 2], sW[4 * j + 3]};
      double mn[4] = {wv[0], wv[1], wv[2], wv[3]};
      double mx[4] = {wv[0], wv[1], wv[2], wv[3]};
      const unsigned short* nbl = sGN;
      const double2* cxy = sGE;
      // entries in batches of 4: the batch's loads are issued together, the
      // sums still run in entry order
      for (int base = e0; base < e1; base += 4) {
        int ob[4];
        double cxb[4], cyb[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int e = base + q;
          bool ok = e < e1;
          double2 c2 = ok ? cxy[e] : make_double2(0.0, 0.0);
          ob[q] = ok ? (int)nbl[e] : 0;
          cxb[q] = c2.x;
          cyb[q] = c2.y;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (base + q >= e1) break;
          int o = ob[q];
          double ov[4] = {sW[4 * o], sW[4 * o + 1], sW[4 * o + 2], sW[4 * o + 3]};
          double cx = cxb[q], cy = cyb[q];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            double d = ov[k] - wv[k];
            gx[k] = gx[k] + cx * d;
            gy[k] = gy[k] + cy * d;
            if (ov[k] < mn[k]) mn[k] = ov[k];
            if (ov[k] > mx[k]) mx[k] = ov[k];
          }
        }
      }
      double dmx[4], dmn[4], tol[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        dmx[k] = mx[k] - wv[k];
        dmn[k] = mn[k] - wv[k];
        tol[k] = 1e-14 * vmax(fabs(wv[k]), 1.0);
      }
      // offsets in batches of 3, of 2 for limiter 3 (larger batches make the
      // 6-wave stages spill)
      const double2* __restrict__ lxy = ba.lent + ls;
      constexpr int kLimBatch = (LIM == 3) ? 2 : 3;
      for (int base = l0; base < l1; base += kLimBatch) {
        double dxb[kLimBatch], dyb[kLimBatch];
#pragma unroll
        for (int q = 0; q < kLimBatch; ++q) {
          int e = base + q;
          bool ok = e < l1;
          double2 d2 = ok ? lxy[e] : make_double2(0.0, 0.0);
          dxb[q] = d2.x;
          dyb[q] = d2.y;
        }
#pragma unroll
        for (int q = 0; q < kLimBatch; ++q) {
        if (base + q >= l1) break;
        double dx = dxb[q], dy = dyb[q];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          limiter_entry_sel<LIM>(gx[k] * dx + gy[k] * dy, dmx[k], dmn[k], tol[k], phi[k]);
        }
        }
      }
    }
    if (f < n_faces) fr = ba.frec[fs + f];
    // the gradients and limiter factors overwrite the entries
    __syncthreads();
    if (t < n_grad) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        sG[8 * t + 2 * k] = gx[k];
        sG[8 * t + 2 * k + 1] = gy[k];
        sP[4 * t + k] = phi[k];
      }
    }
    __syncthreads();
  }
  // P1 (viscous): WLSQ gradients of (u, v, T) of own and ring-1 cells
  // (k_visc_gradient: face entries in LSQ order, then the ghost entries),
  // viscosity and conductivity -> sG[8 j .. 8 j + 7]
  if constexpr (VISC) {
    if (t < n_grad) {
      const int i = cl[t];
      const P4 w = lds4(sW, t);
      const double T = w.p / (w.r * vp.R);
      if (t < n_own && (!finite(T) || !(T > 0.0))) report(status, MCFD_ERR_TEMPERATURE, i);
      double val[3] = {w.u, w.v, T};
      double rx[3] = {0.0, 0.0, 0.0}, ry[3] = {0.0, 0.0, 0.0};
      const unsigned short* nbl = ba.gnb + gs;
      const double2* went = ba.vent + gs;
      for (int base = e0; base < e1; base += 4) {
        int ob[4];
        double2 wb[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const bool ok = base + q < e1;
          ob[q] = ok ? (int)nbl[base + q] : t;
          wb[q] = ok ? went[base + q] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (base + q >= e1) break;
          const P4 o = lds4(sW, ob[q]);
          double d[3] = {o.u - w.u, o.v - w.v, o.p / (o.r * vp.R) - T};
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            rx[k] = rx[k] + wb[q].x * d[k];
            ry[k] = ry[k] + wb[q].y * d[k];
          }
        }
      }
      double out[8];
      visc_gradient_finish(i, w, val, rx, ry, va, fa, vp, status, out);
#pragma unroll
      for (int k = 0; k < 8; ++k) sG[8 * t + k] = out[k];
    }
    __syncthreads();
  }

  // P2: face fluxes, one face per thread (a block touches at most kBlock
  // faces); the results wait in registers until every thread is done with
  // the gradient slots they overwrite.
  F4 fl{0.0, 0.0, 0.0, 0.0};
  double flen = 0.0, fnx = 0.0, fny = 0.0;
  double vf[4] = {0.0, 0.0, 0.0, 0.0};  // viscous flux (VISC)
  if (f < n_faces) {
    if constexpr (!MUSCL) fr = ba.frec[fs + f];
    const int kind = fr.kind_bc & 15;
    const int l = fr.lr & 0xFFFF, r = fr.lr >> 16;
    const double nx = fr.nx, ny = fr.ny;
    fnx = nx;
    fny = ny;
    P4 L = rec_lds(sW, sG, sP, l, fr.lx, fr.ly, muscl);
    P4 R;
    int err = 0;
    if (kind <= MCFD_FACE_PERIODIC) {
      R = rec_lds(sW, sG, sP, r, fr.rx, fr.ry, muscl);
    } else {
      int bi = fr.kind_bc >> 4;
      const double* st = fa.bc_state + 4 * (long)bi;
      if (kind == MCFD_FACE_WALL) {
        double rx = L.u - fa.bc_wall[2 * (long)bi];
        double ry = L.v - fa.bc_wall[2 * (long)bi + 1];
        double un = rx * nx + ry * ny;
        R = L;
        R.u = L.u - 2.0 * un * nx;
        R.v = L.v - 2.0 * un * ny;
      } else if (kind == MCFD_FACE_VELOCITY_INLET) {
        R = P4{st[0], st[1], st[2], st[3]};
      } else if (kind == MCFD_FACE_PRESSURE_OUTLET) {
        R = L;
        R.p = st[3];
      } else if (FARFIELD && kind == MCFD_FACE_FAR_FIELD) {
        R = far_field(L, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, &err);
      } else {
        R = L;
      }
    }
    if constexpr (SCHEME == 1) {
      fl = hllc(L, R, nx, ny, g, kind == MCFD_FACE_INTERIOR);
      if (kind != MCFD_FACE_INTERIOR &&
          !(finite(fl.a) && finite(fl.b) && finite(fl.c) && finite(fl.d)))
        err = MCFD_ERR_HLLC;
    } else if constexpr (SCHEME == 2) {
      fl = roe(L, R, nx, ny, g, &err);
      if (!err && !(finite(fl.a) && finite(fl.b) && finite(fl.c) && finite(fl.d)))
        err = MCFD_ERR_ROE;
    } else {
      fl = rusanov(L, R, nx, ny, g);
    }
    if (err) report(status, err, ba.faces[fs + f]);
    flen = fr.len;
    if constexpr (VISC) {
      // k_visc_flux: face averages on interior and periodic faces, the
      // boundary corrections elsewhere
      const double* go = sG + 8 * l;
      if (kind <= MCFD_FACE_PERIODIC) {
        const double* gr = sG + 8 * r;
        double gq[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) gq[k] = 0.5 * (go[k] + gr[k]);
        const P4 a = lds4(sW, l), c = lds4(sW, r);
        visc_flux(gq, 0.5 * (a.u + c.u), 0.5 * (a.v + c.v), 0.5 * (go[6] + gr[6]), 0.5 * (go[7] + gr[7]), nx, ny,
                  vf);
      } else {
        visc_boundary_flux(fr.kind_bc >> 4, nx, ny, fr.lx, fr.ly, lds4(sW, l), go, va, fa, vp, status, cl[l], vf);
      }
    }
  }
  // P3's per-cell loads, issued before the barrier. P3 runs two threads per
  // own cell: thread t handles components 2*(t&1) and 2*(t&1)+1 of own cell
  // t>>1 (blocks have at most kBlock/2 own cells).
  const int j3 = t >> 1, hf = t & 1;
  int a0 = 0, a1 = 0, d0 = 0, d1 = 0;
  double A = 1.0, h = 0.0;
  double2 u2 = make_double2(0.0, 0.0), z2 = make_double2(0.0, 0.0);
  if (j3 < n_own) {
    const int i = first + j3;
    a0 = ga.acc_ptr[i];
    a1 = ga.acc_ptr[i + 1];
    A = ga.area[i];
    if (DTNEXT && hf == 0) {
      d0 = da.ptr[i];
      d1 = da.ptr[i + 1];
    }
    if (s.op != MCFD_OP_RESIDUAL) {
      h = s.dt_per_cell ? s.dt[i] : s.dt[0];
      // the Jacobian op: u2 = direction v, z2 = rate
      u2 = reinterpret_cast<const double2*>((s.op == kOpJacobian ? s.u1 : U) + 4 * (long)i)[hf];
      if (s.op != MCFD_OP_EULER) z2 = reinterpret_cast<const double2*>(s.u0 + 4 * (long)i)[hf];
    }
  }
  // The block's accumulation (and spectral-radius) items are contiguous in
  // the global lists: loaded here with coalesced loads (up to kStageItems
  // per thread in registers), stored to LDS after the barrier, so that P3's
  // loop reads LDS instead of a chain of dependent global loads.
  constexpr int kStageItems = 3;
  const int abase = inf[12];
  const int n_acc = inf[13];
  double ac_r[kStageItems];
  unsigned short ai_r[kStageItems];
#pragma unroll
  for (int q = 0; q < kStageItems; ++q) {
    const int k = t + q * kBlock;
    ac_r[q] = (k < n_acc) ? ga.acc_coef[abase + k] : 0.0;
    ai_r[q] = (k < n_acc) ? ba.acc_local[abase + k] : 0;
  }
  int dbase = 0, n_dts = 0;
  unsigned short di_r[kStageItems];
  if (DTNEXT) {
    dbase = inf[14];
    n_dts = inf[15];
#pragma unroll
    for (int q = 0; q < kStageItems; ++q) {
      const int k = t + q * kBlock;
      di_r[q] = (k < n_dts) ? ba.dts_local[dbase + k] : 0;
    }
  }
  // own cell state for the body force (W is overwritten by the staging)
  double bw_r = 0.0, bw_u = 0.0, bw_v = 0.0;
  if ((ax != 0.0 || ay != 0.0) && j3 < n_own) {
    bw_r = sW[4 * j3];
    bw_u = sW[4 * j3 + 1];
    bw_v = sW[4 * j3 + 2];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kStageItems; ++q) {
    const int k = t + q * kBlock;
    if (k < n_acc) {
      sAC[k] = ac_r[q];
      sAI[k] = ai_r[q];
    }
    if (DTNEXT && k < n_dts) sDI[k] = di_r[q];
  }
  for (int k = t + kStageItems * kBlock; k < n_acc; k += kBlock) {
    sAC[k] = ga.acc_coef[abase + k];
    sAI[k] = ba.acc_local[abase + k];
  }
  if (DTNEXT)
    for (int k = t + kStageItems * kBlock; k < n_dts; k += kBlock) sDI[k] = ba.dts_local[dbase + k];
  if (f < n_faces) {
    sF[4 * f] = fl.a;
    sF[4 * f + 1] = fl.b;
    sF[4 * f + 2] = fl.c;
    sF[4 * f + 3] = fl.d;
    sL[f] = flen;
    if (DTNEXT) {
      sN[2 * f] = fnx;
      sN[2 * f + 1] = fny;
    }
    if constexpr (VISC) {
#pragma unroll
      for (int k = 0; k < 4; ++k) sV[4 * f + k] = vf[k];
    }
  }
  __syncthreads();

  // P3: gather (solver.py:2501-2704) + body force + stage update, two
  // components per thread in the reference order of each component
  double dt_local = INFINITY;
  if (j3 < n_own) {
    const int i = first + j3;
    double ra = 0.0, rb = 0.0;
    for (int e = a0; e < a1; ++e) {
      int item = sAI[e - abase];
      int slot = item >> 2;
      int mode = item & 3;
      double2 F = reinterpret_cast<const double2*>(sF + 4 * slot)[hf];
      if (mode <= MCFD_ACC_NEIGHBOUR) {
        double c = sAC[e - abase];  // (-L)/A or L/A, precomputed like the reference's CSR data
        ra = ra + c * F.x;
        rb = rb + c * F.y;
        continue;
      }
      double L = sL[slot];
      if (mode == MCFD_ACC_BOUNDARY) {
        ra = ra + (-(F.x * L)) / A;
        rb = rb + (-(F.y * L)) / A;
      } else {
        ra = ra + (F.x * L) / A;
        rb = rb + (F.y * L) / A;
      }
    }
    if constexpr (VISC) {
      // k_update's viscous gather (visc_gather): 0 - S over the interior
      // items, then the periodic and boundary items
      double s0 = 0.0, s1 = 0.0, dv0 = 0.0, dv1 = 0.0;
      bool interior = true;
      const int v0 = vacc.ptr[i], v1 = vacc.ptr[i + 1];
      for (int e = v0; e < v1; ++e) {
        const int it = ba.vacc_local[e];
        const int slot = it >> 2, mode = it & 3;
        const double2 F = reinterpret_cast<const double2*>(sV + 4 * slot)[hf];
        if (interior && mode == 0) {
          const double c = vacc.coef[e];
          s0 = s0 + c * F.x;
          s1 = s1 + c * F.y;
          continue;
        }
        if (interior) {
          dv0 = 0.0 - s0;
          dv1 = 0.0 - s1;
          interior = false;
        }
        const double L = sL[slot];
        const double x0 = F.x * L, x1 = F.y * L;
        if (mode == 1) {
          dv0 = dv0 + x0 / A;
          dv1 = dv1 + x1 / A;
        } else {
          dv0 = dv0 + (-x0) / A;
          dv1 = dv1 + (-x1) / A;
        }
      }
      if (interior) {
        dv0 = 0.0 - s0;
        dv1 = 0.0 - s1;
      }
      ra = ra + dv0;
      rb = rb + dv1;
    }
    if (ax != 0.0 || ay != 0.0) {
      if (hf == 0) {
        rb = rb + bw_r * ax;
      } else {
        ra = ra + bw_r * ay;
        rb = rb + bw_r * (ax * bw_u + ay * bw_v);
      }
    }
    if (!(finite(ra) && finite(rb))) report(status, MCFD_ERR_RESIDUAL, i);
    if (s.r_out) reinterpret_cast<double2*>(s.r_out + 4 * (long)i)[hf] = make_double2(ra, rb);
    double oa = 0.0, ob = 0.0;
    if (s.op == MCFD_OP_EULER) {
      double k = s.c0 * h;
      oa = u2.x + k * ra;
      ob = u2.y + k * rb;
    } else if (s.op == MCFD_OP_SSP) {
      oa = s.c0 * z2.x + s.c1 * (u2.x + h * ra);
      ob = s.c0 * z2.y + s.c1 * (u2.y + h * rb);
    } else if (s.op == MCFD_OP_RK4) {
      double k = s.c2 * h;
      oa = s.c0 * z2.x + s.c1 * u2.x + k * ra;
      ob = s.c0 * z2.y + s.c1 * u2.y + k * rb;
    } else if (s.op == kOpJacobian) {
      // out = v - (h * ((rp - r) / eps)) / sigma   (k_jv_finish, solver.py:1557-1563)
      const double dn = s.jv_norm[0];
      if (dn == 0.0) {
        oa = u2.x;
        ob = u2.y;
      } else {
        const double eps = s.jv_num / dn;
        oa = u2.x - h * ((ra - z2.x) / eps) / s.jv_sigma[2 * hf];
        ob = u2.y - h * ((rb - z2.y) / eps) / s.jv_sigma[2 * hf + 1];
      }
    } else if (s.op == MCFD_OP_RK4_FINAL) {
      double2 y = reinterpret_cast<const double2*>(s.u1 + 4 * (long)i)[hf];
      double2 q = reinterpret_cast<const double2*>(s.r1 + 4 * (long)i)[hf];
      double k2 = s.c2 * h, k4 = s.c4 * h;
      oa = s.c0 * z2.x + s.c1 * y.x + k2 * q.x + s.c3 * u2.x + k4 * ra;
      ob = s.c0 * z2.y + s.c1 * y.y + k2 * q.y + s.c3 * u2.y + k4 * rb;
    }
    if (s.op != MCFD_OP_RESIDUAL)
      reinterpret_cast<double2*>(s.out + 4 * (long)i)[hf] = make_double2(oa, ob);
    if (DTNEXT && s.op != MCFD_OP_RESIDUAL) {
      // the pair's other half of the updated state
      double pa = __shfl_xor(oa, 1), pb = __shfl_xor(ob, 1);
      if (hf == 0) {
        // next step's stable step from the updated state (euler.py:20-40 + solver.py:4132-4343)
        double o0 = oa, o1 = ob, o2 = pa, o3 = pb;
        double rho = o0;
        double uu = o1 / rho;
        double vv = o2 / rho;
        double p = (g - 1.0) * (o3 - 0.5 * (o1 * o1 + o2 * o2) / rho);
        if (!finite(rho) || !finite(uu) || !finite(vv) || !finite(p) || !(rho > 0.0) || !(p > 0.0))
          report(status, MCFD_ERR_STATE, i);
        double dt = block_time_step<FARFIELD>(i, A, P4{rho, uu, vv, p}, d0, d1, sDI, dbase, sN, sL, da, g,
                                    s.cfl, status);
        if (s.dt_next) s.dt_next[i] = dt;
        dt_local = dt;
      }
    }
  }
  if (DTNEXT) {
    // block minimum (wave minimum by lane exchanges, then over the four
    // waves) into the per-block partial; k_min_reduce folds the partials
    __shared__ double wave_min[kBlock / 64];
    double v = dt_local;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    if ((t & 63) == 0) wave_min[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
      double m = wave_min[0];
#pragma unroll
      for (int w = 1; w < kBlock / 64; ++w) m = fmin(m, wave_min[w]);
      s.dt_partial[b] = m;
    }
  }
}


// ---------------------------------------------------------------------------
// Implicit Euler: LU-SGS preconditioner (solver.py:1185-1323), matrix-free
// Jacobian (solver.py:1532-1567) and BLAS-1 helpers for GMRES.
// ---------------------------------------------------------------------------

struct LusgsArrays {
  int n_rec;
  const int* rec_o;      // device cell on the owner side
  const int* rec_n;      // device cell on the neighbour side, -1 for boundary records
  const double* nox;     // normal used with the owner-side state
  const double* noy;
  const double* nnx;     // normal used with the neighbour-side state
  const double* nny;
  const double* len;
  const int* diag_ptr;   // per cell: diagonal items in reference order
  const int* diag_item;  // record*2 + side; -(record*2 + side) - 1: a self-paired periodic entry
  const int* off_ptr;    // per cell: off-diagonal items, lower (natural order) first
  const int* off_split;  // per cell: first upper item
  const int* off_item;   // record*2 + side
  const int* off_nb;     // device cell of the off-diagonal column
  const int* off_t;      // the column cell's item for the same record (transposed coefficient)
  int n_levels_f, n_levels_b;
  const int* lvl_ptr_f;
  const int* lvl_cells_f;
  const int* lvl_ptr_b;
  const int* lvl_cells_b;
  double* coef;          // per record
  double* diag;          // per cell
  double* off_val;       // per off item
  // The sweeps' data in level-position order (position k of a sweep is
  // lvl_cells[k]): per position the number of items, and the first kf / kb
  // items' column cells and coefficients (normal and transposed), padded;
  // the diagonal in position order. Filled by k_lusgs_pack after assembly.
  int kf, kb;
  const int* pcnt_f;
  const int* pcnt_b;
  const int* pnb_f;      // [n * kf] column cell, -1 padding
  const int* pnb_b;      // [n * kb]
  const int* pit_f;      // [n * kf] the off item, -1 padding
  const int* pit_b;
  double* pval_f;        // [n * kf] coefficients
  double* pval_ft;       // transposed coefficients
  double* pval_b;
  double* pval_bt;
  double* pdiag_f;       // [n]
  double* pdiag_b;
  // The one-workgroup sweeps' own values in their level-position order, so
  // a position's data never waits on its cell index: own_f[k] = v at the
  // forward's position k (gathered before the sweep), own_b[k] = y at the
  // backward's position k (written by the forward sweep through fb, the
  // backward position of each forward position's cell).
  const int* fb;         // [n]
  double* own_f;         // [n * 4]
  double* own_b;         // [n * 4]
};

// Per record: 0.5*max(|u_o.n_o|+c_o, |u_n.n_n|+c_n)*L for face pairs, the
// owner's |u.n|+c for boundary records (solver.py:1214-1299).
__global__ void __launch_bounds__(kBlock)
    k_lusgs_coef(LusgsArrays la, const double* __restrict__ W, double g) {
  int r = xcd_block() * blockDim.x + threadIdx.x;
  if (r >= la.n_rec) return;
  P4 wo = load4(W, la.rec_o[r]);
  double ro = fabs(wo.u * la.nox[r] + wo.v * la.noy[r]) + sound(wo, g);
  int n = la.rec_n[r];
  if (n < 0) {
    la.coef[r] = ro;
    return;
  }
  P4 wn = load4(W, n);
  double rn = fabs(wn.u * la.nnx[r] + wn.v * la.nny[r]) + sound(wn, g);
  la.coef[r] = 0.5 * vmax(ro, rn) * la.len[r];
}

__global__ void __launch_bounds__(kBlock)
    k_lusgs_assemble(int n, LusgsArrays la, const double* __restrict__ area,
                     const double* __restrict__ dt, int dt_per_cell) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double h = dt_per_cell ? dt[i] : dt[0];
  double A = area[i];
  double d = 1.0;
  for (int e = la.diag_ptr[i]; e < la.diag_ptr[i + 1]; ++e) {
    const int it = la.diag_item[e];
    if (it < 0) {  // a periodic pair whose faces share this owner: its off-diagonal entry
      d = d + (-(h * la.coef[(-it - 1) >> 1]) / A);
      continue;
    }
    int r = it >> 1;
    if (la.rec_n[r] < 0)
      d = d + h * la.coef[r] * la.len[r] / A;
    else
      d = d + h * la.coef[r] / A;
  }
  la.diag[i] = d;
  for (int e = la.off_ptr[i]; e < la.off_ptr[i + 1]; ++e) {
    int r = la.off_item[e] >> 1;
    la.off_val[e] = -(h * la.coef[r]) / A;
  }
}

// Coefficients and diagonal in the sweeps' position order (after assembly).
__global__ void __launch_bounds__(kBlock) k_lusgs_pack(int n, LusgsArrays la) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  la.pdiag_f[k] = la.diag[la.lvl_cells_f[k]];
  la.pdiag_b[k] = la.diag[la.lvl_cells_b[k]];
  for (int q = 0; q < la.kf; ++q) {
    int e = la.pit_f[(long)k * la.kf + q];
    la.pval_f[(long)k * la.kf + q] = e >= 0 ? la.off_val[e] : 0.0;
    la.pval_ft[(long)k * la.kf + q] = e >= 0 ? la.off_val[la.off_t[e]] : 0.0;
  }
  for (int q = 0; q < la.kb; ++q) {
    int e = la.pit_b[(long)k * la.kb + q];
    la.pval_b[(long)k * la.kb + q] = e >= 0 ? la.off_val[e] : 0.0;
    la.pval_bt[(long)k * la.kb + q] = e >= 0 ? la.off_val[la.off_t[e]] : 0.0;
  }
}

// Level-scheduled triangular sweeps in one workgroup: every cell of a level
// depends only on cells of earlier levels, so the sweep equals the
// sequential natural-order substitution. Forward: y = L^-1 v. Backward (in
// place): x = U^-1 (max(diag, 1e-14) * y).
// With te (transposed application M^T = L^-T D U^-T): the forward sweep
// solves U^T x = v and the backward sweep L^T z = D x, over the same lower /
// upper items with the coefficient of the column cell's item for the same
// record, S[j][i] = off_val[te[e]]. Same levels: S is structurally symmetric.
//
// The per-position data (cell, item count, diagonal, own value, the first
// kLusgsK items' column cells and coefficients) of the next level do not
// depend on the sweep, so they are loaded while the current level computes
// and waits at the barrier; a level then costs one round trip for the
// column values written by earlier levels. Thread t owns positions a + t and
// a + t + 1024 of a level in registers; further positions (levels of more
// than 2048 cells) and items beyond kLusgsK take the direct path.
constexpr int kLusgsK = 4;
constexpr int kLusgsSlots = 2;
// average cells per level from which the sweeps launch a grid per level
constexpr int kLusgsWideLevel = 4096;

struct LusgsSlot {
  int cell, cnt, fb;
  double d;
  P4 v;
  int nb[kLusgsK];
  double c[kLusgsK];
};

// POS: own is in the sweep's position order (indexed by k, not by the cell),
// and the forward slot also carries its cell's backward position.
template <bool FORWARD, bool POS = false>
__device__ __forceinline__ void lusgs_fetch(const LusgsArrays& la, int k, int K, const int* cells,
                                            const int* cnt, const int* pnb, const double* pval,
                                            const double* pdiag, const double* own, LusgsSlot& sl) {
  sl.cell = cells[k];
  sl.cnt = cnt[k];
  sl.d = pdiag[k];
  sl.v = POS ? load4(own, k) : load4(own, sl.cell);
  sl.fb = (POS && FORWARD) ? la.fb[k] : 0;
#pragma unroll
  for (int q = 0; q < kLusgsK; ++q) {
    bool ok = q < K;
    sl.nb[q] = ok ? pnb[(long)k * K + q] : -1;
    sl.c[q] = ok ? pval[(long)k * K + q] : 0.0;
  }
}

// y_j = v_j / d_j: the value a cell of the first forward level (no lower
// items) stores, formed where it is read by the two-level sweep
__device__ __forceinline__ P4 lusgs_first_level(const LusgsArrays& la, const double* v, int j) {
  const P4 o = load4(v, j);
  const double d = la.diag[j];
  return P4{o.r / d, o.u / d, o.v / d, o.p / d};
}

// One position of a sweep: (own - sum c * src[column]) / d forward,
// (max(d, 1e-14) own - sum c * src[column]) / d backward. ONFLY: the column
// values are the first forward level's, formed from src = v.
template <bool FORWARD, bool ONFLY = false>
__device__ __forceinline__ P4 lusgs_solve(const LusgsArrays& la, const LusgsSlot& sl, int K, const double* src,
                                          const int* te) {
  P4 s;
  if (FORWARD) {
    s = sl.v;
  } else {
    double dm = vmax(sl.d, 1e-14);
    s = P4{dm * sl.v.r, dm * sl.v.u, dm * sl.v.v, dm * sl.v.p};
  }
  const int lim = K < kLusgsK ? K : kLusgsK;
#pragma unroll
  for (int q = 0; q < kLusgsK; ++q) {
    if (q >= lim || q >= sl.cnt) break;
    double c = sl.c[q];
    P4 o = ONFLY ? lusgs_first_level(la, src, sl.nb[q]) : load4(src, sl.nb[q]);
    s.r = s.r - c * o.r;
    s.u = s.u - c * o.u;
    s.v = s.v - c * o.v;
    s.p = s.p - c * o.p;
  }
  if (sl.cnt > lim) {  // items beyond the packed ones, in item order
    const int i = sl.cell;
    const int e0 = (FORWARD ? la.off_ptr[i] : la.off_split[i]) + lim;
    const int e1 = FORWARD ? la.off_split[i] : la.off_ptr[i + 1];
    for (int e = e0; e < e1; ++e) {
      double c = te ? la.off_val[te[e]] : la.off_val[e];
      P4 o = ONFLY ? lusgs_first_level(la, src, la.off_nb[e]) : load4(src, la.off_nb[e]);
      s.r = s.r - c * o.r;
      s.u = s.u - c * o.u;
      s.v = s.v - c * o.v;
      s.p = s.p - c * o.p;
    }
  }
  const double d = sl.d;
  return P4{s.r / d, s.u / d, s.v / d, s.p / d};
}

template <bool FORWARD>
__device__ __forceinline__ void lusgs_finish(const LusgsArrays& la, const LusgsSlot& sl, int K,
                                             const double* src, double* dst, const int* te) {
  const P4 o = lusgs_solve<FORWARD>(la, sl, K, src, te);
  store4(dst, sl.cell, o.r, o.u, o.v, o.p);
}

// own_f[k] = v[cell of the forward's position k]
__global__ void __launch_bounds__(kBlock)
    k_lusgs_gather_own(int n, LusgsArrays la, const double* __restrict__ v) {
  const int k = xcd_block() * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const P4 o = load4(v, la.lvl_cells_f[k]);
  store4(la.own_f, k, o.r, o.u, o.v, o.p);
}

// Forward: own = own_f (v by position), src = dst = y, and y by backward
// position into own_b. Backward: own = own_b, src = dst = x (in place).
template <bool FORWARD>
__device__ __forceinline__ void lusgs_sweep_finish(const LusgsArrays& la, const LusgsSlot& sl, int K, double* y,
                                                   const int* te) {
  const P4 o = lusgs_solve<FORWARD>(la, sl, K, y, te);
  store4(y, sl.cell, o.r, o.u, o.v, o.p);
  if (FORWARD) store4(la.own_b, sl.fb, o.r, o.u, o.v, o.p);
}

template <bool FORWARD>
__global__ void __launch_bounds__(1024)
    k_lusgs_sweep(LusgsArrays la, double* y, const int* __restrict__ te) {
  const double* own = FORWARD ? la.own_f : la.own_b;
  const int nl = FORWARD ? la.n_levels_f : la.n_levels_b;
  const int* ptr = FORWARD ? la.lvl_ptr_f : la.lvl_ptr_b;
  const int* cells = FORWARD ? la.lvl_cells_f : la.lvl_cells_b;
  const int* cnt = FORWARD ? la.pcnt_f : la.pcnt_b;
  const int* pnb = FORWARD ? la.pnb_f : la.pnb_b;
  const int K = FORWARD ? la.kf : la.kb;
  const double* pval = FORWARD ? (te ? la.pval_ft : la.pval_f) : (te ? la.pval_bt : la.pval_b);
  const double* pdiag = FORWARD ? la.pdiag_f : la.pdiag_b;
  const int t = threadIdx.x, T = blockDim.x;
  LusgsSlot nxt[kLusgsSlots];
  int a = ptr[0], b = ptr[1];
#pragma unroll
  for (int q = 0; q < kLusgsSlots; ++q)
    if (a + t + q * T < b)
      lusgs_fetch<FORWARD, true>(la, a + t + q * T, K, cells, cnt, pnb, pval, pdiag, own, nxt[q]);
  for (int lev = 0; lev < nl; ++lev) {
    LusgsSlot cur[kLusgsSlots];
#pragma unroll
    for (int q = 0; q < kLusgsSlots; ++q) cur[q] = nxt[q];
    const int ca = a, cb = b;
    if (lev + 1 < nl) {
      a = cb;
      b = ptr[lev + 2];
#pragma unroll
      for (int q = 0; q < kLusgsSlots; ++q)
        if (a + t + q * T < b)
          lusgs_fetch<FORWARD, true>(la, a + t + q * T, K, cells, cnt, pnb, pval, pdiag, own, nxt[q]);
    }
#pragma unroll
    for (int q = 0; q < kLusgsSlots; ++q)
      if (ca + t + q * T < cb) lusgs_sweep_finish<FORWARD>(la, cur[q], K, y, te);
    for (int k = ca + t + kLusgsSlots * T; k < cb; k += T) {
      LusgsSlot sl;
      lusgs_fetch<FORWARD, true>(la, k, K, cells, cnt, pnb, pval, pdiag, own, sl);
      lusgs_sweep_finish<FORWARD>(la, sl, K, y, te);
    }
    // workgroup-scope release / acquire: the one workgroup's waves share
    // the CU's write-through L1, so the next level reads these values (an
    // agent-scope fence would write L2 back at every level)
    __syncthreads();
  }
}

// One level of a sweep across the whole GPU (orderings with few, large
// levels, such as the multicolour one): a thread per position; the levels
// are launched in order on the stream.
template <bool FORWARD>
__global__ void __launch_bounds__(kBlock)
    k_lusgs_level(LusgsArrays la, const double* __restrict__ own, double* y, const int* __restrict__ te, int a,
                  int b) {
  const int k = a + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b) return;
  const int* cells = FORWARD ? la.lvl_cells_f : la.lvl_cells_b;
  const int* cnt = FORWARD ? la.pcnt_f : la.pcnt_b;
  const int* pnb = FORWARD ? la.pnb_f : la.pnb_b;
  const int K = FORWARD ? la.kf : la.kb;
  const double* pval = FORWARD ? (te ? la.pval_ft : la.pval_f) : (te ? la.pval_bt : la.pval_b);
  const double* pdiag = FORWARD ? la.pdiag_f : la.pdiag_b;
  LusgsSlot sl;
  lusgs_fetch<FORWARD>(la, k, K, cells, cnt, pnb, pval, pdiag, own, sl);
  lusgs_finish<FORWARD>(la, sl, K, y, y, te);
}

// Sweeps of two levels whose forward levels are the backward levels swapped
// (a 2-colouring): two launches instead of four, same bits. The first runs
// the forward's second level, reading its lower neighbours' first-level
// values as v / d, and the backward's first level on the result (no upper
// items: x = (max(d, 1e-14) y) / d); the second runs the backward's second
// level with its own first-level value v / d.
__global__ void __launch_bounds__(kBlock)
    k_lusgs_two_first(LusgsArrays la, const double* __restrict__ v, double* x, const int* __restrict__ te, int a,
                      int b) {
  const int k = a + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b) return;
  const double* pval = te ? la.pval_ft : la.pval_f;
  LusgsSlot sl;
  lusgs_fetch<true>(la, k, la.kf, la.lvl_cells_f, la.pcnt_f, la.pnb_f, pval, la.pdiag_f, v, sl);
  const P4 y = lusgs_solve<true, true>(la, sl, la.kf, v, te);
  const double d = sl.d, dm = vmax(d, 1e-14);
  store4(x, sl.cell, (dm * y.r) / d, (dm * y.u) / d, (dm * y.v) / d, (dm * y.p) / d);
}

__global__ void __launch_bounds__(kBlock)
    k_lusgs_two_second(LusgsArrays la, const double* __restrict__ v, double* x, const int* __restrict__ te, int a,
                       int b) {
  const int k = a + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= b) return;
  const double* pval = te ? la.pval_bt : la.pval_b;
  LusgsSlot sl;
  lusgs_fetch<false>(la, k, la.kb, la.lvl_cells_b, la.pcnt_b, la.pnb_b, pval, la.pdiag_b, v, sl);
  const double d = sl.d;
  sl.v = P4{sl.v.r / d, sl.v.u / d, sl.v.v / d, sl.v.p / d};
  lusgs_finish<false>(la, sl, la.kb, x, x, te);
}

// Deterministic dot product: fixed-shape block partials, then one block.
__global__ void __launch_bounds__(kBlock)
    k_dot_partial(long n, const double* __restrict__ a, const double* __restrict__ b,
                  double* __restrict__ partial) {
  __shared__ double red[kBlock];
  double acc = 0.0;
  long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += a[i] * b[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Block b reduces partial[b*nb .. b*nb+nb) into out[b]; bit b of sqrt_mask
// takes the square root.
// Workgroups of the deterministic reductions (fixed, so every dot has one
// order whatever the mesh)
constexpr int kDotBlocks = 1024;
// elements per thread loaded ahead of a kernel's leading reduction
constexpr int kPrefetch = 16;

// Sum of nb block partials, returned to every thread of a kBlock workgroup:
// thread t adds part[t], part[t + kBlock], ... in order, then a fixed tree.
// Every reduction of dot partials goes through this one order, so a result
// is the same bits whichever kernel finishes it.
__device__ double block_sum(const double* __restrict__ part, int nb, double* red) {
  double acc = 0.0;
  for (int j = threadIdx.x; j < nb; j += kBlock) acc += part[j];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double v = red[0];
  __syncthreads();  // red is free for the caller again
  return v;
}

__global__ void __launch_bounds__(kBlock)
    k_dot_final(int nb, const double* __restrict__ partial, double* __restrict__ out, int sqrt_mask,
                double* __restrict__ host_out = nullptr) {
  __shared__ double red[kBlock];
  const double sum = block_sum(partial + (long)blockIdx.x * nb, nb, red);
  if (threadIdx.x == 0) {
    double v = ((sqrt_mask >> blockIdx.x) & 1) ? sqrt(sum) : sum;
    out[blockIdx.x] = v;
    // pinned host mirror: the host reads it after synchronising the stream,
    // without a copy launch
    if (host_out) host_out[blockIdx.x] = v;
  }
}

// (w, w) and (a, w) in one pass: partials of the first at partial[0..grid),
// of the second at partial[grid..2*grid); per-thread order and tree as
// k_dot_partial, so each result is bit-identical to a separate dot.
__global__ void __launch_bounds__(kBlock)
    k_dot2_partial(long n, const double* __restrict__ a, const double* __restrict__ w,
                   double* __restrict__ partial) {
  __shared__ double red[2][kBlock];
  double ww = 0.0, aw = 0.0;
  long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double wi = w[i];
    ww += wi * wi;
    aw += a[i] * wi;
  }
  red[0][threadIdx.x] = ww;
  red[1][threadIdx.x] = aw;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = red[0][0];
    partial[gridDim.x + blockIdx.x] = red[1][0];
  }
}

// Block partials of sum_i (u_i / sigma_(i mod 4))^2 over n = 4 * cells values.
__global__ void __launch_bounds__(kBlock)
    k_scaled_sq_partial(long n, const double* __restrict__ u, double s0, double s1, double s2, double s3,
                        double* __restrict__ partial) {
  __shared__ double red[kBlock];
  double acc = 0.0;
  long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int k = (int)(i & 3);
    const double sig = k == 0 ? s0 : (k == 1 ? s1 : (k == 2 ? s2 : s3));
    double q = u[i] / sig;
    acc += q * q;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// One modified Gram-Schmidt step fused with the next one's dot. Each
// workgroup first finishes the previous dot(s) from their block partials
// (nprev slots of kDotBlocks at prev, bit j of sqrt_mask taking the square
// root; workgroup 0 stores them to out[0..nprev) and the pinned mirror), the
// last of which is the coefficient c; then w -= c * p and the block partial
// of (q, w) over the updated w, or of (w, w) with SELF; per-thread order and
// tree as k_dot_partial.
template <bool SELF>
__global__ void __launch_bounds__(kBlock)
    k_axpy_dot_partial(long n, const double* __restrict__ p, double* __restrict__ w,
                       const double* __restrict__ prev, int nprev, int sqrt_mask, double* __restrict__ out,
                       double* __restrict__ host_out, const double* __restrict__ q,
                       double* __restrict__ partial) {
  __shared__ double red[kBlock];
  // the thread's first kPrefetch elements are loaded before the coefficient
  // is finished, so their latency overlaps the reduction
  const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  double pw[kPrefetch], pp[kPrefetch], pq[kPrefetch];
#pragma unroll
  for (int r = 0; r < kPrefetch; ++r) {
    const long i = i0 + r * stride;
    if (i < n) {
      pw[r] = w[i];
      pp[r] = p[i];
      pq[r] = SELF ? 0.0 : q[i];
    }
  }
  double c = 0.0;
  for (int j = 0; j < nprev; ++j) {
    double v = block_sum(prev + (long)j * kDotBlocks, kDotBlocks, red);
    if ((sqrt_mask >> j) & 1) v = sqrt(v);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      out[j] = v;
      if (host_out) host_out[j] = v;
    }
    c = v;
  }
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < kPrefetch; ++r) {
    const long i = i0 + r * stride;
    if (i < n) {
      double wi = pw[r] - c * pp[r];
      w[i] = wi;
      acc += SELF ? wi * wi : pq[r] * wi;
    }
  }
  for (long i = i0 + kPrefetch * stride; i < n; i += stride) {
    double wi = w[i] - c * p[i];
    w[i] = wi;
    acc += SELF ? wi * wi : q[i] * wi;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// y = a*x + b*y (host scalars)
__global__ void __launch_bounds__(kBlock)
    k_axpby(long n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

__global__ void __launch_bounds__(kBlock) k_fill(long n, double v, double* __restrict__ y) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

// out = x * s over a grid of kDotBlocks workgroups, with the block partials of
// (out, out) in k_dot_partial's order and shape (x may be out)
__global__ void __launch_bounds__(kBlock)
    k_scale_sq_partial(long n, const double* x, double s, double* out, double* __restrict__ partial) {
  __shared__ double red[kBlock];
  double acc = 0.0;
  long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double o = x[i] * s;
    out[i] = o;
    acc += o * o;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int t = kBlock / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Arnoldi normalisation over a grid of kDotBlocks workgroups: h = ||w|| from
// the block partials at prev (square root of block_sum), stored by
// workgroup 0 to out[0] and the pinned mirror; v = w * (1 / h), or v = w on
// breakdown (h <= eps * h0, h0 = ||w|| before the projections, at *h0p);
// partial = the block partials of (v, v) in k_dot_partial's order and shape.
__global__ void __launch_bounds__(kBlock)
    k_arnoldi_scale(long n, const double* __restrict__ w, const double* __restrict__ prev,
                    const double* __restrict__ h0p, double eps, double* __restrict__ out,
                    double* __restrict__ host_out, double* __restrict__ v, double* __restrict__ partial) {
  __shared__ double red[kBlock];
  const long i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long stride = (long)gridDim.x * blockDim.x;
  double pw[kPrefetch];
#pragma unroll
  for (int r = 0; r < kPrefetch; ++r)
    if (i0 + r * stride < n) pw[r] = w[i0 + r * stride];
  const double h = sqrt(block_sum(prev, kDotBlocks, red));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    out[0] = h;
    if (host_out) host_out[0] = h;
  }
  const bool breakdown = h <= eps * h0p[0];
  const double s = 1.0 / h;
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < kPrefetch; ++r) {
    const long i = i0 + r * stride;
    if (i < n) {
      double o = breakdown ? pw[r] : pw[r] * s;
      v[i] = o;
      acc += o * o;
    }
  }
  for (long i = i0 + kPrefetch * stride; i < n; i += stride) {
    double o = breakdown ? w[i] : w[i] * s;
    v[i] = o;
    acc += o * o;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int t = kBlock / 2; t > 0; t >>= 1) {
    if (threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// x += sum_k y_k V_k, k = 0..m-1 in order (y in device memory)
__global__ void __launch_bounds__(kBlock)
    k_combine(long n, int m, const double* __restrict__ V, const double* __restrict__ y, double* __restrict__ x) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int k = 0; k < m; ++k) acc = acc + y[k] * V[(long)k * n + i];
  x[i] = x[i] + acc;
}

// perturbed = s + (eps*sigma_k) * d   (solver.py:1553-1556), eps = num / ||d||
// with ||d|| from the device (solver.py:1549-1552); ||d|| = 0: the state
// itself (the caller's product is then d)
__global__ void __launch_bounds__(kBlock)
    k_perturb(int n, const double* __restrict__ s, const double* __restrict__ d, double num,
              const double* __restrict__ dnorm, double s0, double s1, double s2, double s3,
              double* __restrict__ out) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dn = dnorm[0];
  const double eps = (dn == 0.0) ? 0.0 : num / dn;
  P4 a = load4(s, i), b = load4(d, i);
  store4(out, i, a.r + eps * s0 * b.r, a.u + eps * s1 * b.u, a.v + eps * s2 * b.v,
         a.p + eps * s3 * b.p);
}

// out = d - (h * ((rp - r) / eps)) / sigma   (solver.py:1557-1563); ||d|| = 0:
// out = d (solver.py:1550-1551)
__global__ void __launch_bounds__(kBlock)
    k_jv_finish(int n, const double* __restrict__ d, const double* __restrict__ rp,
                const double* __restrict__ r, double num, const double* __restrict__ dnorm,
                const double* __restrict__ dt, int dt_per_cell, double s0, double s1, double s2,
                double s3, double* __restrict__ out) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dn = dnorm[0];
  P4 a = load4(d, i);
  if (dn == 0.0) {
    store4(out, i, a.r, a.u, a.v, a.p);
    return;
  }
  const double eps = num / dn;
  double h = dt_per_cell ? dt[i] : dt[0];
  P4 p = load4(rp, i), q = load4(r, i);
  store4(out, i, a.r - h * ((p.r - q.r) / eps) / s0, a.u - h * ((p.u - q.u) / eps) / s1,
         a.v - h * ((p.v - q.v) / eps) / s2, a.p - h * ((p.p - q.p) / eps) / s3);
}

// Newton equation (solver.py:1415-1424): scaled = (u - u0 - h*R)/sigma and the
// max-norms of scaled and of (u - u0)/sigma, per block then final.
__global__ void __launch_bounds__(kBlock)
    k_newton_equation(int n, const double* __restrict__ u, const double* __restrict__ u0,
                      const double* __restrict__ R, const double* __restrict__ dt, int dt_per_cell,
                      double s0, double s1, double s2, double s3, double* __restrict__ scaled,
                      double* __restrict__ partial) {
  __shared__ double r1[kBlock];
  __shared__ double r2[kBlock];
  int i = xcd_block() * blockDim.x + threadIdx.x;
  double m1 = 0.0, m2 = 0.0;
  if (i < n) {
    double h = dt_per_cell ? dt[i] : dt[0];
    P4 a = load4(u, i), b = load4(u0, i), r = load4(R, i);
    double e0 = (a.r - b.r - h * r.r) / s0, e1 = (a.u - b.u - h * r.u) / s1;
    double e2 = (a.v - b.v - h * r.v) / s2, e3 = (a.p - b.p - h * r.p) / s3;
    if (scaled) store4(scaled, i, e0, e1, e2, e3);
    m1 = fmax(fmax(fabs(e0), fabs(e1)), fmax(fabs(e2), fabs(e3)));
    m2 = fmax(fmax(fabs((a.r - b.r) / s0), fabs((a.u - b.u) / s1)),
              fmax(fabs((a.v - b.v) / s2), fabs((a.p - b.p) / s3)));
    if (!finite(e0) || !finite(e1) || !finite(e2) || !finite(e3)) m1 = INFINITY;
  }
  r1[threadIdx.x] = m1;
  r2[threadIdx.x] = m2;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + s]);
      r2[threadIdx.x] = fmax(r2[threadIdx.x], r2[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * xcd_block()] = r1[0];
    partial[2 * xcd_block() + 1] = r2[0];
  }
}

__global__ void __launch_bounds__(1024)
    k_max2_final(int nb, const double* __restrict__ partial, double* __restrict__ out) {
  __shared__ double r1[1024];
  __shared__ double r2[1024];
  double m1 = 0.0, m2 = 0.0;
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    m1 = fmax(m1, partial[2 * j]);
    m2 = fmax(m2, partial[2 * j + 1]);
  }
  r1[threadIdx.x] = m1;
  r2[threadIdx.x] = m2;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + s]);
      r2[threadIdx.x] = fmax(r2[threadIdx.x], r2[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = r1[0];
    out[1] = r2[0];
  }
}

// candidate = state + (damping*corr)*sigma   (solver.py:1465-1468)
__global__ void __launch_bounds__(kBlock)
    k_candidate(int n, const double* __restrict__ u, const double* __restrict__ corr,
                double damping, double s0, double s1, double s2, double s3,
                double* __restrict__ out) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  P4 a = load4(u, i), c = load4(corr, i);
  store4(out, i, a.r + damping * c.r * s0, a.u + damping * c.u * s1, a.v + damping * c.v * s2,
         a.p + damping * c.p * s3);
}


// ---------------------------------------------------------------------------
// Adjoint residual (vector-Jacobian product).
//
// Face fluxes are differentiated in forward mode with 8-direction dual
// numbers (4 left + 4 right primitive components); every comparison is made
// on the value part, which is computed with the same arithmetic as the
// forward kernels, so the derivative is that of the branch the forward pass
// took. The reverse sweep through reconstruction, limiter, LSQ gradient and
// primitive conversion is written out by hand as per-cell gathers (no
// atomics: deterministic).
// ---------------------------------------------------------------------------

struct D8 {
  double v;
  double d[8];
};

__device__ __forceinline__ D8 dconst(double v) {
  D8 r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = 0.0;
  return r;
}
__device__ __forceinline__ D8 operator+(const D8& a, const D8& b) {
  D8 r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ D8 operator-(const D8& a, const D8& b) {
  D8 r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ __forceinline__ D8 operator-(const D8& a) {
  D8 r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = -a.d[k];
  return r;
}
__device__ __forceinline__ D8 operator*(const D8& a, const D8& b) {
  D8 r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ D8 operator/(const D8& a, const D8& b) {
  D8 r;
  r.v = a.v / b.v;
  double inv = 1.0 / b.v;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * inv;
  return r;
}
__device__ __forceinline__ D8 operator+(const D8& a, double b) { D8 r = a; r.v = a.v + b; return r; }
__device__ __forceinline__ D8 operator+(double a, const D8& b) { D8 r = b; r.v = a + b.v; return r; }
__device__ __forceinline__ D8 operator-(const D8& a, double b) { D8 r = a; r.v = a.v - b; return r; }
__device__ __forceinline__ D8 operator-(double a, const D8& b) {
  D8 r;
  r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = -b.d[k];
  return r;
}
__device__ __forceinline__ D8 operator*(double a, const D8& b) {
  D8 r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = a * b.d[k];
  return r;
}
__device__ __forceinline__ D8 operator*(const D8& a, double b) {
  D8 r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] * b;
  return r;
}
__device__ __forceinline__ D8 operator/(const D8& a, double b) {
  D8 r;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] / b;
  return r;
}
__device__ __forceinline__ D8 operator/(double a, const D8& b) { return dconst(a) / b; }
__device__ __forceinline__ D8 dsqrt(const D8& a) {
  D8 r;
  r.v = sqrt(a.v);
  double h = 0.5 / r.v;
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] * h;
  return r;
}
__device__ __forceinline__ D8 dfabs(const D8& a) { return (a.v < 0.0) ? -a : a; }
__device__ __forceinline__ D8 dpow(const D8& a, double e) {
  D8 r;
  r.v = pow(a.v, e);
  double dv = e * pow(a.v, e - 1.0);
#pragma unroll
  for (int k = 0; k < 8; ++k) r.d[k] = a.d[k] * dv;
  return r;
}
__device__ __forceinline__ D8 dmin(const D8& a, const D8& b) { return (a.v <= b.v) ? a : b; }
__device__ __forceinline__ D8 dmax(const D8& a, const D8& b) { return (a.v >= b.v) ? a : b; }

struct P4D {
  D8 r, u, v, p;
};
struct F4D {
  D8 a, b, c, d;
};

struct SideD {
  D8 rho, u, v, p, un, E;
  D8 U[4];
  D8 F[4];
};

__device__ __forceinline__ void side_state_d(const P4D& w, double nx, double ny, double g, SideD& s) {
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

__device__ __forceinline__ D8 sound_d(const P4D& w, double g) { return dsqrt(g * w.p / w.r); }

__device__ F4D rusanov_d(const P4D& L, const P4D& R, double nx, double ny, double g) {
  SideD a, b;
  side_state_d(L, nx, ny, g, a);
  side_state_d(R, nx, ny, g, b);
  D8 s = dmax(dfabs(a.un) + sound_d(L, g), dfabs(b.un) + sound_d(R, g));
  F4D f;
  f.a = 0.5 * (a.F[0] + b.F[0] - s * (b.U[0] - a.U[0]));
  f.b = 0.5 * (a.F[1] + b.F[1] - s * (b.U[1] - a.U[1]));
  f.c = 0.5 * (a.F[2] + b.F[2] - s * (b.U[2] - a.U[2]));
  f.d = 0.5 * (a.F[3] + b.F[3] - s * (b.U[3] - a.U[3]));
  return f;
}

__device__ F4D hllc_star_d(const SideD& s, const D8& wave, const D8& star, double nx, double ny) {
  D8 factor = s.rho * (wave - s.un) / (wave - star);
  D8 vx = s.u + (star - s.un) * nx;
  D8 vy = s.v + (star - s.un) * ny;
  D8 energy = factor * (s.E / s.rho + (star - s.un) * (star + s.p / (s.rho * (wave - s.un))));
  F4D f;
  f.a = s.F[0] + wave * (factor - s.U[0]);
  f.b = s.F[1] + wave * (factor * vx - s.U[1]);
  f.c = s.F[2] + wave * (factor * vy - s.U[2]);
  f.d = s.F[3] + wave * (energy - s.U[3]);
  return f;
}

__device__ F4D hllc_d(const P4D& L, const P4D& R, double nx, double ny, double g, bool fused_order) {
  SideD a, b;
  side_state_d(L, nx, ny, g, a);
  side_state_d(R, nx, ny, g, b);
  D8 cl = sound_d(L, g), cr = sound_d(R, g);
  D8 sl = dmin(a.un - cl, b.un - cr);
  D8 sr = dmax(a.un + cl, b.un + cr);
  D8 den = a.rho * (sl - a.un) - b.rho * (sr - b.un);
  if (!(fabs(den.v) > 1e-14)) den = dconst(copysign(1e-14, den.v));
  D8 star = (b.p - a.p + a.rho * a.un * (sl - a.un) - b.rho * b.un * (sr - b.un)) / den;
  if (sl.v >= 0.0) return F4D{a.F[0], a.F[1], a.F[2], a.F[3]};
  if (fused_order) {
    if (sr.v <= 0.0) return F4D{b.F[0], b.F[1], b.F[2], b.F[3]};
    if (star.v >= 0.0) return hllc_star_d(a, sl, star, nx, ny);
    return hllc_star_d(b, sr, star, nx, ny);
  }
  if (star.v >= 0.0) return hllc_star_d(a, sl, star, nx, ny);
  if (sr.v > 0.0) return hllc_star_d(b, sr, star, nx, ny);
  return F4D{b.F[0], b.F[1], b.F[2], b.F[3]};
}

__device__ F4D roe_d(const P4D& L, const P4D& R, double nx, double ny, double g) {
  SideD a, b;
  side_state_d(L, nx, ny, g, a);
  side_state_d(R, nx, ny, g, b);
  D8 sl = dsqrt(L.r), sr = dsqrt(R.r);
  D8 den = sl + sr;
  D8 u = (sl * L.u + sr * R.u) / den;
  D8 v = (sl * L.v + sr * R.v) / den;
  D8 h = (sl * (L.p / L.r + 0.5 * (L.u * L.u + L.v * L.v)) + sr * (R.p / R.r + 0.5 * (R.u * R.u + R.v * R.v))) / den;
  D8 un = u * nx + v * ny;
  D8 ut = u * -ny + v * nx;
  D8 a2 = (g - 1.0) * (h - 0.5 * (u * u + v * v));
  D8 c = dsqrt(a2);
  D8 rho_avg = sl * sr;
  D8 utl = L.u * -ny + L.v * nx;
  D8 utr = R.u * -ny + R.v * nx;
  D8 drho = R.r - L.r, dun = b.un - a.un, dut = utr - utl, dp = R.p - L.p;
  D8 am = (dp - rho_avg * c * dun) / (2.0 * a2);
  D8 ap = (dp + rho_avg * c * dun) / (2.0 * a2);
  D8 ac = drho - dp / a2;
  D8 as = rho_avg * dut;
  D8 fix = 0.1 * c;
  D8 l0 = dfabs(un - c), l1 = dfabs(un), l3 = dfabs(un + c);
  if (l0.v < fix.v) l0 = 0.5 * (l0 * l0 / fix + fix);
  if (l1.v < fix.v) l1 = 0.5 * (l1 * l1 / fix + fix);
  if (l3.v < fix.v) l3 = 0.5 * (l3 * l3 / fix + fix);
  D8 w0 = l0 * am, w1 = l1 * ac, w2 = l1 * as, w3 = l3 * ap;
  D8 d0 = w0 + w1 + w3;
  D8 d1 = (un - c) * w0 + un * w1 + (un + c) * w3;
  D8 d2 = ut * w0 + ut * w1 + w2 + ut * w3;
  D8 d3 = (h - un * c) * w0 + 0.5 * (un * un + ut * ut) * w1 + ut * w2 + (h + un * c) * w3;
  F4D f;
  f.a = 0.5 * (a.F[0] + b.F[0] - d0);
  f.b = 0.5 * (a.F[1] + b.F[1] - (d1 * nx - d2 * ny));
  f.c = 0.5 * (a.F[2] + b.F[2] - (d1 * ny + d2 * nx));
  f.d = 0.5 * (a.F[3] + b.F[3] - d3);
  return f;
}

// Far-field characteristics in duals of both the interior state and the
// free-stream state (the latter seeded for inflow-parameter gradients).
__device__ P4D far_field_d(const P4D& in, const P4D& out, double nx, double ny, double g) {
  D8 un_i = in.u * nx + in.v * ny;
  D8 un_f = out.u * nx + out.v * ny;
  D8 c_i = sound_d(in, g);
  D8 c_f = sound_d(out, g);
  if (un_i.v <= -c_i.v) return out;
  if (un_i.v >= c_i.v) return in;
  D8 r_out = un_i + 2.0 * c_i / (g - 1.0);
  D8 r_in = un_f - 2.0 * c_f / (g - 1.0);
  D8 vb = 0.5 * (r_out + r_in);
  D8 cb = 0.25 * (g - 1.0) * (r_out - r_in);
  bool inflow = vb.v < 0.0;
  P4D src = inflow ? out : in;
  D8 entropy = src.p / dpow(src.r, g);
  D8 rho = dpow(cb * cb / (g * entropy), 1.0 / (g - 1.0));
  D8 p = entropy * dpow(rho, g);
  double tx = -ny, ty = nx;
  D8 vt = src.u * tx + src.v * ty;
  return P4D{rho, vb * nx + vt * tx, vb * ny + vt * ty, p};
}

__device__ __forceinline__ P4D seed(const P4& w, int base) {
  P4D r{dconst(w.r), dconst(w.u), dconst(w.v), dconst(w.p)};
  r.r.d[base] = 1.0;
  r.u.d[base + 1] = 1.0;
  r.v.d[base + 2] = 1.0;
  r.p.d[base + 3] = 1.0;
  return r;
}

// ---------------------------------------------------------------------------
// Reverse mode of the Rusanov and HLLC fluxes: one primal pass that keeps the
// intermediates in registers, one backward pass. (dF/dW_L)^T Fbar and
// (dF/dW_R)^T Fbar cost about three primal fluxes instead of the nine of
// eight-direction duals. Roe keeps the dual path.

struct SideR {
  double r, u, v, p, un, E;
};
struct SideB {
  double r = 0.0, u = 0.0, v = 0.0, p = 0.0, un = 0.0, E = 0.0;
  double nx = 0.0, ny = 0.0;  // adjoint of the face normal through this side's terms
};

__device__ __forceinline__ SideR side_r(const P4& w, double nx, double ny, double g) {
  return SideR{w.r, w.u, w.v, w.p, w.u * nx + w.v * ny,
               w.p / (g - 1.0) + 0.5 * w.r * (w.u * w.u + w.v * w.v)};
}

// F = (r un, r u un + p nx, r v un + p ny, (E + p) un): accumulate F^T fb.
__device__ __forceinline__ void phys_flux_b(const SideR& s, const double* fb, double nx, double ny,
                                            SideB& b) {
  b.r += fb[0] * s.un + fb[1] * s.u * s.un + fb[2] * s.v * s.un;
  b.u += fb[1] * s.r * s.un;
  b.v += fb[2] * s.r * s.un;
  b.p += fb[1] * nx + fb[2] * ny + fb[3] * s.un;
  b.un += fb[0] * s.r + fb[1] * s.r * s.u + fb[2] * s.r * s.v + fb[3] * (s.E + s.p);
  b.E += fb[3] * s.un;
  b.nx += fb[1] * s.p;
  b.ny += fb[2] * s.p;
}

// U = (r, r u, r v, E): accumulate U^T ub.
__device__ __forceinline__ void cons_b(const SideR& s, const double* ub, SideB& b) {
  b.r += ub[0] + ub[1] * s.u + ub[2] * s.v;
  b.u += ub[1] * s.r;
  b.v += ub[2] * s.r;
  b.E += ub[3];
}

// c = sqrt(g p / r)
__device__ __forceinline__ void sound_b(const SideR& s, double c, double cb, double g, SideB& b) {
  double h = cb * 0.5 / c;
  b.p += h * g / s.r;
  b.r -= h * (g * s.p / s.r) / s.r;
}

// un = u nx + v ny and E = p/(g-1) + r (u^2 + v^2)/2 back to (r, u, v, p);
// the normal's adjoint from this side is added to (nbx, nby).
__device__ __forceinline__ P4 side_finish_b(const SideR& s, const SideB& b, double nx, double ny,
                                            double g, double& nbx, double& nby) {
  nbx += b.nx + b.un * s.u;
  nby += b.ny + b.un * s.v;
  return P4{b.r + b.E * 0.5 * (s.u * s.u + s.v * s.v), b.u + b.E * s.r * s.u + b.un * nx,
            b.v + b.E * s.r * s.v + b.un * ny, b.p + b.E / (g - 1.0)};
}

// Star-region flux of one side (hllc_star): F + S (U* - U).
__device__ void hllc_star_b(const SideR& s, double S, double star, const double* fb, double nx,
                            double ny, SideB& sb, double& Sb, double& starb) {
  double q = S - s.un, d = S - star;
  double factor = s.r * q / d;
  double t = star - s.un;
  double vx = s.u + t * nx, vy = s.v + t * ny;
  double e1 = s.E / s.r;
  double rq = s.r * q;
  double e2 = star + s.p / rq;
  double inner = e1 + t * e2;
  double energy = factor * inner;
  phys_flux_b(s, fb, nx, ny, sb);
  Sb += fb[0] * (factor - s.r) + fb[1] * (factor * vx - s.r * s.u) +
        fb[2] * (factor * vy - s.r * s.v) + fb[3] * (energy - s.E);
  double factorb = fb[0] * S + fb[1] * S * vx + fb[2] * S * vy;
  double vxb = fb[1] * S * factor, vyb = fb[2] * S * factor;
  double ub[4] = {-fb[0] * S, -fb[1] * S, -fb[2] * S, -fb[3] * S};
  double energyb = fb[3] * S;
  factorb += energyb * inner;
  double innerb = energyb * factor;
  sb.E += innerb / s.r;
  sb.r -= innerb * e1 / s.r;
  double tb = innerb * e2;
  double e2b = innerb * t;
  starb += e2b;
  sb.p += e2b / rq;
  double rqb = -e2b * (s.p / rq) / rq;
  sb.r += rqb * q;
  double qb = rqb * s.r;
  sb.u += vxb;
  sb.v += vyb;
  sb.nx += vxb * t;
  sb.ny += vyb * t;
  tb += vxb * nx + vyb * ny;
  starb += tb;
  sb.un -= tb;
  sb.r += factorb * q / d;
  qb += factorb * s.r / d;
  double db = -factorb * factor / d;
  Sb += qb + db;
  sb.un -= qb;
  starb -= db;
  cons_b(s, ub, sb);
}

// (dF/dL)^T fb -> lb, (dF/dR)^T fb -> rb for hllc() (same branches).
__device__ void hllc_b(const P4& L, const P4& R, double nx, double ny, double g, bool fused_order,
                       const double* fb, P4& lb, P4& rb, double& nbx, double& nby) {
  SideR a = side_r(L, nx, ny, g), b = side_r(R, nx, ny, g);
  SideB A, B;
  double cl = sqrt(g * L.p / L.r), cr = sqrt(g * R.p / R.r);
  double al = a.un - cl, ar = b.un - cr;
  bool sl_left = al <= ar;
  double sl = sl_left ? al : ar;
  double bl = a.un + cl, br = b.un + cr;
  bool sr_left = bl >= br;
  double sr = sr_left ? bl : br;
  double den = a.r * (sl - a.un) - b.r * (sr - b.un);
  bool clamped = !(fabs(den) > 1e-14);
  if (clamped) den = copysign(1e-14, den);
  double num = b.p - a.p + a.r * a.un * (sl - a.un) - b.r * b.un * (sr - b.un);
  double star = num / den;
  int branch;  // 0 left flux, 1 right flux, 2 left star state, 3 right star state
  if (sl >= 0.0)
    branch = 0;
  else if (fused_order)
    branch = (sr <= 0.0) ? 1 : (star >= 0.0 ? 2 : 3);
  else
    branch = (star >= 0.0) ? 2 : (sr > 0.0 ? 3 : 1);
  double clb = 0.0, crb = 0.0;
  if (branch == 0) {
    phys_flux_b(a, fb, nx, ny, A);
  } else if (branch == 1) {
    phys_flux_b(b, fb, nx, ny, B);
  } else {
    double slb = 0.0, srb = 0.0, starb = 0.0;
    if (branch == 2)
      hllc_star_b(a, sl, star, fb, nx, ny, A, slb, starb);
    else
      hllc_star_b(b, sr, star, fb, nx, ny, B, srb, starb);
    double numb = starb / den;
    double denb = clamped ? 0.0 : -starb * star / den;
    B.p += numb;
    A.p -= numb;
    double t1 = numb * (sl - a.un), q1 = numb * a.r * a.un;
    A.r += t1 * a.un;
    A.un += t1 * a.r - q1;
    slb += q1;
    double t2 = -numb * (sr - b.un), q2 = -numb * b.r * b.un;
    B.r += t2 * b.un;
    B.un += t2 * b.r - q2;
    srb += q2;
    A.r += denb * (sl - a.un);
    slb += denb * a.r;
    A.un -= denb * a.r;
    B.r -= denb * (sr - b.un);
    srb -= denb * b.r;
    B.un += denb * b.r;
    if (sl_left) {
      A.un += slb;
      clb -= slb;
    } else {
      B.un += slb;
      crb -= slb;
    }
    if (sr_left) {
      A.un += srb;
      clb += srb;
    } else {
      B.un += srb;
      crb += srb;
    }
  }
  if (clb != 0.0) sound_b(a, cl, clb, g, A);
  if (crb != 0.0) sound_b(b, cr, crb, g, B);
  lb = side_finish_b(a, A, nx, ny, g, nbx, nby);
  rb = side_finish_b(b, B, nx, ny, g, nbx, nby);
}

// (dF/dL)^T fb, (dF/dR)^T fb for rusanov().
__device__ void rusanov_b(const P4& L, const P4& R, double nx, double ny, double g, const double* fb,
                          P4& lb, P4& rb, double& nbx, double& nby) {
  SideR a = side_r(L, nx, ny, g), b = side_r(R, nx, ny, g);
  SideB A, B;
  double cl = sqrt(g * L.p / L.r), cr = sqrt(g * R.p / R.r);
  double xa = fabs(a.un) + cl, xb = fabs(b.un) + cr;
  bool left = xa >= xb;
  double sp = left ? xa : xb;
  double h[4] = {0.5 * fb[0], 0.5 * fb[1], 0.5 * fb[2], 0.5 * fb[3]};
  phys_flux_b(a, h, nx, ny, A);
  phys_flux_b(b, h, nx, ny, B);
  double spb = -(h[0] * (b.r - a.r) + h[1] * (b.r * b.u - a.r * a.u) + h[2] * (b.r * b.v - a.r * a.v) +
                 h[3] * (b.E - a.E));
  double ulb[4] = {h[0] * sp, h[1] * sp, h[2] * sp, h[3] * sp};
  double urb[4] = {-h[0] * sp, -h[1] * sp, -h[2] * sp, -h[3] * sp};
  cons_b(a, ulb, A);
  cons_b(b, urb, B);
  if (left) {
    A.un += (a.un < 0.0) ? -spb : spb;
    sound_b(a, cl, spb, g, A);
  } else {
    B.un += (b.un < 0.0) ? -spb : spb;
    sound_b(b, cr, spb, g, B);
  }
  lb = side_finish_b(a, A, nx, ny, g, nbx, nby);
  rb = side_finish_b(b, B, nx, ny, g, nbx, nby);
}

// Reverse of far_field(): e-bar -> interior-state, free-stream-state and
// normal adjoints (accumulated). Same branches as the forward.
__device__ void far_field_b(const P4& in, const P4& out, double nx, double ny, double g, const P4& eb,
                            P4& inb, P4& outb, double& nbx, double& nby) {
  double un_i = in.u * nx + in.v * ny;
  double un_f = out.u * nx + out.v * ny;
  double c_i = sound(in, g), c_f = sound(out, g);
  if (un_i <= -c_i) {
    outb.r += eb.r;
    outb.u += eb.u;
    outb.v += eb.v;
    outb.p += eb.p;
    return;
  }
  if (un_i >= c_i) {
    inb.r += eb.r;
    inb.u += eb.u;
    inb.v += eb.v;
    inb.p += eb.p;
    return;
  }
  double r_out = un_i + 2.0 * c_i / (g - 1.0);
  double r_in = un_f - 2.0 * c_f / (g - 1.0);
  double vb = 0.5 * (r_out + r_in);
  double cb = 0.25 * (g - 1.0) * (r_out - r_in);
  bool inflow = vb < 0.0;
  const P4& src = inflow ? out : in;
  double sr_g = pow(src.r, g);
  double entropy = src.p / sr_g;
  double X = cb * cb / (g * entropy);
  double rho = pow(X, 1.0 / (g - 1.0));
  double rho_g = pow(rho, g);
  double tx = -ny, ty = nx;
  double vt = src.u * tx + src.v * ty;
  // e = (rho, vb nx + vt tx, vb ny + vt ty, entropy rho^g)
  double rhob = eb.r + eb.p * entropy * g * rho_g / rho;
  double entb = eb.p * rho_g;
  double vbb = eb.u * nx + eb.v * ny;
  double vtb = eb.u * tx + eb.v * ty;
  nbx += eb.u * vb + eb.v * vt;
  nby += -eb.u * vt + eb.v * vb;
  double Xb = rhob * rho / ((g - 1.0) * X);
  double cbb = Xb * 2.0 * cb / (g * entropy);
  entb -= Xb * X / entropy;
  P4 sb{0.0, 0.0, 0.0, 0.0};
  sb.u += vtb * tx;
  sb.v += vtb * ty;
  nbx += vtb * src.v;
  nby -= vtb * src.u;
  sb.p += entb / sr_g;
  sb.r -= entb * g * entropy / src.r;
  double routb = 0.5 * vbb + 0.25 * (g - 1.0) * cbb;
  double rinb = 0.5 * vbb - 0.25 * (g - 1.0) * cbb;
  double cib = 2.0 * routb / (g - 1.0);
  double cfb = -2.0 * rinb / (g - 1.0);
  inb.u += routb * nx;
  inb.v += routb * ny;
  nbx += routb * in.u;
  nby += routb * in.v;
  outb.u += rinb * nx;
  outb.v += rinb * ny;
  nbx += rinb * out.u;
  nby += rinb * out.v;
  double hi = cib * 0.5 / c_i;
  inb.p += hi * g / in.r;
  inb.r -= hi * (g * in.p / in.r) / in.r;
  double hf = cfb * 0.5 / c_f;
  outb.p += hf * g / out.r;
  outb.r -= hf * (g * out.p / out.r) / out.r;
  P4& tgt = inflow ? outb : inb;
  tgt.r += sb.r;
  tgt.u += sb.u;
  tgt.v += sb.v;
  tgt.p += sb.p;
}

// Face adjoint in reverse mode (Rusanov, HLLC). Same outputs as
// k_flux_adjoint: FADJ[k][0..3] the adjoint of the reconstructed left state
// (through the ghost state on boundary faces), [4..7] that of the right /
// exterior state.
__global__ void __launch_bounds__(kBlock)
    k_flux_adjoint_rev(int nf, FaceArrays fa, const double* __restrict__ W,
                       const double* __restrict__ G, const double* __restrict__ PHI,
                       const double* __restrict__ K, const double* __restrict__ flen,
                       const double* __restrict__ area,
                       const double* __restrict__ lam, int scheme, int muscl, double g,
                       double* __restrict__ FADJ, double* __restrict__ SBAR) {
  int k = xcd_block() * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  int kind = fa.kind[k];
  double nx = fa.nx[k], ny = fa.ny[k];
  int l = fa.left[k];
  int r = fa.right[k];
  double L = flen[k];
  P4 ll = load4(lam, l);
  double cl = -L / area[l];
  double fb[4] = {cl * ll.r, cl * ll.u, cl * ll.v, cl * ll.p};
  if (kind <= MCFD_FACE_PERIODIC) {
    P4 lr = load4(lam, r);
    double cr = L / area[r];
    fb[0] += cr * lr.r;
    fb[1] += cr * lr.u;
    fb[2] += cr * lr.v;
    fb[3] += cr * lr.p;
  }
  P4 wl = K ? reconstruct_q(W, G, K, PHI, l, fa.lx[k], fa.ly[k])
            : reconstruct(W, G, PHI, l, fa.lx[k], fa.ly[k], muscl != 0);
  P4 wr;
  int b = fa.bc[k];
  if (kind <= MCFD_FACE_PERIODIC) {
    wr = K ? reconstruct_q(W, G, K, PHI, r, fa.rx[k], fa.ry[k])
           : reconstruct(W, G, PHI, r, fa.rx[k], fa.ry[k], muscl != 0);
  } else {
    const double* st = fa.bc_state + 4 * (long)b;
    wr = wl;
    if (kind == MCFD_FACE_WALL) {
      double un = (wl.u - fa.bc_wall[2 * (long)b]) * nx + (wl.v - fa.bc_wall[2 * (long)b + 1]) * ny;
      wr.u = wl.u - 2.0 * un * nx;
      wr.v = wl.v - 2.0 * un * ny;
    } else if (kind == MCFD_FACE_VELOCITY_INLET) {
      wr = P4{st[0], st[1], st[2], st[3]};
    } else if (kind == MCFD_FACE_PRESSURE_OUTLET) {
      wr.p = st[3];
    } else if (kind == MCFD_FACE_FAR_FIELD) {
      int err = 0;
      wr = far_field(wl, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, &err);
    } else if (kind >= MCFD_FACE_PRESSURE_INLET) {
      wr = load4(fa.inflow, b - fa.inflow_first);
    }
  }
  P4 lb, rb;
  double nbx = 0.0, nby = 0.0;  // unused here
  if (scheme == 1)
    hllc_b(wl, wr, nx, ny, g, kind == MCFD_FACE_INTERIOR, fb, lb, rb, nbx, nby);
  else
    rusanov_b(wl, wr, nx, ny, g, fb, lb, rb, nbx, nby);
  P4 ext{0.0, 0.0, 0.0, 0.0};
  if (kind > MCFD_FACE_PERIODIC) {
    if (kind == MCFD_FACE_WALL) {
      // mirrored velocity: d(u_R, v_R)/d(u_L, v_L) = I - 2 n n^T
      double un_b = rb.u * nx + rb.v * ny;
      lb.r += rb.r;
      lb.u += rb.u - 2.0 * nx * un_b;
      lb.v += rb.v - 2.0 * ny * un_b;
      lb.p += rb.p;
    } else if (kind == MCFD_FACE_VELOCITY_INLET) {
      ext = rb;
    } else if (kind == MCFD_FACE_PRESSURE_OUTLET) {
      lb.r += rb.r;
      lb.u += rb.u;
      lb.v += rb.v;
    } else if (kind == MCFD_FACE_FAR_FIELD) {
      const double* st = fa.bc_state + 4 * (long)b;
      far_field_b(wl, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, rb, lb, ext, nbx, nby);
    } else if (kind >= MCFD_FACE_PRESSURE_INLET) {
      // the exterior state's adjoint stays in the right slot for
      // k_inflow_adjoint (and the mass-flow-inlet patch pass)
      ext = rb;
    } else {  // transmissive: exterior = interior
      lb.r += rb.r;
      lb.u += rb.u;
      lb.v += rb.v;
      lb.p += rb.p;
    }
    rb = ext;
  }
  double2* q = reinterpret_cast<double2*>(FADJ + 8 * (long)k);
  q[0] = make_double2(lb.r, lb.u);
  q[1] = make_double2(lb.v, lb.p);
  q[2] = make_double2(rb.r, rb.u);
  q[3] = make_double2(rb.v, rb.p);
  // inflow records: no state of their own to differentiate (zero)
  if (SBAR && kind > MCFD_FACE_PERIODIC) {
    if (kind >= MCFD_FACE_PRESSURE_INLET) ext = P4{0.0, 0.0, 0.0, 0.0};
    store4(SBAR, b, ext.r, ext.u, ext.v, ext.p);
  }
}

// Geometry adjoint of the face terms, lambda^T dR/d(face geometry): per
// flux face (nx, ny, L, lx, ly, rx, ry)-bar into out[7k..7k+6]. L enters
// the residual as -L/A_l, +L/A_r; the normal through the flux and the wall's
// mirror state; the offsets through the reconstruction W + phi (g . d).
// Rusanov and HLLC.
__global__ void __launch_bounds__(kBlock)
    k_geom_face(int nf, FaceArrays fa, const double* __restrict__ W, const double* __restrict__ G,
                const double* __restrict__ PHI, const double* __restrict__ flen,
                const double* __restrict__ area, const double* __restrict__ lam, int scheme, int muscl,
                double g, double* __restrict__ out) {
  int k = xcd_block() * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  int kind = fa.kind[k];
  double nx = fa.nx[k], ny = fa.ny[k];
  int l = fa.left[k];
  int r = fa.right[k];
  double L = flen[k];
  P4 ll = load4(lam, l);
  double Al = area[l];
  double cl = -L / Al;
  double fb[4] = {cl * ll.r, cl * ll.u, cl * ll.v, cl * ll.p};
  double lc[4] = {-ll.r / Al, -ll.u / Al, -ll.v / Al, -ll.p / Al};  // dR/dL per flux component
  const bool two = kind <= MCFD_FACE_PERIODIC;
  if (two) {
    P4 lr = load4(lam, r);
    double Ar = area[r];
    double cr = L / Ar;
    fb[0] += cr * lr.r;
    fb[1] += cr * lr.u;
    fb[2] += cr * lr.v;
    fb[3] += cr * lr.p;
    lc[0] += lr.r / Ar;
    lc[1] += lr.u / Ar;
    lc[2] += lr.v / Ar;
    lc[3] += lr.p / Ar;
  }
  P4 wl = reconstruct(W, G, PHI, l, fa.lx[k], fa.ly[k], muscl != 0);
  P4 wr = wl;
  int b = fa.bc[k];
  double wrx = 0.0, wry = 0.0, wun = 0.0;
  if (two) {
    wr = reconstruct(W, G, PHI, r, fa.rx[k], fa.ry[k], muscl != 0);
  } else if (kind == MCFD_FACE_WALL) {
    wrx = wl.u - fa.bc_wall[2 * (long)b];
    wry = wl.v - fa.bc_wall[2 * (long)b + 1];
    wun = wrx * nx + wry * ny;
    wr.u = wl.u - 2.0 * wun * nx;
    wr.v = wl.v - 2.0 * wun * ny;
  } else if (kind == MCFD_FACE_VELOCITY_INLET) {
    const double* st = fa.bc_state + 4 * (long)b;
    wr = P4{st[0], st[1], st[2], st[3]};
  } else if (kind == MCFD_FACE_PRESSURE_OUTLET) {
    wr.p = fa.bc_state[4 * (long)b + 3];
  } else if (kind == MCFD_FACE_FAR_FIELD) {
    const double* st = fa.bc_state + 4 * (long)b;
    int err = 0;
    wr = far_field(wl, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, &err);
  }
  F4 F = (scheme == 1) ? hllc(wl, wr, nx, ny, g, kind == MCFD_FACE_INTERIOR) : rusanov(wl, wr, nx, ny, g);
  double lbar = F.a * lc[0] + F.b * lc[1] + F.c * lc[2] + F.d * lc[3];
  P4 lb, rb;
  double nbx = 0.0, nby = 0.0;
  if (scheme == 1)
    hllc_b(wl, wr, nx, ny, g, kind == MCFD_FACE_INTERIOR, fb, lb, rb, nbx, nby);
  else
    rusanov_b(wl, wr, nx, ny, g, fb, lb, rb, nbx, nby);
  if (!two) {
    if (kind == MCFD_FACE_WALL) {
      // u_R = u_L - 2 un nx, v_R = v_L - 2 un ny, un = (u_L - wu) nx + (v_L - wv) ny
      nbx += rb.u * (-2.0 * (wrx * nx + wun)) + rb.v * (-2.0 * wrx * ny);
      nby += rb.u * (-2.0 * wry * nx) + rb.v * (-2.0 * (wry * ny + wun));
      double un_b = rb.u * nx + rb.v * ny;
      lb.r += rb.r;
      lb.u += rb.u - 2.0 * nx * un_b;
      lb.v += rb.v - 2.0 * ny * un_b;
      lb.p += rb.p;
    } else if (kind == MCFD_FACE_PRESSURE_OUTLET) {
      lb.r += rb.r;
      lb.u += rb.u;
      lb.v += rb.v;
    } else if (kind == MCFD_FACE_FAR_FIELD) {
      const double* st = fa.bc_state + 4 * (long)b;
      P4 outb{0.0, 0.0, 0.0, 0.0};
      far_field_b(wl, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, rb, lb, outb, nbx, nby);
    } else if (kind != MCFD_FACE_VELOCITY_INLET) {
      lb.r += rb.r;
      lb.u += rb.u;
      lb.v += rb.v;
      lb.p += rb.p;
    }
  }
  double dlx = 0.0, dly = 0.0, drx = 0.0, dry = 0.0;
  if (muscl) {
    const double* gl = G + 8 * (long)l;
    P4 pl = load4(PHI, l);
    double a[4] = {lb.r * pl.r, lb.u * pl.u, lb.v * pl.v, lb.p * pl.p};
    for (int c = 0; c < 4; ++c) {
      dlx += a[c] * gl[2 * c];
      dly += a[c] * gl[2 * c + 1];
    }
    if (two) {
      const double* gr = G + 8 * (long)r;
      P4 pr = load4(PHI, r);
      double q[4] = {rb.r * pr.r, rb.u * pr.u, rb.v * pr.v, rb.p * pr.p};
      for (int c = 0; c < 4; ++c) {
        drx += q[c] * gr[2 * c];
        dry += q[c] * gr[2 * c + 1];
      }
    }
  }
  double* o = out + 7 * (long)k;
  o[0] = nbx;
  o[1] = nby;
  o[2] = lbar;
  o[3] = dlx;
  o[4] = dly;
  o[5] = drx;
  o[6] = dry;
}

// Face adjoint: Fbar = sum of the face's residual coefficients times lambda,
// then (dF/dW_L)^T Fbar and (dF/dW_R)^T Fbar -> FADJ[k][0..7].
__global__ void __launch_bounds__(kBlock)
    k_flux_adjoint(int nf, FaceArrays fa, const double* __restrict__ W,
                   const double* __restrict__ G, const double* __restrict__ PHI,
                   const double* __restrict__ K, const double* __restrict__ flen,
                   const double* __restrict__ area,
                   const double* __restrict__ lam, int scheme, int muscl, double g,
                   double* __restrict__ FADJ, double* __restrict__ SBAR, int only_far_field) {
  int k = xcd_block() * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  int kind = fa.kind[k];
  if (only_far_field && kind != MCFD_FACE_FAR_FIELD) return;
  double nx = fa.nx[k], ny = fa.ny[k];
  int l = fa.left[k];
  int r = fa.right[k];
  double L = flen[k];
  // dR/dF: owner/left side -L/A (interior: -L/A * F; boundary: -(F*L)/A has the same derivative)
  P4 ll = load4(lam, l);
  double cl = -L / area[l];
  double fb[4] = {cl * ll.r, cl * ll.u, cl * ll.v, cl * ll.p};
  if (kind <= MCFD_FACE_PERIODIC) {
    P4 lr = load4(lam, r);
    double cr = L / area[r];
    fb[0] += cr * lr.r;
    fb[1] += cr * lr.u;
    fb[2] += cr * lr.v;
    fb[3] += cr * lr.p;
  }
  P4 wl = K ? reconstruct_q(W, G, K, PHI, l, fa.lx[k], fa.ly[k])
            : reconstruct(W, G, PHI, l, fa.lx[k], fa.ly[k], muscl != 0);
  P4D DL = seed(wl, 0);
  P4D DR;
  if (kind <= MCFD_FACE_PERIODIC) {
    P4 wr = K ? reconstruct_q(W, G, K, PHI, r, fa.rx[k], fa.ry[k])
              : reconstruct(W, G, PHI, r, fa.rx[k], fa.ry[k], muscl != 0);
    DR = seed(wr, 4);
  } else {
    int b = fa.bc[k];
    const double* st = fa.bc_state + 4 * (long)b;
    if (kind == MCFD_FACE_WALL) {
      D8 rx = DL.u - fa.bc_wall[2 * (long)b];
      D8 ry = DL.v - fa.bc_wall[2 * (long)b + 1];
      D8 un = rx * nx + ry * ny;
      DR = DL;
      DR.u = DL.u - 2.0 * un * nx;
      DR.v = DL.v - 2.0 * un * ny;
    } else if (kind == MCFD_FACE_VELOCITY_INLET) {
      // directions 4-7 carry the record's exterior state (inflow gradients)
      DR = seed(P4{st[0], st[1], st[2], st[3]}, 4);
    } else if (kind == MCFD_FACE_PRESSURE_OUTLET) {
      DR = DL;
      DR.p = dconst(st[3]);
    } else if (kind == MCFD_FACE_FAR_FIELD) {
      DR = far_field_d(DL, seed(P4{st[0], st[1], st[2], st[3]}, 4), nx, ny, g);
    } else if (kind >= MCFD_FACE_PRESSURE_INLET) {
      // directions 4-7: the inflow record's exterior state (k_inflow_adjoint)
      DR = seed(load4(fa.inflow, b - fa.inflow_first), 4);
    } else {
      DR = DL;
    }
  }
  F4D f;
  if (scheme == 1)
    f = hllc_d(DL, DR, nx, ny, g, kind == MCFD_FACE_INTERIOR);
  else if (scheme == 2)
    f = roe_d(DL, DR, nx, ny, g);
  else
    f = rusanov_d(DL, DR, nx, ny, g);
  double out[8];
#pragma unroll
  for (int j = 0; j < 8; ++j)
    out[j] = fb[0] * f.a.d[j] + fb[1] * f.b.d[j] + fb[2] * f.c.d[j] + fb[3] * f.d.d[j];
  double2* q = reinterpret_cast<double2*>(FADJ + 8 * (long)k);
  q[0] = make_double2(out[0], out[1]);
  q[1] = make_double2(out[2], out[3]);
  q[2] = make_double2(out[4], out[5]);
  q[3] = make_double2(out[6], out[7]);
  if (SBAR && kind > MCFD_FACE_PERIODIC) {
    // lambda^T dR/d(exterior state) of the face's boundary record
    bool ext = kind == MCFD_FACE_VELOCITY_INLET || kind == MCFD_FACE_FAR_FIELD;
    store4(SBAR, fa.bc[k], ext ? out[4] : 0.0, ext ? out[5] : 0.0, ext ? out[6] : 0.0,
           ext ? out[7] : 0.0);
  }
}

// Cell pass 1: gather face adjoints (reconstruction backward), limiter
// backward, LSQ-gradient backward. Writes per cell Wbar without the
// neighbours' contributions, and for each LSQ entry e (i -> j) the 4-vector
// that cell i's gradient and extremum adjoints contribute to W_j,
// c_e . gbar_i (+ the extremum adjoint if j holds i's extremum), into
// ebar[rev(e)], the slot of j's entry for i.
struct AdjArrays {
  const int* acc_ptr;
  const int* acc_item;
  const double* flx;   // face offsets, for the side of each item
  const double* fly;
  const double* frx;
  const double* fry;
  const int* grad_ptr;
  const int* grad_nb;
  const double* grad_cx;
  const double* grad_cy;
  const int* grad_rev;  // reverse entry of each LSQ entry
  const int* lim_ptr;
  const double* lim_dx;
  const double* lim_dy;
  double* wbar;   // (C,4)
  double* ebar;   // (n_grad,4) per LSQ entry
  unsigned* arg;  // (C,4) limiting / extremum entries of the forward limiter (k_gradient_t<true>)
  double* cbar;   // (n_grad,2) LSQ coefficient adjoints, or NULL
  double* dlbar;  // (n_lim,2) limiter offset adjoints, or NULL
  // MUSCL parabolic (all NULL otherwise): curvatures, two-ring fit entries,
  // their reverse entries and the per-entry contributions
  const double* K;
  const int* pfull;
  const int* par_ptr;
  const int* par_nb;
  const double* par_coef;
  const int* par_rev;
  double* ebar_par;
  // viscous terms (NULL otherwise): per LSQ entry the (u, v, T) adjoint a
  // cell's WLSQ gradient gives its neighbour, stored at the reverse entry
  double* vbar;
};

__device__ __forceinline__ double limiter_value(double ratio, int code, double& dfdr) {
  double f;
  dfdr = 0.0;
  if (code <= 1) {
    if (ratio < 1.0) {
      f = ratio;
      dfdr = 1.0;
    } else {
      f = 1.0;
    }
  } else if (code == 2) {
    if (ratio > 0.0) {
      double t = 2.0 * ratio / (1.0 + ratio);
      if (t < 1.0) {
        f = t;
        dfdr = 2.0 / ((1.0 + ratio) * (1.0 + ratio));
      } else {
        f = 1.0;
      }
    } else {
      f = 0.0;
    }
  } else {
    const double eps = 1e-12;
    double num = ratio * ratio + 2.0 * ratio + eps;
    double den = ratio * ratio + ratio + 2.0 + eps;
    f = num / den;
    if (f < 1.0) {
      dfdr = ((2.0 * ratio + 2.0) * den - num * (2.0 * ratio + 1.0)) / (den * den);
    } else {
      f = 1.0;
    }
  }
  return f;
}

__global__ void __launch_bounds__(kBlock)
    k_cell_adjoint1(int n, const double* __restrict__ W, const double* __restrict__ G,
                    const double* __restrict__ PHI, const double* __restrict__ FADJ, AdjArrays aa,
                    int muscl, int code) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double wb[4] = {0.0, 0.0, 0.0, 0.0};
  double pb[4] = {0.0, 0.0, 0.0, 0.0};
  double gb[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double gv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  double pv[4] = {1.0, 1.0, 1.0, 1.0};
  const bool quad = aa.K != nullptr;
  double kv[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // curvature [k][xx, xy, yy]
  double kb[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (quad)
    for (int a = 0; a < 12; ++a) kv[a] = aa.K[12 * (long)i + a];
  if (muscl) {
    const double2* gq = reinterpret_cast<const double2*>(G + 8 * (long)i);
    for (int k = 0; k < 4; ++k) {
      double2 t = gq[k];
      gv[2 * k] = t.x;
      gv[2 * k + 1] = t.y;
    }
    P4 f = load4(PHI, i);
    pv[0] = f.r;
    pv[1] = f.u;
    pv[2] = f.v;
    pv[3] = f.p;
  }
  for (int e = aa.acc_ptr[i]; e < aa.acc_ptr[i + 1]; ++e) {
    int item = aa.acc_item[e];
    int f = item >> 2;
    int mode = item & 3;
    bool right = (mode == MCFD_ACC_NEIGHBOUR) || (mode == MCFD_ACC_PERIODIC_B);
    const double2* q = reinterpret_cast<const double2*>(FADJ + 8 * (long)f + (right ? 4 : 0));
    double2 q0 = q[0], q1 = q[1];
    double a[4] = {q0.x, q0.y, q1.x, q1.y};
    double dx = right ? aa.frx[f] : aa.flx[f];
    double dy = right ? aa.fry[f] : aa.fly[f];
    for (int k = 0; k < 4; ++k) {
      wb[k] += a[k];
      if (muscl) {
        double inc = gv[2 * k] * dx + gv[2 * k + 1] * dy;
        if (quad)
          inc += 0.5 * kv[3 * k] * (dx * dx) + kv[3 * k + 1] * (dx * dy) + 0.5 * kv[3 * k + 2] * (dy * dy);
        pb[k] += a[k] * inc;
        double ap = a[k] * pv[k];
        gb[2 * k] += ap * dx;
        gb[2 * k + 1] += ap * dy;
        if (quad) {
          kb[3 * k] += ap * 0.5 * (dx * dx);
          kb[3 * k + 1] += ap * (dx * dy);
          kb[3 * k + 2] += ap * 0.5 * (dy * dy);
        }
      }
    }
  }
  if (muscl) {
    P4 w = load4(W, i);
    double wv[4] = {w.r, w.u, w.v, w.p};
    const int g0 = aa.grad_ptr[i];
    const bool full = quad && aa.pfull[i];
    const int p0 = quad ? aa.par_ptr[i] : 0;
    uint4 av = *reinterpret_cast<const uint4*>(aa.arg + 4 * (long)i);
    unsigned arg[4] = {av.x, av.y, av.z, av.w};
    // extremum adjoints: xb[k] goes to the entry holding the extremum
    // (0xFF: the cell itself); at most one per variable is non-zero
    double xb[4] = {0.0, 0.0, 0.0, 0.0};
    int xe[4] = {-1, -1, -1, -1};
    if (aa.dlbar)
      for (int e = aa.lim_ptr[i]; e < aa.lim_ptr[i + 1]; ++e) {
        aa.dlbar[2 * (long)e] = 0.0;
        aa.dlbar[2 * (long)e + 1] = 0.0;
      }
    // limiter backward at the entry that set phi (the forward's first strict minimum)
    for (int k = 0; k < 4; ++k) {
      unsigned le = arg[k] & 0xFFu;
      if (pb[k] == 0.0 || le == 0xFFu) continue;
      int e = aa.lim_ptr[i] + (int)le;
      double dx = aa.lim_dx[e], dy = aa.lim_dy[e];
      double inc = gv[2 * k] * dx + gv[2 * k + 1] * dy;
      if (quad)
        inc += 0.5 * kv[3 * k] * (dx * dx) + kv[3 * k + 1] * (dx * dy) + 0.5 * kv[3 * k + 2] * (dy * dy);
      double tol = 1e-14 * vmax(fabs(wv[k]), 1.0);
      if (!(inc > tol) && !(inc < -tol)) continue;  // ratio 1: no dependence
      bool up = inc > tol;
      unsigned xa = up ? ((arg[k] >> 8) & 0xFFu) : ((arg[k] >> 16) & 0xFFu);
      // extremum entries count in the two-ring list of full-rank cells
      int xn = (xa == 0xFFu) ? -1 : (full ? p0 : g0) + (int)xa;
      double ext = wv[k];
      if (xn >= 0) {
        P4 o = load4(W, full ? aa.par_nb[xn] : aa.grad_nb[xn]);
        double ov[4] = {o.r, o.u, o.v, o.p};
        ext = ov[k];
      }
      double ratio = (ext - wv[k]) / inc;
      if (0.0 > ratio) continue;  // clamped at 0
      double dfdr;
      (void)limiter_value(ratio, code, dfdr);
      if (dfdr == 0.0) continue;
      double rb = pb[k] * dfdr;      // d phi / d ratio
      double ext_b = rb / inc;       // d ratio / d extremum
      wb[k] -= ext_b;                // d ratio / d W_i
      double incb = -rb * ratio / inc;
      gb[2 * k] += incb * dx;
      gb[2 * k + 1] += incb * dy;
      if (quad) {
        kb[3 * k] += incb * 0.5 * (dx * dx);
        kb[3 * k + 1] += incb * (dx * dy);
        kb[3 * k + 2] += incb * 0.5 * (dy * dy);
      }
      if (aa.dlbar) {
        aa.dlbar[2 * (long)e] += incb * gv[2 * k];
        aa.dlbar[2 * (long)e + 1] += incb * gv[2 * k + 1];
      }
      if (xn < 0)
        wb[k] += ext_b;
      else {
        xb[k] = ext_b;
        xe[k] = xn;
      }
    }
    if (quad) {
      // quadratic fit backward (full rank): (grad, curvature)_i = sum_e P_e (W_j - W_i)
      for (int e = p0; e < aa.par_ptr[i + 1]; ++e) {
        double c[4] = {0.0, 0.0, 0.0, 0.0};
        if (full) {
          const double* pe = aa.par_coef + 5 * (long)e;
          for (int k = 0; k < 4; ++k) {
            double t = pe[0] * gb[2 * k] + pe[1] * gb[2 * k + 1] + pe[2] * kb[3 * k] +
                       pe[3] * kb[3 * k + 1] + pe[4] * kb[3 * k + 2];
            wb[k] -= t;
            if (xe[k] == e) t += xb[k];
            c[k] = t;
          }
        }
        store4(aa.ebar_par, aa.par_rev[e], c[0], c[1], c[2], c[3]);
      }
    }
    // LSQ gradient backward: grad_i = sum_e c_e (W_j - W_i)
    for (int e = g0; e < aa.grad_ptr[i + 1]; ++e) {
      double cx = aa.grad_cx[e], cy = aa.grad_cy[e];
      double c[4];
      for (int k = 0; k < 4; ++k) {
        double t = full ? 0.0 : cx * gb[2 * k] + cy * gb[2 * k + 1];
        wb[k] -= t;
        if (!full && xe[k] == e) t += xb[k];
        c[k] = t;
      }
      store4(aa.ebar, aa.grad_rev[e], c[0], c[1], c[2], c[3]);  // read back by cell j in entry order
      if (aa.cbar) {  // grad_i = sum_e c_e (W_j - W_i)
        P4 o = load4(W, aa.grad_nb[e]);
        double d[4] = {o.r - wv[0], o.u - wv[1], o.v - wv[2], o.p - wv[3]};
        double bx = 0.0, by = 0.0;
        for (int k = 0; k < 4; ++k) {
          bx += gb[2 * k] * d[k];
          by += gb[2 * k + 1] * d[k];
        }
        aa.cbar[2 * (long)e] = bx;
        aa.cbar[2 * (long)e + 1] = by;
      }
    }
  }
  store4(aa.wbar, i, wb[0], wb[1], wb[2], wb[3]);
}

// ---------------------------------------------------------------------------
// Adjoint of the laminar viscous terms (reverse of k_visc_gradient,
// k_visc_flux and visc_gather). Per flux face and side a 12-double record
// {gbar[6] (cell gradient of u, v, T), mubar, kbar (cell viscosity and
// conductivity), Wbar[4] (direct primitive adjoint)}; k_visc_cell_adjoint
// gathers them per cell and reverses the WLSQ gradient with its ghost
// entries, adding to wbar and to the per-entry neighbour slots vbar.

__device__ __forceinline__ double visc_dmu_dT(double T, double mu, const ViscParams& v) {
  if (v.model == MCFD_VISCOSITY_CONSTANT) return 0.0;
  return mu * (1.5 / T - 1.0 / (T + v.S));
}

// Reverse of visc_flux: F-bar (components 1..3) -> g-bar, u, v, mu, k adjoints.
__device__ __forceinline__ void visc_flux_b(const double g[6], double u, double v, double mu, double k,
                                            double nx, double ny, const double Fb[4], double gb[6],
                                            double& ub, double& vb, double& mub, double& kb) {
  double div = g[0] + g[3];
  double txx = mu * (2.0 * g[0] - (2.0 / 3.0) * div);
  double tyy = mu * (2.0 * g[3] - (2.0 / 3.0) * div);
  double txy = mu * (g[1] + g[2]);
  double tx = txx * nx + txy * ny;
  double ty = txy * nx + tyy * ny;
  double txb = Fb[1] + Fb[3] * u;
  double tyb = Fb[2] + Fb[3] * v;
  ub += Fb[3] * tx;
  vb += Fb[3] * ty;
  double qb = Fb[3];
  double txxb = txb * nx, txyb = txb * ny + tyb * nx, tyyb = tyb * ny;
  kb += qb * (g[4] * nx + g[5] * ny);
  gb[4] += qb * k * nx;
  gb[5] += qb * k * ny;
  gb[0] += txxb * (4.0 / 3.0) * mu + tyyb * (-2.0 / 3.0) * mu;
  gb[3] += txxb * (-2.0 / 3.0) * mu + tyyb * (4.0 / 3.0) * mu;
  gb[1] += txyb * mu;
  gb[2] += txyb * mu;
  mub += txxb * (2.0 * g[0] - (2.0 / 3.0) * div) + tyyb * (2.0 * g[3] - (2.0 / 3.0) * div) +
         txyb * (g[1] + g[2]);
}

// (u, v, T)-bar of a cell state -> primitive (rho, u, v, p)-bar, T = p/(rho R)
__device__ __forceinline__ void uvT_to_prim_b(const P4& w, double R, double ub, double vb, double Tb,
                                              double out[4]) {
  out[0] += -Tb * w.p / (w.r * w.r * R);
  out[1] += ub;
  out[2] += vb;
  out[3] += Tb / (w.r * R);
}

// Exterior (u, v, T) adjoint of visc_inlet -> interior primitive adjoint
// (far field through far_field_b; a velocity inlet's state is fixed).
// Reverse of visc_inlet: the boundary values' adjoint xb (u, v, T) into the
// owner state (far field) and, when sb is given, into the boundary record's
// state adjoint (rho, u, v, p), added to what the convective faces stored.
__device__ __forceinline__ void visc_inlet_b(int code, const P4& w, const double* st, double nx, double ny,
                                             const ViscParams& v, const double xb[3], double wb[4],
                                             double* sb, int kind = 0, const double* par = nullptr,
                                             double proj = 0.0, double* vib = nullptr) {
  if (code == kVisInflow) {
    // pressure / mass-flow inlet: st is the record's exterior state from the
    // owner cell w; its (u, v, T) adjoint goes through inflow_state_b into w,
    // and for a mass-flow inlet also into the record's slot of vib for the
    // patch pass (k_mass_flow_inlet_visc_adjoint)
    const P4 e{st[0], st[1], st[2], st[3]};
    const P4 eb{-xb[2] * e.p / (e.r * e.r * v.R), xb[0], xb[1], xb[2] / (e.r * v.R)};
    P4 inb{0.0, 0.0, 0.0, 0.0};
    inflow_state_b(kind, w, e, nx, ny, par, proj, v.gamma, v.R, eb, inb);
    wb[0] += inb.r;
    wb[1] += inb.u;
    wb[2] += inb.v;
    wb[3] += inb.p;
    if (kind == MCFD_FACE_MASS_FLOW_INLET && vib) {
      vib[0] += eb.r;
      vib[1] += eb.u;
      vib[2] += eb.v;
      vib[3] += eb.p;
    }
    return;
  }
  if (code == kVisInlet) {
    if (sb) {
      sb[0] += -xb[2] * st[3] / (st[0] * st[0] * v.R);
      sb[1] += xb[0];
      sb[2] += xb[1];
      sb[3] += xb[2] / (st[0] * v.R);
    }
    return;
  }
  if (code != kVisFarField) return;
  int err = 0;
  P4 out{st[0], st[1], st[2], st[3]};
  P4 e = far_field(w, out, nx, ny, v.gamma, &err);
  P4 eb{-xb[2] * e.p / (e.r * e.r * v.R), xb[0], xb[1], xb[2] / (e.r * v.R)};
  P4 inb{0.0, 0.0, 0.0, 0.0}, outb{0.0, 0.0, 0.0, 0.0};
  double nbx = 0.0, nby = 0.0;
  far_field_b(w, out, nx, ny, v.gamma, eb, inb, outb, nbx, nby);
  wb[0] += inb.r;
  wb[1] += inb.u;
  wb[2] += inb.v;
  wb[3] += inb.p;
  if (sb) {
    sb[0] += outb.r;
    sb[1] += outb.u;
    sb[2] += outb.v;
    sb[3] += outb.p;
  }
}

__global__ void __launch_bounds__(kBlock)
    k_visc_face_adjoint(int nf, FaceArrays fa, const double* __restrict__ W,
                        const double* __restrict__ VG, ViscArrays va, ViscParams vp,
                        const double* __restrict__ flen, const double* __restrict__ area,
                        const double* __restrict__ lam, double* __restrict__ VFA,
                        double* __restrict__ SBAR, InflowArrays ia, double* __restrict__ VIB) {
  int f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int kind = fa.kind[f];
  double nx = fa.nx[f], ny = fa.ny[f];
  int o = fa.left[f];
  const double* go = VG + 8 * (long)o;
  double L = flen[f];
  // the viscous terms enter the owner as +(F L)/A and the other side as -(F L)/A
  P4 lo = load4(lam, o);
  double co = L / area[o];
  double Fb[4] = {co * lo.r, co * lo.u, co * lo.v, co * lo.p};
  double rec[2][12];
#pragma unroll
  for (int a = 0; a < 12; ++a) rec[0][a] = rec[1][a] = 0.0;
  if (kind == MCFD_FACE_INTERIOR || kind == MCFD_FACE_PERIODIC) {
    int r = fa.right[f];
    const double* gr = VG + 8 * (long)r;
    P4 lr = load4(lam, r);
    double cr = L / area[r];
    Fb[0] -= cr * lr.r;
    Fb[1] -= cr * lr.u;
    Fb[2] -= cr * lr.v;
    Fb[3] -= cr * lr.p;
    double g[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) g[k] = 0.5 * (go[k] + gr[k]);
    P4 a = load4(W, o), c = load4(W, r);
    double gb[6] = {0, 0, 0, 0, 0, 0}, ub = 0.0, vb = 0.0, mub = 0.0, kb = 0.0;
    visc_flux_b(g, 0.5 * (a.u + c.u), 0.5 * (a.v + c.v), 0.5 * (go[6] + gr[6]), 0.5 * (go[7] + gr[7]), nx,
                ny, Fb, gb, ub, vb, mub, kb);
    for (int sd = 0; sd < 2; ++sd) {
#pragma unroll
      for (int k = 0; k < 6; ++k) rec[sd][k] = 0.5 * gb[k];
      rec[sd][6] = 0.5 * mub;
      rec[sd][7] = 0.5 * kb;
      rec[sd][9] = 0.5 * ub;
      rec[sd][10] = 0.5 * vb;
    }
  } else {
    int b = fa.bc[f];
    int code = va.bv_code[b];
    const double* bg = va.bv_geom + 4 * (long)b;
    double dn = bg[0], tx = bg[1], ty = bg[2];
    double offx = fa.lx[f], offy = fa.ly[f];
    P4 w = load4(W, o);
    double T = w.p / (w.r * vp.R);
    int flags = va.bv_flags[b];
    int thermal = (flags >> 1) & 3;
    // forward (as k_visc_flux), keeping what the reverse needs
    double g[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) g[k] = go[k];
    double bv[3] = {w.u, w.v, T};
    double x[3] = {0.0, 0.0, 0.0};
    if (code == kVisWall) {
      if (flags & 1) {
        const double* wv = fa.bc_wall + 2 * (long)b;
        for (int c = 0; c < 2; ++c) {
          double nd = (wv[c] - bv[c] - (g[2 * c] * tx + g[2 * c + 1] * ty)) / dn;
          double sc = nd - (g[2 * c] * nx + g[2 * c + 1] * ny);
          g[2 * c] = g[2 * c] + sc * nx;
          g[2 * c + 1] = g[2 * c + 1] + sc * ny;
          bv[c] = wv[c];
        }
      } else {
        double un = bv[0] * nx + bv[1] * ny;
        bv[0] = bv[0] + (g[0] * offx + g[1] * offy);
        bv[1] = bv[1] + (g[2] * offx + g[3] * offy);
        double un2 = bv[0] * nx + bv[1] * ny;
        bv[0] = bv[0] - un2 * nx;
        bv[1] = bv[1] - un2 * ny;
        double gnx = nx * g[0] + ny * g[2];
        double gny = nx * g[1] + ny * g[3];
        double nd = -(un + (gnx * tx + gny * ty)) / dn;
        double vn0 = g[0] * nx + g[1] * ny;
        double vn1 = g[2] * nx + g[3] * ny;
        g[0] = g[0] + (nd * nx * nx - vn0 * nx);
        g[1] = g[1] + (nd * nx * ny - vn0 * ny);
        g[2] = g[2] + (nd * ny * nx - vn1 * nx);
        g[3] = g[3] + (nd * ny * ny - vn1 * ny);
      }
      if (thermal == 1) {
        double nd = (bg[3] - bv[2] - (g[4] * tx + g[5] * ty)) / dn;
        bv[2] = bg[3];
        double sc = nd - (g[4] * nx + g[5] * ny);
        g[4] = g[4] + sc * nx;
        g[5] = g[5] + sc * ny;
      } else {
        double nd = bg[3] / go[7];
        bv[2] = bv[2] + (g[4] * offx + g[5] * offy);
        double sc = nd - (g[4] * nx + g[5] * ny);
        g[4] = g[4] + sc * nx;
        g[5] = g[5] + sc * ny;
      }
    } else if (code == kVisSymmetry) {
      double un = bv[0] * nx + bv[1] * ny;
      double dv[2] = {-un / dn * nx, -un / dn * ny};
      for (int c = 0; c < 2; ++c) {
        double sc = dv[c] - (g[2 * c] * nx + g[2 * c + 1] * ny);
        g[2 * c] = g[2 * c] + sc * nx;
        g[2 * c + 1] = g[2 * c + 1] + sc * ny;
      }
      double gt = g[4] * nx + g[5] * ny;
      g[4] = g[4] - gt * nx;
      g[5] = g[5] - gt * ny;
      bv[0] = bv[0] - un * nx;
      bv[1] = bv[1] - un * ny;
    } else if (code == kVisInlet || code == kVisInflow || code == kVisFarField) {
      int dummy = 0;
      visc_inlet(code, w, visc_inlet_state(code, b, va, fa), nx, ny, vp, x, &dummy, o);
      double cv[3] = {w.u, w.v, T};
      for (int c = 0; c < 3; ++c) {
        bv[c] = x[c];
        double nd = (bv[c] - cv[c] - (g[2 * c] * tx + g[2 * c + 1] * ty)) / dn;
        double sc = nd - (g[2 * c] * nx + g[2 * c + 1] * ny);
        g[2 * c] = g[2 * c] + sc * nx;
        g[2 * c + 1] = g[2 * c + 1] + sc * ny;
      }
    } else {
      for (int c = 0; c < 3; ++c) {
        double gn = g[2 * c] * nx + g[2 * c + 1] * ny;
        g[2 * c] = g[2 * c] - gn * nx;
        g[2 * c + 1] = g[2 * c + 1] - gn * ny;
        bv[c] = bv[c] + (g[2 * c] * offx + g[2 * c + 1] * offy);
      }
    }
    int dummy = 0;
    double fmu = visc_mu(bv[2], vp, &dummy, o);
    double fk = visc_k(fmu, vp);
    // reverse
    if (code == kVisWall && thermal == 2) Fb[3] = 0.0;  // F[3] = prescribed heat flux
    double gb[6] = {0, 0, 0, 0, 0, 0}, bvb[3] = {0.0, 0.0, 0.0}, fmub = 0.0, fkb = 0.0;
    visc_flux_b(g, bv[0], bv[1], fmu, fk, nx, ny, Fb, gb, bvb[0], bvb[1], fmub, fkb);
    fmub += fkb * vp.gamma * vp.R / ((vp.gamma - 1.0) * vp.Pr);
    bvb[2] += fmub * visc_dmu_dT(bv[2], fmu, vp);
    double wb[4] = {0.0, 0.0, 0.0, 0.0};   // direct primitive adjoint of the owner
    double uvTb[3] = {0.0, 0.0, 0.0};      // adjoint of the owner's (u, v, T)
    double kob = 0.0;                      // adjoint of the owner's conductivity
    if (code == kVisWall) {
      // thermal part (applied last in the forward)
      if (thermal == 1) {
        double sb = gb[4] * nx + gb[5] * ny;
        gb[4] += sb * (-tx / dn - nx);
        gb[5] += sb * (-ty / dn - ny);
        uvTb[2] += -sb / dn;  // bv[2] before the reset was T
      } else {
        double sb = gb[4] * nx + gb[5] * ny;
        gb[4] -= sb * nx;
        gb[5] -= sb * ny;
        kob += -sb * bg[3] / (go[7] * go[7]);
        uvTb[2] += bvb[2];
        gb[4] += bvb[2] * offx;
        gb[5] += bvb[2] * offy;
      }
      if (flags & 1) {
        for (int c = 0; c < 2; ++c) {
          double sb = gb[2 * c] * nx + gb[2 * c + 1] * ny;
          gb[2 * c] += sb * (-tx / dn - nx);
          gb[2 * c + 1] += sb * (-ty / dn - ny);
          uvTb[c] += -sb / dn;
        }
      } else {
        // bv'' = (I - n n^T) bv'
        double pb = bvb[0] * nx + bvb[1] * ny;
        double b0 = bvb[0] - nx * pb, b1 = bvb[1] - ny * pb;
        double ndb = gb[0] * nx * nx + gb[1] * nx * ny + gb[2] * ny * nx + gb[3] * ny * ny;
        double vn0b = -(gb[0] * nx + gb[1] * ny);
        double vn1b = -(gb[2] * nx + gb[3] * ny);
        gb[0] += vn0b * nx;
        gb[1] += vn0b * ny;
        gb[2] += vn1b * nx;
        gb[3] += vn1b * ny;
        double unb = -ndb / dn;
        double gnxb = -ndb * tx / dn, gnyb = -ndb * ty / dn;
        gb[0] += gnxb * nx;
        gb[2] += gnxb * ny;
        gb[1] += gnyb * nx;
        gb[3] += gnyb * ny;
        uvTb[0] += b0;
        gb[0] += b0 * offx;
        gb[1] += b0 * offy;
        uvTb[1] += b1;
        gb[2] += b1 * offx;
        gb[3] += b1 * offy;
        uvTb[0] += unb * nx;
        uvTb[1] += unb * ny;
      }
    } else if (code == kVisSymmetry) {
      double unb = -(bvb[0] * nx + bvb[1] * ny);
      uvTb[0] += bvb[0];
      uvTb[1] += bvb[1];
      uvTb[2] += bvb[2];
      double gtb = gb[4] * nx + gb[5] * ny;
      gb[4] -= gtb * nx;
      gb[5] -= gtb * ny;
      for (int c = 0; c < 2; ++c) {
        double sb = gb[2 * c] * nx + gb[2 * c + 1] * ny;
        gb[2 * c] -= sb * nx;
        gb[2 * c + 1] -= sb * ny;
        unb += -sb * (c == 0 ? nx : ny) / dn;
      }
      uvTb[0] += unb * nx;
      uvTb[1] += unb * ny;
    } else if (code == kVisInlet || code == kVisInflow || code == kVisFarField) {
      double xb[3];
      for (int c = 0; c < 3; ++c) {
        double sb = gb[2 * c] * nx + gb[2 * c + 1] * ny;
        gb[2 * c] += sb * (-tx / dn - nx);
        gb[2 * c + 1] += sb * (-ty / dn - ny);
        xb[c] = sb / dn + bvb[c];
        uvTb[c] += -sb / dn;
      }
      const int ir = b - va.inf_first;
      const bool rec = code == kVisInflow;
      visc_inlet_b(code, w, visc_inlet_state(code, b, va, fa), nx, ny, vp, xb, wb,
                   SBAR ? SBAR + 4 * (long)b : nullptr, kind, rec ? ia.params + 4 * (long)ir : nullptr,
                   rec ? ia.proj[ir] : 0.0, rec && VIB ? VIB + 4 * (long)ir : nullptr);
    } else {
      for (int c = 0; c < 3; ++c) {
        gb[2 * c] += bvb[c] * offx;
        gb[2 * c + 1] += bvb[c] * offy;
        uvTb[c] += bvb[c];
        double gnb = gb[2 * c] * nx + gb[2 * c + 1] * ny;
        gb[2 * c] -= gnb * nx;
        gb[2 * c + 1] -= gnb * ny;
      }
    }
    uvT_to_prim_b(w, vp.R, uvTb[0], uvTb[1], uvTb[2], wb);
#pragma unroll
    for (int k = 0; k < 6; ++k) rec[0][k] = gb[k];
    rec[0][7] = kob;
#pragma unroll
    for (int k = 0; k < 4; ++k) rec[0][8 + k] = wb[k];
  }
  double2* q = reinterpret_cast<double2*>(VFA + 24 * (long)f);
#pragma unroll
  for (int a = 0; a < 6; ++a) q[a] = make_double2(rec[0][2 * a], rec[0][2 * a + 1]);
#pragma unroll
  for (int a = 0; a < 6; ++a) q[6 + a] = make_double2(rec[1][2 * a], rec[1][2 * a + 1]);
}

// Per cell: gather the face records, reverse the viscosity law and the WLSQ
// gradient with its ghost entries; adds to wbar and writes the neighbours'
// (u, v, T) adjoints into vbar at the reverse entries.
__global__ void __launch_bounds__(kBlock)
    k_visc_cell_adjoint(int n, const double* __restrict__ W, const double* __restrict__ VG,
                        const int* __restrict__ acc_ptr, const int* __restrict__ acc_item,
                        const int* __restrict__ grad_ptr, const int* __restrict__ grad_rev,
                        ViscArrays va, FaceArrays fa, ViscParams vp, const double* __restrict__ VFA,
                        double* __restrict__ wbar, double* __restrict__ vbar, double* __restrict__ SBAR,
                        InflowArrays ia, double* __restrict__ VIB) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double G[6] = {0, 0, 0, 0, 0, 0}, mub = 0.0, kb = 0.0, wb[4] = {0.0, 0.0, 0.0, 0.0};
  for (int e = acc_ptr[i]; e < acc_ptr[i + 1]; ++e) {
    int item = acc_item[e];
    int f = item >> 2, mode = item & 3;
    bool right = (mode == MCFD_ACC_NEIGHBOUR) || (mode == MCFD_ACC_PERIODIC_B);
    const double* r = VFA + 24 * (long)f + (right ? 12 : 0);
#pragma unroll
    for (int k = 0; k < 6; ++k) G[k] += r[k];
    mub += r[6];
    kb += r[7];
#pragma unroll
    for (int k = 0; k < 4; ++k) wb[k] += r[8 + k];
  }
  P4 w = load4(W, i);
  double T = w.p / (w.r * vp.R);
  const double* vg = VG + 8 * (long)i;
  double mu = vg[6], kc = vg[7];
  const double* m = va.wl_cell + 4 * (long)i;
  double xx = m[0], xy = m[1], yy = m[2], inv = m[3];
  double rxb[3] = {0.0, 0.0, 0.0}, ryb[3] = {0.0, 0.0, 0.0};
  if (inv != 0.0)
    for (int k = 0; k < 3; ++k) {
      rxb[k] = (G[2 * k] * yy - G[2 * k + 1] * xy) * inv;
      ryb[k] = (G[2 * k + 1] * xx - G[2 * k] * xy) * inv;
    }
  double valb[3] = {0.0, 0.0, 0.0};
  for (int e = grad_ptr[i]; e < grad_ptr[i + 1]; ++e) {
    double wx = va.wl_wx[e], wy = va.wl_wy[e];
    double d[3];
    for (int k = 0; k < 3; ++k) {
      d[k] = wx * rxb[k] + wy * ryb[k];
      valb[k] -= d[k];
    }
    double* vq = vbar + 3 * (long)grad_rev[e];
    vq[0] = d[0];
    vq[1] = d[1];
    vq[2] = d[2];
  }
  for (int e = va.gh_ptr[i]; e < va.gh_ptr[i + 1]; ++e) {
    int f = va.gh_face[e];
    int b = fa.bc[f];
    int code = va.bv_code[b];
    double nx = fa.nx[f], ny = fa.ny[f];
    double wx = va.gh_wx[e], wy = va.gh_wy[e];
    double ghb[3];
    for (int k = 0; k < 3; ++k) {
      ghb[k] = wx * rxb[k] + wy * ryb[k];
      valb[k] -= ghb[k];
    }
    if (code == kVisWall) {
      int flags = va.bv_flags[b];
      const double* bg = va.bv_geom + 4 * (long)b;
      if (flags & 1) {
        valb[0] -= ghb[0];
        valb[1] -= ghb[1];
      } else {
        double t = 2.0 * (ghb[0] * nx + ghb[1] * ny);
        valb[0] += ghb[0] - t * nx;
        valb[1] += ghb[1] - t * ny;
      }
      if (((flags >> 1) & 3) == 1) {
        valb[2] -= ghb[2];
      } else {
        valb[2] += ghb[2];
        kb += -ghb[2] * 2.0 * bg[0] * bg[3] / (kc * kc);
      }
    } else if (code == kVisSymmetry) {
      double t = 2.0 * (ghb[0] * nx + ghb[1] * ny);
      valb[0] += ghb[0] - t * nx;
      valb[1] += ghb[1] - t * ny;
      valb[2] += ghb[2];
    } else if (code == kVisInlet || code == kVisInflow || code == kVisFarField) {
      double xb[3] = {2.0 * ghb[0], 2.0 * ghb[1], 2.0 * ghb[2]};
      for (int k = 0; k < 3; ++k) valb[k] -= ghb[k];
      const int ir = b - va.inf_first;
      const bool rec = code == kVisInflow;
      visc_inlet_b(code, w, visc_inlet_state(code, b, va, fa), nx, ny, vp, xb, wb,
                   SBAR ? SBAR + 4 * (long)b : nullptr, fa.kind[f], rec ? ia.params + 4 * (long)ir : nullptr,
                   rec ? ia.proj[ir] : 0.0, rec && VIB ? VIB + 4 * (long)ir : nullptr);
    } else {
      for (int k = 0; k < 3; ++k) valb[k] += ghb[k];
    }
  }
  mub += kb * vp.gamma * vp.R / ((vp.gamma - 1.0) * vp.Pr);
  valb[2] += mub * visc_dmu_dT(T, mu, vp);
  uvT_to_prim_b(w, vp.R, valb[0], valb[1], valb[2], wb);
  P4 cur = load4(wbar, i);
  store4(wbar, i, cur.r + wb[0], cur.u + wb[1], cur.v + wb[2], cur.p + wb[3]);
}

// Cell pass 2: the neighbours' contributions (stored by each neighbour at
// the cell's own entry for it, so they are read contiguously in entry
// order), body force, and the primitive -> conservative chain rule. Writes
// Ubar.
__global__ void __launch_bounds__(kBlock)
    k_cell_adjoint2(int n, const double* __restrict__ U, const double* __restrict__ W,
                    const double* __restrict__ lam, AdjArrays aa, int muscl, double g, double ax,
                    double ay, double vR, double* __restrict__ ubar) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  P4 wb4 = load4(aa.wbar, i);
  double wb[4] = {wb4.r, wb4.u, wb4.v, wb4.p};
  if (muscl) {
    for (int e = aa.grad_ptr[i]; e < aa.grad_ptr[i + 1]; ++e) {
      P4 c = load4(aa.ebar, e);  // written by the neighbour for its entry j -> i
      wb[0] += c.r;
      wb[1] += c.u;
      wb[2] += c.v;
      wb[3] += c.p;
    }
    if (aa.ebar_par)
      for (int e = aa.par_ptr[i]; e < aa.par_ptr[i + 1]; ++e) {
        P4 c = load4(aa.ebar_par, e);
        wb[0] += c.r;
        wb[1] += c.u;
        wb[2] += c.v;
        wb[3] += c.p;
      }
  }
  if (aa.vbar) {
    // the neighbours' WLSQ adjoints of this cell's (u, v, T), T = p/(rho R)
    double ub = 0.0, vb = 0.0, Tb = 0.0;
    for (int e = aa.grad_ptr[i]; e < aa.grad_ptr[i + 1]; ++e) {
      const double* vq = aa.vbar + 3 * (long)e;
      ub += vq[0];
      vb += vq[1];
      Tb += vq[2];
    }
    P4 q = load4(U, i);
    double rho = q.r;
    double pr = (g - 1.0) * (q.p - 0.5 * (q.u * q.u + q.v * q.v) / rho);
    wb[0] += -Tb * pr / (rho * rho * vR);
    wb[1] += ub;
    wb[2] += vb;
    wb[3] += Tb / (rho * vR);
  }
  // primitive velocity of the cell, as k_primitive computes it
  P4 q = load4(U, i);
  P4 w{q.r, q.u / q.r, q.v / q.r, 0.0};
  if (ax != 0.0 || ay != 0.0) {
    P4 l = load4(lam, i);
    wb[0] += l.u * ax + l.v * ay + l.p * (ax * w.u + ay * w.v);
    wb[1] += l.p * w.r * ax;
    wb[2] += l.p * w.r * ay;
  }
  // W = prim(U): rho = U0, u = U1/U0, v = U2/U0, p = (g-1)(U3 - 0.5(U1^2+U2^2)/U0)
  double rho = w.r, u = w.u, v = w.v;
  double u0 = wb[0] - wb[1] * u / rho - wb[2] * v / rho + wb[3] * (g - 1.0) * 0.5 * (u * u + v * v);
  double u1 = wb[1] / rho - wb[3] * (g - 1.0) * u;
  double u2 = wb[2] / rho - wb[3] * (g - 1.0) * v;
  double u3 = wb[3] * (g - 1.0);
  store4(ubar, i, u0, u1, u2, u3);
  (void)W;
}


// ---------------------------------------------------------------------------
// Fused VJP, part 1: one workgroup per fused-stage block does the reverse of
// the block's residual up to the per-cell results of k_cell_adjoint1, out of
// LDS: W of own + ring cells, lambda/A and the gradient, limiter factors
// and limiter arguments of own + ring-1 cells, the face adjoints of every
// face touching an own cell. k_cell_adjoint2 then adds the neighbours'
// contributions. Rusanov and HLLC (reverse mode).
//   P0  W of local cells, lambda/A of own + ring-1 cells       -> LDS
//   P1  gradient + limiter (+ limiting / extremum entries)     -> LDS (+ registers)
//   P2  face adjoints, one face per thread                     -> LDS
//   P3  per own cell: reconstruction, limiter and LSQ backward -> wbar, ebar
template <int SCHEME, bool MUSCL, int LIM>
__global__ void __launch_bounds__(kBlock)
    k_vjp_fused(int max_local, int max_grad, int max_faces, int max_acc, BlockArrays ba, GeomArrays ga,
                FaceArrays fa, const double* __restrict__ U, const double* __restrict__ lam, double g,
                const int* __restrict__ grad_rev, double* __restrict__ wbar, double* __restrict__ ebar,
                double* __restrict__ SBAR, int* __restrict__ status) {
  extern __shared__ double smem[];
  const int b = xcd_block();
  const int* inf = ba.info + kBlockInfoDev * b;
  const int cs = inf[0], n_own = inf[1], n_r1 = inf[2], n_r2 = inf[3];
  const int fs = inf[4], n_faces = inf[5], gs = inf[6], gps = inf[7];
  const int first = inf[8], ls = inf[9];
  const int n_loc = n_own + n_r1 + n_r2;
  const int n_grad = n_own + n_r1;
  double* sW = smem;                       // [4 max_local]
  double* sG = sW + 4 * max_local;         // [8 max_grad]
  double* sP = sG + 8 * max_grad;          // [4 max_grad]
  double* sLam = sP + 4 * max_grad;        // [4 max_grad] lambda / A, until P2 ends
  double* sA = sLam;                       // [12 max_faces] face adjoints and offsets, after P2
  unsigned short* sAI = reinterpret_cast<unsigned short*>(
      sLam + (4 * max_grad > 12 * max_faces ? 4 * max_grad : 12 * max_faces));  // [max_acc] gather items
  (void)max_acc;
  const int t = threadIdx.x;
  int e0 = 0, e1 = 0, l0 = 0, l1 = 0;
  if (MUSCL && t < n_grad) {
    e0 = ba.gptr[gps + t];
    e1 = ba.gptr[gps + t + 1];
    l0 = ba.lptr[gps + t];
    l1 = ba.lptr[gps + t + 1];
  }
  // P0
  for (int j = t; j < n_loc; j += kBlock) {
    int c = (j < n_own) ? first + j : ba.cells[cs + j];
    P4 q = load4(U, c);
    double rho = q.r;
    double u = q.u / rho;
    double v = q.v / rho;
    double p = (g - 1.0) * (q.p - 0.5 * (q.u * q.u + q.v * q.v) / rho);
    if (j < n_own && (!finite(rho) || !finite(u) || !finite(v) || !finite(p) || !(rho > 0.0) || !(p > 0.0)))
      report(status, MCFD_ERR_STATE, c);
    sW[4 * j] = rho;
    sW[4 * j + 1] = u;
    sW[4 * j + 2] = v;
    sW[4 * j + 3] = p;
    if (j < n_grad) {
      P4 l = load4(lam, c);
      double A = ga.area[c];
      sLam[4 * j] = l.r / A;
      sLam[4 * j + 1] = l.u / A;
      sLam[4 * j + 2] = l.v / A;
      sLam[4 * j + 3] = l.p / A;
    }
  }
  __syncthreads();
  // P1 (as k_gradient_t<true>)
  unsigned arg[4] = {0xFFFFFFu, 0xFFFFFFu, 0xFFFFFFu, 0xFFFFFFu};
  if constexpr (MUSCL) {
    if (t < n_grad) {
      const int j = t;
      double wv[4] = {sW[4 * j], sW[4 * j + 1], sW[4 * j + 2], sW[4 * j + 3]};
      double gx[4] = {0.0, 0.0, 0.0, 0.0};
      double gy[4] = {0.0, 0.0, 0.0, 0.0};
      double mn[4] = {wv[0], wv[1], wv[2], wv[3]};
      double mx[4] = {wv[0], wv[1], wv[2], wv[3]};
      // entries in batches of 4: the batch's loads are issued together, the
      // sums and min / max still run in entry order
      for (int base = e0; base < e1; base += 4) {
        int ob[4];
        double2 cb[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          bool ok = base + q < e1;
          ob[q] = ok ? (int)ba.gnb[gs + base + q] : 0;
          cb[q] = ok ? ba.gent[gs + base + q] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int e = base + q;
          if (e >= e1) break;
          const int o = ob[q];
          const double2 c2 = cb[q];
          double ov[4] = {sW[4 * o], sW[4 * o + 1], sW[4 * o + 2], sW[4 * o + 3]};
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            double d = ov[k] - wv[k];
            gx[k] = gx[k] + c2.x * d;
            gy[k] = gy[k] + c2.y * d;
            if (ov[k] < mn[k]) {
              mn[k] = ov[k];
              arg[k] = (arg[k] & 0x00FFFFu) | ((unsigned)(e - e0) << 16);
            }
            if (ov[k] > mx[k]) {
              mx[k] = ov[k];
              arg[k] = (arg[k] & 0xFF00FFu) | ((unsigned)(e - e0) << 8);
            }
          }
        }
      }
      double phi[4] = {1.0, 1.0, 1.0, 1.0};
      for (int e = l0; e < l1; ++e) {
        double2 d2 = ba.lent[ls + e];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          double before = phi[k];
          limiter_entry(gx[k] * d2.x + gy[k] * d2.y, wv[k], mx[k], mn[k], LIM, phi[k]);
          if (phi[k] < before) arg[k] = (arg[k] & 0xFFFF00u) | (unsigned)(e - l0);
        }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        sG[8 * j + 2 * k] = gx[k];
        sG[8 * j + 2 * k + 1] = gy[k];
        sP[4 * j + k] = phi[k];
      }
    }
    __syncthreads();
  }
  // P3's gather items (contiguous for the block's own cells), loaded now and
  // stored to LDS after P2, and the own cell's item range
  constexpr int kStageItems = 3;
  const int abase = ga.acc_ptr[first];
  const int n_acc = ga.acc_ptr[first + n_own] - abase;
  unsigned short ai_r[kStageItems];
#pragma unroll
  for (int q = 0; q < kStageItems; ++q) {
    const int k = t + q * kBlock;
    ai_r[q] = (k < n_acc) ? ba.acc_local[abase + k] : 0;
  }
  int a0 = 0, a1 = 0;
  if (t < n_own) {
    a0 = ga.acc_ptr[first + t];
    a1 = ga.acc_ptr[first + t + 1];
  }
  // P2: face adjoints (as k_flux_adjoint_rev); results stay in registers
  // until every thread has read lambda; the face's reconstruction offsets
  // go to LDS beside them for P3
  P4 alb{0.0, 0.0, 0.0, 0.0}, arb{0.0, 0.0, 0.0, 0.0};
  double olx = 0.0, oly = 0.0, orx = 0.0, ory = 0.0;
  if (t < n_faces) {
    const FaceRec fr = ba.frec[fs + t];
    olx = fr.lx;
    oly = fr.ly;
    orx = fr.rx;
    ory = fr.ry;
    const int kind = fr.kind_bc & 15;
    const int l = fr.lr & 0xFFFF, r = fr.lr >> 16;
    const double nx = fr.nx, ny = fr.ny, L = fr.len;
    double fb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) fb[k] = -L * sLam[4 * l + k];
    P4 wl = rec_lds(sW, sG, sP, l, fr.lx, fr.ly, MUSCL);
    P4 wr = wl;
    int bi = fr.kind_bc >> 4;
    if (kind <= MCFD_FACE_PERIODIC) {
#pragma unroll
      for (int k = 0; k < 4; ++k) fb[k] += L * sLam[4 * r + k];
      wr = rec_lds(sW, sG, sP, r, fr.rx, fr.ry, MUSCL);
    } else if (kind == MCFD_FACE_WALL) {
      double un = (wl.u - fa.bc_wall[2 * (long)bi]) * nx + (wl.v - fa.bc_wall[2 * (long)bi + 1]) * ny;
      wr.u = wl.u - 2.0 * un * nx;
      wr.v = wl.v - 2.0 * un * ny;
    } else if (kind == MCFD_FACE_VELOCITY_INLET) {
      const double* st = fa.bc_state + 4 * (long)bi;
      wr = P4{st[0], st[1], st[2], st[3]};
    } else if (kind == MCFD_FACE_PRESSURE_OUTLET) {
      wr.p = fa.bc_state[4 * (long)bi + 3];
    } else if (kind == MCFD_FACE_FAR_FIELD) {
      const double* st = fa.bc_state + 4 * (long)bi;
      int err = 0;
      wr = far_field(wl, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, &err);
    }
    double nbx = 0.0, nby = 0.0;  // unused here
    if constexpr (SCHEME == 1)
      hllc_b(wl, wr, nx, ny, g, kind == MCFD_FACE_INTERIOR, fb, alb, arb, nbx, nby);
    else
      rusanov_b(wl, wr, nx, ny, g, fb, alb, arb, nbx, nby);
    if (kind > MCFD_FACE_PERIODIC) {
      P4 ext{0.0, 0.0, 0.0, 0.0};
      if (kind == MCFD_FACE_WALL) {
        double un_b = arb.u * nx + arb.v * ny;
        alb.r += arb.r;
        alb.u += arb.u - 2.0 * nx * un_b;
        alb.v += arb.v - 2.0 * ny * un_b;
        alb.p += arb.p;
      } else if (kind == MCFD_FACE_VELOCITY_INLET) {
        ext = arb;
      } else if (kind == MCFD_FACE_PRESSURE_OUTLET) {
        alb.r += arb.r;
        alb.u += arb.u;
        alb.v += arb.v;
      } else if (kind == MCFD_FACE_FAR_FIELD) {
        const double* st = fa.bc_state + 4 * (long)bi;
        far_field_b(wl, P4{st[0], st[1], st[2], st[3]}, nx, ny, g, arb, alb, ext, nbx, nby);
      } else {
        alb.r += arb.r;
        alb.u += arb.u;
        alb.v += arb.v;
        alb.p += arb.p;
      }
      arb = ext;
      if (SBAR) store4(SBAR, bi, ext.r, ext.u, ext.v, ext.p);
    }
  }
  __syncthreads();
  if (t < n_faces) {
    double* q = sA + 12 * t;
    q[0] = alb.r;
    q[1] = alb.u;
    q[2] = alb.v;
    q[3] = alb.p;
    q[4] = arb.r;
    q[5] = arb.u;
    q[6] = arb.v;
    q[7] = arb.p;
    q[8] = olx;
    q[9] = oly;
    q[10] = orx;
    q[11] = ory;
  }
#pragma unroll
  for (int q = 0; q < kStageItems; ++q) {
    const int k = t + q * kBlock;
    if (k < n_acc) sAI[k] = ai_r[q];
  }
  for (int k = t + kStageItems * kBlock; k < n_acc; k += kBlock) sAI[k] = ba.acc_local[abase + k];
  __syncthreads();
  // P3: own cell t (as k_cell_adjoint1)
  if (t >= n_own) return;
  const int i = first + t;
  double wb[4] = {0.0, 0.0, 0.0, 0.0};
  double pb[4] = {0.0, 0.0, 0.0, 0.0};
  double gb[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double gv[8], pv[4];
#pragma unroll
  for (int k = 0; k < 8; ++k) gv[k] = MUSCL ? sG[8 * t + k] : 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) pv[k] = MUSCL ? sP[4 * t + k] : 1.0;
  for (int e = a0; e < a1; ++e) {
    int item = sAI[e - abase];
    int slot = item >> 2;
    int mode = item & 3;
    bool right = (mode == MCFD_ACC_NEIGHBOUR) || (mode == MCFD_ACC_PERIODIC_B);
    const double* q = sA + 12 * slot + (right ? 4 : 0);
    double a[4] = {q[0], q[1], q[2], q[3]};
#pragma unroll
    for (int k = 0; k < 4; ++k) wb[k] += a[k];
    if (MUSCL) {
      const double* o = sA + 12 * slot + (right ? 10 : 8);
      double2 d = make_double2(o[0], o[1]);  // (rx, ry) or (lx, ly)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double inc = gv[2 * k] * d.x + gv[2 * k + 1] * d.y;
        pb[k] += a[k] * inc;
        gb[2 * k] += a[k] * pv[k] * d.x;
        gb[2 * k + 1] += a[k] * pv[k] * d.y;
      }
    }
  }
  if constexpr (MUSCL) {
    double wv[4] = {sW[4 * t], sW[4 * t + 1], sW[4 * t + 2], sW[4 * t + 3]};
    double xb[4] = {0.0, 0.0, 0.0, 0.0};
    int xe[4] = {-1, -1, -1, -1};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      unsigned le = arg[k] & 0xFFu;
      if (pb[k] == 0.0 || le == 0xFFu) continue;
      double2 d2 = ba.lent[ls + l0 + (int)le];
      double inc = gv[2 * k] * d2.x + gv[2 * k + 1] * d2.y;
      double tol = 1e-14 * vmax(fabs(wv[k]), 1.0);
      if (!(inc > tol) && !(inc < -tol)) continue;
      bool up = inc > tol;
      unsigned xa = up ? ((arg[k] >> 8) & 0xFFu) : ((arg[k] >> 16) & 0xFFu);
      int xn = (xa == 0xFFu) ? -1 : e0 + (int)xa;
      double ext = (xn < 0) ? wv[k] : sW[4 * (int)ba.gnb[gs + xn] + k];
      double ratio = (ext - wv[k]) / inc;
      if (0.0 > ratio) continue;
      double dfdr;
      (void)limiter_value(ratio, LIM, dfdr);
      if (dfdr == 0.0) continue;
      double rb = pb[k] * dfdr;
      double ext_b = rb / inc;
      wb[k] -= ext_b;
      double incb = -rb * ratio / inc;
      gb[2 * k] += incb * d2.x;
      gb[2 * k + 1] += incb * d2.y;
      if (xn < 0) {
        wb[k] += ext_b;
      } else {
        xb[k] = ext_b;
        xe[k] = xn;
      }
    }
    const int gbase = ga.grad_ptr[i] - e0;
    for (int e = e0; e < e1; ++e) {
      double2 c2 = ba.gent[gs + e];
      double c[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        double tv = c2.x * gb[2 * k] + c2.y * gb[2 * k + 1];
        wb[k] -= tv;
        if (xe[k] == e) tv += xb[k];
        c[k] = tv;
      }
      store4(ebar, grad_rev[gbase + e], c[0], c[1], c[2], c[3]);
    }
  }
  store4(wbar, i, wb[0], wb[1], wb[2], wb[3]);
}

// Batched cases: per-case CFL step = min over the case's cell range, clipped
// to end_time - time_case, then broadcast to the case's cells. One block per
// case; clock = [time, dt] per case.
__global__ void __launch_bounds__(1024)
    k_batch_dt(int n_cases, const int* __restrict__ case_ptr, double end_time,
               double* __restrict__ clock, double* __restrict__ dt_cell, int* __restrict__ status) {
  __shared__ double red[1024];
  __shared__ double shared_dt;
  int c = blockIdx.x;
  if (c >= n_cases) return;
  int a = case_ptr[c], b = case_ptr[c + 1];
  double m = INFINITY;
  for (int i = a + threadIdx.x; i < b; i += blockDim.x) m = fmin(m, dt_cell[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double t = clock[2 * c];
    double rest = end_time - t;
    double dt = (rest < red[0]) ? rest : red[0];
    if (!finite(dt) || !(dt > 0.0)) report(status, MCFD_ERR_GLOBAL_DT, c);
    clock[2 * c + 1] = dt;
    clock[2 * c] = t + dt;
    shared_dt = dt;
  }
  __syncthreads();
  double dt = shared_dt;
  for (int i = a + threadIdx.x; i < b; i += blockDim.x) dt_cell[i] = dt;
}

// Spectral-radius sum and local step (solver.py:4132-4343, inviscid part).
__global__ void __launch_bounds__(kBlock)
    k_time_steps(int n, const double* __restrict__ W, DtArrays d, double g, double cfl,
                 double* __restrict__ dt_cell, double* __restrict__ partial,
                 int* __restrict__ status) {
  __shared__ double red[kBlock];
  int i = xcd_block() * blockDim.x + threadIdx.x;
  double dt = INFINITY;
  if (i < n) {
    dt = cell_time_step(i, load4(W, i), d, g, cfl, status);
    dt_cell[i] = dt;
  }
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[xcd_block()] = red[0];
}

// Viscous flow: spectral sum + 2 * diffusion sum (solver.py:4246-4332).
__global__ void __launch_bounds__(kBlock)
    k_time_steps_viscous(int n, const double* __restrict__ W, DtArrays d, ViscDt vd, double g,
                         double cfl, double* __restrict__ dt_cell, double* __restrict__ partial,
                         int* __restrict__ status) {
  __shared__ double red[kBlock];
  int i = xcd_block() * blockDim.x + threadIdx.x;
  double dt = INFINITY;
  if (i < n) {
    double sum = cell_spectral_sum(i, load4(W, i), d, g, status);
    sum = sum + 2.0 * visc_dt_sum(i, vd);
    dt = cfl * d.area[i] / sum;
    if (!finite(dt) || !(dt > 0.0)) report(status, MCFD_ERR_LOCAL_DT, i);
    dt_cell[i] = dt;
  }
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[xcd_block()] = red[0];
}

// out[0] = min(partial[0:n]); one block of 1024 threads: four independent
// loads in flight per thread, then lane exchanges and one LDS step.
__global__ void __launch_bounds__(1024)
    k_min_reduce(int n, const double* __restrict__ partial, double* __restrict__ out) {
  __shared__ double wave_min[16];
  const int t = threadIdx.x;
  double m0 = INFINITY, m1 = INFINITY, m2 = INFINITY, m3 = INFINITY;
  int j = t;
  for (; j + 3 * 1024 < n; j += 4 * 1024) {
    m0 = fmin(m0, partial[j]);
    m1 = fmin(m1, partial[j + 1024]);
    m2 = fmin(m2, partial[j + 2048]);
    m3 = fmin(m3, partial[j + 3072]);
  }
  for (; j < n; j += 1024) m0 = fmin(m0, partial[j]);
  double m = fmin(fmin(m0, m1), fmin(m2, m3));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off));
  if ((t & 63) == 0) wave_min[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    double r = wave_min[0];
    for (int w = 1; w < 16; ++w) r = fmin(r, wave_min[w]);
    out[0] = r;
  }
}

__global__ void k_clock(double* __restrict__ src, double fixed_dt, double end_time,
                        double* __restrict__ clock, int reset, int* __restrict__ status) {
  if (threadIdx.x != 0) return;
  double t = clock[0];
  double v = src ? src[0] : fixed_dt;
  if (src && reset) src[0] = INFINITY;
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
  int i = xcd_block() * blockDim.x + threadIdx.x;
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
  if (threadIdx.x < 4) partial[4 * (long)xcd_block() + threadIdx.x] = red[threadIdx.x][0];
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
  double* acc_coef = nullptr;
  int* dts_ptr = nullptr;
  double* dts_nx = nullptr;
  double* dts_ny = nullptr;
  double* dts_len = nullptr;
  int* dts_bc = nullptr;
  int* bc_kind = nullptr;
  DtArrays dts{};
  // scratch
  double* W = nullptr;
  double* G = nullptr;
  double* PHI = nullptr;
  double* FLUX = nullptr;
  double* partial = nullptr;
  int* status = nullptr;
  bool has_far_field = false;
  // fused-stage blocks
  bool fused = false;
  int nb = 0, block_cells = 0, max_local = 0, max_grad = 0, max_faces = 0;
  int max_acc = 0, max_dts = 0;  // most accumulation / spectral-radius items of a block's own cells
  int max_lsq = 0;  // most LSQ entries of a block's own + ring-1 cells
  double* block_partial = nullptr;  // [nb] per-block minima of the fused final stage
  bool fused_visc = false;  // per-block viscous arrays uploaded (mcfd_mesh_set_block_viscous)
  long block_n_gnb = 0;     // LSQ entries of the block layout
  long n_vacc_visc = 0;     // viscous accumulation items (mcfd_mesh_set_viscous)
  BlockArrays blocks{};
  std::vector<void*> block_allocations;
  // implicit
  bool implicit = false;
  LusgsArrays lusgs{};
  std::vector<void*> implicit_allocations;
  std::vector<int> host_lvl_f, host_lvl_b;  // level pointers of the LU-SGS sweeps
  // two levels each, the backward's the forward's swapped (a 2-colouring)
  bool lusgs_two = false;
  int lusgs_path = -1;  // -1 undecided, 0 one-workgroup sweeps, 1 one launch per level
  int gmres_m = 0;
  double* gm_V = nullptr;       // (m+1) * 4C Krylov basis
  double* gm_w = nullptr;       // 4C
  double* gm_r = nullptr;       // 4C
  double* gm_av = nullptr;      // 4C
  double* gm_pert = nullptr;    // 4C
  double* gm_rp = nullptr;      // 4C
  double* gm_scalars = nullptr; // 64 fixed slots, then 3 + m per-column dots, then m coefficients
  double* gm_y_host = nullptr;  // pinned, m coefficients of the basis combination
  double* gm_scalars_host = nullptr;  // pinned mirror of gm_scalars, written by k_dot_final
  size_t gm_scalars_n = 0;
  double* gm_partial = nullptr; // dot partials
  // MUSCL parabolic
  bool parabolic = false;
  int n_par = 0;
  int* par_ptr = nullptr;
  int* par_nb = nullptr;
  double* par_coef = nullptr;
  int* par_full = nullptr;
  double* K = nullptr;
  std::vector<void*> parabolic_allocations;
  // adjoint
  bool adjoint = false;
  int* grad_rev = nullptr;
  double* FADJ = nullptr;
  double* VFA = nullptr;   // viscous face adjoint records [n_faces*24]
  double* VBAR = nullptr;  // viscous neighbour adjoints [n_grad*3]
  AdjArrays adj{};
  std::vector<void*> adjoint_allocations;
  // laminar viscous terms
  bool viscous = false;
  ViscArrays visc{};
  ViscAcc vacc{};
  ViscDt vdt{};
  double* VG = nullptr;       // [C*8] (u, v, T) gradients, mu, k
  double* VF = nullptr;       // [n_faces*4] viscous face fluxes
  double* EFF = nullptr;      // [C] effective diffusivity
  std::vector<void*> viscous_allocations;
  // pressure-inlet and mass-flow records
  bool needs_inflow = false;  // the mesh has boundary records of kinds 7-9
  bool inflow = false;        // mcfd_mesh_set_inflow done
  InflowArrays inf{};
  double* inf_ext_v = nullptr;  // [4 * inflow records] exterior states for the viscous terms
  double* inf_vbar = nullptr;   // [4 * inflow records] their adjoints (viscous mass-flow-inlet patch pass)
  int inflow_patches = 0, inflow_max_n = 0, inflow_max_leaves = 0;
  std::vector<void*> inflow_allocations;
  std::vector<int> host_face_bc;  // [n_faces] for validating the inflow records
  std::vector<int> host_bc_kind;  // [n_bc]
  std::vector<int> host_acc_ptr, host_dts_ptr;  // [n_cells + 1], for the fused blocks' LDS sizes
  int max_cell_entries = 0;       // longest LSQ or limiter list of a cell
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
int alloc_scratch(McfdMesh* m, T** dst, size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
  m->allocations.push_back(p);
  *dst = static_cast<T*>(p);
  return 0;
}

void free_blocks(McfdMesh* m) {
  for (void* p : m->block_allocations) (void)hipFree(p);
  m->block_allocations.clear();
  m->fused = false;
  m->fused_visc = false;
  m->blocks = BlockArrays{};
  m->block_partial = nullptr;
}

template <typename T>
int upload_block(McfdMesh* m, T** dst, const T* src, size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
  m->block_allocations.push_back(p);
  if (count) {
    if (!src) return fail("mcfd_mesh_set_blocks: missing array");
    HIP_CHECK(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
  }
  *dst = static_cast<T*>(p);
  return 0;
}

// W of the local cells + one region shared by (gradients, limiter factors)
// and then (face fluxes, lengths, normals) and the final block reduction,
// then the own cells' accumulation coefficients and items and their
// spectral-radius items, staged for the gather.
// The staging is written after P2, over W when it fits there.
// LSQ coefficients and (dword-aligned, 16-bit) LSQ neighbour ids of a
// block, copied into the region at kernel start.
size_t fused_entry_doubles(const McfdMesh* m) {
  return 2 * (size_t)m->max_lsq + ((size_t)m->max_lsq + 2 + 3) / 4;
}

size_t fused_region(const McfdMesh* m, bool muscl, bool visc = false) {
  size_t region = 7 * (size_t)m->max_faces;
  if (muscl) region = std::max(region, std::max(12 * (size_t)m->max_grad, fused_entry_doubles(m)));
  // viscous: the gradients, viscosity and conductivity of own + ring-1
  // cells, then per face the convective flux, length and viscous flux
  if (visc) region = std::max(9 * (size_t)m->max_faces, 8 * (size_t)m->max_grad);
  return std::max(region, (size_t)kBlock);
}

size_t fused_stage_doubles(const McfdMesh* m) {
  size_t shorts = (size_t)m->max_acc + (size_t)m->max_dts;
  return (size_t)m->max_acc + (shorts * sizeof(unsigned short) + 7) / 8;
}

int fused_stage_offset(const McfdMesh* m, bool muscl, bool visc = false) {
  if (fused_stage_doubles(m) <= 4 * (size_t)m->max_local) return 0;
  return (int)(4 * (size_t)m->max_local + fused_region(m, muscl, visc));
}

size_t fused_lds_bytes(const McfdMesh* m, bool muscl, bool visc = false) {
  size_t doubles = 4 * (size_t)m->max_local + fused_region(m, muscl, visc);
  if (fused_stage_offset(m, muscl, visc) != 0) doubles += fused_stage_doubles(m);
  return doubles * sizeof(double);
}

void free_implicit(McfdMesh* m) {
  for (void* p : m->implicit_allocations) (void)hipFree(p);
  m->implicit_allocations.clear();
  if (m->gm_y_host) (void)hipHostFree(m->gm_y_host);
  m->gm_y_host = nullptr;
  if (m->gm_scalars_host) (void)hipHostFree(m->gm_scalars_host);
  m->gm_scalars_host = nullptr;
  m->implicit = false;
  m->lusgs = LusgsArrays{};
  m->gmres_m = 0;
}

template <typename T>
int upload_implicit(McfdMesh* m, T** dst, const T* src, size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
  m->implicit_allocations.push_back(p);
  if (count && src) HIP_CHECK(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
  *dst = static_cast<T*>(p);
  return 0;
}

void free_adjoint(McfdMesh* m) {
  for (void* p : m->adjoint_allocations) (void)hipFree(p);
  m->adjoint_allocations.clear();
  m->adjoint = false;
  m->adj = AdjArrays{};
}

template <typename T>
int alloc_adjoint(McfdMesh* m, T** dst, const T* src, size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
  m->adjoint_allocations.push_back(p);
  if (count && src) HIP_CHECK(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
  *dst = static_cast<T*>(p);
  return 0;
}

void free_parabolic(McfdMesh* m) {
  for (void* p : m->parabolic_allocations) (void)hipFree(p);
  m->parabolic_allocations.clear();
  m->parabolic = false;
}

template <typename T>
int alloc_parabolic(McfdMesh* m, T** dst, const T* src, size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
  m->parabolic_allocations.push_back(p);
  if (count && src) HIP_CHECK(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
  *dst = static_cast<T*>(p);
  return 0;
}

void free_viscous(McfdMesh* m) {
  for (void* p : m->viscous_allocations) (void)hipFree(p);
  m->viscous_allocations.clear();
  m->viscous = false;
  m->fused_visc = false;  // the fused kernel reads the viscous lists
  m->visc = ViscArrays{};
  m->vacc = ViscAcc{};
  m->vdt = ViscDt{};
}

template <typename T>
int alloc_viscous(McfdMesh* m, const T** dst, const T* src, size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
  m->viscous_allocations.push_back(p);
  if (count) {
    if (!src) return fail("mcfd_mesh_set_viscous: missing array");
    HIP_CHECK(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
  }
  *dst = static_cast<const T*>(p);
  return 0;
}

int alloc_viscous_scratch(McfdMesh* m, double** dst, size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(double) * (count ? count : 1)));
  m->viscous_allocations.push_back(p);
  *dst = static_cast<double*>(p);
  return 0;
}

ViscParams visc_params(const McfdSettings* st) {
  return ViscParams{st->viscosity,
                    st->gamma,
                    st->gas_constant,
                    st->prandtl_number,
                    st->dynamic_viscosity,
                    st->sutherland_reference_temperature,
                    st->sutherland_reference_viscosity,
                    st->sutherland_temperature};
}

void free_inflow(McfdMesh* m) {
  for (void* p : m->inflow_allocations) (void)hipFree(p);
  m->inflow_allocations.clear();
  m->inflow = false;
  m->inf = InflowArrays{};
  m->inf_ext_v = nullptr;
  m->inf_vbar = nullptr;
  m->faces.inflow = nullptr;
  m->dts.inflow = nullptr;
  m->inflow_patches = m->inflow_max_n = m->inflow_max_leaves = 0;
}

template <typename T>
int alloc_inflow(McfdMesh* m, T** dst, const T* src, size_t count) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
  m->inflow_allocations.push_back(p);
  if (count && src) HIP_CHECK(hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice));
  *dst = static_cast<T*>(p);
  return 0;
}

void free_mesh(McfdMesh* m) {
  free_inflow(m);
  free_viscous(m);
  free_parabolic(m);
  free_adjoint(m);
  free_implicit(m);
  free_blocks(m);
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
  if (s->reconstruction < 0 || s->reconstruction > 2)
    return fail("Unsupported reconstruction code for the GPU residual.");
  if (s->limiter < 0 || s->limiter > 3) return fail("Unsupported limiter code.");
  if (s->viscosity < 0 || s->viscosity > 2) return fail("Unsupported viscosity model.");
  if (s->viscosity && !(s->gas_constant > 0.0 && s->prandtl_number > 0.0))
    return fail("Viscous settings need a positive gas constant and Prandtl number.");
  return 0;
}

int check_viscous(const McfdMesh* m, const McfdSettings* s) {
  if (s->viscosity && !m->viscous)
    return fail("Viscous settings need mcfd_mesh_set_viscous first.");
  return 0;
}


template <int SCHEME, bool MUSCL, int LIM, bool FARFIELD, bool DTNEXT>
void launch_fused_t(McfdMesh* m, size_t lds, hipStream_t s, const GeomArrays& ga,
                    const McfdSettings* st, const StageArgs& a) {
  auto kern = k_stage_fused<SCHEME, MUSCL, LIM, FARFIELD, DTNEXT>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
  hipLaunchKernelGGL(kern, dim3(m->nb), dim3(kBlock), lds, s, m->max_acc, fused_stage_offset(m, MUSCL), m->max_local,
                     m->max_grad, m->max_faces, m->max_lsq, m->blocks, ga, m->faces, m->dts, st->gamma,
                     st->acceleration_x, st->acceleration_y, a, m->status, ViscArrays{}, ViscParams{}, ViscAcc{});
}

template <int SCHEME, bool MUSCL, int LIM>
void launch_fused_l(McfdMesh* m, bool far, bool dtn, size_t lds, hipStream_t s,
                    const GeomArrays& ga, const McfdSettings* st, const StageArgs& a) {
  if (far) {
    if (dtn) launch_fused_t<SCHEME, MUSCL, LIM, true, true>(m, lds, s, ga, st, a);
    else launch_fused_t<SCHEME, MUSCL, LIM, true, false>(m, lds, s, ga, st, a);
  } else {
    if (dtn) launch_fused_t<SCHEME, MUSCL, LIM, false, true>(m, lds, s, ga, st, a);
    else launch_fused_t<SCHEME, MUSCL, LIM, false, false>(m, lds, s, ga, st, a);
  }
}

template <int SCHEME>
void launch_fused_s(McfdMesh* m, bool muscl, int lim, bool far, bool dtn, size_t lds,
                    hipStream_t s, const GeomArrays& ga, const McfdSettings* st,
                    const StageArgs& a) {
  if (!muscl) return launch_fused_l<SCHEME, false, 0>(m, far, dtn, lds, s, ga, st, a);
  // Barth-Jespersen and Minmod share one limiter function
  if (lim <= 1) return launch_fused_l<SCHEME, true, 0>(m, far, dtn, lds, s, ga, st, a);
  if (lim == 2) return launch_fused_l<SCHEME, true, 2>(m, far, dtn, lds, s, ga, st, a);
  return launch_fused_l<SCHEME, true, 3>(m, far, dtn, lds, s, ga, st, a);
}

// First-order viscous stages (no next-step dt): one launch per stage.
template <int SCHEME, bool FARFIELD>
void launch_fused_visc_t(McfdMesh* m, size_t lds, hipStream_t s, const GeomArrays& ga, const McfdSettings* st,
                         const StageArgs& a) {
  auto kern = k_stage_fused<SCHEME, false, 0, FARFIELD, false, true>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
  hipLaunchKernelGGL(kern, dim3(m->nb), dim3(kBlock), lds, s, m->max_acc, fused_stage_offset(m, false, true),
                     m->max_local, m->max_grad, m->max_faces, m->max_lsq, m->blocks, ga, m->faces,
                     m->dts, st->gamma, st->acceleration_x, st->acceleration_y, a, m->status, m->visc,
                     visc_params(st), m->vacc);
}

void launch_fused_visc(McfdMesh* m, int scheme, bool far, size_t lds, hipStream_t s, const GeomArrays& ga,
                       const McfdSettings* st, const StageArgs& a) {
  if (scheme == 1) {
    if (far) launch_fused_visc_t<1, true>(m, lds, s, ga, st, a);
    else launch_fused_visc_t<1, false>(m, lds, s, ga, st, a);
  } else if (scheme == 2) {
    if (far) launch_fused_visc_t<2, true>(m, lds, s, ga, st, a);
    else launch_fused_visc_t<2, false>(m, lds, s, ga, st, a);
  } else {
    if (far) launch_fused_visc_t<0, true>(m, lds, s, ga, st, a);
    else launch_fused_visc_t<0, false>(m, lds, s, ga, st, a);
  }
}

void launch_fused(McfdMesh* m, int scheme, bool muscl, bool far, bool dtn, size_t lds,
                  hipStream_t s, const GeomArrays& ga, const McfdSettings* st,
                  const StageArgs& a) {
  if (scheme == 1) return launch_fused_s<1>(m, muscl, st->limiter, far, dtn, lds, s, ga, st, a);
  if (scheme == 2) return launch_fused_s<2>(m, muscl, st->limiter, far, dtn, lds, s, ga, st, a);
  return launch_fused_s<0>(m, muscl, st->limiter, far, dtn, lds, s, ga, st, a);
}

int check_inflow(const McfdMesh* m, const McfdSettings* st) {
  if (!m->needs_inflow) return 0;
  if (!m->inflow) return fail("Pressure-inlet and mass-flow boundaries need mcfd_mesh_set_inflow first.");
  if (!(st->gas_constant > 0.0)) return fail("Inflow boundaries need a positive gas constant.");
  return 0;
}

// Exterior states of the inflow records from the reconstructed face states
// (residual) or the owner cells' primitive states in m->W (time steps).
int inflow_states(McfdMesh* m, const McfdSettings* st, hipStream_t s, bool reconstructed, bool para,
                  bool muscl, double* ext = nullptr) {
  if (!m->inflow || m->inf.n == 0) return 0;
  InflowArrays ia = m->inf;
  if (ext) ia.ext = ext;
  hipLaunchKernelGGL(k_inflow_records, dim3(blocks(ia.n)), dim3(kBlock), 0, s, ia, m->faces,
                     m->W, m->G, m->PHI, (reconstructed && para) ? m->K : nullptr,
                     (reconstructed && muscl) ? 1 : 0, reconstructed ? 1 : 0, st->gamma,
                     st->gas_constant, m->status);
  if (m->inflow_patches) {
    size_t lds = sizeof(double) * ((size_t)m->inflow_max_n + 9 * (size_t)m->inflow_max_leaves + 65);
    hipLaunchKernelGGL(k_mass_flow_inlet, dim3(m->inflow_patches), dim3(256), lds, s, ia,
                       m->face_len, m->inflow_max_n, m->inflow_max_leaves, st->gamma,
                       st->gas_constant, m->status);
  }
  HIP_CHECK(hipGetLastError());
  return 0;
}

int residual_front(McfdMesh* m, const McfdSettings* st, const double* u, hipStream_t s) {
  int n = m->n_cells;
  if (check_inflow(m, st)) return -1;
  {
    TimedLaunch t(m, s, MCFD_KERNEL_PRIMITIVE);
    hipLaunchKernelGGL(k_primitive, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, m->W, st->gamma,
                       m->status);
  }
  bool para = st->reconstruction == 2 && m->n_grad > 0;
  if (para && !m->parabolic) return fail("MUSCL parabolic needs mcfd_mesh_set_parabolic first");
  bool muscl = (st->reconstruction == 1 || para) && m->n_grad > 0;
  if (para) {
    TimedLaunch t(m, s, MCFD_KERNEL_GRADIENT);
    hipLaunchKernelGGL(k_gradient_parabolic_t<false>, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W,
                       m->grad_ptr, m->grad_nb, m->grad_cx, m->grad_cy, m->lim_ptr, m->lim_dx,
                       m->lim_dy, st->limiter, m->par_ptr, m->par_nb, m->par_coef, m->par_full,
                       m->G, m->K, m->PHI, nullptr);
  } else if (muscl) {
    TimedLaunch t(m, s, MCFD_KERNEL_GRADIENT);
    hipLaunchKernelGGL(k_gradient_t<false>, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->grad_ptr,
                       m->grad_nb, m->grad_cx, m->grad_cy, m->lim_ptr, m->lim_dx, m->lim_dy,
                       st->limiter, m->G, m->PHI, nullptr);
  }
  if (m->n_faces) {
    TimedLaunch t(m, s, MCFD_KERNEL_FLUX);
    if (inflow_states(m, st, s, true, para, muscl)) return -1;
    hipLaunchKernelGGL(k_flux, dim3(blocks(m->n_faces)), dim3(kBlock), 0, s, m->n_faces,
                       m->faces, m->W, m->G, m->PHI, para ? m->K : nullptr, st->flux,
                       muscl ? 1 : 0, st->gamma, m->FLUX, m->status);
  }
  if (st->viscosity) {
    TimedLaunch t(m, s, MCFD_KERNEL_VISCOUS);
    ViscParams vp = visc_params(st);
    ViscArrays va = m->visc;
    if (m->inflow && m->inf.n) {
      // pressure / mass-flow inlets: the viscous boundary values are their
      // exterior states from the owner cells' states (solver.py:2749-2772)
      if (inflow_states(m, st, s, false, false, false, m->inf_ext_v)) return -1;
      va.inf_ext = m->inf_ext_v;
      va.inf_first = m->inf.first;
    }
    hipLaunchKernelGGL(k_visc_gradient, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W,
                       m->grad_ptr, m->grad_nb, va, m->faces, vp, m->VG, m->status);
    if (m->n_faces)
      hipLaunchKernelGGL(k_visc_flux, dim3(blocks(m->n_faces)), dim3(kBlock), 0, s, m->n_faces,
                         m->faces, m->W, m->VG, va, vp, m->VF, m->status);
  }
  HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace


namespace {

// The pinned host mirror of a GMRES scalar slot (nullptr outside gm_scalars).
double* host_slot(McfdMesh* m, const double* dev) {
  if (!m->gm_scalars_host || dev < m->gm_scalars || dev >= m->gm_scalars + m->gm_scalars_n) return nullptr;
  return m->gm_scalars_host + (dev - m->gm_scalars);
}

int dot_dev(McfdMesh* m, long n, const double* a, const double* b, double* out, bool sqrt_out,
            hipStream_t s) {
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(kBlock), 0, s, n, a, b, m->gm_partial);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kBlock), 0, s, kDotBlocks, m->gm_partial, out,
                     sqrt_out ? 1 : 0, host_slot(m, out));
  HIP_CHECK(hipGetLastError());
  return 0;
}

// Scalars written by k_dot_final into gm_scalars are read from the pinned
// mirror after a stream synchronisation (no copy launch); others by a copy.
int read_scalars(McfdMesh* m, double* host, const double* dev, int count, hipStream_t s) {
  const double* h = host_slot(m, dev);
  if (h && host_slot(m, dev + count - 1)) {
    HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy(host, h, sizeof(double) * count);
    return 0;
  }
  HIP_CHECK(hipMemcpyAsync(host, dev, sizeof(double) * count, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return 0;
}

int lblocks(long n) { return (int)((n + kBlock - 1) / kBlock); }

// g->state_norm, computed on the device (||state / sigma||_2) when the caller
// passes a negative value; the result is stored back into g.
int resolve_state_norm(McfdMesh* m, McfdGmres* g, hipStream_t s) {
  if (g->state_norm >= 0.0) return 0;
  for (int k = 0; k < 4; ++k)
    if (!(g->sigma[k] > 0.0)) return fail("McfdGmres.sigma must be positive to compute state_norm");
  double* out = m->gm_scalars + 63;
  hipLaunchKernelGGL(k_scaled_sq_partial, dim3(kDotBlocks), dim3(kBlock), 0, s, 4L * m->n_cells, g->state,
                     g->sigma[0], g->sigma[1], g->sigma[2], g->sigma[3], m->gm_partial);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kBlock), 0, s, kDotBlocks, m->gm_partial, out, 1,
                     host_slot(m, out));
  HIP_CHECK(hipGetLastError());
  return read_scalars(m, &g->state_norm, out, 1, s);
}

void lusgs_launch(McfdMesh* m, const double* v, double* out, hipStream_t s, const int* te, bool wide) {
  if (wide && m->lusgs_two) {
    const std::vector<int>& pf = m->host_lvl_f;
    const std::vector<int>& pb = m->host_lvl_b;
    if (pf[2] > pf[1])
      hipLaunchKernelGGL(k_lusgs_two_first, dim3(blocks(pf[2] - pf[1])), dim3(kBlock), 0, s, m->lusgs, v, out, te,
                         pf[1], pf[2]);
    if (pb[2] > pb[1])
      hipLaunchKernelGGL(k_lusgs_two_second, dim3(blocks(pb[2] - pb[1])), dim3(kBlock), 0, s, m->lusgs, v, out, te,
                         pb[1], pb[2]);
  } else if (wide) {
    // few, large levels: one launch per level over the whole GPU
    const std::vector<int>& pf = m->host_lvl_f;
    for (int l = 0; l < m->lusgs.n_levels_f; ++l)
      if (pf[l + 1] > pf[l])
        hipLaunchKernelGGL(k_lusgs_level<true>, dim3(blocks(pf[l + 1] - pf[l])), dim3(kBlock), 0, s, m->lusgs, v,
                         out, te, pf[l], pf[l + 1]);
    const std::vector<int>& pb = m->host_lvl_b;
    for (int l = 0; l < m->lusgs.n_levels_b; ++l)
      if (pb[l + 1] > pb[l])
        hipLaunchKernelGGL(k_lusgs_level<false>, dim3(blocks(pb[l + 1] - pb[l])), dim3(kBlock), 0, s, m->lusgs,
                         out, out, te, pb[l], pb[l + 1]);
  } else {
    hipLaunchKernelGGL(k_lusgs_gather_own, dim3(blocks(m->n_cells)), dim3(kBlock), 0, s, m->n_cells, m->lusgs, v);
    hipLaunchKernelGGL(k_lusgs_sweep<true>, dim3(1), dim3(1024), 0, s, m->lusgs, out, te);
    hipLaunchKernelGGL(k_lusgs_sweep<false>, dim3(1), dim3(1024), 0, s, m->lusgs, out, te);
  }
}

// Both launch paths compute every cell with the same fetch and finish, so
// they give identical bits and the choice only decides speed. Many small
// levels favour the one-workgroup sweeps (a level costs a barrier, not a
// launch), few large ones a launch per level; between the two bounds the
// first application on a mesh times both (one warm-up and two timed
// applications each) and keeps the faster.
constexpr int kLusgsTimedLevel = 384;  // average cells per level from which both are timed

int lusgs_choose_path(McfdMesh* m, const double* v, double* out, hipStream_t s, const int* te) {
  const int n = m->n_cells;
  const int lf = m->lusgs.n_levels_f, lb = m->lusgs.n_levels_b;
  if (lf <= 0 || lb <= 0) {
    m->lusgs_path = 0;
    return 0;
  }
  const int per = std::min(n / lf, n / lb);
  if (per >= kLusgsWideLevel || per < kLusgsTimedLevel) {
    m->lusgs_path = per >= kLusgsWideLevel ? 1 : 0;
    return 0;
  }
  if (v == out) return 0;  // an in-place call is not repeatable: decide on a later call
  hipEvent_t ev[2];
  for (int k = 0; k < 2; ++k) HIP_CHECK(hipEventCreate(&ev[k]));
  float ms[2] = {0.0f, 0.0f};
  for (int wide = 0; wide < 2; ++wide) {
    lusgs_launch(m, v, out, s, te, wide == 1);
    HIP_CHECK(hipEventRecord(ev[0], s));
    lusgs_launch(m, v, out, s, te, wide == 1);
    lusgs_launch(m, v, out, s, te, wide == 1);
    HIP_CHECK(hipEventRecord(ev[1], s));
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventSynchronize(ev[1]));
    HIP_CHECK(hipEventElapsedTime(&ms[wide], ev[0], ev[1]));
  }
  for (int k = 0; k < 2; ++k) HIP_CHECK(hipEventDestroy(ev[k]));
  m->lusgs_path = ms[1] < ms[0] ? 1 : 0;
  return 0;
}

int lusgs_apply(McfdMesh* m, const double* v, double* out, hipStream_t s, bool transpose = false) {
  const int* te = transpose ? m->lusgs.off_t : nullptr;
  if (m->lusgs_path < 0 && lusgs_choose_path(m, v, out, s, te)) return -1;
  const int n = m->n_cells;
  bool wide = m->lusgs_path == 1;
  if (m->lusgs_path < 0)  // undecided (in-place call): the size rule
    wide = m->lusgs.n_levels_f > 0 && n / m->lusgs.n_levels_f >= kLusgsWideLevel &&
           n / m->lusgs.n_levels_b >= kLusgsWideLevel;
  lusgs_launch(m, v, out, s, te, wide);
  HIP_CHECK(hipGetLastError());
  return 0;
}

// gm_scalars: slots 0-3 (||b||, ||M b||, ||r||, ||M r||), 40 (||v|| of the
// Jacobian product), 63 (||state / sigma||); from kGmScalarCols the column's
// ||w|| and projections (col + 3), after them the m combination coefficients.
constexpr int kGmScalarCols = 64;

int ensure_gmres(McfdMesh* m, int krylov) {
  if (krylov <= m->gmres_m) return 0;
  long N = 4L * m->n_cells;
  double* bufs[7] = {nullptr};
  size_t sizes[7] = {(size_t)(krylov + 1) * N, (size_t)N, (size_t)N, (size_t)N, (size_t)N, (size_t)N,
                     (size_t)kGmScalarCols + 2 * (size_t)krylov + 3};
  for (int k = 0; k < 7; ++k) {
    if (upload_implicit(m, &bufs[k], (const double*)nullptr, sizes[k])) return -1;
  }
  // dot partials: two alternating 2-slot halves (each MGS kernel reads its
  // predecessor's while writing its own), then the Krylov vector's (v, v)
  if (upload_implicit(m, &m->gm_partial, (const double*)nullptr, 5 * (size_t)kDotBlocks)) return -1;
  m->gm_V = bufs[0];
  m->gm_w = bufs[1];
  m->gm_r = bufs[2];
  m->gm_av = bufs[3];
  m->gm_pert = bufs[4];
  m->gm_rp = bufs[5];
  m->gm_scalars = bufs[6];
  if (m->gm_y_host) (void)hipHostFree(m->gm_y_host);
  m->gm_y_host = nullptr;
  HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&m->gm_y_host), sizeof(double) * krylov));
  if (m->gm_scalars_host) (void)hipHostFree(m->gm_scalars_host);
  m->gm_scalars_host = nullptr;
  HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&m->gm_scalars_host), sizeof(double) * sizes[6]));
  m->gm_scalars_n = sizes[6];
  m->gmres_m = krylov;
  return 0;
}

struct JacobianOp {
  McfdMesh* m;
  const McfdSettings* st;
  const double* state;
  const double* rate;
  const double* dt;
  int per_cell;
  double sigma[4];
  double base;  // ||state / sigma||
  // adjoint: A^T v = v - J^T (dt v), J the exact Jacobian of the residual
  // with settings st (reverse kernels); preconditioner the transposed LU-SGS
  bool adjoint = false;
};

int mcfd_residual_vjp_impl(McfdMesh* m, const McfdSettings* st, const double* u,
                           const double* lam, double* ubar, double* sbar, hipStream_t s);

__global__ void __launch_bounds__(kBlock)
    k_scale_dt(int n, const double* __restrict__ v, const double* __restrict__ dt, int per_cell,
               double* __restrict__ w) {
  int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double h = per_cell ? dt[i] : dt[0];
  P4 a = load4(v, i);
  store4(w, i, h * a.r, h * a.u, h * a.v, h * a.p);
}

// out = A^T v = v - J^T (dt v)
int adjoint_apply(JacobianOp& op, const double* v, double* out, hipStream_t s) {
  McfdMesh* m = op.m;
  long N = 4L * m->n_cells;
  hipLaunchKernelGGL(k_scale_dt, dim3(blocks(m->n_cells)), dim3(kBlock), 0, s, m->n_cells, v, op.dt,
                     op.per_cell, m->gm_pert);
  if (mcfd_residual_vjp_impl(m, op.st, op.state, m->gm_pert, m->gm_rp, nullptr, s)) return -1;
  HIP_CHECK(hipMemcpyAsync(out, v, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_axpby, dim3(lblocks(N)), dim3(kBlock), 0, s, N, -1.0, m->gm_rp, 1.0, out);
  HIP_CHECK(hipGetLastError());
  return 0;
}

int jacobian_apply(JacobianOp& op, const double* v, double* out, hipStream_t s,
                   const double* vpart = nullptr);

// vpart: block partials of (v, v) already formed by the kernel that wrote v
// (the Jacobian product's ||v|| then needs only their final sum)
int op_apply(JacobianOp& op, const double* v, double* out, hipStream_t s, const double* vpart = nullptr) {
  return op.adjoint ? adjoint_apply(op, v, out, s) : jacobian_apply(op, v, out, s, vpart);
}

int op_precond(JacobianOp& op, const double* v, double* out, hipStream_t s) {
  return lusgs_apply(op.m, v, out, s, op.adjoint);
}

// out = v - dt * J v / sigma with J v = (R1(s + eps*sigma*v) - R1(s)) / eps
int jacobian_apply(JacobianOp& op, const double* v, double* out, hipStream_t s, const double* vpart) {
  McfdMesh* m = op.m;
  long N = 4L * m->n_cells;
  // ||v|| stays on the device: eps = (fd * max(base, 1)) / ||v|| is formed
  // by the perturbation and finishing kernels (same operations), so the
  // product needs no host round trip
  double* dn = m->gm_scalars + 40;
  if (!vpart) {
    hipLaunchKernelGGL(k_dot_partial, dim3(kDotBlocks), dim3(kBlock), 0, s, N, v, v, m->gm_partial);
    vpart = m->gm_partial;
  }
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kBlock), 0, s, kDotBlocks, vpart, dn, 1, (double*)nullptr);
  const double fd = 1.4901161193847656e-08;  // sqrt(2^-52)
  const double num = fd * (op.base > 1.0 ? op.base : 1.0);
  const McfdSettings* st = op.st;
  if (m->fused && !m->needs_inflow && st->reconstruction == 0 && (!st->viscosity || m->fused_visc)) {
    // perturbation, first-order residual and finish in one fused launch
    StageArgs a{};
    a.op = kOpJacobian;
    a.dt_per_cell = op.per_cell;
    a.u = op.state;
    a.u0 = op.rate;
    a.u1 = v;
    a.dt = op.dt;
    a.out = out;
    a.cfl = st->cfl;
    a.jv_norm = dn;
    a.jv_num = num;
    for (int k = 0; k < 4; ++k) a.jv_sigma[k] = op.sigma[k];
    if (check_viscous(m, st)) return -1;
    GeomArrays ga{m->area,   m->grad_ptr, m->grad_cx,  m->grad_cy, m->lim_ptr,
                  m->lim_dx, m->lim_dy,   m->acc_ptr,  m->face_len, m->acc_coef};
    if (st->viscosity)
      launch_fused_visc(m, st->flux, m->has_far_field, fused_lds_bytes(m, false, true), s, ga, st, a);
    else
      launch_fused(m, st->flux, false, m->has_far_field, false, fused_lds_bytes(m, false), s, ga, st, a);
    HIP_CHECK(hipGetLastError());
    return 0;
  }
  int nb = blocks(m->n_cells);
  hipLaunchKernelGGL(k_perturb, dim3(nb), dim3(kBlock), 0, s, m->n_cells, op.state, v, num, dn,
                     op.sigma[0], op.sigma[1], op.sigma[2], op.sigma[3], m->gm_pert);
  if (mcfd_residual(m, op.st, m->gm_pert, m->gm_rp, s)) return -1;
  hipLaunchKernelGGL(k_jv_finish, dim3(nb), dim3(kBlock), 0, s, m->n_cells, v, m->gm_rp, op.rate,
                     num, dn, op.dt, op.per_cell, op.sigma[0], op.sigma[1], op.sigma[2], op.sigma[3],
                     out);
  HIP_CHECK(hipGetLastError());
  return 0;
}

// LAPACK dlartg (3.10+): [c s; -s c] [f; g] = [r; 0]
void lartg(double f, double g, double& c, double& sn, double& r) {
  if (g == 0.0) {
    c = 1.0;
    sn = 0.0;
    r = f;
  } else if (f == 0.0) {
    c = 0.0;
    sn = g > 0 ? 1.0 : -1.0;
    r = fabs(g);
  } else {
    double d = sqrt(f * f + g * g);
    c = fabs(f) / d;
    r = copysign(d, f);
    sn = g / r;
  }
}

// One call of restarted, left-preconditioned GMRES with maxiter = 1 (the
// algorithm of scipy.sparse.linalg.gmres, SciPy 1.15): MGS + Givens, inner
// stop on the preconditioned residual, success on the true residual.
int gmres_cycle(JacobianOp& op, const double* b, double* x, double rtol, int restart,
                bool x_nonzero, int& inner, int& info, double& rnorm, hipStream_t s) {
  McfdMesh* m = op.m;
  long N = 4L * m->n_cells;
  int nbl = lblocks(N);
  double* V = m->gm_V;
  double* sc = m->gm_scalars;
  double* cs = sc + kGmScalarCols;  // per-column scalars
  std::vector<double> host((size_t)restart + 3);
  inner = 0;
  if (dot_dev(m, N, b, b, sc + 0, true, s)) return -1;
  double bnrm2 = 0.0;
  if (read_scalars(m, &bnrm2, sc, 1, s)) return -1;
  if (bnrm2 == 0.0) {
    HIP_CHECK(hipMemcpyAsync(x, b, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    info = 0;
    rnorm = 0.0;
    return 0;
  }
  double atol = rtol * bnrm2;
  // ptol = ||M b|| * min(1, atol / bnrm2)
  if (op_precond(op, b, m->gm_w, s)) return -1;
  if (dot_dev(m, N, m->gm_w, m->gm_w, sc + 1, true, s)) return -1;
  double mb = 0.0;
  if (read_scalars(m, &mb, sc + 1, 1, s)) return -1;
  double ratio = atol / bnrm2;
  double ptol = mb * (ratio < 1.0 ? ratio : 1.0);
  const double eps = 2.220446049250313e-16;
  // r = b - A x (or b when x == 0)
  if (x_nonzero) {
    if (op_apply(op, x, m->gm_av, s)) return -1;
    HIP_CHECK(hipMemcpyAsync(m->gm_r, b, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_axpby, dim3(nbl), dim3(kBlock), 0, s, N, -1.0, m->gm_av, 1.0, m->gm_r);
    if (dot_dev(m, N, m->gm_r, m->gm_r, sc + 2, true, s)) return -1;
    double r0 = 0.0;
    if (read_scalars(m, &r0, sc + 2, 1, s)) return -1;
    if (r0 < atol) {
      info = 0;
      rnorm = r0;
      return 0;
    }
  } else {
    HIP_CHECK(hipMemcpyAsync(m->gm_r, b, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  }
  // v0 = M r / ||M r||
  if (op_precond(op, m->gm_r, V, s)) return -1;
  if (dot_dev(m, N, V, V, sc + 3, true, s)) return -1;
  double t0 = 0.0;
  if (read_scalars(m, &t0, sc + 3, 1, s)) return -1;
  // partials of each Krylov vector's (v, v), for the product's ||v||
  double* vpart = m->gm_partial + 4L * kDotBlocks;
  hipLaunchKernelGGL(k_scale_sq_partial, dim3(kDotBlocks), dim3(kBlock), 0, s, N, V, 1.0 / t0, V, vpart);
  std::vector<double> h((size_t)restart * (restart + 1), 0.0);  // h[col*(restart+1) + row]
  std::vector<double> S(restart + 1, 0.0), gc(restart, 0.0), gs(restart, 0.0);
  S[0] = t0;
  bool breakdown = false;
  double presid = 0.0;
  int col = 0;
  for (col = 0; col < restart; ++col) {
    double* vc = V + (long)col * N;
    double* vn = V + (long)(col + 1) * N;
    if (op_apply(op, vc, m->gm_av, s, vpart)) return -1;
    if (op_precond(op, m->gm_av, m->gm_w, s)) return -1;
    // h0 = ||w||, then modified Gram-Schmidt with device coefficients: each
    // projection w -= h_k v_k runs in the pass that forms h_{k+1} = (v_{k+1}, w)
    // (and the last in the pass forming ||w||), same arithmetic as one dot and
    // one axpy per step. Each kernel finishes its predecessor's dot from the
    // block partials (alternating halves of gm_partial), and the normalisation
    // finishes the last, so the column takes col + 3 launches and one read.
    double* part[2] = {m->gm_partial, m->gm_partial + 2L * kDotBlocks};
    hipLaunchKernelGGL(k_dot2_partial, dim3(kDotBlocks), dim3(kBlock), 0, s, N, V, m->gm_w, part[0]);
    for (int k = 1; k <= col; ++k) {
      // k = 1 also finishes h0 (square root) into cs[0]
      double* o = k == 1 ? cs : cs + k;
      hipLaunchKernelGGL(k_axpy_dot_partial<false>, dim3(kDotBlocks), dim3(kBlock), 0, s, N,
                         V + (long)(k - 1) * N, m->gm_w, part[(k - 1) & 1], k == 1 ? 2 : 1, k == 1 ? 1 : 0, o,
                         host_slot(m, o), V + (long)k * N, part[k & 1]);
    }
    {
      double* o = col == 0 ? cs : cs + 1 + col;
      hipLaunchKernelGGL(k_axpy_dot_partial<true>, dim3(kDotBlocks), dim3(kBlock), 0, s, N, V + (long)col * N,
                         m->gm_w, part[col & 1], col == 0 ? 2 : 1, col == 0 ? 1 : 0, o, host_slot(m, o),
                         (const double*)nullptr, part[(col + 1) & 1]);
    }
    // h1 = ||w||; v_{col+1} = w / h1, or w on breakdown (h1 <= eps h0), with
    // the partials of its (v, v) for the next product
    hipLaunchKernelGGL(k_arnoldi_scale, dim3(kDotBlocks), dim3(kBlock), 0, s, N, m->gm_w, part[(col + 1) & 1],
                       cs, eps, cs + col + 2, host_slot(m, cs + col + 2), vn, vpart);
    HIP_CHECK(hipGetLastError());
    if (read_scalars(m, host.data(), cs, col + 3, s)) return -1;
    double h0 = host[0];
    double* hc = h.data() + (size_t)col * (restart + 1);
    for (int k = 0; k <= col; ++k) hc[k] = host[1 + k];
    double h1 = host[col + 2];
    hc[col + 1] = h1;
    if (h1 <= eps * h0) {
      hc[col + 1] = 0.0;
      breakdown = true;
    }
    for (int k = 0; k < col; ++k) {
      double c = gc[k], sn = gs[k];
      double n0 = hc[k], n1 = hc[k + 1];
      hc[k] = c * n0 + sn * n1;
      hc[k + 1] = -sn * n0 + c * n1;
    }
    double c, sn, mag;
    lartg(hc[col], hc[col + 1], c, sn, mag);
    gc[col] = c;
    gs[col] = sn;
    hc[col] = mag;
    hc[col + 1] = 0.0;
    double tmp = -sn * S[col];
    S[col] = c * S[col];
    S[col + 1] = tmp;
    presid = fabs(tmp);
    inner += 1;
    if (presid <= ptol || breakdown) break;
  }
  if (col == restart) col = restart - 1;
  // solve the upper-triangular system
  if (h[(size_t)col * (restart + 1) + col] == 0.0) S[col] = 0.0;
  // y in pinned host memory: the stream synchronises on the residual norm
  // below, before the next cycle writes y again
  double* y = m->gm_y_host;
  for (int k = 0; k <= col; ++k) y[k] = S[k];
  for (int k = col; k > 0; --k) {
    if (y[k] != 0.0) {
      y[k] /= h[(size_t)k * (restart + 1) + k];
      double t = y[k];
      for (int i = 0; i < k; ++i) y[i] -= t * h[(size_t)k * (restart + 1) + i];
    }
  }
  if (y[0] != 0.0) y[0] /= h[0];
  double* y_dev = cs + restart + 3;
  HIP_CHECK(hipMemcpyAsync(y_dev, y, sizeof(double) * (col + 1), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_combine, dim3(nbl), dim3(kBlock), 0, s, N, col + 1, V, y_dev, x);
  // true residual
  if (op_apply(op, x, m->gm_av, s)) return -1;
  HIP_CHECK(hipMemcpyAsync(m->gm_r, b, sizeof(double) * N, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_axpby, dim3(nbl), dim3(kBlock), 0, s, N, -1.0, m->gm_av, 1.0, m->gm_r);
  if (dot_dev(m, N, m->gm_r, m->gm_r, sc + 2, true, s)) return -1;
  if (read_scalars(m, &rnorm, sc + 2, 1, s)) return -1;
  info = (rnorm <= atol) ? 0 : 1;
  HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

namespace {

template <int SCHEME, bool MUSCL, int LIM>
void launch_vjp_fused_t(McfdMesh* m, size_t lds, hipStream_t s, const GeomArrays& ga, const double* u,
                        const double* lam, double g, double* sbar) {
  auto kern = k_vjp_fused<SCHEME, MUSCL, LIM>;
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
  hipLaunchKernelGGL(kern, dim3(m->nb), dim3(kBlock), lds, s, m->max_local, m->max_grad, m->max_faces, m->max_acc,
                     m->blocks, ga, m->faces, u, lam, g, m->grad_rev, m->adj.wbar, m->adj.ebar, sbar,
                     m->status);
}

template <int SCHEME>
void launch_vjp_fused_s(McfdMesh* m, bool muscl, int lim, size_t lds, hipStream_t s, const GeomArrays& ga,
                        const double* u, const double* lam, double g, double* sbar) {
  if (!muscl) return launch_vjp_fused_t<SCHEME, false, 0>(m, lds, s, ga, u, lam, g, sbar);
  if (lim <= 1) return launch_vjp_fused_t<SCHEME, true, 0>(m, lds, s, ga, u, lam, g, sbar);
  if (lim == 2) return launch_vjp_fused_t<SCHEME, true, 2>(m, lds, s, ga, u, lam, g, sbar);
  return launch_vjp_fused_t<SCHEME, true, 3>(m, lds, s, ga, u, lam, g, sbar);
}

// LDS of k_vjp_fused: W (4 max_local), G and phi (12 max_grad), then
// lambda/A (4 max_grad) aliased by the face adjoints and offsets
// (12 max_faces), then the gather items (max_acc shorts).
size_t vjp_fused_lds_bytes(const McfdMesh* m) {
  size_t region = std::max(4 * (size_t)m->max_grad, 12 * (size_t)m->max_faces);
  size_t items = ((size_t)m->max_acc * sizeof(unsigned short) + 7) & ~(size_t)7;
  return (4 * (size_t)m->max_local + 12 * (size_t)m->max_grad + region) * sizeof(double) + items;
}

int mcfd_residual_vjp_impl(McfdMesh* m, const McfdSettings* st, const double* u,
                           const double* lam, double* ubar, double* sbar, hipStream_t s) {
  if (!m || !u || !lam || !ubar) return fail("mcfd_residual_vjp: NULL argument");
  if (!m->adjoint) return fail("mcfd_residual_vjp: call mcfd_mesh_set_adjoint first");
  if (check_settings(st)) return -1;
  const bool para = st->reconstruction == 2 && m->n_grad > 0;
  if (para && !(m->parabolic && m->adj.par_rev))
    return fail("mcfd_residual_vjp: MUSCL parabolic needs mcfd_mesh_set_parabolic_adjoint first");
  if (st->viscosity && !(m->viscous && m->VFA))
    return fail("mcfd_residual_vjp: viscous settings need mcfd_mesh_set_viscous before mcfd_mesh_set_adjoint");
  if (check_inflow(m, st)) return -1;
  int n = m->n_cells;
  bool muscl = (st->reconstruction == 1 || para) && m->n_grad > 0;
  size_t vjp_lds = m->fused ? vjp_fused_lds_bytes(m) : 0;
  if (m->fused && !para && !st->viscosity && st->flux != 2 && !m->needs_inflow && vjp_lds <= 64 * 1024) {
    // one workgroup per block (reverse-mode fluxes), then the neighbour pass
    GeomArrays ga{m->area,   m->grad_ptr, m->grad_cx,  m->grad_cy, m->lim_ptr,
                  m->lim_dx, m->lim_dy,   m->acc_ptr,  m->face_len, m->acc_coef};
    if (st->flux == 1)
      launch_vjp_fused_s<1>(m, muscl, st->limiter, vjp_lds, s, ga, u, lam, st->gamma, sbar);
    else
      launch_vjp_fused_s<0>(m, muscl, st->limiter, vjp_lds, s, ga, u, lam, st->gamma, sbar);
    hipLaunchKernelGGL(k_cell_adjoint2, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, m->W, lam,
                       m->adj, muscl ? 1 : 0, st->gamma, st->acceleration_x, st->acceleration_y, st->gas_constant,
                       ubar);
    HIP_CHECK(hipGetLastError());
    return 0;
  }
  hipLaunchKernelGGL(k_primitive, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, m->W, st->gamma,
                     m->status);
  AdjArrays aa = m->adj;
  if (para) {
    hipLaunchKernelGGL(k_gradient_parabolic_t<true>, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W,
                       m->grad_ptr, m->grad_nb, m->grad_cx, m->grad_cy, m->lim_ptr, m->lim_dx,
                       m->lim_dy, st->limiter, m->par_ptr, m->par_nb, m->par_coef, m->par_full,
                       m->G, m->K, m->PHI, m->adj.arg);
    aa.K = m->K;
    aa.pfull = m->par_full;
    aa.par_ptr = m->par_ptr;
    aa.par_nb = m->par_nb;
    aa.par_coef = m->par_coef;
  } else {
    aa.par_rev = nullptr;
    aa.ebar_par = nullptr;
    if (muscl)
      hipLaunchKernelGGL(k_gradient_t<true>, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->grad_ptr,
                         m->grad_nb, m->grad_cx, m->grad_cy, m->lim_ptr, m->lim_dx, m->lim_dy,
                         st->limiter, m->G, m->PHI, m->adj.arg);
  }
  const bool inflow = m->inflow && m->inf.n > 0;
  // pressure / mass-flow boundaries: the forward exterior states from the
  // reconstructed face states (what the residual's fluxes saw)
  if (inflow && inflow_states(m, st, s, true, para, muscl)) return -1;
  if (m->n_faces) {
    // Rusanov and HLLC in reverse mode; Roe with duals
    bool reverse = st->flux != 2;
    const double* Kq = para ? m->K : nullptr;
    if (reverse)
      hipLaunchKernelGGL(k_flux_adjoint_rev, dim3(blocks(m->n_faces)), dim3(kBlock), 0, s,
                         m->n_faces, m->faces, m->W, m->G, m->PHI, Kq, m->face_len, m->area, lam,
                         st->flux, muscl ? 1 : 0, st->gamma, m->FADJ, sbar);
    else
      hipLaunchKernelGGL(k_flux_adjoint, dim3(blocks(m->n_faces)), dim3(kBlock), 0, s, m->n_faces,
                         m->faces, m->W, m->G, m->PHI, Kq, m->face_len, m->area, lam, st->flux,
                         muscl ? 1 : 0, st->gamma, m->FADJ, sbar, 0);
    if (inflow) {
      // exterior-state adjoints (right slots) -> reconstructed left states
      hipLaunchKernelGGL(k_inflow_adjoint, dim3(blocks(m->inf.n)), dim3(kBlock), 0, s, m->inf, m->faces, m->W,
                         m->G, m->PHI, Kq, muscl ? 1 : 0, st->gamma, st->gas_constant, m->FADJ);
      if (m->inflow_patches)
        hipLaunchKernelGGL(k_mass_flow_inlet_adjoint, dim3(m->inflow_patches), dim3(256), 0, s, m->inf,
                           m->face_len, st->gamma, st->gas_constant, m->FADJ);
    }
  }
  hipLaunchKernelGGL(k_cell_adjoint1, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->G, m->PHI,
                     m->FADJ, aa, muscl ? 1 : 0, st->limiter);
  aa.vbar = nullptr;
  if (st->viscosity) {
    ViscParams vp = visc_params(st);
    ViscArrays va = m->visc;
    double* vib = nullptr;
    if (inflow) {
      // the viscous boundary values of the inlets: exterior states from the
      // owner cells, as in the forward residual
      if (inflow_states(m, st, s, false, false, false, m->inf_ext_v)) return -1;
      va.inf_ext = m->inf_ext_v;
      va.inf_first = m->inf.first;
      vib = m->inf_vbar;
      hipLaunchKernelGGL(k_fill, dim3(lblocks(4L * m->inf.n)), dim3(kBlock), 0, s, 4L * m->inf.n, 0.0, vib);
    }
    hipLaunchKernelGGL(k_visc_gradient, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->grad_ptr,
                       m->grad_nb, va, m->faces, vp, m->VG, m->status);
    if (m->n_faces)
      hipLaunchKernelGGL(k_visc_face_adjoint, dim3(blocks(m->n_faces)), dim3(kBlock), 0, s, m->n_faces,
                         m->faces, m->W, m->VG, va, vp, m->face_len, m->area, lam, m->VFA, sbar, m->inf, vib);
    hipLaunchKernelGGL(k_visc_cell_adjoint, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->VG,
                       m->acc_ptr, m->acc_item, m->grad_ptr, m->grad_rev, va, m->faces, vp, m->VFA,
                       m->adj.wbar, m->VBAR, sbar, m->inf, vib);
    if (inflow && m->inflow_patches)
      hipLaunchKernelGGL(k_mass_flow_inlet_visc_adjoint, dim3(1), dim3(256), 0, s, m->inf, m->inflow_patches,
                         m->inf_ext_v, m->face_len, m->faces.left, vib, st->gamma, st->gas_constant,
                         m->adj.wbar);
    aa.vbar = m->VBAR;
  }
  hipLaunchKernelGGL(k_cell_adjoint2, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, m->W, lam,
                     aa, muscl ? 1 : 0, st->gamma, st->acceleration_x, st->acceleration_y, st->gas_constant,
                     ubar);
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
  for (int b = 0; b < d->n_bc; ++b) {
    int kind = d->bc_kind ? d->bc_kind[b] : -1;
    if (kind < MCFD_FACE_WALL || kind > MCFD_FACE_MASS_FLOW_OUTLET) {
      delete m;
      return fail("mcfd_mesh_create: unknown boundary record kind");
    }
    if (kind == MCFD_FACE_FAR_FIELD) m->has_far_field = true;
    if (kind >= MCFD_FACE_PRESSURE_INLET) m->needs_inflow = true;
  }
  if (d->n_bc > 0) m->host_bc_kind.assign(d->bc_kind, d->bc_kind + d->n_bc);
  for (int i = 0; i < d->n_cells && d->grad_ptr && d->lim_ptr; ++i) {
    m->max_cell_entries = std::max(m->max_cell_entries, d->grad_ptr[i + 1] - d->grad_ptr[i]);
    m->max_cell_entries = std::max(m->max_cell_entries, d->lim_ptr[i + 1] - d->lim_ptr[i]);
  }
  if (d->n_faces > 0 && d->face_bc) m->host_face_bc.assign(d->face_bc, d->face_bc + d->n_faces);
  if (d->acc_ptr) m->host_acc_ptr.assign(d->acc_ptr, d->acc_ptr + (size_t)d->n_cells + 1);
  if (d->dts_ptr) m->host_dts_ptr.assign(d->dts_ptr, d->dts_ptr + (size_t)d->n_cells + 1);
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
  rc |= upload(m, &m->acc_coef, d->acc_coef, d->n_acc);
  rc |= upload(m, &m->dts_ptr, d->dts_ptr, C + 1);
  rc |= upload(m, &m->dts_nx, d->dts_nx, d->n_dts);
  rc |= upload(m, &m->dts_ny, d->dts_ny, d->n_dts);
  rc |= upload(m, &m->dts_len, d->dts_len, d->n_dts);
  rc |= upload(m, &m->dts_bc, d->dts_bc, d->n_dts);
  rc |= upload(m, &m->bc_kind, d->bc_kind, d->n_bc);
  rc |= upload(m, const_cast<double**>(&fa.bc_state), d->bc_state, 4 * (size_t)d->n_bc);
  rc |= upload(m, const_cast<double**>(&fa.bc_wall), d->bc_wall, 2 * (size_t)d->n_bc);
  m->dts = DtArrays{m->dts_ptr, m->dts_nx, m->dts_ny, m->dts_len,
                    m->dts_bc,  m->bc_kind, fa.bc_state, m->area};
  rc |= alloc_scratch(m, &m->W, 4 * C);
  rc |= alloc_scratch(m, &m->G, 8 * C);
  rc |= alloc_scratch(m, &m->PHI, 4 * C);
  rc |= alloc_scratch(m, &m->FLUX, 4 * F);
  rc |= alloc_scratch(m, &m->partial, 4 * (size_t)blocks(d->n_cells) + 4);
  rc |= alloc_scratch(m, &m->status, 2);
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

int mcfd_mesh_set_blocks(McfdMesh* m, const McfdBlockDesc* d) {
  if (!m) return fail("mcfd_mesh_set_blocks: NULL mesh");
  (void)hipDeviceSynchronize();
  free_blocks(m);
  if (!d) return 0;
  if (d->n_blocks <= 0 || d->block_cells <= 0 || !d->info)
    return fail("mcfd_mesh_set_blocks: no blocks");
  {
    // own cells of block b are [info[b][8], info[b][8] + n_own) and tile the mesh;
    // a block has at most kBlock/2 own cells, kBlock own + ring-1 cells and
    // touches at most kBlock flux faces (one thread per item in each phase)
    long next = 0;
    for (int b = 0; b < d->n_blocks; ++b) {
      const int32_t* inf = d->info + (size_t)kBlockInfo * b;
      if (inf[8] != next || inf[1] <= 0 || inf[1] > kBlock / 2 || inf[2] < 0 ||
          inf[1] + inf[2] > kBlock || inf[5] < 0 || inf[5] > kBlock)
        return fail("mcfd_mesh_set_blocks: blocks do not tile the mesh or touch too many faces");
      next += inf[1];
    }
    if (next != m->n_cells) return fail("mcfd_mesh_set_blocks: blocks do not cover the mesh");
    if (d->max_faces > kBlock) return fail("mcfd_mesh_set_blocks: max_faces above the block size");
  }
  m->nb = d->n_blocks;
  m->block_cells = d->block_cells;
  m->max_local = d->max_local;
  m->max_grad = d->max_grad;
  m->max_faces = d->max_faces;
  {
    // the gather stages each block's accumulation and spectral-radius items
    const std::vector<int>& aptr = m->host_acc_ptr;
    const std::vector<int>& dptr = m->host_dts_ptr;
    if (aptr.size() != (size_t)m->n_cells + 1 || dptr.size() != (size_t)m->n_cells + 1)
      return fail("mcfd_mesh_set_blocks: the mesh has no accumulation lists");
    m->max_acc = 0;
    m->max_dts = 0;
    for (int b = 0; b < d->n_blocks; ++b) {
      const int32_t* inf = d->info + (size_t)kBlockInfo * b;
      m->max_acc = std::max(m->max_acc, aptr[inf[8] + inf[1]] - aptr[inf[8]]);
      m->max_dts = std::max(m->max_dts, dptr[inf[8] + inf[1]] - dptr[inf[8]]);
    }
    // entries per block: the block-relative offsets after its last gradient cell
    m->max_lsq = 0;
    for (int b = 0; b < d->n_blocks; ++b) {
      const int32_t* inf = d->info + (size_t)kBlockInfo * b;
      const long end = (long)inf[7] + inf[1] + inf[2];
      if (end >= d->n_gptr) return fail("mcfd_mesh_set_blocks: gptr too short");
      m->max_lsq = std::max(m->max_lsq, d->gptr[end]);
    }
  }
  if (fused_lds_bytes(m, true) > 64 * 1024) {
    free_blocks(m);
    return fail("mcfd_mesh_set_blocks: block needs more than 64 KiB of LDS; use smaller blocks");
  }
  int rc = 0;
  BlockArrays& b = m->blocks;
  {
    // device block records and the local-cell list at a fixed stride per
    // block (padded with each block's first cell), so that the fused kernel
    // addresses a block's cells from its block id alone
    const int nb = d->n_blocks;
    const int stride = std::max(d->max_local, kBlock);
    std::vector<int32_t> info((size_t)kBlockInfoDev * nb);
    std::vector<int32_t> cells((size_t)stride * nb);
    const std::vector<int>& aptr = m->host_acc_ptr;
    const std::vector<int>& dptr = m->host_dts_ptr;
    for (int k = 0; k < nb; ++k) {
      const int32_t* in = d->info + (size_t)kBlockInfo * k;
      int32_t* out = info.data() + (size_t)kBlockInfoDev * k;
      const int n_loc = in[1] + in[2] + in[3];
      if (in[0] < 0 || n_loc > d->max_local || (long)in[0] + n_loc > d->n_local) {
        rc = fail("mcfd_mesh_set_blocks: local-cell lists out of range");
        break;
      }
      std::copy(in, in + kBlockInfo, out);
      out[0] = stride * k;
      const long end = (long)in[7] + in[1] + in[2];
      out[10] = d->gptr[end];
      out[11] = d->lptr[end];
      out[12] = aptr[in[8]];
      out[13] = aptr[in[8] + in[1]] - aptr[in[8]];
      out[14] = dptr[in[8]];
      out[15] = dptr[in[8] + in[1]] - dptr[in[8]];
      int32_t* c = cells.data() + (size_t)stride * k;
      std::copy(d->cells + in[0], d->cells + in[0] + n_loc, c);
      std::fill(c + n_loc, c + stride, d->cells[in[0]]);
    }
    b.cstride = stride;
    if (!rc) rc |= upload_block(m, const_cast<int**>(&b.info), info.data(), info.size());
    if (!rc) rc |= upload_block(m, const_cast<int**>(&b.cells), cells.data(), cells.size());
  }
  rc |= upload_block(m, const_cast<int**>(&b.gptr), d->gptr, (size_t)d->n_gptr);
  {
    // two spare entries: a block's ids are copied as dwords from the even
    // entry at or before its first one to the dword holding its last one
    std::vector<uint16_t> gnb(d->gnb, d->gnb + d->n_gnb);
    gnb.resize((size_t)d->n_gnb + 2, 0);
    rc |= upload_block(m, const_cast<unsigned short**>(&b.gnb), gnb.data(), gnb.size());
  }
  rc |= upload_block(m, reinterpret_cast<double**>(const_cast<double2**>(&b.gent)), d->gent,
                     2 * (size_t)d->n_gnb);
  rc |= upload_block(m, const_cast<int**>(&b.lptr), d->lptr, (size_t)d->n_gptr);
  rc |= upload_block(m, reinterpret_cast<double**>(const_cast<double2**>(&b.lent)), d->lent,
                     2 * (size_t)d->n_lent);
  rc |= upload_block(m, const_cast<int**>(&b.faces), d->faces, (size_t)d->n_block_faces);
  rc |= upload_block(m, reinterpret_cast<double**>(const_cast<FaceRec**>(&b.frec)),
                     static_cast<const double*>(d->frec), 8 * (size_t)d->n_block_faces);
  rc |= upload_block(m, const_cast<unsigned short**>(&b.acc_local), d->acc_local, (size_t)m->n_acc);
  rc |= upload_block(m, const_cast<unsigned short**>(&b.dts_local), d->dts_local, (size_t)m->n_dts);
  if (!rc) {
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(double) * (size_t)d->n_blocks) != hipSuccess) {
      rc = fail("mcfd_mesh_set_blocks: out of device memory");
    } else {
      m->block_allocations.push_back(p);
      m->block_partial = static_cast<double*>(p);
    }
  }
  if (rc) {
    std::string msg = g_last_error;
    free_blocks(m);
    return fail(msg);
  }
  m->block_n_gnb = d->n_gnb;
  m->fused = true;
  return 0;
}

int mcfd_mesh_set_block_viscous(McfdMesh* m, const double* vent, int64_t n_vent, const uint16_t* vacc_local,
                                int64_t n_vacc) {
  if (!m) return fail("mcfd_mesh_set_block_viscous: NULL mesh");
  if (!m->fused) return fail("mcfd_mesh_set_block_viscous: call mcfd_mesh_set_blocks first");
  if (!m->viscous) return fail("mcfd_mesh_set_block_viscous: call mcfd_mesh_set_viscous first");
  if (!vent || !vacc_local) return fail("mcfd_mesh_set_block_viscous: missing array");
  if (n_vent != m->block_n_gnb || n_vacc != m->n_vacc_visc)
    return fail("mcfd_mesh_set_block_viscous: array lengths do not match the blocks and viscous lists");
  if (fused_lds_bytes(m, false, true) > 64 * 1024)
    return fail("mcfd_mesh_set_block_viscous: blocks too large for the viscous fused kernel's LDS");
  (void)hipDeviceSynchronize();
  BlockArrays& b = m->blocks;
  int rc = 0;
  rc |= upload_block(m, reinterpret_cast<double**>(const_cast<double2**>(&b.vent)), vent, 2 * (size_t)n_vent);
  rc |= upload_block(m, const_cast<unsigned short**>(&b.vacc_local), vacc_local, (size_t)n_vacc);
  if (rc) return -1;
  m->fused_visc = true;
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
  StageArgs a{sg->op,   sg->dt_per_cell, sg->u,    sg->u0,   sg->u1,   sg->r1,
              sg->dt,   sg->out,         sg->r_out, sg->c[0], sg->c[1], sg->c[2],
              sg->c[3], sg->c[4],        sg->dt_next,
              sg->dt_min ? m->block_partial : nullptr, st->cfl};
  if ((sg->dt_next || sg->dt_min) &&
      (!m->fused || m->needs_inflow || !sg->dt_min || sg->op == MCFD_OP_RESIDUAL))
    return fail("mcfd_stage: dt_next needs the fused path, dt_min and a state update");
  if (check_viscous(m, st)) return -1;
  if (sg->dt_min && (st->viscosity || st->reconstruction == 2))
    return fail("mcfd_stage: the next step's dt needs the fused stage (inviscid, no MUSCL parabolic)");
  if (m->fused && m->fused_visc && st->viscosity && st->reconstruction == 0 && !m->needs_inflow &&
      !a.dt_partial) {
    // first-order viscous residual or stage (the Jacobian products of the
    // implicit path): one fused launch
    size_t lds = fused_lds_bytes(m, false, true);
    GeomArrays ga{m->area,   m->grad_ptr, m->grad_cx,  m->grad_cy, m->lim_ptr,
                  m->lim_dx, m->lim_dy,   m->acc_ptr,  m->face_len, m->acc_coef};
    {
      TimedLaunch t(m, s, MCFD_KERNEL_STAGE);
      launch_fused_visc(m, st->flux, m->has_far_field, lds, s, ga, st, a);
    }
    HIP_CHECK(hipGetLastError());
    return 0;
  }
  if (m->fused && st->reconstruction != 2 && !st->viscosity && !m->needs_inflow) {
    bool muscl = st->reconstruction == 1 && m->n_grad > 0;
    size_t lds = fused_lds_bytes(m, muscl);
    GeomArrays ga{m->area,   m->grad_ptr, m->grad_cx,  m->grad_cy, m->lim_ptr,
                  m->lim_dx, m->lim_dy,   m->acc_ptr,  m->face_len, m->acc_coef};
    {
      TimedLaunch t(m, s, MCFD_KERNEL_STAGE);
      launch_fused(m, st->flux, muscl, m->has_far_field, a.dt_partial != nullptr, lds, s, ga, st, a);
    }
    if (sg->dt_min) {
      TimedLaunch t(m, s, MCFD_KERNEL_REDUCE);
      hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(1024), 0, s, m->nb, m->block_partial, sg->dt_min);
    }
    HIP_CHECK(hipGetLastError());
    return 0;
  }
  if (residual_front(m, st, sg->u, s)) return -1;
  {
    TimedLaunch t(m, s, MCFD_KERNEL_UPDATE);
    hipLaunchKernelGGL(k_update, dim3(blocks(m->n_cells)), dim3(kBlock), 0, s, m->n_cells,
                       m->acc_ptr, m->acc_item, m->acc_coef, m->FLUX, m->face_len, m->area, m->W,
                       st->acceleration_x, st->acceleration_y, a,
                       st->viscosity ? m->vacc : ViscAcc{}, m->status);
  }
  HIP_CHECK(hipGetLastError());
  return 0;
}

namespace {
// Stable steps from the primitive states already in m->W.
int time_steps_from_w(McfdMesh* m, const McfdSettings* st, double* dt_cell, double* dt_min, hipStream_t s) {
  int n = m->n_cells;
  if (m->inflow) {
    TimedLaunch t(m, s, MCFD_KERNEL_TIME_STEP);
    if (inflow_states(m, st, s, false, false, false)) return -1;
  }
  if (st->viscosity) {
    if (check_viscous(m, st)) return -1;
    TimedLaunch t(m, s, MCFD_KERNEL_TIME_STEP);
    hipLaunchKernelGGL(k_visc_diffusivity, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W,
                       visc_params(st), m->EFF, m->status);
    hipLaunchKernelGGL(k_time_steps_viscous, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W,
                       m->dts, m->vdt, st->gamma, st->cfl, dt_cell, m->partial, m->status);
  } else {
    TimedLaunch t(m, s, MCFD_KERNEL_TIME_STEP);
    hipLaunchKernelGGL(k_time_steps, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->dts,
                       st->gamma, st->cfl, dt_cell, m->partial, m->status);
  }
  if (dt_min) {
    TimedLaunch t(m, s, MCFD_KERNEL_REDUCE);
    hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(1024), 0, s, blocks(n), m->partial, dt_min);
  }
  HIP_CHECK(hipGetLastError());
  return 0;
}
}  // namespace

int mcfd_time_steps(McfdMesh* m, const McfdSettings* st, const double* u, double* dt_cell,
                    double* dt_min, void* stream) {
  if (!m || !u || !dt_cell) return fail("mcfd_time_steps: NULL argument");
  if (check_settings(st)) return -1;
  if (check_inflow(m, st)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int n = m->n_cells;
  {
    TimedLaunch t(m, s, MCFD_KERNEL_PRIMITIVE);
    hipLaunchKernelGGL(k_primitive, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, m->W, st->gamma,
                       m->status);
  }
  return time_steps_from_w(m, st, dt_cell, dt_min, s);
}

int mcfd_time_steps_primitive(McfdMesh* m, const McfdSettings* st, const double* w, double* dt_cell,
                              double* dt_min, void* stream) {
  if (!m || !w || !dt_cell) return fail("mcfd_time_steps_primitive: NULL argument");
  if (check_settings(st)) return -1;
  if (check_inflow(m, st)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_CHECK(hipMemcpyAsync(m->W, w, sizeof(double) * 4 * (size_t)m->n_cells, hipMemcpyDeviceToDevice, s));
  return time_steps_from_w(m, st, dt_cell, dt_min, s);
}

int mcfd_check_state(McfdMesh* m, const McfdSettings* st, const double* u, void* stream) {
  if (!m || !u) return fail("mcfd_check_state: NULL argument");
  if (check_settings(st)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int n = m->n_cells;
  hipLaunchKernelGGL(k_check_state, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, st->gamma, m->status);
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_advance_clock(McfdMesh* m, const double* dt_source, double fixed_dt, double end_time,
                       double* clock, int32_t reset, void* stream) {
  if (!m || !clock) return fail("mcfd_advance_clock: NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_clock, dim3(1), dim3(64), 0, s, const_cast<double*>(dt_source), fixed_dt,
                     end_time, clock, reset, m->status);
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

// ---------------------------------------------------------------------------
// Allocation-free residual norms: the block partials of k_sumsq live in a
// static device array instead of a mesh's scratch. Each GPU holds its own
// copy of the array (2 MiB) until the process ends; its address is looked up
// once per device, with that device current, on the first call there.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxCells = 1 << 24;                 // 16,777,216 cells
constexpr int kNoAllocBlocks = kNoAllocMaxCells / kBlock;  // block partials
static __device__ __align__(256) double scratch[4 * kNoAllocBlocks];

namespace {

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// The device address of `scratch` on the current device (host code must
// never pass the array's name itself to a kernel).
int scratch_address(double** out) {
  int dev = -1;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("mcfd_sum_squares_noalloc: device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return 0;
}

}  // namespace

int mcfd_sum_squares_noalloc(const double* r, int32_t n_cells, double* sums, void* stream) {
  if (!r || !sums) return fail("mcfd_sum_squares_noalloc: NULL argument");
  if (n_cells < 0 || n_cells > kNoAllocMaxCells)
    return fail("mcfd_sum_squares_noalloc: n_cells must be in [0, 16777216], the capacity of its static scratch");
  double* partial = nullptr;
  if (scratch_address(&partial)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = blocks(n_cells);
  // every partial that k_sumsq_final reads is written first by k_sumsq in
  // stream order, so nothing is cleared
  if (nb > 0) hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(kBlock), 0, s, n_cells, r, partial);
  hipLaunchKernelGGL(k_sumsq_final, dim3(1), dim3(256), 0, s, nb, partial, sums);
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


int mcfd_mesh_set_implicit(McfdMesh* m, const McfdImplicitDesc* d) {
  if (!m) return fail("mcfd_mesh_set_implicit: NULL mesh");
  (void)hipDeviceSynchronize();
  free_implicit(m);
  if (!d) return 0;
  if (!d->off_t) return fail("mcfd_mesh_set_implicit: off_t is required");
  if (d->launch_path < 0 || d->launch_path > 2) return fail("mcfd_mesh_set_implicit: launch_path must be 0, 1 or 2");
  int n = m->n_cells;
  // Level schedules of the natural-order triangular solves: a cell's level
  // is one more than the largest level among the cells it depends on.
  std::vector<int> order(d->natural_order, d->natural_order + n);  // position -> device cell
  std::vector<int> lev_f(n, 0), lev_b(n, 0);
  int nlf = 0, nlb = 0;
  for (int k = 0; k < n; ++k) {
    int i = order[k];
    int l = 0;
    for (int e = d->off_ptr[i]; e < d->off_split[i]; ++e) l = std::max(l, lev_f[d->off_nb[e]] + 1);
    lev_f[i] = l;
    nlf = std::max(nlf, l + 1);
  }
  for (int k = n - 1; k >= 0; --k) {
    int i = order[k];
    int l = 0;
    for (int e = d->off_split[i]; e < d->off_ptr[i + 1]; ++e)
      l = std::max(l, lev_b[d->off_nb[e]] + 1);
    lev_b[i] = l;
    nlb = std::max(nlb, l + 1);
  }
  auto bucket = [n](const std::vector<int>& lev, int nl, std::vector<int>& ptr,
                    std::vector<int>& cells) {
    ptr.assign(nl + 1, 0);
    for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
    for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
    cells.assign(n, 0);
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i) cells[fill[lev[i]]++] = i;
  };
  std::vector<int> pf, cf, pb, cb;
  bucket(lev_f, nlf, pf, cf);
  bucket(lev_b, nlb, pb, cb);
  // per position: item count and the first kf / kb items (column cell, item)
  int kf = 1, kb = 1;
  for (int i = 0; i < n; ++i) {
    kf = std::max(kf, d->off_split[i] - d->off_ptr[i]);
    kb = std::max(kb, d->off_ptr[i + 1] - d->off_split[i]);
  }
  kf = std::min(kf, kLusgsK);
  kb = std::min(kb, kLusgsK);
  std::vector<int> cnt_f(n), cnt_b(n), nb_f((size_t)n * kf, -1), it_f((size_t)n * kf, -1),
      nb_b((size_t)n * kb, -1), it_b((size_t)n * kb, -1);
  for (int k = 0; k < n; ++k) {
    int i = cf[k];
    cnt_f[k] = d->off_split[i] - d->off_ptr[i];
    for (int q = 0; q < kf && q < cnt_f[k]; ++q) {
      nb_f[(size_t)k * kf + q] = d->off_nb[d->off_ptr[i] + q];
      it_f[(size_t)k * kf + q] = d->off_ptr[i] + q;
    }
    int j = cb[k];
    cnt_b[k] = d->off_ptr[j + 1] - d->off_split[j];
    for (int q = 0; q < kb && q < cnt_b[k]; ++q) {
      nb_b[(size_t)k * kb + q] = d->off_nb[d->off_split[j] + q];
      it_b[(size_t)k * kb + q] = d->off_split[j] + q;
    }
  }
  LusgsArrays& la = m->lusgs;
  la.kf = kf;
  la.kb = kb;
  m->host_lvl_f = pf;
  m->host_lvl_b = pb;
  m->lusgs_two = nlf == 2 && nlb == 2;
  for (int i = 0; i < n && m->lusgs_two; ++i) m->lusgs_two = lev_b[i] == 1 - lev_f[i];
  m->lusgs_path = d->launch_path - 1;  // -1 undecided
  la.n_rec = d->n_records;
  la.n_levels_f = nlf;
  la.n_levels_b = nlb;
  int rc = 0;
  size_t R = d->n_records, NO = (size_t)d->off_ptr[n], ND = (size_t)d->diag_ptr[n];
  rc |= upload_implicit(m, const_cast<int**>(&la.rec_o), d->rec_o, R);
  rc |= upload_implicit(m, const_cast<int**>(&la.rec_n), d->rec_n, R);
  rc |= upload_implicit(m, const_cast<double**>(&la.nox), d->rec_nox, R);
  rc |= upload_implicit(m, const_cast<double**>(&la.noy), d->rec_noy, R);
  rc |= upload_implicit(m, const_cast<double**>(&la.nnx), d->rec_nnx, R);
  rc |= upload_implicit(m, const_cast<double**>(&la.nny), d->rec_nny, R);
  rc |= upload_implicit(m, const_cast<double**>(&la.len), d->rec_len, R);
  rc |= upload_implicit(m, const_cast<int**>(&la.diag_ptr), d->diag_ptr, (size_t)n + 1);
  rc |= upload_implicit(m, const_cast<int**>(&la.diag_item), d->diag_item, ND);
  rc |= upload_implicit(m, const_cast<int**>(&la.off_ptr), d->off_ptr, (size_t)n + 1);
  rc |= upload_implicit(m, const_cast<int**>(&la.off_split), d->off_split, (size_t)n);
  rc |= upload_implicit(m, const_cast<int**>(&la.off_item), d->off_item, NO);
  rc |= upload_implicit(m, const_cast<int**>(&la.off_nb), d->off_nb, NO);
  rc |= upload_implicit(m, const_cast<int**>(&la.off_t), d->off_t, NO);
  rc |= upload_implicit(m, const_cast<int**>(&la.lvl_ptr_f), pf.data(), pf.size());
  rc |= upload_implicit(m, const_cast<int**>(&la.lvl_cells_f), cf.data(), cf.size());
  rc |= upload_implicit(m, const_cast<int**>(&la.lvl_ptr_b), pb.data(), pb.size());
  rc |= upload_implicit(m, const_cast<int**>(&la.lvl_cells_b), cb.data(), cb.size());
  rc |= upload_implicit(m, &la.coef, (const double*)nullptr, R);
  rc |= upload_implicit(m, &la.diag, (const double*)nullptr, (size_t)n);
  rc |= upload_implicit(m, &la.off_val, (const double*)nullptr, NO);
  rc |= upload_implicit(m, const_cast<int**>(&la.pcnt_f), cnt_f.data(), (size_t)n);
  rc |= upload_implicit(m, const_cast<int**>(&la.pcnt_b), cnt_b.data(), (size_t)n);
  rc |= upload_implicit(m, const_cast<int**>(&la.pnb_f), nb_f.data(), nb_f.size());
  rc |= upload_implicit(m, const_cast<int**>(&la.pnb_b), nb_b.data(), nb_b.size());
  rc |= upload_implicit(m, const_cast<int**>(&la.pit_f), it_f.data(), it_f.size());
  rc |= upload_implicit(m, const_cast<int**>(&la.pit_b), it_b.data(), it_b.size());
  rc |= upload_implicit(m, &la.pval_f, (const double*)nullptr, nb_f.size());
  rc |= upload_implicit(m, &la.pval_ft, (const double*)nullptr, nb_f.size());
  rc |= upload_implicit(m, &la.pval_b, (const double*)nullptr, nb_b.size());
  rc |= upload_implicit(m, &la.pval_bt, (const double*)nullptr, nb_b.size());
  rc |= upload_implicit(m, &la.pdiag_f, (const double*)nullptr, (size_t)n);
  rc |= upload_implicit(m, &la.pdiag_b, (const double*)nullptr, (size_t)n);
  {
    std::vector<int> posb(n), fbv(n);
    for (int k = 0; k < n; ++k) posb[cb[k]] = k;
    for (int k = 0; k < n; ++k) fbv[k] = posb[cf[k]];
    rc |= upload_implicit(m, const_cast<int**>(&la.fb), fbv.data(), (size_t)n);
  }
  rc |= upload_implicit(m, &la.own_f, (const double*)nullptr, 4 * (size_t)n);
  rc |= upload_implicit(m, &la.own_b, (const double*)nullptr, 4 * (size_t)n);
  if (rc) {
    std::string msg = g_last_error;
    free_implicit(m);
    return fail(msg);
  }
  m->implicit = true;
  return 0;
}

int mcfd_lusgs_assemble(McfdMesh* m, const McfdSettings* st, const double* u, const double* dt,
                        int32_t dt_per_cell, void* stream) {
  if (!m || !u || !dt) return fail("mcfd_lusgs_assemble: NULL argument");
  if (!m->implicit) return fail("mcfd_lusgs_assemble: call mcfd_mesh_set_implicit first");
  if (check_settings(st)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int n = m->n_cells;
  hipLaunchKernelGGL(k_primitive, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, m->W, st->gamma,
                     m->status);
  hipLaunchKernelGGL(k_lusgs_coef, dim3(blocks(m->lusgs.n_rec)), dim3(kBlock), 0, s, m->lusgs,
                     m->W, st->gamma);
  hipLaunchKernelGGL(k_lusgs_assemble, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->lusgs, m->area,
                     dt, dt_per_cell);
  hipLaunchKernelGGL(k_lusgs_pack, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->lusgs);
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_lusgs_apply(McfdMesh* m, const double* v, double* out, void* stream) {
  if (!m || !v || !out) return fail("mcfd_lusgs_apply: NULL argument");
  if (!m->implicit) return fail("mcfd_lusgs_apply: call mcfd_mesh_set_implicit first");
  return lusgs_apply(m, v, out, static_cast<hipStream_t>(stream));
}

int mcfd_jacobian_apply(McfdMesh* m, const McfdSettings* jst, McfdGmres* g, const double* v,
                        double* out, void* stream) {
  if (!m || !g || !v || !out) return fail("mcfd_jacobian_apply: NULL argument");
  if (check_settings(jst)) return -1;
  if (ensure_gmres(m, 1)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (resolve_state_norm(m, g, s)) return -1;
  JacobianOp op{m, jst, g->state, g->rate, g->dt, g->dt_per_cell,
                {g->sigma[0], g->sigma[1], g->sigma[2], g->sigma[3]}, g->state_norm};
  return jacobian_apply(op, v, out, s);
}

int mcfd_gmres(McfdMesh* m, const McfdSettings* jst, McfdGmres* g, void* stream) {
  if (!m || !g || !g->state || !g->rate || !g->rhs || !g->x || !g->dt)
    return fail("mcfd_gmres: NULL argument");
  if (!m->implicit) return fail("mcfd_gmres: call mcfd_mesh_set_implicit first");
  if (check_settings(jst)) return -1;
  if (g->krylov <= 0 || g->max_iterations <= 0) return fail("mcfd_gmres: bad iteration limits");
  if (ensure_gmres(m, g->krylov)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  long N = 4L * m->n_cells;
  if (resolve_state_norm(m, g, s)) return -1;
  HIP_CHECK(hipMemsetAsync(g->x, 0, sizeof(double) * N, s));
  JacobianOp op{m, jst, g->state, g->rate, g->dt, g->dt_per_cell,
                {g->sigma[0], g->sigma[1], g->sigma[2], g->sigma[3]}, g->state_norm};
  int used = 0;
  bool nonzero = false;
  g->info = 1;
  g->iterations = 0;
  while (used < g->max_iterations) {
    int restart = std::min(g->krylov, g->max_iterations - used);
    int inner = 0, info = 1;
    double rnorm = 0.0;
    if (gmres_cycle(op, g->rhs, g->x, g->tolerance, restart, nonzero, inner, info, rnorm, s))
      return -1;
    nonzero = true;
    used += inner;
    g->iterations = used;
    g->residual = rnorm;
    if (info == 0) {
      g->info = 0;
      return 0;
    }
    if (inner == 0) break;
  }
  g->info = 1;
  return 0;
}

int mcfd_newton_equation(McfdMesh* m, const double* u, const double* u0, const double* r,
                         const double* dt, int32_t dt_per_cell, const double* sigma,
                         double* scaled, double* norms, void* stream) {
  if (!m || !u || !u0 || !r || !dt || !sigma || !norms)
    return fail("mcfd_newton_equation: NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int nb = blocks(m->n_cells);
  hipLaunchKernelGGL(k_newton_equation, dim3(nb), dim3(kBlock), 0, s, m->n_cells, u, u0, r, dt,
                     dt_per_cell, sigma[0], sigma[1], sigma[2], sigma[3], scaled, m->partial);
  hipLaunchKernelGGL(k_max2_final, dim3(1), dim3(1024), 0, s, nb, m->partial, norms);
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_newton_candidate(McfdMesh* m, const double* u, const double* corr, double damping,
                          const double* sigma, double* out, void* stream) {
  if (!m || !u || !corr || !sigma || !out) return fail("mcfd_newton_candidate: NULL argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_candidate, dim3(blocks(m->n_cells)), dim3(kBlock), 0, s, m->n_cells, u,
                     corr, damping, sigma[0], sigma[1], sigma[2], sigma[3], out);
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_norm2(McfdMesh* m, const double* a, double* out, void* stream) {
  if (!m || !a || !out) return fail("mcfd_norm2: NULL argument");
  if (ensure_gmres(m, 1)) return -1;
  return dot_dev(m, 4L * m->n_cells, a, a, out, true, static_cast<hipStream_t>(stream));
}


int mcfd_mesh_set_adjoint(McfdMesh* m, const int32_t* grad_rev) {
  if (!m) return fail("mcfd_mesh_set_adjoint: NULL mesh");
  (void)hipDeviceSynchronize();
  free_adjoint(m);
  if (!grad_rev) return 0;
  if (m->max_cell_entries >= 255)
    return fail("mcfd_mesh_set_adjoint: a cell has 255 or more LSQ or limiter entries");
  size_t C = m->n_cells;
  int rc = 0;
  AdjArrays& a = m->adj;
  rc |= alloc_adjoint(m, &m->grad_rev, grad_rev, (size_t)m->n_grad);
  rc |= alloc_adjoint(m, &m->FADJ, (const double*)nullptr, 8 * (size_t)m->n_faces);
  rc |= alloc_adjoint(m, &a.wbar, (const double*)nullptr, 4 * C);
  rc |= alloc_adjoint(m, &a.ebar, (const double*)nullptr, 4 * (size_t)m->n_grad);
  if (m->viscous) {
    rc |= alloc_adjoint(m, &m->VFA, (const double*)nullptr, 24 * (size_t)m->n_faces);
    rc |= alloc_adjoint(m, &m->VBAR, (const double*)nullptr, 3 * (size_t)m->n_grad);
  }
  rc |= alloc_adjoint(m, &a.arg, (const unsigned*)nullptr, 4 * C);
  if (rc) {
    std::string msg = g_last_error;
    free_adjoint(m);
    return fail(msg);
  }
  a.acc_ptr = m->acc_ptr;
  a.acc_item = m->acc_item;
  a.flx = m->faces.lx;
  a.fly = m->faces.ly;
  a.frx = m->faces.rx;
  a.fry = m->faces.ry;
  a.grad_ptr = m->grad_ptr;
  a.grad_nb = m->grad_nb;
  a.grad_cx = m->grad_cx;
  a.grad_cy = m->grad_cy;
  a.grad_rev = m->grad_rev;
  a.lim_ptr = m->lim_ptr;
  a.lim_dx = m->lim_dx;
  a.lim_dy = m->lim_dy;
  m->adjoint = true;
  return 0;
}


int mcfd_mesh_set_parabolic(McfdMesh* m, int32_t n_entries, const int32_t* ptr, const int32_t* nb,
                            const double* coef, const int32_t* full) {
  if (!m) return fail("mcfd_mesh_set_parabolic: NULL mesh");
  (void)hipDeviceSynchronize();
  free_parabolic(m);
  if (!ptr) return 0;
  size_t C = m->n_cells;
  int rc = 0;
  rc |= alloc_parabolic(m, &m->par_ptr, ptr, C + 1);
  rc |= alloc_parabolic(m, &m->par_nb, nb, (size_t)n_entries);
  rc |= alloc_parabolic(m, &m->par_coef, coef, 5 * (size_t)n_entries);
  rc |= alloc_parabolic(m, &m->par_full, full, C);
  rc |= alloc_parabolic(m, &m->K, (const double*)nullptr, 12 * C);
  if (rc) {
    std::string msg = g_last_error;
    free_parabolic(m);
    return fail(msg);
  }
  m->n_par = n_entries;
  m->parabolic = true;
  return 0;
}

int mcfd_mesh_set_viscous(McfdMesh* m, const McfdViscousDesc* d) {
  if (!m) return fail("mcfd_mesh_set_viscous: NULL mesh");
  (void)hipDeviceSynchronize();
  free_viscous(m);
  if (!d) return 0;
  if (d->n_ghost < 0 || d->n_vacc < 0 || d->n_vdt < 0)
    return fail("mcfd_mesh_set_viscous: negative count");
  size_t C = m->n_cells;
  size_t B = m->n_bc;
  ViscArrays& v = m->visc;
  int rc = 0;
  rc |= alloc_viscous(m, &v.wl_wx, d->wl_wx, (size_t)m->n_grad);
  rc |= alloc_viscous(m, &v.wl_wy, d->wl_wy, (size_t)m->n_grad);
  rc |= alloc_viscous(m, &v.wl_cell, d->wl_cell, 4 * C);
  rc |= alloc_viscous(m, &v.gh_ptr, d->gh_ptr, C + 1);
  rc |= alloc_viscous(m, &v.gh_face, d->gh_face, (size_t)d->n_ghost);
  rc |= alloc_viscous(m, &v.gh_wx, d->gh_wx, (size_t)d->n_ghost);
  rc |= alloc_viscous(m, &v.gh_wy, d->gh_wy, (size_t)d->n_ghost);
  rc |= alloc_viscous(m, &v.bv_code, d->bv_code, B);
  rc |= alloc_viscous(m, &v.bv_flags, d->bv_flags, B);
  rc |= alloc_viscous(m, &v.bv_geom, d->bv_geom, 4 * B);
  rc |= alloc_viscous(m, &m->vacc.ptr, d->va_ptr, C + 1);
  rc |= alloc_viscous(m, &m->vacc.item, d->va_item, (size_t)d->n_vacc);
  rc |= alloc_viscous(m, &m->vacc.coef, d->va_coef, (size_t)d->n_vacc);
  m->n_vacc_visc = d->n_vacc;
  rc |= alloc_viscous(m, &m->vdt.ptr, d->vd_ptr, C + 1);
  rc |= alloc_viscous(m, &m->vdt.nb, d->vd_nb, (size_t)d->n_vdt);
  rc |= alloc_viscous(m, &m->vdt.len, d->vd_len, (size_t)d->n_vdt);
  rc |= alloc_viscous(m, &m->vdt.dist, d->vd_dist, (size_t)d->n_vdt);
  rc |= alloc_viscous_scratch(m, &m->VG, 8 * C);
  rc |= alloc_viscous_scratch(m, &m->VF, 4 * (size_t)m->n_faces);
  rc |= alloc_viscous_scratch(m, &m->EFF, C);
  if (rc) {
    std::string msg = g_last_error;
    free_viscous(m);
    return fail(msg);
  }
  m->vacc.VF = m->VF;
  m->vdt.eff = m->EFF;
  m->viscous = true;
  return 0;
}

int mcfd_mesh_set_inflow(McfdMesh* m, const McfdInflowDesc* d) {
  if (!m) return fail("mcfd_mesh_set_inflow: NULL mesh");
  (void)hipDeviceSynchronize();
  free_inflow(m);
  if (!d) return 0;
  const int n = d->n_records, first = d->first_record, np = d->n_patches;
  if (n < 0 || np < 0 || d->n_leaves < 0 || d->n_program < 0 || first < 0 || first + n > m->n_bc)
    return fail("mcfd_mesh_set_inflow: record range outside the boundary records");
  if (n > 0 && (!d->face || !d->params || !d->projection))
    return fail("mcfd_mesh_set_inflow: missing array");
  if (np > 0 && (!d->patch_rec || !d->leaf_ptr || !d->leaf || !d->prog_ptr || !d->program))
    return fail("mcfd_mesh_set_inflow: missing patch array");
  // the records are exactly the boundary records of kinds 7-9, each on its own flux face
  int count = 0;
  for (int b = 0; b < m->n_bc; ++b) count += m->host_bc_kind[b] >= MCFD_FACE_PRESSURE_INLET;
  if (count != n) return fail("mcfd_mesh_set_inflow: records do not match the inflow boundary records");
  for (int r = 0; r < n; ++r) {
    int k = d->face[r];
    if (k < 0 || k >= m->n_faces || m->host_face_bc[k] != first + r ||
        m->host_bc_kind[first + r] < MCFD_FACE_PRESSURE_INLET)
      return fail("mcfd_mesh_set_inflow: record face does not carry the record");
  }
  // patches: contiguous mass-flow-inlet records covering all of them; leaves
  // inside the patch; programs that leave one value on a stack of <= 64
  int max_n = 0, max_leaves = 0, covered = 0;
  for (int p = 0; p < np; ++p) {
    int a = d->patch_rec[p], b = d->patch_rec[p + 1];
    if (a < 0 || b <= a || b > n)
      return fail("mcfd_mesh_set_inflow: bad patch record range");
    if (b - a > 4096) return fail("Mass-flow inlet patches of more than 4096 faces are not supported on the GPU.");
    for (int r = a; r < b; ++r)
      if (m->host_bc_kind[first + r] != MCFD_FACE_MASS_FLOW_INLET)
        return fail("mcfd_mesh_set_inflow: patch record is not a mass-flow inlet");
    covered += b - a;
    int l0 = d->leaf_ptr[p], l1 = d->leaf_ptr[p + 1];
    if (l0 < 0 || l1 <= l0 || l1 > d->n_leaves || l1 - l0 > 128)
      return fail("mcfd_mesh_set_inflow: bad leaf range");
    for (int l = l0; l < l1; ++l) {
      int st = d->leaf[2 * l], sz = d->leaf[2 * l + 1];
      if (st < 0 || sz <= 0 || st + sz > b - a) return fail("mcfd_mesh_set_inflow: leaf outside its patch");
    }
    int q0 = d->prog_ptr[p], q1 = d->prog_ptr[p + 1];
    if (q0 < 0 || q1 <= q0 || q1 > d->n_program) return fail("mcfd_mesh_set_inflow: bad program range");
    int depth = 0;
    for (int q = q0; q < q1; ++q) {
      int tok = d->program[q];
      if (tok >= 0) {
        if (tok >= l1 - l0 || ++depth > 64) return fail("mcfd_mesh_set_inflow: bad summation program");
      } else {
        if (tok != -1 || depth < 2) return fail("mcfd_mesh_set_inflow: bad summation program");
        --depth;
      }
    }
    if (depth != 1) return fail("mcfd_mesh_set_inflow: bad summation program");
    max_n = std::max(max_n, b - a);
    max_leaves = std::max(max_leaves, l1 - l0);
  }
  int n_mass = 0;
  for (int r = 0; r < n; ++r) n_mass += m->host_bc_kind[first + r] == MCFD_FACE_MASS_FLOW_INLET;
  if (covered != n_mass) return fail("mcfd_mesh_set_inflow: patches do not cover the mass-flow inlets");
  InflowArrays& ia = m->inf;
  int rc = 0;
  int* face = nullptr;
  double *params = nullptr, *proj = nullptr;
  int *patch_rec = nullptr, *leaf_ptr = nullptr, *leaf = nullptr, *prog_ptr = nullptr, *program = nullptr;
  rc |= alloc_inflow(m, &face, d->face, (size_t)n);
  rc |= alloc_inflow(m, &params, d->params, 4 * (size_t)n);
  rc |= alloc_inflow(m, &proj, d->projection, (size_t)n);
  rc |= alloc_inflow(m, &patch_rec, np ? d->patch_rec : nullptr, (size_t)np + 1);
  rc |= alloc_inflow(m, &leaf_ptr, np ? d->leaf_ptr : nullptr, (size_t)np + 1);
  rc |= alloc_inflow(m, &leaf, d->leaf, 2 * (size_t)d->n_leaves);
  rc |= alloc_inflow(m, &prog_ptr, np ? d->prog_ptr : nullptr, (size_t)np + 1);
  rc |= alloc_inflow(m, &program, d->program, (size_t)d->n_program);
  rc |= alloc_inflow(m, &ia.ext, (const double*)nullptr, 4 * (size_t)n);
  rc |= alloc_inflow(m, &m->inf_ext_v, (const double*)nullptr, 4 * (size_t)n);
  rc |= alloc_inflow(m, &m->inf_vbar, (const double*)nullptr, 4 * (size_t)n);
  if (rc) {
    std::string msg = g_last_error;
    free_inflow(m);
    return fail(msg);
  }
  ia.first = first;
  ia.n = n;
  ia.face = face;
  ia.params = params;
  ia.proj = proj;
  ia.patch_rec = patch_rec;
  ia.leaf_ptr = leaf_ptr;
  ia.leaf = leaf;
  ia.prog_ptr = prog_ptr;
  ia.program = program;
  m->inflow_patches = np;
  m->inflow_max_n = max_n;
  m->inflow_max_leaves = max_leaves;
  m->faces.inflow = ia.ext;
  m->faces.inflow_first = first;
  m->dts.inflow = ia.ext;
  m->dts.inflow_first = first;
  m->inflow = true;
  return 0;
}

int mcfd_residual_geometry_vjp(McfdMesh* m, const McfdSettings* st, const double* u, const double* lam,
                               double* face_bar, double* grad_bar, double* lim_bar, void* stream) {
  if (!m || !u || !lam || !face_bar || !grad_bar || !lim_bar)
    return fail("mcfd_residual_geometry_vjp: NULL argument");
  if (!m->adjoint) return fail("mcfd_residual_geometry_vjp: call mcfd_mesh_set_adjoint first");
  if (check_settings(st)) return -1;
  if (st->flux == 2 || st->reconstruction == 2 || st->viscosity || m->needs_inflow)
    return fail("Geometry gradients cover the inviscid Rusanov and HLLC residual, first order or MUSCL "
                "linear, without pressure-inlet or mass-flow boundaries.");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int n = m->n_cells;
  bool muscl = st->reconstruction == 1 && m->n_grad > 0;
  hipLaunchKernelGGL(k_primitive, dim3(blocks(n)), dim3(kBlock), 0, s, n, u, m->W, st->gamma, m->status);
  if (muscl)
    hipLaunchKernelGGL(k_gradient_t<true>, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->grad_ptr,
                       m->grad_nb, m->grad_cx, m->grad_cy, m->lim_ptr, m->lim_dx, m->lim_dy,
                       st->limiter, m->G, m->PHI, m->adj.arg);
  if (m->n_faces) {
    hipLaunchKernelGGL(k_flux_adjoint_rev, dim3(blocks(m->n_faces)), dim3(kBlock), 0, s, m->n_faces,
                       m->faces, m->W, m->G, m->PHI, nullptr, m->face_len, m->area, lam, st->flux,
                       muscl ? 1 : 0, st->gamma, m->FADJ, nullptr);
    hipLaunchKernelGGL(k_geom_face, dim3(blocks(m->n_faces)), dim3(kBlock), 0, s, m->n_faces, m->faces,
                       m->W, m->G, m->PHI, m->face_len, m->area, lam, st->flux, muscl ? 1 : 0,
                       st->gamma, face_bar);
  }
  if (muscl) {
    AdjArrays aa = m->adj;
    aa.cbar = grad_bar;
    aa.dlbar = lim_bar;
    aa.par_rev = nullptr;
    aa.ebar_par = nullptr;
    hipLaunchKernelGGL(k_cell_adjoint1, dim3(blocks(n)), dim3(kBlock), 0, s, n, m->W, m->G, m->PHI,
                       m->FADJ, aa, 1, st->limiter);
  } else {
    long ng = 2L * m->n_grad, nl = 2L * m->n_lim;
    if (ng) hipLaunchKernelGGL(k_fill, dim3(lblocks(ng)), dim3(kBlock), 0, s, ng, 0.0, grad_bar);
    if (nl) hipLaunchKernelGGL(k_fill, dim3(lblocks(nl)), dim3(kBlock), 0, s, nl, 0.0, lim_bar);
  }
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_mesh_set_parabolic_adjoint(McfdMesh* m, const int32_t* par_rev) {
  if (!m) return fail("mcfd_mesh_set_parabolic_adjoint: NULL mesh");
  if (!m->adjoint || !m->parabolic)
    return fail("mcfd_mesh_set_parabolic_adjoint: call mcfd_mesh_set_parabolic and mcfd_mesh_set_adjoint first");
  if (!par_rev) return fail("mcfd_mesh_set_parabolic_adjoint: NULL reverse map");
  for (int e = 0; e < m->n_par; ++e)
    if (par_rev[e] < 0 || par_rev[e] >= m->n_par)
      return fail("mcfd_mesh_set_parabolic_adjoint: reverse entry out of range");
  (void)hipDeviceSynchronize();
  int* rev = nullptr;
  double* eb = nullptr;
  int rc = 0;
  rc |= alloc_adjoint(m, &rev, par_rev, (size_t)m->n_par);
  rc |= alloc_adjoint(m, &eb, (const double*)nullptr, 4 * (size_t)m->n_par);
  if (rc) return -1;
  m->adj.par_rev = rev;
  m->adj.ebar_par = eb;
  return 0;
}

int mcfd_residual_vjp(McfdMesh* m, const McfdSettings* st, const double* u, const double* lam,
                      double* ubar, double* sbar, void* stream) {
  return mcfd_residual_vjp_impl(m, st, u, lam, ubar, sbar, static_cast<hipStream_t>(stream));
}

int mcfd_lusgs_apply_transpose(McfdMesh* m, const double* v, double* out, void* stream) {
  if (!m || !v || !out) return fail("mcfd_lusgs_apply_transpose: NULL argument");
  if (!m->implicit) return fail("mcfd_lusgs_apply_transpose: call mcfd_mesh_set_implicit first");
  return lusgs_apply(m, v, out, static_cast<hipStream_t>(stream), true);
}

int mcfd_adjoint_apply(McfdMesh* m, const McfdSettings* st, McfdGmres* g, const double* v, double* out,
                       void* stream) {
  if (!m || !g || !g->state || !g->dt || !v || !out) return fail("mcfd_adjoint_apply: NULL argument");
  if (!m->adjoint) return fail("mcfd_adjoint_apply: call mcfd_mesh_set_adjoint first");
  if (check_settings(st)) return -1;
  if (ensure_gmres(m, 1)) return -1;
  JacobianOp op{m, st, g->state, nullptr, g->dt, g->dt_per_cell, {1.0, 1.0, 1.0, 1.0}, 0.0};
  op.adjoint = true;
  return adjoint_apply(op, v, out, static_cast<hipStream_t>(stream));
}

int mcfd_adjoint_gmres(McfdMesh* m, const McfdSettings* st, McfdGmres* g, void* stream) {
  if (!m || !g || !g->state || !g->rhs || !g->x || !g->dt)
    return fail("mcfd_adjoint_gmres: NULL argument");
  if (!m->implicit) return fail("mcfd_adjoint_gmres: call mcfd_mesh_set_implicit first");
  if (!m->adjoint) return fail("mcfd_adjoint_gmres: call mcfd_mesh_set_adjoint first");
  if (check_settings(st)) return -1;
  if (g->krylov <= 0 || g->max_iterations <= 0) return fail("mcfd_adjoint_gmres: bad iteration limits");
  if (ensure_gmres(m, g->krylov)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  long N = 4L * m->n_cells;
  HIP_CHECK(hipMemsetAsync(g->x, 0, sizeof(double) * N, s));
  JacobianOp op{m, st, g->state, nullptr, g->dt, g->dt_per_cell, {1.0, 1.0, 1.0, 1.0}, 0.0};
  op.adjoint = true;
  int used = 0;
  bool nonzero = false;
  g->info = 1;
  g->iterations = 0;
  while (used < g->max_iterations) {
    int restart = std::min(g->krylov, g->max_iterations - used);
    int inner = 0, info = 1;
    double rnorm = 0.0;
    if (gmres_cycle(op, g->rhs, g->x, g->tolerance, restart, nonzero, inner, info, rnorm, s))
      return -1;
    nonzero = true;
    used += inner;
    g->iterations = used;
    g->residual = rnorm;
    if (info == 0) {
      g->info = 0;
      return 0;
    }
    if (inner == 0) break;
  }
  g->info = 1;
  return 0;
}


int mcfd_batch_time_steps(McfdMesh* m, const McfdSettings* st, const double* u, int32_t n_cases,
                          const int32_t* case_ptr, double end_time, double* clock, double* dt_cell,
                          void* stream) {
  if (!m || !u || !case_ptr || !clock || !dt_cell || n_cases <= 0)
    return fail("mcfd_batch_time_steps: NULL argument or no cases");
  if (mcfd_time_steps(m, st, u, dt_cell, nullptr, stream)) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  {
    TimedLaunch t(m, s, MCFD_KERNEL_REDUCE);
    hipLaunchKernelGGL(k_batch_dt, dim3(n_cases), dim3(1024), 0, s, n_cases, case_ptr, end_time,
                       clock, dt_cell, m->status);
  }
  HIP_CHECK(hipGetLastError());
  return 0;
}

int mcfd_batch_case_steps(McfdMesh* m, int32_t n_cases, const int32_t* case_ptr, double end_time,
                          double* clock, double* dt_cell, void* stream) {
  if (!m || !case_ptr || !clock || !dt_cell || n_cases <= 0)
    return fail("mcfd_batch_case_steps: NULL argument or no cases");
  hipStream_t s = static_cast<hipStream_t>(stream);
  {
    TimedLaunch t(m, s, MCFD_KERNEL_REDUCE);
    hipLaunchKernelGGL(k_batch_dt, dim3(n_cases), dim3(1024), 0, s, n_cases, case_ptr, end_time,
                       clock, dt_cell, m->status);
  }
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
