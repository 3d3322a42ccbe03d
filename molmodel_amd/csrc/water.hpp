// water.hpp -- rigid-water constraint solver shared by the integrator and
// the replica-exchange velocity moves: two O-H length constraints
// (value 0.5 (r^2 - l^2), constraint.cpp:30-66) and one H-O-H angle
// constraint (constraint.cpp:68-112, geograd.cpp:59-103), solved with the
// reference's Newton iteration (constraint.cpp:295-358) in FP64.
#pragma once
#include <hip/hip_runtime.h>

namespace mmamd {

struct WaterC {
  double val[3];
  double g[3][9];
};

// gradient entries that can be nonzero: the O-H1 length reaches O and H1
// (coordinates 0-5), the O-H2 length O and H2 (0-2, 6-8), the angle all
// nine.  Products with the other entries are exact zeros, and adding a zero
// leaves a sum unchanged, so the loops below skip them with results equal
// to the full sums (for finite inputs).
__host__ __device__ constexpr bool water_reach(int a, int d) {
  return a == 2 || (a == 0 ? d < 6 : (d < 3 || d >= 6));
}

// constraint values and gradients (constraint.cpp:30-112, geograd.cpp:59-103)
__device__ inline void water_constraints(const double *x, double bond, double t0, WaterC &c) {
  for (int a = 0; a < 3; a++)
    for (int d = 0; d < 9; d++) c.g[a][d] = 0;
  double r01[3], r02[3];
  for (int d = 0; d < 3; d++) { r01[d] = x[d] - x[3 + d]; r02[d] = x[d] - x[6 + d]; }
  c.val[0] = 0.5 * ((r01[0] * r01[0] + r01[1] * r01[1] + r01[2] * r01[2]) - bond * bond);
  c.val[1] = 0.5 * ((r02[0] * r02[0] + r02[1] * r02[1] + r02[2] * r02[2]) - bond * bond);
  for (int d = 0; d < 3; d++) {
    c.g[0][d] = r01[d]; c.g[0][3 + d] = -r01[d];
    c.g[1][d] = r02[d]; c.g[1][6 + d] = -r02[d];
  }
  // angle H1-O-H2: a = H1 - O, b = H2 - O
  double a[3], b[3];
  for (int d = 0; d < 3; d++) { a[d] = x[3 + d] - x[d]; b[d] = x[6 + d] - x[d]; }
  const double a2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  const double b2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  const double ma2 = 1.0 / a2, mb2 = 1.0 / b2, ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  const double ma = sqrt(ma2), mb = sqrt(mb2);
  const double t0a = ma * mb, t3 = -ma * ma2 * mb * ab, t4 = -ab * ma * mb * mb2;
  double ct = ab * t0a;
  ct = ct > 1 ? 1 : (ct < -1 ? -1 : ct);
  const double s2t = 1 - ct * ct;
  double dt, t;
  if (s2t < 64.0 * 2.220446049250313e-16) { dt = 0; t = ct > 0 ? 0 : M_PI; }
  else { dt = -1 / sqrt(s2t); t = acos(ct); }
  for (int d = 0; d < 3; d++) {
    const double g1 = (b[d] * t0a + a[d] * t3) * dt;
    const double g3 = (a[d] * t0a + b[d] * t4) * dt;
    c.g[2][3 + d] = g1;
    c.g[2][6 + d] = g3;
    c.g[2][d] = -g1 - g3;
  }
  c.val[2] = t - t0;
}

// 3x3 Gaussian elimination with partial pivoting
__device__ inline bool solve3(double s[3][3], double *x) {
  for (int c = 0; c < 3; c++) {
    int piv = c;
    for (int r = c + 1; r < 3; r++)
      if (fabs(s[r][c]) > fabs(s[piv][c])) piv = r;
    if (s[piv][c] == 0) return false;
    if (piv != c) {
      for (int k = 0; k < 3; k++) { const double t = s[c][k]; s[c][k] = s[piv][k]; s[piv][k] = t; }
      const double t = x[c]; x[c] = x[piv]; x[piv] = t;
    }
    for (int r = c + 1; r < 3; r++) {
      const double m = s[r][c] / s[c][c];
      for (int k = c; k < 3; k++) s[r][k] -= m * s[c][k];
      x[r] -= m * x[c];
    }
  }
  for (int c = 2; c >= 0; c--) {
    double t = x[c];
    for (int k = c + 1; k < 3; k++) t -= s[c][k] * x[k];
    x[c] = t / s[c][c];
  }
  return true;
}

// constraint.cpp:295-335; returns false if not converged
__device__ inline bool shake_water(double *x, double *v, const double *x0, const double *m, double dt,
                            double bond, double t0, double tol, int maxit) {
  WaterC c0;
  water_constraints(x0, bond, t0, c0);
  double sci0[3][9];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int d = 0; d < 9; d++) sci0[a][d] = water_reach(a, d) ? c0.g[a][d] / m[d / 3] : 0.0;
  const double invdt = dt > 0 ? 1 / dt : 0;
  for (int it = 0; it < maxit; it++) {
    WaterC c;
    water_constraints(x, bond, t0, c);
    double xs[3];
    bool sat = true;
    for (int a = 0; a < 3; a++) { xs[a] = -c.val[a]; if (fabs(xs[a]) > tol) sat = false; }
    if (sat) return true;
    double s[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        double t = 0;
#pragma unroll
        for (int d = 0; d < 9; d++)
          if (water_reach(a, d) && water_reach(b, d)) t += c.g[a][d] * sci0[b][d];
        s[a][b] = t;
      }
    if (!solve3(s, xs)) return false;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int d = 0; d < 9; d++) {
        if (!water_reach(a, d)) continue;
        const double dx = xs[a] * sci0[a][d];
        x[d] += dx;
        v[d] += dx * invdt;
      }
  }
  return false;
}

// constraint.cpp:336-358
__device__ inline void rattle_water(const double *x, double *v, const double *m, double bond, double t0) {
  WaterC c;
  water_constraints(x, bond, t0, c);
  double sci0[3][9], xs[3], s[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    xs[a] = 0;
#pragma unroll
    for (int d = 0; d < 9; d++) {
      sci0[a][d] = 0.0;
      if (!water_reach(a, d)) continue;
      sci0[a][d] = c.g[a][d] / m[d / 3];
      xs[a] -= c.g[a][d] * v[d];
    }
  }
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = a; b < 3; b++) {
      double t = 0;
#pragma unroll
      for (int d = 0; d < 9; d++)
        if (water_reach(a, d) && water_reach(b, d)) t += c.g[a][d] * sci0[b][d];
      s[a][b] = s[b][a] = t;
    }
  if (!solve3(s, xs)) return;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int d = 0; d < 9; d++)
      if (water_reach(a, d)) v[d] += xs[a] * sci0[a][d];
}

}  // namespace mmamd
