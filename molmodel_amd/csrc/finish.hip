// finish.hip -- per-group completion of a force evaluation and the rigid
// water velocity-Verlet step.
//
// Per sorted group (one thread):
//   * real-space pair results + k-space/P3M field: F = q E + F_LJ, phi
//     (ElecLJ::add_to_energy_and_forces, eleclj.cpp:794-805);
//   * excluded intramolecular pairs: -erf(eta r)/r with minimum image
//     (rspace.c:492-521), FP64;
//   * self term -2 eta/sqrt(pi) q and neutralising background
//     (eleclj.cpp:961-976), FP64;
//   * energies 1/2 q.phi and LJ, the virial, and the largest / second
//     largest first-site displacement since the last neighbour-list build
//     (nbrlist.cpp:64-76), reduced per system in block order (deterministic);
//   * optional dynamics (mdyn.cpp:142-160, msys.cpp:1275-1285):
//       KICK : v += F dt/2m, velocity constraints (constraint.cpp:336-358)
//       DRIFT: v += F dt/2m, r += v dt, position constraints with the
//              reference's Newton iteration (constraint.cpp:295-335)
//     for waters (O,H,H) with two O-H length constraints and one H-O-H
//     angle constraint, as the reference sets them up; single-site groups
//     move freely.
#include "engine.hpp"

namespace mmamd {

constexpr int FIN_TB = 256;
constexpr int NPART = 14;   // elec, lj, kin, vir[9], top1, top2

struct WaterC {
  double val[3];
  double g[3][9];
};

// constraint values and gradients (constraint.cpp:30-112, geograd.cpp:59-103)
__device__ void water_constraints(const double *x, double bond, double t0, WaterC &c) {
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
__device__ bool solve3(double s[3][3], double *x) {
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
__device__ bool shake_water(double *x, double *v, const double *x0, const double *m, double dt,
                            double bond, double t0, double tol, int maxit) {
  WaterC c0;
  water_constraints(x0, bond, t0, c0);
  double sci0[3][9];
  for (int a = 0; a < 3; a++)
    for (int d = 0; d < 9; d++) sci0[a][d] = c0.g[a][d] / m[d / 3];
  const double invdt = dt > 0 ? 1 / dt : 0;
  for (int it = 0; it < maxit; it++) {
    WaterC c;
    water_constraints(x, bond, t0, c);
    double xs[3];
    bool sat = true;
    for (int a = 0; a < 3; a++) { xs[a] = -c.val[a]; if (fabs(xs[a]) > tol) sat = false; }
    if (sat) return true;
    double s[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double t = 0;
        for (int d = 0; d < 9; d++) t += c.g[a][d] * sci0[b][d];
        s[a][b] = t;
      }
    if (!solve3(s, xs)) return false;
    for (int a = 0; a < 3; a++)
      for (int d = 0; d < 9; d++) {
        const double dx = xs[a] * sci0[a][d];
        x[d] += dx;
        v[d] += dx * invdt;
      }
  }
  return false;
}

// constraint.cpp:336-358
__device__ void rattle_water(const double *x, double *v, const double *m, double bond, double t0) {
  WaterC c;
  water_constraints(x, bond, t0, c);
  double sci0[3][9], xs[3], s[3][3];
  for (int a = 0; a < 3; a++) {
    xs[a] = 0;
    for (int d = 0; d < 9; d++) { sci0[a][d] = c.g[a][d] / m[d / 3]; xs[a] -= c.g[a][d] * v[d]; }
  }
  for (int a = 0; a < 3; a++)
    for (int b = a; b < 3; b++) {
      double t = 0;
      for (int d = 0; d < 9; d++) t += c.g[a][d] * sci0[b][d];
      s[a][b] = s[b][a] = t;
    }
  if (!solve3(s, xs)) return;
  for (int a = 0; a < 3; a++)
    for (int d = 0; d < 9; d++) v[d] += xs[a] * sci0[a][d];
}

__device__ __forceinline__ void top2_merge(double &a1, double &a2, double b1, double b2) {
  if (b1 > a1) { a2 = fmax(a1, b2); a1 = b1; }
  else if (b1 > a2) a2 = b1;
}

__global__ void __launch_bounds__(FIN_TB) k_finish(
    const SysParams *__restrict__ sp, const int *__restrict__ order, const int *__restrict__ gfirst,
    const int *__restrict__ gns, const int *__restrict__ gexcl, const double *__restrict__ q,
    const double *__restrict__ mass, const double *__restrict__ pf, const double *__restrict__ gacc,
    const double *__restrict__ kres, double *__restrict__ r, double *__restrict__ v,
    double *__restrict__ f, double *__restrict__ phi_out, const double *__restrict__ last,
    int N, int mode, int want_virial, double bond, double angle_rad, double tol,
    double *__restrict__ part, int *__restrict__ done, int maxblk, double *__restrict__ acc,
    int *__restrict__ rebuild, int *__restrict__ err) {
  __shared__ double red[FIN_TB / 64][NPART];
  __shared__ int islast;
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int nblk = (P.ngroup + FIN_TB - 1) / FIN_TB;
  if ((int)blockIdx.x >= nblk) return;
  const int i = blockIdx.x * FIN_TB + threadIdx.x;
  const bool active = i < P.ngroup;
  double pe = 0, pl = 0, pk = 0, w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double top1 = 0, top2 = 0;
  if (active) {
    const int gs = P.group0 + i;
    const int g = order[gs];
    const int s0 = gfirst[g], ns = gns[g];
    double F[MAXS][3], ph[MAXS], qq[MAXS], x[MAXS][3];
    for (int k = 0; k < MAXS; k++) {
      if (k >= ns) break;
      const int sk = s0 + k;
      const double *o = pf + ((long)gs * MAXS + k) * 4;
      const double *kr = kres + 4L * sk;
      qq[k] = q[sk];
      F[k][0] = o[0] + qq[k] * kr[1];
      F[k][1] = o[1] + qq[k] * kr[2];
      F[k][2] = o[2] + qq[k] * kr[3];
      ph[k] = o[3] + kr[0] - P.self_coef * qq[k] + P.bg_phi;
      x[k][0] = r[sk]; x[k][1] = r[N + sk]; x[k][2] = r[2 * N + sk];
    }
    // excluded pairs inside the group
    const int ex = gexcl[g];
    if (ex) {
      const double L = P.L, invL = P.invL, eta = P.eta;
      for (int k = 0; k < ns; k++)
        for (int l = k + 1; l < ns; l++) {
          if (!((ex >> (k * MAXS + l)) & 1)) continue;
          double d0 = x[k][0] - x[l][0], d1 = x[k][1] - x[l][1], d2 = x[k][2] - x[l][2];
          d0 -= L * floor(d0 * invL + 0.5);
          d1 -= L * floor(d1 * invL + 0.5);
          d2 -= L * floor(d2 * invL + 0.5);
          const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
          const double rm = sqrt(r2), er = eta * rm;
          const double b0 = -erf(er) / rm;
          const double b1 = (b0 + eta * 1.1283791670955126 * exp(-er * er)) / r2;
          ph[k] += qq[l] * b0;
          ph[l] += qq[k] * b0;
          const double fq = qq[k] * qq[l] * b1;
          F[k][0] += fq * d0; F[k][1] += fq * d1; F[k][2] += fq * d2;
          F[l][0] -= fq * d0; F[l][1] -= fq * d1; F[l][2] -= fq * d2;
          if (want_virial) {
            const double d[3] = {d0, d1, d2};
            for (int a = 0; a < 3; a++)
              for (int b = 0; b < 3; b++) w[3 * a + b] += d[a] * fq * d[b];
          }
        }
    }
    for (int k = 0; k < ns; k++) pe += 0.5 * qq[k] * ph[k];
    const double *ga = gacc + (long)gs * 10;
    pl = ga[0];
    if (want_virial)
      for (int c = 0; c < 9; c++) w[c] += ga[1 + c];
    for (int k = 0; k < ns; k++) {
      const int sk = s0 + k;
      f[sk] = F[k][0]; f[N + sk] = F[k][1]; f[2 * N + sk] = F[k][2];
      if (phi_out) phi_out[sk] = ph[k];
    }
    if (mode) {
      const double dt = P.dt;
      double vv[MAXS][3], m[MAXS];
      for (int k = 0; k < ns; k++) {
        const int sk = s0 + k;
        vv[k][0] = v[sk]; vv[k][1] = v[N + sk]; vv[k][2] = v[2 * N + sk];
        m[k] = mass[sk];
      }
      const bool water = (ns == 3);
      if (mode & FIN_KICK) {
        for (int k = 0; k < ns; k++)
          for (int d = 0; d < 3; d++) vv[k][d] += F[k][d] * (dt / 2 / m[k]);
        if (water) rattle_water(&x[0][0], &vv[0][0], m, bond, angle_rad);
        for (int k = 0; k < ns; k++)
          pk += 0.5 * m[k] * (vv[k][0] * vv[k][0] + vv[k][1] * vv[k][1] + vv[k][2] * vv[k][2]);
      }
      if (mode & FIN_DRIFT) {
        double x0[MAXS][3];
        for (int k = 0; k < ns; k++)
          for (int d = 0; d < 3; d++) {
            x0[k][d] = x[k][d];
            vv[k][d] += F[k][d] * (dt / 2 / m[k]);
            x[k][d] += vv[k][d] * dt;
          }
        if (water && !shake_water(&x[0][0], &vv[0][0], &x0[0][0], m, dt, bond, angle_rad, tol, 500))
          atomicOr(err, ERR_CONSTRAINT);
        for (int k = 0; k < ns; k++) {
          const int sk = s0 + k;
          r[sk] = x[k][0]; r[N + sk] = x[k][1]; r[2 * N + sk] = x[k][2];
        }
      }
      for (int k = 0; k < ns; k++) {
        const int sk = s0 + k;
        v[sk] = vv[k][0]; v[N + sk] = vv[k][1]; v[2 * N + sk] = vv[k][2];
      }
    }
    // nbrlist.cpp:64-73 on the first site
    const double dx = x[0][0] - last[3 * g], dy = x[0][1] - last[3 * g + 1],
                 dz = x[0][2] - last[3 * g + 2];
    top1 = sqrt(dx * dx + dy * dy + dz * dz);
  }
  // block reduction
  double vals[NPART] = {pe, pl, pk, w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8], top1, top2};
  for (int o = 32; o >= 1; o >>= 1) {
    for (int c = 0; c < 12; c++) vals[c] += __shfl_xor(vals[c], o);
    const double b1 = __shfl_xor(vals[12], o), b2 = __shfl_xor(vals[13], o);
    top2_merge(vals[12], vals[13], b1, b2);
  }
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int c = 0; c < NPART; c++) red[wid][c] = vals[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[NPART];
    for (int c = 0; c < NPART; c++) t[c] = 0;
    for (int ww = 0; ww < FIN_TB / 64; ww++) {
      for (int c = 0; c < 12; c++) t[c] += red[ww][c];
      top2_merge(t[12], t[13], red[ww][12], red[ww][13]);
    }
    double *pp = part + ((long)s * maxblk + blockIdx.x) * NPART;
    for (int c = 0; c < NPART; c++) pp[c] = t[c];
    __threadfence();
    islast = (atomicAdd(&done[s], 1) == nblk - 1);
  }
  __syncthreads();
  if (islast && threadIdx.x == 0) {
    __threadfence();
    double t[NPART];
    for (int c = 0; c < NPART; c++) t[c] = 0;
    for (int b = 0; b < nblk; b++) {
      const volatile double *pp = part + ((long)s * maxblk + b) * NPART;
      for (int c = 0; c < 12; c++) t[c] += pp[c];
      top2_merge(t[12], t[13], pp[12], pp[13]);
    }
    double *a = acc + s * NACC;
    a[ACC_ELEC] = t[0];
    a[ACC_LJ] = t[1];
    if (mode & FIN_KICK) a[ACC_KIN] = t[2];
    for (int c = 0; c < 9; c++) a[ACC_VIR + c] = t[3 + c];
    a[ACC_TOP1] = t[12];
    a[ACC_TOP2] = t[13];
    rebuild[s] = (t[12] + t[13] > P.skin) ? 1 : 0;   // nbrlist.cpp:75-76
    if (!(t[0] == t[0]) || !(t[1] == t[1])) atomicOr(err, ERR_NAN);
    done[s] = 0;
  }
}

// initial position constraints (mdyn.cpp:120-121: constrain_positions(0))
__global__ void k_constrain0(const SysParams *__restrict__ sp, const int *__restrict__ gfirst,
                             const int *__restrict__ gns, const double *__restrict__ mass,
                             double *__restrict__ r, int N, double bond, double angle_rad,
                             double tol, int *__restrict__ err) {
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.ngroup) return;
  const int g = P.group0 + i;
  if (gns[g] != 3) return;
  const int s0 = gfirst[g];
  double x[9], x0[9], vv[9], m[3];
  for (int k = 0; k < 3; k++) {
    for (int d = 0; d < 3; d++) { x[3 * k + d] = r[d * N + s0 + k]; vv[3 * k + d] = 0; }
    m[k] = mass[s0 + k];
  }
  for (int c = 0; c < 9; c++) x0[c] = x[c];
  if (!shake_water(x, vv, x0, m, 0.0, bond, angle_rad, tol, 10000)) atomicOr(err, ERR_CONSTRAINT);
  for (int k = 0; k < 3; k++)
    for (int d = 0; d < 3; d++) r[d * N + s0 + k] = x[3 * k + d];
}

// RSpaceEwald's own outputs (rspace.c:382-558): pair field + LJ force from
// the field-mode pair kernel plus the excluded-pair loop, in site order
__global__ void k_field_out(const SysParams *__restrict__ sp, const int *__restrict__ order,
                            const int *__restrict__ gfirst, const int *__restrict__ gns,
                            const int *__restrict__ gexcl, const double *__restrict__ q,
                            const double *__restrict__ pf, const double *__restrict__ plj,
                            const double *__restrict__ gacc, const double *__restrict__ r, int N,
                            double *__restrict__ out, double *__restrict__ ulj) {
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.ngroup) return;
  const int gs = P.group0 + i, g = order[gs];
  const int s0 = gfirst[g], ns = gns[g];
  double E[MAXS][3], ph[MAXS], x[MAXS][3];
  for (int k = 0; k < ns; k++) {
    const double *o = pf + ((long)gs * MAXS + k) * 4;
    E[k][0] = o[0]; E[k][1] = o[1]; E[k][2] = o[2]; ph[k] = o[3];
    const int sk = s0 + k;
    x[k][0] = r[sk]; x[k][1] = r[N + sk]; x[k][2] = r[2 * N + sk];
  }
  const int ex = gexcl[g];
  for (int k = 0; k < ns; k++)
    for (int l = k + 1; l < ns; l++) {
      if (!((ex >> (k * MAXS + l)) & 1)) continue;
      double d0 = x[k][0] - x[l][0], d1 = x[k][1] - x[l][1], d2 = x[k][2] - x[l][2];
      d0 -= P.L * floor(d0 * P.invL + 0.5);
      d1 -= P.L * floor(d1 * P.invL + 0.5);
      d2 -= P.L * floor(d2 * P.invL + 0.5);
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2, rm = sqrt(r2), er = P.eta * rm;
      const double b0 = -erf(er) / rm;
      const double b1 = (b0 + P.eta * 1.1283791670955126 * exp(-er * er)) / r2;
      const double qk = q[s0 + k], ql = q[s0 + l];
      ph[k] += ql * b0; ph[l] += qk * b0;
      E[k][0] += ql * b1 * d0; E[k][1] += ql * b1 * d1; E[k][2] += ql * b1 * d2;
      E[l][0] -= qk * b1 * d0; E[l][1] -= qk * b1 * d1; E[l][2] -= qk * b1 * d2;
    }
  for (int k = 0; k < ns; k++) {
    const int sk = s0 + k;
    const double *lj = plj + ((long)gs * MAXS + k) * 3;
    out[7L * sk] = ph[k];
    out[7L * sk + 1] = E[k][0]; out[7L * sk + 2] = E[k][1]; out[7L * sk + 3] = E[k][2];
    out[7L * sk + 4] = lj[0]; out[7L * sk + 5] = lj[1]; out[7L * sk + 6] = lj[2];
  }
  atomicAdd(&ulj[s], gacc[(long)gs * 10]);
}

int launch_field_out(Engine &e, double *out, double *ulj) {
  const int tb = 256;
  hipLaunchKernelGGL(k_field_out, dim3((e.maxg + tb - 1) / tb, e.nsys), dim3(tb), 0, e.stream,
                     e.d_sp, e.d_order, e.d_gfirst, e.d_gns, e.d_gexcl, e.d_q, e.d_pf, e.d_plj,
                     e.d_gacc, e.d_r, e.nsite, out, ulj);
  MM_HIP(hipGetLastError());
  return 0;
}

int launch_finish_ex(Engine &e, int mode, int want_virial, double *phi_out) {
  dim3 grid((e.maxg + FIN_TB - 1) / FIN_TB, e.nsys);
  hipLaunchKernelGGL(k_finish, grid, dim3(FIN_TB), 0, e.stream, e.d_sp, e.d_order, e.d_gfirst,
                     e.d_gns, e.d_gexcl, e.d_q, e.d_mass, e.d_pf, e.d_gacc, e.d_kres, e.d_r,
                     e.d_v, e.d_f, phi_out, e.d_last, e.nsite, mode, want_virial, e.bond,
                     e.angle * (M_PI / 180.0), e.tol, e.d_part, e.d_done, e.nblk_part, e.d_acc,
                     e.d_rebuild, e.d_err);
  MM_HIP(hipGetLastError());
  return 0;
}

int launch_finish(Engine &e, int mode) { return launch_finish_ex(e, mode, 0, nullptr); }

int launch_constrain_initial(Engine &e) {
  const int tb = 256;
  hipLaunchKernelGGL(k_constrain0, dim3((e.maxg + tb - 1) / tb, e.nsys), dim3(tb), 0, e.stream,
                     e.d_sp, e.d_gfirst, e.d_gns, e.d_mass, e.d_r, e.nsite, e.bond,
                     e.angle * (M_PI / 180.0), e.tol, e.d_err);
  MM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mmamd
