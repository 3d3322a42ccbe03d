// finish.hip -- per-group completion of a force evaluation and the rigid
// water velocity-Verlet step.
//
// Per group (one thread, groups in site order):
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
#include "reduce.hpp"
#include "water.hpp"

namespace mmamd {

// reduced per system: elec, lj, kin, [vir[9]], top1, top2
template <bool VIR> struct FinSlots {
  static constexpr int N = VIR ? 14 : 5;
  static constexpr int TOP = N - 2;
};
static_assert(FinSlots<false>::N == FIN_NPART && FinSlots<false>::TOP == FIN_TOP,
              "the neighbour-list check reads this layout");

template <int MODE, bool VIR, bool LAZY>
__global__ void __launch_bounds__(FIN_TB) k_finish(
    const SysParams *__restrict__ sp, const int *__restrict__ order, const int *__restrict__ gfirst,
    const int *__restrict__ gns, const int *__restrict__ gexcl, const double *__restrict__ q,
    const double *__restrict__ mass, const double *__restrict__ pf, const double *__restrict__ gacc,
    const double *__restrict__ kres, double *__restrict__ r, double *__restrict__ v,
    double *__restrict__ f, double *__restrict__ phi_out, const double *__restrict__ last,
    int N, int want_virial, double bond, double angle_rad, double tol,
    double *__restrict__ part, int *__restrict__ done, int maxblk, double *__restrict__ acc,
    int *__restrict__ rebuild, int *__restrict__ err, double4 *__restrict__ gpos,
    float4 *__restrict__ spk, const int *__restrict__ ginfo, const int *__restrict__ gslot_inv) {
  constexpr int NPART = FinSlots<VIR>::N, TOP = FinSlots<VIR>::TOP;
  __shared__ double red[FIN_TB / 64][NPART];
  __shared__ int islast;
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int nblk = (P.ngroup + FIN_TB - 1) / FIN_TB;
  if ((int)blockIdx.x >= nblk) return;
  const int i = blockIdx.x * FIN_TB + threadIdx.x;
  const bool active = i < P.ngroup;
  double vals[NPART];
#pragma unroll
  for (int c = 0; c < NPART; c++) vals[c] = 0;
  if (active) {
    // threads take groups in site order, so the site arrays (positions,
    // velocities, forces, k-space results) are read and written coalesced;
    // the pair results and packed inputs sit at the group's slot
    const int g = P.group0 + i;
    const int gs = gslot_inv[g];
    const int s0 = gfirst[g], ns = gns[g];
    constexpr int K = 3;
    double F[K][3], ph[K], qq[K], x[K][3];
    // every load is issued here, before any arithmetic: one wave per SIMD
    // leaves no other wave to hide a late load's latency
    double vv[K][3], m[K];
    const int gin = ginfo[g];
    const double l0 = last[3 * g], l1 = last[3 * g + 1], l2 = last[3 * g + 2];
#pragma unroll
    for (int k = 0; k < K; k++) {
      F[k][0] = F[k][1] = F[k][2] = ph[k] = qq[k] = 0;
      x[k][0] = x[k][1] = x[k][2] = 0;
      vv[k][0] = vv[k][1] = vv[k][2] = 0;
      m[k] = 1;
      if (k < ns) {
        const int sk = s0 + k;
        if (MODE) {
          vv[k][0] = v[sk]; vv[k][1] = v[N + sk]; vv[k][2] = v[2 * N + sk];
          m[k] = mass[sk];
        }
        const double *o = pf + ((long)gs * MAXS + k) * 4;
        const double *kr = kres + 4L * sk;
        qq[k] = q[sk];
        F[k][0] = o[0] + qq[k] * kr[1];
        F[k][1] = o[1] + qq[k] * kr[2];
        F[k][2] = o[2] + qq[k] * kr[3];
        // lazy (inner MD) steps: the pair kernel formed forces only
        if (!LAZY) ph[k] = o[3] + kr[0] - P.self_coef * qq[k] + P.bg_phi;
        x[k][0] = r[sk]; x[k][1] = r[N + sk]; x[k][2] = r[2 * N + sk];
      }
    }
    // excluded pairs inside the group (rspace.c:492-521)
    const int ex = gexcl[g];
    if (ex) {
      const double L = P.L, invL = P.invL, eta = P.eta;
#pragma unroll
      for (int k = 0; k < K; k++)
#pragma unroll
        for (int l = k + 1; l < K; l++) {
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
          if constexpr (VIR) {
            vals[3] += d0 * fq * d0; vals[4] += d0 * fq * d1; vals[5] += d0 * fq * d2;
            vals[6] += d1 * fq * d0; vals[7] += d1 * fq * d1; vals[8] += d1 * fq * d2;
            vals[9] += d2 * fq * d0; vals[10] += d2 * fq * d1; vals[11] += d2 * fq * d2;
          }
        }
    }
    const double *ga = gacc + (long)gs * 10;
    if constexpr (LAZY) {
      // no energies: slot 0 carries the force sum, so a non-finite force
      // still raises ERR_NAN below
#pragma unroll
      for (int k = 0; k < K; k++) vals[0] += F[k][0] + F[k][1] + F[k][2];
    } else {
#pragma unroll
      for (int k = 0; k < K; k++) vals[0] += 0.5 * qq[k] * ph[k];
      vals[1] = ga[0];
    }
    if constexpr (VIR) {
#pragma unroll
      for (int c = 0; c < 9; c++) vals[3 + c] += ga[1 + c];
    }
#pragma unroll
    for (int k = 0; k < K; k++)
      if (k < ns) {
        const int sk = s0 + k;
        f[sk] = F[k][0]; f[N + sk] = F[k][1]; f[2 * N + sk] = F[k][2];
        if (!LAZY && phi_out) phi_out[sk] = ph[k];
      }
    if (MODE) {
      const double dt = P.dt;
      const bool water = (ns == 3);
      if (MODE & FIN_KICK) {
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
          for (int d = 0; d < 3; d++) vv[k][d] += F[k][d] * (dt / 2 / m[k]);
        if (water) rattle_water(&x[0][0], &vv[0][0], m, bond, angle_rad);
#pragma unroll
        for (int k = 0; k < K; k++)
          if (k < ns) vals[2] += 0.5 * m[k] * (vv[k][0] * vv[k][0] + vv[k][1] * vv[k][1] + vv[k][2] * vv[k][2]);
      }
      if (MODE & FIN_DRIFT) {
        double x0[K][3];
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
          for (int d = 0; d < 3; d++) {
            x0[k][d] = x[k][d];
            vv[k][d] += F[k][d] * (dt / 2 / m[k]);
            x[k][d] += vv[k][d] * dt;
          }
        if (water && !shake_water(&x[0][0], &vv[0][0], &x0[0][0], m, dt, bond, angle_rad, tol, 500))
          atomicOr(err, ERR_CONSTRAINT);
#pragma unroll
        for (int k = 0; k < K; k++)
          if (k < ns) {
            const int sk = s0 + k;
            r[sk] = x[k][0]; r[N + sk] = x[k][1]; r[2 * N + sk] = x[k][2];
          }
        // the pair kernel's packed inputs for this sorted slot, as k_pack
        // forms them; a neighbour-list rebuild re-packs in its new order
        gpos[gs] = make_double4(x[0][0], x[0][1], x[0][2], 0.0);
#pragma unroll
        for (int k = 0; k < MAXS; k++) {
          float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
          if (k < K && k < ns)
            o = make_float4((float)(x[k][0] - x[0][0]), (float)(x[k][1] - x[0][1]),
                            (float)(x[k][2] - x[0][2]),
                            ((gin >> (8 * k)) & SITE_HAS_Q) ? (float)qq[k] : 0.f);
          spk[(long)gs * MAXS + k] = o;
        }
      }
#pragma unroll
      for (int k = 0; k < K; k++)
        if (k < ns) {
          const int sk = s0 + k;
          v[sk] = vv[k][0]; v[N + sk] = vv[k][1]; v[2 * N + sk] = vv[k][2];
        }
    }
    // nbrlist.cpp:64-73 on the first site
    const double dx = x[0][0] - l0, dy = x[0][1] - l1, dz = x[0][2] - l2;
    vals[TOP] = sqrt(dx * dx + dy * dy + dz * dz);
  }
  if constexpr (LAZY) {
    // partials only; a non-finite block sum is a non-finite total
    store_partials<NPART, TOP, FIN_TB>(vals, part + (long)s * maxblk * NPART, blockIdx.x, red);
    if (threadIdx.x == 0) {
      if (!(vals[0] == vals[0]) || !(vals[1] == vals[1])) atomicOr(err, ERR_NAN);
      if (blockIdx.x == 0) rebuild[s] = 0;   // the next check decides
    }
    return;
  }
  if (finish_partials<NPART, TOP, FIN_TB>(vals, part + (long)s * maxblk * NPART, nblk, blockIdx.x,
                                         &done[s], red, &islast)) {
    double *a = acc + s * NACC;
    a[ACC_ELEC] = vals[0];
    a[ACC_LJ] = vals[1];
    if (MODE & FIN_KICK) a[ACC_KIN] = vals[2];
    for (int c = 0; c < 9; c++) a[ACC_VIR + c] = VIR ? vals[3 + c] : 0.0;
    a[ACC_TOP1] = vals[TOP];
    a[ACC_TOP2] = vals[TOP + 1];
    rebuild[s] = (vals[TOP] + vals[TOP + 1] > P.skin) ? 1 : 0;   // nbrlist.cpp:75-76
    if (!(vals[0] == vals[0]) || !(vals[1] == vals[1])) atomicOr(err, ERR_NAN);
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
                            double *__restrict__ out, double *__restrict__ ulj, int min_image) {
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
      if (min_image) {   // RSpaceEwald does (rspace.c:501-505), RSpaceEwaldPhi not (:747-751)
        d0 -= P.L * floor(d0 * P.invL + 0.5);
        d1 -= P.L * floor(d1 * P.invL + 0.5);
        d2 -= P.L * floor(d2 * P.invL + 0.5);
      }
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

int launch_field_out(Engine &e, double *out, double *ulj, int min_image) {
  const int tb = 256;
  hipLaunchKernelGGL(k_field_out, dim3((e.maxg + tb - 1) / tb, e.nsys), dim3(tb), 0, e.stream,
                     e.d_sp, e.d_order, e.d_gfirst, e.d_gns, e.d_gexcl, e.d_q, e.d_pf, e.d_plj,
                     e.d_gacc, e.d_r, e.nsite, out, ulj, min_image);
  MM_HIP(hipGetLastError());
  return 0;
}

int launch_finish_ex(Engine &e, int mode, int want_virial, double *phi_out) {
  dim3 grid((e.maxg + FIN_TB - 1) / FIN_TB, e.nsys);
#define FIN_ARGS e.d_sp, e.d_order, e.d_gfirst, e.d_gns, e.d_gexcl, e.d_q, e.d_mass, e.d_pf, \
    e.d_gacc, e.d_kres, e.d_r, e.d_v, e.d_f, phi_out, e.d_last, e.nsite, want_virial, e.bond,  \
    e.angle * (M_PI / 180.0), e.tol, e.d_fpart, e.d_done, e.nblk_part, e.d_acc, e.d_rebuild,  \
    e.d_err, e.d_gpos, e.d_spk, e.d_ginfo, e.d_gslot_inv
#define FIN_CASE(M)                                                                            \
  if (want_virial)                                                                             \
    hipLaunchKernelGGL((k_finish<M, true, false>), grid, dim3(FIN_TB), 0, e.stream, FIN_ARGS);  \
  else                                                                                         \
    hipLaunchKernelGGL((k_finish<M, false, false>), grid, dim3(FIN_TB), 0, e.stream, FIN_ARGS);
  if (mode & FIN_LAZY) {
    if (want_virial || (mode & ~FIN_LAZY) != (FIN_KICK | FIN_DRIFT)) {
      set_last_error("partial-sum finish is built for the MD step without virial only");
      return -1;
    }
    hipLaunchKernelGGL((k_finish<FIN_KICK | FIN_DRIFT, false, true>), grid, dim3(FIN_TB), 0,
                       e.stream, FIN_ARGS);
  } else {
    switch (mode) {
      case 0: FIN_CASE(0) break;
      case 1: FIN_CASE(1) break;
      case 2: FIN_CASE(2) break;
      default: FIN_CASE(3) break;
    }
  }
#undef FIN_CASE
#undef FIN_ARGS
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
