// adjoint.hip -- gradients of the potential energy and of a force-matching
// loss with respect to per-site charges and per-type-pair LJ coefficients,
// with every pair term recomputed (nothing is stored per pair).
//
// For L = sum_i |F_i - F_ref,i|^2 and R_i = 2 (F_i - F_ref,i):
//   dL/dq_j = chi_j,   chi_j = sum over every force term that carries q_j
//                      of R . dF/dq_j;
//   dL/da_tu, dL/db_tu for the pair coefficients a = 4 e s^12, b = 4 e s^6
//   of each LJ type pair (U_LJ = a/r^12 - b/r^6);
//   dE/da_tu, dE/db_tu likewise for the energy (dE/dq_j = phi_j comes with
//   the force evaluation).
// Force terms and their charge/coefficient derivatives, as the forward path
// forms them (rspace.c:104-156, 258-297, 479-489, 492-521):
//   * site pair k in g, l in h within the group-pair cutoff: force on k
//     q_k q_l b1(r) r S + (12 a/r^14 - 6 b/r^8) r S, on l its negative;
//   * switching force on the first sites: dsm D tmps on g's, -dsm D tmps
//     on h's, tmps = sum_kl (q_k q_l b0 + a/r^12 - b/r^6);
//   * excluded pairs inside a group: q_k q_l b1e(d) d on k, negative on l.
// So each site pair gives w_kl = (R_k - R_l).r b1 S + (R_g0 - R_h0).D dsm b0
// with chi_k += w_kl q_l (and chi_l += w_kl q_k, from the other side of the
// full neighbour row), and per LJ type pair (half from each side)
//   dL/da += (R_k - R_l).r 12 S/r^14 + (R_g0 - R_h0).D dsm / r^12,
//   dL/db -= (R_k - R_l).r 6 S/r^8 + (R_g0 - R_h0).D dsm / r^6,
//   dE/da += S/r^12, dE/db -= S/r^6.
// The reciprocal-space part (F_i = q_i E_i, E linear in the charges) adds
// R_j . E_j (host) and psi_j = sum_i q_i R_i . dE_i/dq_j: for Ewald from
// adjoint structure factors A_k = sum_i q_i (R_i.k) sin(k.r_i),
// B_k = sum_i q_i (R_i.k) cos(k.r_i), psi_j = sum_k 2 a_k (A_k cos - B_k sin)
// (k_ks_field with (S, C) -> (-B, A)); for P3M by spreading the adjoint
// source q_i (R_i T) with derivative weights, the same convolution and an
// ordinary gather (pppm.hip, launch_p3m_adjoint); for P3M with the field
// formed on the mesh (ik) by three forward ik passes (launch_adj_p3m_ik).
// Pair arithmetic is FP32 with exact erfc/exp (the forward's table stands
// for the same function), accumulation FP64.
#include "engine.hpp"

namespace mmamd {

constexpr int ADJ_TB = 128;
constexpr int ADJ_MAXT = 8;    // LJ types per system (engine limit)

// One thread per sorted group slot of system P; its full neighbour row.
// chi: per site (global index), written for the group's sites.
// part: per block, ADJ_MAXT^2 x 4 (dLa, dLb, dEa, dEb) partial sums, the
// type pair (t, u) with t the type of this group's site.
__global__ void __launch_bounds__(ADJ_TB) k_adj_pair(
    const SysParams *__restrict__ sp, const int4 *__restrict__ gslot,
    const double4 *__restrict__ gpos, const float4 *__restrict__ spk,
    const int2 *__restrict__ gmeta, const int *__restrict__ nbr, const int *__restrict__ nnb,
    int nbcap, const double *__restrict__ R, const double *__restrict__ r, const double *__restrict__ q,
    int N, double *__restrict__ chi, double *__restrict__ part) {
  __shared__ double lj[ADJ_TB][3][ADJ_MAXT][4];
  const SysParams &P = sp[0];
  const int nty = P.ntype;
  const int i = blockIdx.x * ADJ_TB + threadIdx.x;
  const bool active = i < P.ngroup;
  for (int k = 0; k < 3; k++)
    for (int u = 0; u < ADJ_MAXT; u++)
      for (int c = 0; c < 4; c++) lj[threadIdx.x][k][u][c] = 0.0;
  if (active) {
    const int gs = P.group0 + i;
    const int4 sl = gslot[gs];
    const int s0 = sl.y, ns = sl.z;
    const double4 pi = gpos[gs];
    const int2 mi = gmeta[gs];
    float4 oi[3];
    float3 Ri[3];
    int mki[3], tyi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      oi[k] = spk[(long)gs * MAXS + k];
      const int b = (mi.y >> (8 * k)) & 0xff;
      mki[k] = k < ns ? (b & 3) : 0;
      tyi[k] = b >> 2;
      Ri[k] = k < ns ? make_float3((float)R[s0 + k], (float)R[N + s0 + k], (float)R[2 * N + s0 + k])
                     : make_float3(0.f, 0.f, 0.f);
    }
    double ch[3] = {0, 0, 0};
    const double L = P.L, invL = P.invL;
    const float eta = (float)P.eta, c2 = (float)(P.eta * 1.1283791670955126);
    const int cnt = nnb[gs];
    for (int t = 0; t < cnt; t++) {
      const int hs = nbr[(long)gs * nbcap + t];
      const double4 pj = gpos[hs];
      double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      dx -= L * floor(dx * invL + 0.5);
      dy -= L * floor(dy * invL + 0.5);
      dz -= L * floor(dz * invL + 0.5);
      const double r2g = dx * dx + dy * dy + dz * dz;
      if (r2g > P.rhi2) continue;
      double sm = 1.0, dsm = 0.0;
      if (r2g > P.rlo2) {   // rspace.c:52-58
        const double z = (r2g - P.rlo2) * P.inv_w2;
        sm = ((-6 * z + 15) * z - 10) * z * z * z + 1;
        dsm = ((60 * z - 120) * z + 60) * z * z * P.inv_w2;
      }
      const float S = (float)sm;
      const int4 slj = gslot[hs];
      const int2 mj = gmeta[hs];
      const int t0 = slj.y, nsj = slj.z;
      float3 Rj[3];
      float4 oj[3];
#pragma unroll
      for (int l = 0; l < 3; l++) {
        oj[l] = spk[(long)hs * MAXS + l];
        Rj[l] = l < nsj ? make_float3((float)R[t0 + l], (float)R[N + t0 + l], (float)R[2 * N + t0 + l])
                        : make_float3(0.f, 0.f, 0.f);
      }
      const float Dx = (float)dx, Dy = (float)dy, Dz = (float)dz;
      // (R_g0 - R_h0) . D dsm
      const float rdD = ((Ri[0].x - Rj[0].x) * Dx + (Ri[0].y - Rj[0].y) * Dy +
                         (Ri[0].z - Rj[0].z) * Dz) * (float)dsm;
#pragma unroll
      for (int l = 0; l < 3; l++) {
        if (l >= nsj) break;
        const int bj = (mj.y >> (8 * l)) & 0xff;
        const int mkj = bj & 3, tyj = bj >> 2;
        const float ql = (float)q[t0 + l];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const int m = mki[k] & mkj;
          if (!m) continue;
          const float rx = Dx - oj[l].x + oi[k].x, ry = Dy - oj[l].y + oi[k].y,
                      rz = Dz - oj[l].z + oi[k].z;
          const float rr2 = rx * rx + ry * ry + rz * rz;
          const float dRr = (Ri[k].x - Rj[l].x) * rx + (Ri[k].y - Rj[l].y) * ry +
                            (Ri[k].z - Rj[l].z) * rz;
          if (m & SITE_HAS_Q) {
            const float rm = sqrtf(rr2), er = eta * rm;
            const float b0 = erfcf(er) / rm;
            const float b1 = (b0 + c2 * expf(-er * er)) / rr2;
            const float w = dRr * b1 * S + rdD * b0;
            ch[k] += (double)(w * ql);
          }
          if (m & SITE_HAS_LJ) {
            const float ir2 = 1.f / rr2, ir6 = ir2 * ir2 * ir2, ir12 = ir6 * ir6;
            double *a = lj[threadIdx.x][k][tyj];
            a[0] += 0.5 * (double)(dRr * 12.f * ir12 * ir2 * S + rdD * ir12);
            a[1] -= 0.5 * (double)(dRr * 6.f * ir6 * ir2 * S + rdD * ir6);
            a[2] += 0.5 * (double)(S * ir12);
            a[3] -= 0.5 * (double)(S * ir6);
          }
        }
      }
    }
    // excluded pairs inside the group (rspace.c:492-521), FP64 as the forward
    const int ex = sl.w;
    if (ex) {
      double x[3][3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int sk = s0 + (k < ns ? k : 0);
        x[k][0] = r[sk]; x[k][1] = r[N + sk]; x[k][2] = r[2 * N + sk];
      }
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = k + 1; l < 3; l++) {
          if (l >= ns || !((ex >> (k * MAXS + l)) & 1)) continue;
          double d0 = x[k][0] - x[l][0], d1 = x[k][1] - x[l][1], d2 = x[k][2] - x[l][2];
          d0 -= L * floor(d0 * invL + 0.5);
          d1 -= L * floor(d1 * invL + 0.5);
          d2 -= L * floor(d2 * invL + 0.5);
          const double rr2 = d0 * d0 + d1 * d1 + d2 * d2, rm = sqrt(rr2), er = P.eta * rm;
          const double b0 = -erf(er) / rm;
          const double b1 = (b0 + P.eta * 1.1283791670955126 * exp(-er * er)) / rr2;
          const double w = ((R[s0 + k] - R[s0 + l]) * d0 + (R[N + s0 + k] - R[N + s0 + l]) * d1 +
                            (R[2 * N + s0 + k] - R[2 * N + s0 + l]) * d2) * b1;
          ch[k] += w * q[s0 + l];
          ch[l] += w * q[s0 + k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
      if (k < ns) chi[s0 + k] = ch[k];
  }
  __syncthreads();
  // per block: sum the threads' LJ slots in thread order into (t, u) slots
  const int nslot = nty * nty * 4;
  for (int e = threadIdx.x; e < nslot; e += ADJ_TB) {
    const int t = e / (nty * 4), u = (e / 4) % nty, c = e % 4;
    double s = 0;
    for (int th = 0; th < ADJ_TB; th++) {
      const int ii = blockIdx.x * ADJ_TB + th;
      if (ii >= P.ngroup) break;
      const int2 mi = gmeta[P.group0 + ii];
      for (int k = 0; k < 3; k++)
        if (k < mi.x && ((mi.y >> (8 * k)) & SITE_HAS_LJ) && (((mi.y >> (8 * k)) & 0xff) >> 2) == t)
          s += lj[th][k][u][c];
    }
    part[(long)blockIdx.x * ADJ_MAXT * ADJ_MAXT * 4 + (t * ADJ_MAXT + u) * 4 + c] = s;
  }
}

// Ewald adjoint structure factors of system P: (S', C') = (-B_k, A_k) with
// A_k = sum_i q_i (R_i.k) sin(k.r_i), B_k = sum_i q_i (R_i.k) cos(k.r_i)
__global__ void __launch_bounds__(256) k_adj_ks_sf(const SysParams *__restrict__ sp,
                                                   const double4 *__restrict__ kvec,
                                                   const double *__restrict__ r,
                                                   const double *__restrict__ q,
                                                   const double *__restrict__ R, int N,
                                                   double *__restrict__ ksc) {
  const SysParams &P = sp[0];
  const int k = blockIdx.x;
  if (P.kmode != 0 || k >= P.nk) return;
  const double4 kv = kvec[P.k0 + k];
  double a = 0, b = 0;
  for (int i = threadIdx.x; i < P.nsite; i += blockDim.x) {
    const int si = P.site0 + i;
    const double kr = kv.x * r[si] + kv.y * r[N + si] + kv.z * r[2 * N + si];
    const double rk = kv.x * R[si] + kv.y * R[N + si] + kv.z * R[2 * N + si];
    double sn, cs;
    sincos(kr, &sn, &cs);
    a += q[si] * rk * sn;
    b += q[si] * rk * cs;
  }
  __shared__ double ra[4], rb[4];
  for (int o = 32; o >= 1; o >>= 1) { a += __shfl_xor(a, o); b += __shfl_xor(b, o); }
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { ra[wid] = a; rb[wid] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0, sb = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) { sa += ra[w]; sb += rb[w]; }
    ksc[2 * (P.k0 + k)] = -sb;
    ksc[2 * (P.k0 + k) + 1] = sa;
  }
}

// P3M adjoint source per site of system P: scale * q_i (R_i T), T = diag(n) L^-1
// in tensor_t layout: src_d = sum_c R_c T[3 c + d]
__global__ void k_adj_p3m_src(const SysParams *__restrict__ sp, const double *__restrict__ q,
                              const double *__restrict__ R, int N, double scale,
                              float4 *__restrict__ src) {
  const SysParams &P = sp[0];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nsite) return;
  const int si = P.site0 + i;
  const double Rx = R[si], Ry = R[N + si], Rz = R[2 * N + si], qq = q[si] * scale;
  const double *T = P.T;
  src[si] = make_float4((float)(qq * (Rx * T[0] + Ry * T[3] + Rz * T[6])),
                        (float)(qq * (Rx * T[1] + Ry * T[4] + Rz * T[7])),
                        (float)(qq * (Rx * T[2] + Ry * T[5] + Rz * T[8])), 0.f);
}

// ik P3M adjoint, component c: the charges of one forward ik pass,
// qsrc_i = scale q_i R_i,c
__global__ void k_adj_ik_src(const SysParams *__restrict__ sp, const double *__restrict__ q,
                             const double *__restrict__ R, int N, double scale, int c,
                             double *__restrict__ qsrc) {
  const SysParams &P = sp[0];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nsite) return;
  const int si = P.site0 + i;
  qsrc[si] = scale * q[si] * R[(long)c * N + si];
}

// psi_i = -sum_c (field component c of pass c at site i), summed in c order
__global__ void k_adj_ik_acc(const SysParams *__restrict__ sp, const double *__restrict__ kres,
                             int c, double *__restrict__ psi) {
  const SysParams &P = sp[0];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nsite) return;
  const int si = P.site0 + i;
  const double v = -kres[4L * si + 1 + c];
  psi[4L * si] = c == 0 ? v : psi[4L * si] + v;
}

int adj_part_slots() { return ADJ_MAXT * ADJ_MAXT * 4; }
int adj_blocks(const Engine &e, int sys) { return (e.hsp[sys].ngroup + ADJ_TB - 1) / ADJ_TB; }

int launch_adj_pair(Engine &e, int sys, const double *R, double *chi, double *part) {
  if (e.hsp[sys].ntype > ADJ_MAXT) {
    set_last_error("more LJ types than the adjoint kernel holds");
    return -1;
  }
  hipLaunchKernelGGL(k_adj_pair, dim3(adj_blocks(e, sys)), dim3(ADJ_TB), 0, e.stream,
                     e.d_sp + sys, e.d_gslot, e.d_gpos, e.d_spk, e.d_gmeta, e.d_nbr, e.d_nnb,
                     e.nbcap, R, e.d_r, e.d_q, e.nsite, chi, part);
  MM_HIP(hipGetLastError());
  return 0;
}

// psi of an Ewald system into kres_adj[4 i] (k_ks_field with the adjoint
// structure factors; its energy and virial slots go to scratch)
int launch_adj_ewald(Engine &e, int sys, const double *R, double *ksc_adj, double *kres_adj,
                     double *acc_scratch) {
  const SysParams &P = e.hsp[sys];
  if (P.kmode != 0) return 0;
  hipLaunchKernelGGL(k_adj_ks_sf, dim3(P.nk), dim3(256), 0, e.stream, e.d_sp + sys,
                     (const double4 *)e.d_kvec, e.d_r, e.d_q, R, e.nsite, ksc_adj);
  return launch_kspace_field_only(e, sys, ksc_adj, kres_adj, acc_scratch);
}

// psi of an ik P3M system into psi[4 i]: F_i = q_i E_i with E_a = W^T M_a W q
// and M_a^T = -M_a (M_a: the mesh operator of multiplier -i k_a G, odd), so
// psi_j = sum_i q_i R_i . dE_i/dq_j = -sum_a [E of a forward pass whose
// charges are q_i R_i,a]_a at site j: three forward ik passes
int launch_adj_p3m_ik(Engine &e, int sys, const double *R, double scale, double *qsrc,
                      double *psi) {
  const SysParams &P = e.hsp[sys];
  if (P.kmode != 1 || !P.ik) return 0;
  const dim3 g((P.nsite + 255) / 256);
  for (int c = 0; c < 3; c++) {
    hipLaunchKernelGGL(k_adj_ik_src, g, dim3(256), 0, e.stream, e.d_sp + sys, e.d_q, R,
                       e.nsite, scale, c, qsrc);
    const int rc = launch_p3m_ik_pass(e, sys, qsrc, e.stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_adj_ik_acc, g, dim3(256), 0, e.stream, e.d_sp + sys, e.d_kres, c, psi);
  }
  MM_HIP(hipGetLastError());
  return 0;
}

int launch_adj_p3m_src(Engine &e, int sys, const double *R, double scale, float4 *src) {
  const SysParams &P = e.hsp[sys];
  if (P.kmode != 1) return 0;
  hipLaunchKernelGGL(k_adj_p3m_src, dim3((P.nsite + 255) / 256), dim3(256), 0, e.stream,
                     e.d_sp + sys, e.d_q, R, e.nsite, scale, src);
  MM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mmamd
