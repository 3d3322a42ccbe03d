// kspace.hip -- plain Ewald reciprocal-space sum, FP64 (kspace.c:87-174).
//
// Half-space integer vectors |n| <= kc (kspace.c:35-58), a_k =
// (4 pi / V) exp(-k^2 / 4 eta^2) / k^2 (kspace.c:96-103), structure factors
// S_k = sum q sin(k.r), C_k = sum q cos(k.r) (kspace.c:147-155), then
// phi_i += sum 2 a_k (S s_i + C c_i) and E_i += sum 2 a_k (C s_i - S c_i) k
// (kspace.c:156-173).  Trigonometric values come from FP64 sincos of k.r
// directly rather than the reference's angle-addition recurrences; both
// are FP64-accurate.  One block per k-vector forms S_k, C_k; one thread per
// site then sums over k.  The k-sum energy and virial
// (W_ab = sum a_k |rho_k|^2 (delta_ab - 2 k_a k_b (1/k^2 + 1/(4 eta^2))))
// are formed by the first block of the field kernel.
#include "engine.hpp"

namespace mmamd {

__global__ void __launch_bounds__(256) k_ks_sf(const SysParams *__restrict__ sp,
                                               const double4 *__restrict__ kvec,
                                               const double *__restrict__ r, const double *__restrict__ q,
                                               int N, double *__restrict__ ksc) {
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int k = blockIdx.x;
  if (P.kmode != 0 || k >= P.nk) return;
  const double4 kv = kvec[P.k0 + k];
  double ss = 0, cc = 0;
  for (int i = threadIdx.x; i < P.nsite; i += blockDim.x) {
    const int si = P.site0 + i;
    const double kr = kv.x * r[si] + kv.y * r[N + si] + kv.z * r[2 * N + si];
    double sn, cs;
    sincos(kr, &sn, &cs);
    ss += q[si] * sn;
    cc += q[si] * cs;
  }
  __shared__ double rs[4], rc[4];
  for (int o = 32; o >= 1; o >>= 1) { ss += __shfl_xor(ss, o); cc += __shfl_xor(cc, o); }
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { rs[wid] = ss; rc[wid] = cc; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) { a += rs[w]; b += rc[w]; }
    ksc[2 * (P.k0 + k)] = a;
    ksc[2 * (P.k0 + k) + 1] = b;
  }
}

// Field kernel: a block serves KS_SITES sites; its KS_SPLIT waves split the
// k-vectors (wave w takes k = w, w + KS_SPLIT, ...), and the per-wave sums
// are added in wave order (deterministic).
constexpr int KS_SITES = 64, KS_SPLIT = 16, KS_TB = KS_SITES * KS_SPLIT;

__global__ void __launch_bounds__(KS_TB) k_ks_field(const SysParams *__restrict__ sp,
                                                    const double4 *__restrict__ kvec,
                                                    const double *__restrict__ ksc,
                                                    const double *__restrict__ r, int N,
                                                    double *__restrict__ kres,
                                                    double *__restrict__ acc, int want_virial) {
  __shared__ double part[KS_SPLIT][4][KS_SITES];
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  if (P.kmode != 0) return;
  if ((int)(blockIdx.x * KS_SITES) >= P.nsite) return;
  if (blockIdx.x == 0) {
    // k-sum energy and virial of this system
    double e = 0, w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double c4 = 0.25 / (P.eta * P.eta);
    for (int k = threadIdx.x; k < P.nk; k += KS_TB) {
      const double4 kv = kvec[P.k0 + k];
      const double S = ksc[2 * (P.k0 + k)], C = ksc[2 * (P.k0 + k) + 1];
      const double ek = kv.w * (S * S + C * C);
      e += ek;
      if (want_virial) {
        const double k2 = kv.x * kv.x + kv.y * kv.y + kv.z * kv.z;
        const double f = 2 * (1 / k2 + c4);
        const double kk[3] = {kv.x, kv.y, kv.z};
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) w[3 * a + b] += ek * ((a == b ? 1.0 : 0.0) - f * kk[a] * kk[b]);
      }
    }
    __shared__ double red[KS_TB / 64][10];
    for (int o = 32; o >= 1; o >>= 1) {
      e += __shfl_xor(e, o);
      for (int c = 0; c < 9; c++) w[c] += __shfl_xor(w[c], o);
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[wid][0] = e;
      for (int c = 0; c < 9; c++) red[wid][1 + c] = w[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double t[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int ww = 0; ww < KS_TB / 64; ww++)
        for (int c = 0; c < 10; c++) t[c] += red[ww][c];
      acc[s * NACC + ACC_KSPACE_E] = t[0];
      // reciprocal virial kept beside the per-site results; finish adds it
      for (int c = 0; c < 9; c++) kres[(long)4 * N + 9 * s + c] = want_virial ? t[1 + c] : 0.0;
    }
  }
  const int lane = threadIdx.x & (KS_SITES - 1), split = threadIdx.x / KS_SITES;
  const int i = blockIdx.x * KS_SITES + lane;
  const bool on = i < P.nsite;
  const int si = P.site0 + (on ? i : 0);
  const double x = r[si], y = r[N + si], z = r[2 * N + si];
  double p = 0, gx = 0, gy = 0, gz = 0;
  for (int k = split; k < P.nk; k += KS_SPLIT) {
    const double4 kv = kvec[P.k0 + k];
    const double S = ksc[2 * (P.k0 + k)], C = ksc[2 * (P.k0 + k) + 1];
    double sn, cs;
    sincos(kv.x * x + kv.y * y + kv.z * z, &sn, &cs);
    const double two_a = 2 * kv.w;
    const double t = two_a * (C * sn - S * cs);
    gx += t * kv.x; gy += t * kv.y; gz += t * kv.z;
    p += two_a * (S * sn + C * cs);
  }
  part[split][0][lane] = p;
  part[split][1][lane] = gx;
  part[split][2][lane] = gy;
  part[split][3][lane] = gz;
  __syncthreads();
  if (split < 4 && on) {   // four waves finish one component each
    double t = 0;
    for (int w = 0; w < KS_SPLIT; w++) t += part[w][split][lane];
    kres[4 * si + split] = t;
  }
}

int launch_kspace(Engine &e, int want_virial, hipStream_t st) {
  int maxk = 0, maxn = 0;
  for (auto &p : e.hsp)
    if (p.kmode == 0) { maxk = std::max(maxk, p.nk); maxn = std::max(maxn, p.nsite); }
  if (maxk == 0) return 0;
  hipLaunchKernelGGL(k_ks_sf, dim3(maxk, e.nsys), dim3(256), 0, st, e.d_sp,
                     (const double4 *)e.d_kvec, e.d_r, e.d_q, e.nsite, e.d_ksc);
  hipLaunchKernelGGL(k_ks_field, dim3((maxn + KS_SITES - 1) / KS_SITES, e.nsys), dim3(KS_TB), 0, st,
                     e.d_sp, (const double4 *)e.d_kvec, e.d_ksc, e.d_r, e.nsite, e.d_kres,
                     e.d_acc, want_virial);
  MM_HIP(hipGetLastError());
  return 0;
}

// Structure factors and field of one system whose parameters, wave vectors,
// SoA positions and result buffers the caller provides (kspace_field_noalloc
// keeps them in its own device array): ksc 2 nk, kres 4 nsite + 9, acc NACC.
int launch_kspace_raw(const SysParams *sp, const double4 *kvec, int nk, int nsite,
                      const double *r, const double *q, double *ksc, double *kres, double *acc,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_ks_sf, dim3(nk, 1), dim3(256), 0, st, sp, kvec, r, q, nsite, ksc);
  hipLaunchKernelGGL(k_ks_field, dim3((nsite + KS_SITES - 1) / KS_SITES, 1), dim3(KS_TB), 0, st,
                     sp, kvec, ksc, r, nsite, kres, acc, 0);
  MM_HIP(hipGetLastError());
  return 0;
}

// the field kernel of one system with other structure factors and outputs
// (the adjoint pass): kres_out needs 4 N + 9 entries, acc_out NACC
int launch_kspace_field_only(Engine &e, int sys, const double *ksc, double *kres_out,
                             double *acc_out) {
  const SysParams &P = e.hsp[sys];
  if (P.kmode != 0) return 0;
  hipLaunchKernelGGL(k_ks_field, dim3((P.nsite + KS_SITES - 1) / KS_SITES, 1), dim3(KS_TB), 0,
                     e.stream, e.d_sp + sys, (const double4 *)e.d_kvec, ksc, e.d_r, e.nsite,
                     kres_out, acc_out, 0);
  MM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mmamd
