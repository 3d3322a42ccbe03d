// adjoint.cpp -- gradients of the potential energy and of a force-matching
// loss with respect to per-type charges and Lennard-Jones parameters.
//
// No reference counterpart; the nearest is ElecLJ::set_lambda's linear
// interpolation of q, sigma, epsilon between two parameter sets
// (eleclj.cpp:273-319).  Method: with site types t (distinct (q, sigma,
// epsilon, mask) among the system's sites) and type charges Q_t,
//   * every force term is bilinear in the charges (pair fields, the
//     excluded-pair correction, the reciprocal-space field and the
//     switching force on first sites, whose energy factor is itself a
//     charge product) and linear in the pair coefficients a_tu = 4 e s^12,
//     b_tu = 4 e s^6 of each LJ type pair (U = a/r^12 - b/r^6, the
//     switching force included);
//   * so F = sum_{t,u} Q_t Q_u B^{tu} + sum_{tu} (a_tu F^a_tu + b_tu F^b_tu)
//     exactly, with each basis force one device force evaluation at the
//     current positions (charges set to indicator vectors, LJ tables set
//     to unit entries): the pair terms are recomputed, nothing is stored
//     per pair;
//   * dE/dQ_t = sum_{i in t} phi_i (phi of the evaluation at the actual
//     parameters); dE/da, dE/db are the LJ energies of the unit passes plus
//     the tail correction's coefficients (eleclj.cpp:1247-1266);
//   * with R = 2 (F - F_ref): dL/dQ_t = 2 sum_u Q_u R . B^{tu},
//     dL/da_tu = R . F^a_tu, dL/db_tu = R . F^b_tu;
//   * sigma and epsilon enter through Lorentz-Berthelot (or geometric)
//     mixing (rspace.c:267-270), differentiated on the host.
// The basis is checked by reassembling F at the actual parameters.
#include "../../include/molmodel_amd.h"
#include "handle.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

using namespace mmamd;

#define TRY_ADJ(x)        \
  do {                    \
    int rc_ = (x);        \
    if (rc_) return rc_;  \
  } while (0)

namespace {

int put_q(Engine &e, int s, const std::vector<double> &q) {
  const SysParams &P = e.hsp[s];
  MM_HIP(hipMemcpy(e.d_q + P.site0, q.data(), sizeof(double) * P.nsite, hipMemcpyHostToDevice));
  e.need_pack = 1;   // the packed site records carry the charges
  return 0;
}

int put_lj(Engine &e, int s, const std::vector<float2> &t) {
  const SysParams &P = e.hsp[s];
  MM_HIP(hipMemcpy(e.d_ljt + P.lj0, t.data(), sizeof(float2) * t.size(), hipMemcpyHostToDevice));
  return 0;
}

// forces (AoS, the system's sites) and the pair LJ energy after one evaluation
int evaluate(mm_engine *m, int s, std::vector<double> &f, double *elj, std::vector<double> *phi) {
  Engine &e = m->e;
  const SysParams &P = e.hsp[s];
  if (int rc = handle_compute(m)) return rc;
  std::vector<double> soa(3 * (size_t)e.nsite);
  MM_HIP(hipMemcpy(soa.data(), e.d_f, sizeof(double) * soa.size(), hipMemcpyDeviceToHost));
  f.assign(3 * (size_t)P.nsite, 0.0);
  for (int i = 0; i < P.nsite; i++)
    for (int c = 0; c < 3; c++) f[3 * i + c] = soa[(size_t)c * e.nsite + P.site0 + i];
  if (elj) {
    double acc[NACC];
    MM_HIP(hipMemcpy(acc, e.d_acc + (size_t)s * NACC, sizeof(acc), hipMemcpyDeviceToHost));
    *elj = acc[ACC_LJ];
  }
  if (phi) {
    phi->assign(P.nsite, 0.0);
    MM_HIP(hipMemcpy(phi->data(), e.d_phi + P.site0, sizeof(double) * P.nsite,
                     hipMemcpyDeviceToHost));
  }
  return 0;
}

double dot(const std::vector<double> &a, const std::vector<double> &b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}

}  // namespace

extern "C" int mm_param_gradients_basis(mm_engine *m, int sys, const double *f_ref, int maxtypes,
                                        int *ntype_out, int *site_type, double *type_params,
                                        double *dE, double *dL, double *loss, double *basis_err) {
  if (!m || !ntype_out || sys < 0 || sys >= m->e.nsys || maxtypes <= 0) {
    set_last_error("mm_param_gradients_basis: bad arguments");
    return MM_ERR_ARG;
  }
  Engine &e = m->e;
  const SystemInput &S = e.in[sys];
  const SysParams &P = e.hsp[sys];
  const int N = S.nsite;
  // site types: distinct (q, sigma, epsilon, mask)
  std::map<std::tuple<double, double, double, int>, int> tmap;
  std::vector<int> ty(N);
  for (int i = 0; i < N; i++) {
    auto key = std::make_tuple(S.q[i], S.sig[i], S.eps[i], S.mask[i] & 3);
    auto it = tmap.find(key);
    if (it == tmap.end()) it = tmap.emplace(key, (int)tmap.size()).first;
    ty[i] = it->second;
  }
  const int nt = (int)tmap.size();
  *ntype_out = nt;
  if (nt > maxtypes) {
    set_last_error("mm_param_gradients: more site types than maxtypes");
    return MM_ERR_ARG;
  }
  std::vector<double> Q(nt), sg(nt), ep(nt);
  std::vector<int> hasq(nt), haslj(nt);
  for (auto &kv : tmap) {
    const int t = kv.second;
    Q[t] = std::get<0>(kv.first);
    sg[t] = std::get<1>(kv.first);
    ep[t] = std::get<2>(kv.first);
    hasq[t] = std::get<3>(kv.first) & SITE_HAS_Q;
    haslj[t] = (std::get<3>(kv.first) & SITE_HAS_LJ) ? 1 : 0;
  }
  // the engine's LJ type of each site type (distinct (sigma, epsilon))
  std::map<std::pair<double, double>, int> ljtypes;
  for (int i = 0; i < N; i++)
    if (S.mask[i] & SITE_HAS_LJ) {
      auto key = std::make_pair(S.sig[i], S.eps[i]);
      if (!ljtypes.count(key)) ljtypes.emplace(key, (int)ljtypes.size());
    }
  const int nlj = P.ntype;
  std::vector<int> owner(nlj, -1);
  for (int t = 0; t < nt; t++)
    if (haslj[t]) {
      const int lt = ljtypes.at(std::make_pair(sg[t], ep[t]));
      if (owner[lt] >= 0) {
        set_last_error("mm_param_gradients: two site types share LJ parameters but differ in charge");
        return MM_ERR_ARG;
      }
      owner[lt] = t;
    }
  if (site_type)
    for (int i = 0; i < N; i++) site_type[i] = ty[i];
  if (type_params)
    for (int t = 0; t < nt; t++) {
      type_params[3 * t] = Q[t];
      type_params[3 * t + 1] = sg[t];
      type_params[3 * t + 2] = ep[t];
    }

  // evaluation at the actual parameters
  std::vector<double> f0, phi0;
  if (int rc = evaluate(m, sys, f0, nullptr, &phi0)) return rc;
  std::vector<double> R(3 * (size_t)N, 0.0);
  double L = 0;
  if (f_ref) {
    for (size_t k = 0; k < R.size(); k++) {
      const double d = f0[k] - f_ref[k];
      L += d * d;
      R[k] = 2 * d;
    }
  }
  if (loss) *loss = L;

  std::vector<double> gE(3 * nt, 0.0), gL(3 * nt, 0.0);
  for (int i = 0; i < N; i++) gE[3 * ty[i]] += phi0[i];   // dE/dQ_t

  // saved device parameters
  std::vector<double> qsave(S.q);
  std::vector<float2> ljsave(nlj * nlj);
  MM_HIP(hipMemcpy(ljsave.data(), e.d_ljt + P.lj0, sizeof(float2) * ljsave.size(),
                   hipMemcpyDeviceToHost));
  std::vector<float2> ljzero(nlj * nlj, make_float2(0.f, 0.f));
  std::vector<double> fre(3 * (size_t)N, 0.0);   // reassembled forces
  int rc = 0;
  do {
    // charge basis: B^{tt} = F(chi_t), B^{tu} = (F(chi_t + chi_u) - B^tt - B^uu) / 2
    if ((rc = put_lj(e, sys, ljzero))) break;
    std::vector<std::vector<double>> Btt(nt);
    for (int t = 0; t < nt && !rc; t++) {
      if (!hasq[t]) continue;
      std::vector<double> qv(N, 0.0);
      for (int i = 0; i < N; i++) qv[i] = ty[i] == t ? 1.0 : 0.0;
      if ((rc = put_q(e, sys, qv))) break;
      rc = evaluate(m, sys, Btt[t], nullptr, nullptr);
    }
    if (rc) break;
    for (int t = 0; t < nt && !rc; t++) {
      if (!hasq[t]) continue;
      for (int u = t; u < nt; u++) {
        if (!hasq[u]) continue;
        std::vector<double> B;
        if (u == t) {
          B = Btt[t];
        } else {
          std::vector<double> qv(N, 0.0), Fs;
          for (int i = 0; i < N; i++) qv[i] = (ty[i] == t || ty[i] == u) ? 1.0 : 0.0;
          if ((rc = put_q(e, sys, qv)) || (rc = evaluate(m, sys, Fs, nullptr, nullptr))) break;
          B.resize(Fs.size());
          for (size_t k = 0; k < Fs.size(); k++) B[k] = 0.5 * (Fs[k] - Btt[t][k] - Btt[u][k]);
        }
        const double mult = (u == t) ? 1.0 : 2.0;   // Q_t Q_u and Q_u Q_t
        for (size_t k = 0; k < B.size(); k++) fre[k] += mult * Q[t] * Q[u] * B[k];
        if (f_ref) {   // dF/dQ_t = 2 sum_u Q_u B^{tu}
          const double rb = dot(R, B);
          gL[3 * t] += 2.0 * Q[u] * rb;
          if (u != t) gL[3 * u] += 2.0 * Q[t] * rb;
        }
      }
    }
    if (rc) break;
    // LJ basis: unit a or b on one unordered engine type pair, no charges
    if ((rc = put_q(e, sys, std::vector<double>(N, 0.0)))) break;
    double ca = 0, cb = 0;
    if (S.include_lj_correction) lj_tail_factors(S, &ca, &cb);
    std::vector<double> nsites(nt, 0.0);
    for (int i = 0; i < N; i++)
      if (S.mask[i] & SITE_HAS_LJ) nsites[ty[i]] += 1;
    for (int lt = 0; lt < nlj && !rc; lt++)
      for (int lu = lt; lu < nlj && !rc; lu++) {
        const int t = owner[lt], u = owner[lu];
        if (t < 0 || u < 0) continue;
        // mixed parameters and their derivatives (rspace.c:267-270)
        const double e_tu = sqrt(ep[t] * ep[u]);
        const double s_tu = S.geometric_sigma ? sqrt(sg[t] * sg[u]) : 0.5 * (sg[t] + sg[u]);
        const double s6 = pow(s_tu, 6);
        const double a = 4 * e_tu * s6 * s6, b = 4 * e_tu * s6;
        // d e_tu / d e_t, d s_tu / d s_t for t and u (same type: both)
        double de[2], ds[2];
        const int who[2] = {t, u};
        for (int k = 0; k < 2; k++) {
          const int w = who[k], o = who[1 - k];
          de[k] = (t == u) ? 0.5 : 0.5 * (ep[w] > 0 ? sqrt(ep[o] / ep[w]) : 0.0);
          ds[k] = S.geometric_sigma ? ((t == u) ? 0.5 : 0.5 * sqrt(sg[o] / sg[w])) : 0.5;
        }
        for (int coef = 0; coef < 2 && !rc; coef++) {
          std::vector<float2> tab(ljzero);
          const float2 unit = coef == 0 ? make_float2(0.f, 1.f) : make_float2(1.f, 0.f);
          tab[lt * nlj + lu] = unit;
          tab[lu * nlj + lt] = unit;
          std::vector<double> F;
          double elj = 0;
          if ((rc = put_lj(e, sys, tab)) || (rc = evaluate(m, sys, F, &elj, nullptr))) break;
          // energy and loss derivatives with respect to this coefficient
          const double mult = (lt == lu) ? 1.0 : 2.0;
          const double ntail = nsites[t] * nsites[u] * mult;
          // tail: U = ca totA + cb totB, totA = sum n n a, totB = -sum n n b
          const double dEdc = elj + (coef == 0 ? ca * ntail : -cb * ntail);
          const double dLdc = f_ref ? dot(R, F) : 0.0;
          const double cval = coef == 0 ? a : b;
          for (size_t k = 0; k < F.size(); k++) fre[k] += cval * F[k];
          // chain rule: c = 4 e s^p with p = 12 (a) or 6 (b)
          const double p = coef == 0 ? 12.0 : 6.0;
          for (int k = 0; k < 2; k++) {
            const int w = who[k];
            const double dc_ds = cval * p / s_tu * ds[k];
            const double dc_de = (e_tu > 0) ? cval / e_tu * de[k] : 0.0;
            gE[3 * w + 1] += dEdc * dc_ds;
            gE[3 * w + 2] += dEdc * dc_de;
            gL[3 * w + 1] += dLdc * dc_ds;
            gL[3 * w + 2] += dLdc * dc_de;
          }
        }
      }
  } while (0);
  // restore the parameters and the state of the actual evaluation
  int rc2 = put_q(e, sys, qsave);
  if (!rc2) rc2 = put_lj(e, sys, ljsave);
  if (!rc2) rc2 = handle_compute(m);
  if (rc) return rc;
  if (rc2) return rc2;
  if (basis_err) {
    double num = 0, den = 0;
    for (size_t k = 0; k < f0.size(); k++) {
      num += (fre[k] - f0[k]) * (fre[k] - f0[k]);
      den += f0[k] * f0[k];
    }
    *basis_err = den > 0 ? sqrt(num / den) : 0.0;
  }
  if (dE)
    for (int k = 0; k < 3 * nt; k++) dE[k] = gE[k];
  if (dL)
    for (int k = 0; k < 3 * nt; k++) dL[k] = gL[k];
  return MM_OK;
}

// ------------------------------------------------------------ adjoint path
namespace {

template <class T>
int ensure(T **p, size_t n) {
  if (*p) return 0;
  MM_HIP(hipMalloc((void **)p, (n ? n : 1) * sizeof(T)));
  MM_HIP(hipMemset(*p, 0, (n ? n : 1) * sizeof(T)));
  return 0;
}

}  // namespace

// Gradients with every pair term recomputed by the adjoint kernels
// (adjoint.hip): one pass over the neighbour rows for the real-space and
// excluded-pair terms, the reciprocal-space adjoint (Ewald structure factors,
// one P3M spread + convolution + gather, or three forward ik passes),
// reduced per site type on the host.  Same site types, outputs and chain
// rule through the mixing rule as mm_param_gradients_basis, which evaluates
// a basis of forward passes instead.  ms: device time of the adjoint passes
// (may be NULL).
extern "C" int mm_param_gradients(mm_engine *m, int sys, const double *f_ref, int maxtypes,
                                  int *ntype_out, int *site_type, double *type_params,
                                  double *dE, double *dL, double *loss, double *ms) {
  if (!m || !ntype_out || sys < 0 || sys >= m->e.nsys || maxtypes <= 0) {
    set_last_error("mm_param_gradients: bad arguments");
    return MM_ERR_ARG;
  }
  Engine &e = m->e;
  const SystemInput &S = e.in[sys];
  const SysParams &P = e.hsp[sys];
  const int N = S.nsite, NT = e.nsite;
  // site types: distinct (q, sigma, epsilon, mask), in order of appearance
  std::map<std::tuple<double, double, double, int>, int> tmap;
  std::vector<int> ty(N);
  for (int i = 0; i < N; i++) {
    auto key = std::make_tuple(S.q[i], S.sig[i], S.eps[i], S.mask[i] & 3);
    auto it = tmap.find(key);
    if (it == tmap.end()) it = tmap.emplace(key, (int)tmap.size()).first;
    ty[i] = it->second;
  }
  const int nt = (int)tmap.size();
  *ntype_out = nt;
  if (nt > maxtypes) {
    set_last_error("mm_param_gradients: more site types than maxtypes");
    return MM_ERR_ARG;
  }
  std::vector<double> Q(nt), sg(nt), ep(nt);
  std::vector<int> haslj(nt);
  for (auto &kv : tmap) {
    const int t = kv.second;
    Q[t] = std::get<0>(kv.first);
    sg[t] = std::get<1>(kv.first);
    ep[t] = std::get<2>(kv.first);
    haslj[t] = (std::get<3>(kv.first) & SITE_HAS_LJ) ? 1 : 0;
  }
  std::map<std::pair<double, double>, int> ljtypes;   // engine LJ types (engine.cpp)
  for (int i = 0; i < N; i++)
    if (S.mask[i] & SITE_HAS_LJ) {
      auto key = std::make_pair(S.sig[i], S.eps[i]);
      if (!ljtypes.count(key)) ljtypes.emplace(key, (int)ljtypes.size());
    }
  const int nlj = P.ntype;
  std::vector<int> owner(nlj, -1);
  for (int t = 0; t < nt; t++)
    if (haslj[t]) {
      const int lt = ljtypes.at(std::make_pair(sg[t], ep[t]));
      if (owner[lt] >= 0) {
        set_last_error("mm_param_gradients: two site types share LJ parameters but differ in charge");
        return MM_ERR_ARG;
      }
      owner[lt] = t;
    }
  if (site_type)
    for (int i = 0; i < N; i++) site_type[i] = ty[i];
  if (type_params)
    for (int t = 0; t < nt; t++) {
      type_params[3 * t] = Q[t];
      type_params[3 * t + 1] = sg[t];
      type_params[3 * t + 2] = ep[t];
    }

  // forward evaluation: forces, potentials, reciprocal-space fields
  if (int rc = handle_compute(m)) return rc;
  std::vector<double> f0, phi0;
  if (int rc = evaluate(m, sys, f0, nullptr, &phi0)) return rc;
  std::vector<double> kres(4 * (size_t)N);
  MM_HIP(hipMemcpy(kres.data(), e.d_kres + 4L * P.site0, sizeof(double) * kres.size(),
                   hipMemcpyDeviceToHost));
  // R = 2 (F - F_ref), SoA over all sites of the engine (zero elsewhere)
  std::vector<double> R(3 * (size_t)N, 0.0), Rsoa(3 * (size_t)NT, 0.0);
  double L = 0;
  if (f_ref)
    for (int i = 0; i < N; i++)
      for (int c = 0; c < 3; c++) {
        const double d = f0[3 * i + c] - f_ref[3 * i + c];
        L += d * d;
        R[3 * i + c] = 2 * d;
        Rsoa[(size_t)c * NT + P.site0 + i] = 2 * d;
      }
  if (loss) *loss = L;
  const int nblk = adj_blocks(e, sys), nslot = adj_part_slots();
  TRY_ADJ(ensure(&e.d_adjR, 3 * (size_t)NT));
  TRY_ADJ(ensure(&e.d_chi, (size_t)NT));
  TRY_ADJ(ensure(&e.d_adjpart, (size_t)std::max(e.ngroup, 1) * nslot));
  TRY_ADJ(ensure(&e.d_adjksc, 2 * (size_t)std::max(e.nk_total, 1)));
  TRY_ADJ(ensure(&e.d_adjkres, 4 * (size_t)NT + 16));
  TRY_ADJ(ensure(&e.d_adjacc, (size_t)NACC));
  TRY_ADJ(ensure(&e.d_adjsrc, (size_t)NT));
  TRY_ADJ(ensure(&e.d_adjq, (size_t)NT));
  MM_HIP(hipMemcpy(e.d_adjR, Rsoa.data(), sizeof(double) * Rsoa.size(), hipMemcpyHostToDevice));
  // P3M: the adjoint source is scaled to at most 1 per site so that the
  // spreading's fixed-point sums keep the forward's range
  double scale = 1.0;
  if (P.kmode == 1 && f_ref) {
    double mx = 0;
    for (int i = 0; i < N; i++)
      for (int d = 0; d < 3; d++) {
        // ik: the source components are q_i R_i; otherwise q_i (R_i T)
        double v = P.ik ? R[3 * i + d] : 0.0;
        if (!P.ik)
          for (int c = 0; c < 3; c++) v += R[3 * i + c] * P.T[3 * c + d];
        mx = std::max(mx, fabs(S.q[i] * v));
      }
    scale = mx > 0 ? 1.0 / mx : 1.0;
  }
  hipEvent_t t0, t1;
  MM_HIP(hipEventCreate(&t0));
  MM_HIP(hipEventCreate(&t1));
  MM_HIP(hipEventRecord(t0, e.stream));
  int rc = launch_adj_pair(e, sys, e.d_adjR, e.d_chi, e.d_adjpart);
  if (!rc && f_ref && P.kmode == 0)
    rc = launch_adj_ewald(e, sys, e.d_adjR, e.d_adjksc, e.d_adjkres, e.d_adjacc);
  if (!rc && f_ref && P.kmode == 1 && P.ik)
    rc = launch_adj_p3m_ik(e, sys, e.d_adjR, scale, e.d_adjq, e.d_adjkres);
  if (!rc && f_ref && P.kmode == 1 && !P.ik) {
    rc = launch_adj_p3m_src(e, sys, e.d_adjR, scale, e.d_adjsrc);
    if (!rc) rc = launch_p3m_adjoint(e, sys, e.d_adjsrc, e.stream);
  }
  MM_HIP(hipEventRecord(t1, e.stream));
  MM_HIP(hipEventSynchronize(t1));
  float tms = 0;
  MM_HIP(hipEventElapsedTime(&tms, t0, t1));
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  // the reciprocal-space buffers now hold the adjoint: the next evaluation
  // starts from scratch
  m->pos_dirty = 1;
  if (rc) return rc;
  if (ms) *ms = tms;
  std::vector<double> chi(N), part((size_t)nblk * nslot), psi(4 * (size_t)N, 0.0);
  MM_HIP(hipMemcpy(chi.data(), e.d_chi + P.site0, sizeof(double) * N, hipMemcpyDeviceToHost));
  MM_HIP(hipMemcpy(part.data(), e.d_adjpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost));
  if (f_ref && (P.kmode == 0 || (P.kmode == 1 && P.ik)))
    MM_HIP(hipMemcpy(psi.data(), e.d_adjkres + 4L * P.site0, sizeof(double) * psi.size(),
                     hipMemcpyDeviceToHost));
  if (f_ref && P.kmode == 1 && !P.ik)
    MM_HIP(hipMemcpy(psi.data(), e.d_kres + 4L * P.site0, sizeof(double) * psi.size(),
                     hipMemcpyDeviceToHost));
  {
    int flags = 0;
    MM_HIP(hipMemcpy(&flags, e.d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (flags & ERR_BIN_OVERFLOW) {   // cannot happen: the forward grew the bins
      set_last_error("mm_param_gradients: P3M bin overflow in the adjoint pass");
      return MM_ERR_DEVICE;
    }
  }
  std::vector<double> gE(3 * nt, 0.0), gL(3 * nt, 0.0);
  for (int i = 0; i < N; i++) {
    gE[3 * ty[i]] += phi0[i];   // dE/dq_i = phi_i
    if (!f_ref) continue;
    // chi_i: pair and excluded terms, R_i . E_i and psi_i of the mesh or k sum
    const double *E = &kres[4 * (size_t)i];
    const double rdotE = R[3 * i] * E[1] + R[3 * i + 1] * E[2] + R[3 * i + 2] * E[3];
    gL[3 * ty[i]] += chi[i] + (P.kmode <= 1 ? rdotE + psi[4 * (size_t)i] / scale : 0.0);
  }
  // LJ: per engine type pair, (dL/da, dL/db, dE/da, dE/db) summed over the
  // blocks in block order, both orientations of an unordered pair together
  std::vector<double> lj(8 * 8 * 4, 0.0);
  for (int b = 0; b < nblk; b++)
    for (int k = 0; k < nslot; k++) lj[k] += part[(size_t)b * nslot + k];
  double ca = 0, cb = 0;
  if (S.include_lj_correction) lj_tail_factors(S, &ca, &cb);
  std::vector<double> nsites(nt, 0.0);
  for (int i = 0; i < N; i++)
    if (S.mask[i] & SITE_HAS_LJ) nsites[ty[i]] += 1;
  for (int lt = 0; lt < nlj; lt++)
    for (int lu = lt; lu < nlj; lu++) {
      const int t = owner[lt], u = owner[lu];
      if (t < 0 || u < 0) continue;
      double d[4];
      for (int c = 0; c < 4; c++)
        d[c] = lj[(lt * 8 + lu) * 4 + c] + (lt != lu ? lj[(lu * 8 + lt) * 4 + c] : 0.0);
      const double e_tu = sqrt(ep[t] * ep[u]);
      const double s_tu = S.geometric_sigma ? sqrt(sg[t] * sg[u]) : 0.5 * (sg[t] + sg[u]);
      const double s6 = pow(s_tu, 6);
      const double a = 4 * e_tu * s6 * s6, b = 4 * e_tu * s6;
      double de[2], ds[2];
      const int who[2] = {t, u};
      for (int k = 0; k < 2; k++) {
        const int w = who[k], o = who[1 - k];
        de[k] = (t == u) ? 0.5 : 0.5 * (ep[w] > 0 ? sqrt(ep[o] / ep[w]) : 0.0);
        ds[k] = S.geometric_sigma ? ((t == u) ? 0.5 : 0.5 * sqrt(sg[o] / sg[w])) : 0.5;
      }
      const double mult = (lt == lu) ? 1.0 : 2.0;
      const double ntail = nsites[t] * nsites[u] * mult;
      for (int coef = 0; coef < 2; coef++) {
        // tail: U = ca totA + cb totB, totA = sum n n a, totB = -sum n n b
        const double dEdc = d[2 + coef] + (coef == 0 ? ca * ntail : -cb * ntail);
        const double dLdc = f_ref ? d[coef] : 0.0;
        const double cval = coef == 0 ? a : b;
        const double p = coef == 0 ? 12.0 : 6.0;   // c = 4 e s^p
        for (int k = 0; k < 2; k++) {
          const int w = who[k];
          const double dc_ds = cval * p / s_tu * ds[k];
          const double dc_de = (e_tu > 0) ? cval / e_tu * de[k] : 0.0;
          gE[3 * w + 1] += dEdc * dc_ds;
          gE[3 * w + 2] += dEdc * dc_de;
          gL[3 * w + 1] += dLdc * dc_ds;
          gL[3 * w + 2] += dLdc * dc_de;
        }
      }
    }
  if (dE)
    for (int k = 0; k < 3 * nt; k++) dE[k] = gE[k];
  if (dL)
    for (int k = 0; k < 3 * nt; k++) dL[k] = gL[k];
  return MM_OK;
}

// ------------------------------------------------------------ parameters
extern "C" int mm_set_site_params(mm_engine *m, int sys, const double *q, const double *sigma,
                                  const double *epsilon) {
  if (!m || sys < 0 || sys >= m->e.nsys || !q || !sigma || !epsilon) {
    set_last_error("mm_set_site_params: bad arguments");
    return MM_ERR_ARG;
  }
  if (int rc = engine_set_params(m->e, sys, q, sigma, epsilon)) return rc;
  m->pos_dirty = 1;   // forces of the old parameters are stale
  return MM_OK;
}

extern "C" int mm_set_perturbation(mm_engine *m, int sys, const double *qB, const double *sigB,
                                   const double *epsB) {
  if (!m || sys < 0 || sys >= m->e.nsys || !qB || !sigB || !epsB) {
    set_last_error("mm_set_perturbation: bad arguments");
    return MM_ERR_ARG;
  }
  Engine &e = m->e;
  const SystemInput &S = e.in[sys];
  const size_t n = S.nsite;
  if (e.pert.size() != (size_t)e.nsys) e.pert.resize(e.nsys);
  std::vector<double> &p = e.pert[sys];
  p.assign(6 * n, 0.0);
  for (size_t i = 0; i < n; i++) {
    p[i] = S.q[i];
    p[n + i] = S.sig[i];
    p[2 * n + i] = S.eps[i];
    p[3 * n + i] = qB[i];
    p[4 * n + i] = sigB[i];
    p[5 * n + i] = epsB[i];
  }
  return MM_OK;
}

// ElecLJ::set_lambda (eleclj.cpp:273-308, Site::set_lambda in eleclj.hpp):
// q = (1 - l) qA + l qB for every site; sigma and epsilon likewise for the
// sites with LJ (the others keep theirs); then the LJ types, tables and the
// tail correction over the "%12.6g" keys of the new parameters.
extern "C" int mm_set_lambda(mm_engine *m, int sys, double lambda) {
  if (!m || sys < 0 || sys >= m->e.nsys || !std::isfinite(lambda)) {
    set_last_error("mm_set_lambda: bad arguments");
    return MM_ERR_ARG;
  }
  Engine &e = m->e;
  if (e.pert.size() != (size_t)e.nsys || e.pert[sys].empty()) {
    set_last_error("mm_set_lambda: mm_set_perturbation first");
    return MM_ERR_ARG;
  }
  const SystemInput &S = e.in[sys];
  const size_t n = S.nsite;
  const std::vector<double> &p = e.pert[sys];
  const double tmp = 1 - lambda;
  std::vector<double> q(n), sg(n), ep(n);
  for (size_t i = 0; i < n; i++) {
    q[i] = tmp * p[i] + lambda * p[3 * n + i];
    if (S.mask[i] & SITE_HAS_LJ) {
      sg[i] = tmp * p[n + i] + lambda * p[4 * n + i];
      ep[i] = tmp * p[2 * n + i] + lambda * p[5 * n + i];
    } else {
      sg[i] = S.sig[i];
      ep[i] = S.eps[i];
    }
  }
  return mm_set_site_params(m, sys, q.data(), sg.data(), ep.data());
}
