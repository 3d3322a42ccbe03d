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

extern "C" int mm_param_gradients(mm_engine *m, int sys, const double *f_ref, int maxtypes,
                                  int *ntype_out, int *site_type, double *type_params,
                                  double *dE, double *dL, double *loss, double *basis_err) {
  if (!m || !ntype_out || sys < 0 || sys >= m->e.nsys || maxtypes <= 0) {
    set_last_error("mm_param_gradients: bad arguments");
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
