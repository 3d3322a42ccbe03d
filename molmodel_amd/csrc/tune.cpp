// tune.cpp -- per-system choice of the reciprocal-space method and its
// parameters for batched runs, where systems of different sizes share one
// engine and each needs its own mesh.
//
// The reference leaves eta, kc and the mesh to the user, or minimises
// chi_r^2 + chi_k^2 over eta for a fixed mesh (p3m_optimize,
// pppm.c:667-716).  This tuner instead fixes the accuracy and minimises the
// device cost:
//   1. eta from the real-space estimate chi_r = tol / sqrt(2), with the
//      Kolafa-Perram form the reference uses (eleclj.cpp:134-138);
//   2. Ewald: the smallest integer kc whose estimate (eleclj.cpp:128-132)
//      is within tol / sqrt(2);
//   3. P3M, for each assignment order: the smallest 2357-smooth cubic mesh
//      whose chi (the reference's estimate, pppm.c:313-393) is within
//      tol / sqrt(2), found by bisection over the candidate sizes;
//   4. the candidate with the lowest modelled device time.
// The cost model's coefficients are per-unit times of this library's
// kernels on MI355X (spreading + gathering per assignment point, FFT per
// mesh point per log2 size, Ewald per site per wave vector, plus fixed
// launch costs).
#include "../../include/molmodel_amd.h"
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace mmamd;

namespace {

// device cost model, microseconds
constexpr double C_P3M_POINT = 6.0e-6;    // spread + gather per site per mesh point
constexpr double C_FFT = 1.0e-5;          // per mesh point per log2(M), 4 passes + x/G pass
constexpr double C_P3M_FIXED = 45.0;      // binning + 9 launches
constexpr double C_EWALD = 2.0e-5;        // structure factor + field, per site per k
constexpr double C_EWALD_FIXED = 10.0;

double err_rspace(double rc, double eta, double L) {   // eleclj.cpp:134-138
  return 2 / sqrt(rc / L) * exp(-(rc * eta) * (rc * eta));
}

double err_ewald(double kc, double eta, double L) {    // eleclj.cpp:128-132
  return eta * L / M_PI * sqrt(8 / kc) * exp(-(M_PI * kc / (eta * L)) * (M_PI * kc / (eta * L)));
}

double sincpi(double x) { return x == 0 ? 1.0 : sin(M_PI * x) / (M_PI * x); }

// chi of p3m_init (pppm.c:313-393, real-space differentiation) for a cubic
// box and cubic mesh.  In a cubic box |k_p|^2 = c (px^2 + py^2 + pz^2) and
// exp(-|k_p|^2 / 4 eta^2) factorises, so the per-axis aliasing factors and
// exponentials are tabulated once per mesh index.
double chi_p3m(int n, double L, double eta, int order, int nalias) {
  const double c = (2 * M_PI / L) * (2 * M_PI / L), inv_v = 1 / (L * L * L);
  const int na = 2 * nalias + 1;
  std::vector<double> u2(n * na), ex(n * na), p2(n * na);
  for (int i = 0; i < n; i++) {
    const int j = i > n / 2 ? i - n : i;
    for (int q = -nalias; q <= nalias; q++) {
      const double p = j + q * n;
      u2[i * na + q + nalias] = pow(sincpi((double)j / n + q), 2 * order);
      ex[i * na + q + nalias] = exp(-c * p * p / (4 * eta * eta));
      p2[i * na + q + nalias] = p * p;
    }
  }
  double chi = 0;
  for (int ix = 0; ix < n; ix++)
    for (int iy = 0; iy < n; iy++)
      for (int iz = 0; iz <= n / 2; iz++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        double numer = 0, d1 = 0, d2 = 0, s = 0;
        for (int a = 0; a < na; a++)
          for (int b = 0; b < na; b++) {
            const double uab = u2[ix * na + a] * u2[iy * na + b];
            const double eab = ex[ix * na + a] * ex[iy * na + b];
            const double pab = p2[ix * na + a] + p2[iy * na + b];
            for (int z = 0; z < na; z++) {
              const double kp2 = c * (pab + p2[iz * na + z]);
              if (kp2 == 0) continue;
              const double u = uab * u2[iz * na + z];
              const double g = 4 * M_PI / kp2 * eab * ex[iz * na + z];
              numer += g * kp2 * u;
              d1 += u;
              d2 += u * kp2;
              s += g * g * kp2;
            }
          }
        chi += s - numer * numer / (d1 * d2);
      }
  return pow(inv_v, 1.0 / 3.0) * sqrt(2 * fabs(chi));
}

long count_kvectors(double kc) {   // half space of kspace.c:41-57
  const double nsqmax = kc * kc + 1e-8;
  const int nm = (int)floor(kc);
  long nk = 0;
  for (int nx = 0; nx <= nm; nx++)
    for (int ny = (nx == 0) ? 0 : -nm; ny <= nm; ny++)
      for (int nz = (nx == 0 && ny == 0) ? 1 : -nm; nz <= nm; nz++)
        if (nx * nx + ny * ny + nz * nz <= nsqmax) nk++;
  return nk;
}

std::vector<int> smooth_sizes() {
  std::vector<int> v;
  for (int n = 8; n <= 128; n++) {
    int m = n;
    for (int p : {2, 3, 5, 7})
      while (m % p == 0) m /= p;
    if (m == 1) v.push_back(n);
  }
  return v;
}

}  // namespace

extern "C" int mm_tune_kspace(double boxl, int nsite, double rc, double tol, int allow_ewald,
                              mm_kspace_choice *out) {
  if (!out || !(boxl > 0) || nsite <= 0 || !(rc > 0) || !(tol > 0) || !(tol < 1)) {
    set_last_error("mm_tune_kspace: bad arguments");
    return MM_ERR_ARG;
  }
  const double target = tol / sqrt(2.0);
  // eta with err_rspace == target (err_rspace decreases in eta)
  const double pre = 2 / sqrt(rc / boxl);
  if (pre <= target) {
    set_last_error("mm_tune_kspace: tolerance is met by any eta");
    return MM_ERR_ARG;
  }
  const double eta = sqrt(log(pre / target)) / rc;
  mm_kspace_choice best;
  best.cost = INFINITY;
  best.kmode = -1;
  if (allow_ewald) {
    for (int kc = 1; kc <= 30; kc++) {
      if (err_ewald(kc, eta, boxl) > target) continue;
      const double cost = C_EWALD_FIXED + C_EWALD * nsite * (double)count_kvectors(kc);
      if (cost < best.cost) {
        best = mm_kspace_choice{};
        best.kmode = 0;
        best.eta = eta;
        best.kc = kc;
        best.err_rspace = err_rspace(rc, eta, boxl);
        best.err_kspace = err_ewald(kc, eta, boxl);
        best.cost = cost;
      }
      break;
    }
  }
  const std::vector<int> sizes = smooth_sizes();
  // orders the reference can run: its assignment weights stop at order 6
  // (order 7 falls through to assert(0), pppm.c:147-171), so a choice of 7
  // would have no reference result to be compared with
  for (int order = 3; order <= 6; order++) {
    // smallest size meeting the target (chi falls with the mesh size):
    // galloping search upwards from the smallest size, then bisection
    const int last = (int)sizes.size() - 1;
    int lo = 0, hi = 0, step = 1;
    while (chi_p3m(sizes[hi], boxl, eta, order, 1) > target) {
      if (hi == last) { hi = -1; break; }
      lo = hi + 1;
      hi = std::min(hi + step, last);
      step *= 2;
    }
    if (hi < 0) continue;
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (chi_p3m(sizes[mid], boxl, eta, order, 1) <= target) hi = mid;
      else lo = mid + 1;
    }
    const int n = sizes[lo];
    int bx, by;
    if (!choose_spread_blocks(n, n, n, 4096, 4096, bx, by)) continue;
    const double M = (double)n * n * n;
    const double cost = C_P3M_FIXED + C_P3M_POINT * nsite * order * order * order +
                        C_FFT * M * log2(M);
    if (cost < best.cost) {
      best = mm_kspace_choice{};
      best.kmode = 1;
      best.eta = eta;
      best.mesh[0] = best.mesh[1] = best.mesh[2] = n;
      best.order = order;
      best.err_rspace = err_rspace(rc, eta, boxl);
      best.err_kspace = chi_p3m(n, boxl, eta, order, 2);
      best.cost = cost;
    }
  }
  if (best.kmode < 0) {
    set_last_error("mm_tune_kspace: no method meets the tolerance");
    return MM_ERR_ARG;
  }
  *out = best;
  return MM_OK;
}
