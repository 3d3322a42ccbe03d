// engine.cpp -- construction of the device engine and the host-side tables
// it needs (all FP64 on the host, rounded once to the device formats).
#include "engine.hpp"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <tuple>

namespace mmamd {

static thread_local std::string g_last_error;
void set_last_error(const std::string &s) { g_last_error = s; }
const char *last_error() { return g_last_error.c_str(); }

#define MM_REQ(cond, msg)                    \
  do {                                       \
    if (!(cond)) {                           \
      set_last_error(msg);                   \
      return -1;                             \
    }                                        \
  } while (0)

static double sq(double x) { return x * x; }

// ---- neighbor-list geometry (nbrlist.cpp:11, 96-102, 105-120) ----
// cells per side for a mean of cell_groups groups per cell; 2 is the
// reference's count (nbrlist.cpp:11)
int nl_ncell(int ngroup, int cell_groups) {
  return (int)floor(cbrt((double)(ngroup / cell_groups + 3)));
}

int nl_nmax(int ncell, double boxl, double rc, double skin) {
  const double lc = boxl * (1.0 / ncell);
  const double maxd = sqrt(3.0) * lc;
  const double vol = fabs(lc * lc * lc);
  const double a = sqrt(0.0 + 0.0 + sq(lc * lc));
  const double mind = vol / a;
  const int nmax = (int)ceil((rc + skin + maxd) / mind);
  const int cap = ncell / 2 + 1;
  return nmax < cap ? nmax : cap;
}

// Distinct cell offsets (as residues mod ncell) that the reference's cell
// pair rows reach, plus the zero offset for same-cell pairs; offsets whose
// cells cannot hold a pair closer than rc+skin are dropped (they add no
// pairs), which leaves the pair set unchanged.
// Index of (x, y, z) along the Hilbert curve through a 2^bits cube, by
// J. Skilling's transpose method ("Programming the Hilbert curve", AIP
// Conf. Proc. 707, 381 (2004)): the axes are transformed to the curve's
// transposed index, whose bits are then interleaved.
static unsigned long long hilbert_index(unsigned x, unsigned y, unsigned z, int bits) {
  unsigned X[3] = {x, y, z};
  const unsigned M = 1u << (bits - 1);
  for (unsigned Q = M; Q > 1; Q >>= 1) {
    const unsigned P = Q - 1;
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const unsigned t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
  unsigned t = 0;
  for (unsigned Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; i++) X[i] ^= t;
  unsigned long long key = 0;
  for (int b = bits - 1; b >= 0; b--)
    for (int i = 0; i < 3; i++) key = (key << 1) | ((X[i] >> b) & 1u);
  return key;
}

// storage position of every cell index of an n^3 grid along the Hilbert
// curve (rank), and the cell index at every position (unrank)
void hilbert_cell_order(int n, std::vector<int> &rank, std::vector<int> &unrank) {
  int bits = 1;
  while ((1 << bits) < n) bits++;
  const int nc = n * n * n;
  std::vector<std::pair<unsigned long long, int>> key(nc);
  for (int c = 0; c < nc; c++)
    key[c] = {hilbert_index((unsigned)(c / (n * n)), (unsigned)((c / n) % n), (unsigned)(c % n), bits), c};
  std::sort(key.begin(), key.end());
  rank.assign(nc, 0);
  unrank.assign(nc, 0);
  for (int k = 0; k < nc; k++) {
    rank[key[k].second] = k;
    unrank[k] = key[k].second;
  }
}

static std::vector<int4> nl_stencil(int ncell, double boxl, double rc, double skin) {
  const double lc = boxl * (1.0 / ncell);
  const double inv = 1 / lc;
  const double c2 = sq(rc + skin + sqrt(3.0) * lc);
  const int nmax = nl_nmax(ncell, boxl, rc, skin);
  std::set<std::tuple<int, int, int>> res;
  auto m = [&](int a) { return ((a % ncell) + ncell) % ncell; };
  for (int jx = 0; jx <= nmax; jx++)
    for (int jy = (jx == 0 ? 0 : -nmax); jy <= nmax; jy++)
      for (int jz = (jx == 0 && jy == 0 ? 1 : -nmax); jz <= nmax; jz++) {
        double x = lc * jx + 0.0 * jy + 0.0 * jz;
        double y = 0.0 * jx + lc * jy + 0.0 * jz;
        double z = 0.0 * jx + 0.0 * jy + lc * jz;
        x -= lc * floor(x * inv + 0.5);
        y -= lc * floor(y * inv + 0.5);
        z -= lc * floor(z * inv + 0.5);
        if (!(x * x + y * y + z * z < c2)) continue;
        auto a = std::make_tuple(m(jx), m(jy), m(jz));
        auto b = std::make_tuple(m(-jx), m(-jy), m(-jz));
        if (a != std::make_tuple(0, 0, 0)) res.insert(a);
        if (b != std::make_tuple(0, 0, 0)) res.insert(b);
      }
  std::vector<int4> out;
  out.push_back(make_int4(0, 0, 0, 1));
  const double cut2 = sq(rc + skin) * (1 + 1e-9) + 1e-9;
  for (auto &t : res) {
    int o[3] = {std::get<0>(t), std::get<1>(t), std::get<2>(t)};
    double lb = 0;
    for (int d = 0; d < 3; d++) {
      const int k = std::min(o[d], ncell - o[d]);
      lb += sq(std::max(0, k - 1) * lc);
    }
    if (lb > cut2) continue;
    out.push_back(make_int4(o[0], o[1], o[2], 0));
  }
  return out;
}

// ---- erfc(eta r)/r table (cheby.c:7-33 fit; rspace.c:25-37 ranges) ----
void cheb_fit(double a, double b, double (*f)(double), double *c4) {
  double g[50], c[4];
  const double fac = 2.0 / 50, bma = 0.5 * (b - a), bpa = 0.5 * (b + a);
  for (int k = 0; k < 50; k++) g[k] = f(cos(M_PI * (k + 0.5) / 50) * bma + bpa);
  for (int j = 0; j < 4; j++) {
    double s = 0;
    for (int k = 0; k < 50; k++) s += g[k] * cos(M_PI * j * (k + 0.5) / 50);
    c[j] = fac * s;
  }
  c4[0] = c[0] / 2 - c[1] + c[2] - c[3];
  c4[1] = 2 * c[1] - 8 * c[2] + 18 * c[3];
  c4[2] = 8 * c[2] - 48 * c[3];
  c4[3] = 32 * c[3];
}

static thread_local double tab_eta;
static double erfc_over_r(double r2) {
  const double r = sqrt(r2);
  return erfc(tab_eta * r) / r;
}

static void build_table(SysParams &P, double rc, double eta, std::vector<float4> &tab) {
  tab_eta = eta;
  const double r2max = 1.2 * sq(rc);
  const double r2min = r2max > 4 ? 4 : 0.9 * r2max;
  // the reference's interval width, widened only if the extended range
  // (site pairs of groups within rc reach about rc + 2 group radii) would
  // not fit in LDS
  const double bmax = sq(rc + 3.5);
  const double dxB = std::max((r2max - r2min) / 512, (bmax - r2min) / 1280);
  const int nB = std::max(1, (int)ceil((bmax - r2min) / dxB));
  const double amin = std::min(0.81, 0.5 * r2min);
  const int nA = 128;
  const double dxA = (r2min - amin) / nA;
  P.tab0 = (int)tab.size();
  P.ntabA = nA;
  P.ntabB = nB;
  P.tabAmin = (float)amin;
  P.tabMid = (float)r2min;
  P.tabAinv = (float)(1.0 / dxA);
  P.tabBinv = (float)(1.0 / dxB);
  // the float table range ends where the last float interval ends
  P.tabBmax = (float)(r2min + nB * dxB) * (1.0f - 4e-7f);
  double c[4];
  for (int i = 0; i < nA; i++) {
    cheb_fit(amin + i * dxA, amin + (i + 1) * dxA, erfc_over_r, c);
    tab.push_back(make_float4((float)c[0], (float)c[1], (float)c[2], (float)c[3]));
  }
  for (int i = 0; i < nB; i++) {
    cheb_fit(i * dxB + r2min, (i + 1) * dxB + r2min, erfc_over_r, c);
    tab.push_back(make_float4((float)c[0], (float)c[1], (float)c[2], (float)c[3]));
  }
}

// ---- tensors in tensor_t layout (cart.h:73-92) ----
void tensor_inverse(double *a, const double *b) {
  const double xx = b[0], yx = b[1], zx = b[2], xy = b[3], yy = b[4], zy = b[5], xz = b[6],
               yz = b[7], zz = b[8];
  const double det = -xz * yy * zx + xy * yz * zx + xz * yx * zy - xx * yz * zy - xy * yx * zz + xx * yy * zz;
  const double id = 1.0 / det;
  a[0] = id * (-yz * zy + yy * zz);
  a[3] = id * (xz * zy - xy * zz);
  a[6] = id * (-xz * yy + xy * yz);
  a[1] = id * (yz * zx - yx * zz);
  a[4] = id * (-xz * zx + xx * zz);
  a[7] = id * (xz * yx - xx * yz);
  a[2] = id * (-yy * zx + yx * zy);
  a[5] = id * (xy * zx - xx * zy);
  a[8] = id * (-xy * yx + xx * yy);
}

static double sincpi(double x) {   // pppm.c:56-59
  const double y = M_PI * x;
  return fabs(y) < 1e-16 ? 1 - sq(y) / 6 : sin(y) / y;
}

// optimal influence function at signed index (pppm.c:331-388): for the
// field by -i k on the mesh (ik) the numerator sums g u^2 (k . k_p) and the
// denominator is (sum u^2)^2 k^2; for real-space differentiation they are
// sum g u^2 k_p^2 and (sum u^2)(sum u^2 k_p^2)
double green_at(int jx, int jy, int jz, int nx, int ny, int nz, const double *rl,
                double inv_4eta2, double inv_v, int nalias, int order, int ik, double *chi_acc) {
  if (jx == 0 && jy == 0 && jz == 0) return 0;
  double ux[16], uy[16], uz[16];
  for (int qq = -nalias; qq <= nalias; qq++) {
    ux[qq + nalias] = pow(sincpi((double)jx / nx + qq), 2 * order);
    uy[qq + nalias] = pow(sincpi((double)jy / ny + qq), 2 * order);
    uz[qq + nalias] = pow(sincpi((double)jz / nz + qq), 2 * order);
  }
  const double kx = rl[0] * jx + rl[3] * jy + rl[6] * jz;
  const double ky = rl[1] * jx + rl[4] * jy + rl[7] * jz;
  const double kz = rl[2] * jx + rl[5] * jy + rl[8] * jz;
  double numer = 0, d1 = 0, d2 = 0, chi = 0;
  for (int px = -nalias; px <= nalias; px++) {
    const double pxx = jx + px * nx;
    for (int py = -nalias; py <= nalias; py++) {
      const double pyy = jy + py * ny;
      for (int pz = -nalias; pz <= nalias; pz++) {
        const double pzz = jz + pz * nz;
        if (fabs(pxx) < 1e-16 && fabs(pyy) < 1e-16 && fabs(pzz) < 1e-16) continue;
        const double k0 = rl[0] * pxx + rl[3] * pyy + rl[6] * pzz;
        const double k1 = rl[1] * pxx + rl[4] * pyy + rl[7] * pzz;
        const double k2 = rl[2] * pxx + rl[5] * pyy + rl[8] * pzz;
        const double kp2 = k0 * k0 + k1 * k1 + k2 * k2;
        const double u2 = ux[px + nalias] * uy[py + nalias] * uz[pz + nalias];
        const double gg = 4 * M_PI / kp2 * exp(-kp2 * inv_4eta2);
        if (ik) {
          numer += gg * u2 * (kx * k0 + ky * k1 + kz * k2);
        } else {
          numer += gg * kp2 * u2;
          d2 += u2 * kp2;
        }
        d1 += u2;
        chi += gg * gg * kp2;
      }
    }
  }
  const double denom = ik ? sq(d1) * (kx * kx + ky * ky + kz * kz) : d1 * d2;
  if (chi_acc) *chi_acc += chi - sq(numer) / denom;
  return inv_v * numer / denom;
}

static bool radix_ok(int n) {
  if (n < 1 || n > 128) return false;
  for (int p : {2, 3, 5, 7})
    while (n % p == 0) n /= p;
  return n == 1;
}

template <class T>
static int dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  MM_HIP(hipMalloc((void **)p, n * sizeof(T)));
  MM_HIP(hipMemset(*p, 0, n * sizeof(T)));
  return 0;
}

template <class T>
static int dput(T *d, const std::vector<T> &h) {
  if (h.empty()) return 0;
  MM_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

#define MM_TRY(x)          \
  do {                     \
    int rc_ = (x);         \
    if (rc_) return rc_;   \
  } while (0)

// LJ types of one system: distinct (sigma, epsilon) among its LJ sites,
// numbered in order of first appearance (stype per site, 0 for sites
// without LJ), and the type-pair table (b, a) with Lorentz-Berthelot or
// geometric mixing (rspace.c:267-270) appended to ljt.
static int lj_types(const SystemInput &S, std::vector<int> &stype, std::vector<float2> &ljt,
                    int &ntype) {
  std::map<std::pair<double, double>, int> types;
  stype.assign(S.nsite, 0);
  for (int i = 0; i < S.nsite; i++)
    if (S.mask[i] & SITE_HAS_LJ) {
      auto key = std::make_pair(S.sig[i], S.eps[i]);
      auto it = types.find(key);
      if (it == types.end()) it = types.emplace(key, (int)types.size()).first;
      stype[i] = it->second;
    }
  const int nt = std::max(1, (int)types.size());
  MM_REQ(nt <= 8, "at most 8 distinct LJ parameter pairs per system on the device");
  std::vector<std::pair<double, double>> tp(nt, {0.0, 0.0});
  for (auto &t : types) tp[t.second] = t.first;
  for (int a = 0; a < nt; a++)
    for (int b = 0; b < nt; b++) {
      const double s2 = S.geometric_sigma ? tp[a].first * tp[b].first
                                          : sq(0.5 * (tp[a].first + tp[b].first));
      const double s6 = s2 * s2 * s2;
      const double bb = 4 * sqrt(tp[a].second * tp[b].second) * s6;
      ljt.push_back(make_float2((float)bb, (float)(bb * s6)));
    }
  ntype = nt;
  return 0;
}

// per-group packed (mask | type << 2) bytes of one system's sites
static int group_info(const SystemInput &S, const std::vector<int> &stype, int g) {
  int info = 0;
  for (int k = 0; k < S.gns[g]; k++) {
    const int i = S.gfirst[g] + k;
    info |= ((S.mask[i] & 3) | (stype[i] << 2)) << (8 * k);
  }
  return info;
}

static double background_phi(const SystemInput &S, double eta, double volume) {
  double qsum = 0;
  for (double x : S.q) qsum += x;
  // eleclj.cpp:970-975
  return fabs(qsum) < 64.0 * 2.220446049250313e-16 ? 0.0 : -M_PI / (eta * eta * volume) * qsum;
}

// The tail energy is linear in totA = sum n_a n_b a_ab and
// totB = -sum n_a n_b b_ab: U = ca totA + cb totB (eleclj.cpp:1259-1264)
void lj_tail_factors(const SystemInput &s, double *ca, double *cb) {
  const double V = s.boxl * s.boxl * s.boxl;
  const double rhi = s.rc, rlo = rhi - s.width;
  const double rlo2 = rlo * rlo, rlo3 = rlo2 * rlo, rlo4 = rlo3 * rlo;
  const double rlo5 = rlo4 * rlo, rlo6 = rlo5 * rlo, rlo7 = rlo6 * rlo, rlo9 = rlo7 * rlo2;
  const double rhi2 = rhi * rhi, rhi3 = rhi2 * rhi, rhi4 = rhi3 * rhi, rhi5 = rhi4 * rhi;
  const double rhi6 = rhi5 * rhi, rhi7 = rhi6 * rhi;
  const double w = (rhi - rlo) / (63 * rhi3 * rlo9 * pow(rhi + rlo, 5));
  *ca = 2 * M_PI / V *
        (1 / (9 * rlo9) - w * (7 * rhi7 + 42 * rhi6 * rlo + 112 * rhi5 * rlo2 + 182 * rhi4 * rlo3 +
                               217 * rhi3 * rlo4 + 224 * rhi2 * rlo5 + 192 * rhi * rlo6 + 32 * rlo7));
  *cb = 2 * M_PI / V *
        (1 / (3 * rlo3) - w * 3 * rhi3 * rlo6 *
                              (7 * rhi4 + 42 * rhi3 * rlo + 112 * rhi2 * rlo2 + 150 * rhi * rlo3 + 25 * rlo4));
}

// eleclj.cpp:279-308 + 1247-1266: LJ tail over distinct "%12.6g" keys
static double lj_tail(const SystemInput &s) {
  std::map<std::string, int> keys;
  for (int i = 0; i < s.nsite; i++) {
    if (!(s.mask[i] & SITE_HAS_LJ)) continue;
    char key[128];
    snprintf(key, sizeof key, "%12.6g %12.6g", s.sig[i], s.eps[i]);
    keys[key]++;
  }
  double totA = 0, totB = 0;
  for (auto &a : keys) {
    double si, ei;
    sscanf(a.first.c_str(), "%lf%lf", &si, &ei);
    for (auto &b : keys) {
      double sj, ej;
      sscanf(b.first.c_str(), "%lf%lf", &sj, &ej);
      const double s6 = pow(s.geometric_sigma ? sqrt(si * sj) : 0.5 * (si + sj), 6);
      const double bb = 4 * sqrt(ei * ej) * a.second * b.second * s6;
      totB -= bb;
      totA += bb * s6;
    }
  }
  const double V = s.boxl * s.boxl * s.boxl;
  const double rhi = s.rc, rlo = rhi - s.width;
  const double rlo2 = rlo * rlo, rlo3 = rlo2 * rlo, rlo4 = rlo3 * rlo;
  const double rlo5 = rlo4 * rlo, rlo6 = rlo5 * rlo, rlo7 = rlo6 * rlo, rlo9 = rlo7 * rlo2;
  const double rhi2 = rhi * rhi, rhi3 = rhi2 * rhi, rhi4 = rhi3 * rhi, rhi5 = rhi4 * rhi;
  const double rhi6 = rhi5 * rhi, rhi7 = rhi6 * rhi;
  return 2 * M_PI / V *
         (totA / (9 * rlo9) + totB / (3 * rlo3) -
          (rhi - rlo) / (63 * rhi3 * rlo9 * pow(rhi + rlo, 5)) *
              (3 * totB * rhi3 * rlo6 * (7 * rhi4 + 42 * rhi3 * rlo + 112 * rhi2 * rlo2 + 150 * rhi * rlo3 + 25 * rlo4) +
               totA * (7 * rhi7 + 42 * rhi6 * rlo + 112 * rhi5 * rlo2 + 182 * rhi4 * rlo3 + 217 * rhi3 * rlo4 +
                       224 * rhi2 * rlo5 + 192 * rhi * rlo6 + 32 * rlo7)));
}

// Host half of engine creation: every system's parameters and its slices of
// the concatenated device buffers (sites, groups, cells, stencils, tables,
// wave vectors, meshes, twiddles, spreading bins), and the host copies of
// the buffers' contents.  Nothing here touches a device.
int engine_layout(Engine &e, const std::vector<SystemInput> &in, HostLayout &h) {
  e.in = in;
  e.nsys = (int)in.size();
  MM_REQ(e.nsys > 0, "no systems");
  std::vector<int> &gfirst = h.gfirst, &gns = h.gns, &gsys = h.gsys, &ginfo = h.ginfo,
                   &gexcl = h.gexcl;
  std::vector<double> &q = h.q, &mass = h.mass;
  std::vector<float4> &tab = h.tab;
  std::vector<float2> &ljt = h.ljt;
  std::vector<int4> &sten = h.sten;
  std::vector<double> &kvec = h.kvec;
  std::vector<float> &green = h.green;
  std::vector<float2> &twid = h.twid;
  int site0 = 0, group0 = 0, cell0 = 0, k0 = 0, bin0 = 0;
  long mesh0 = 0, cmesh0 = 0, green0 = 0, halo0 = 0;
  int maxnb = 64;
  for (int s = 0; s < e.nsys; s++) {
    const SystemInput &S = in[s];
    SysParams P;
    memset(&P, 0, sizeof P);
    MM_REQ(S.nsite > 0 && S.ngroup > 0, "empty system");
    MM_REQ(S.boxl > 0, "box length must be positive");
    MM_REQ(S.rc > 0 && S.rc <= 0.49 * S.boxl,
           "rspace cutoff must be > 0 and <= 0.49 * box length (eleclj.cpp:164-170)");
    MM_REQ(S.width >= 0 && S.width < S.rc, "smoothing width must be less than the cutoff (eleclj.cpp:180-182)");
    MM_REQ(S.eta > 0, "Ewald screening parameter must be positive");
    MM_REQ(S.skin > 0, "neighbour list width must be positive");
    P.L = S.boxl;
    P.invL = 1 / S.boxl;
    P.volume = S.boxl * S.boxl * S.boxl;
    P.rhi2 = sq(S.rc);
    P.rlo2 = sq(S.rc - S.width);
    P.inv_w2 = P.rhi2 > P.rlo2 ? 1.0 / (P.rhi2 - P.rlo2) : 0.0;
    P.nl_c2 = sq(S.rc + S.skin);
    P.skin = S.skin;
    P.eta = S.eta;
    P.self_coef = S.eta * 1.12837916709551257390;   // M_2_SQRTPI
    P.site0 = site0;
    P.nsite = S.nsite;
    P.group0 = group0;
    P.ngroup = S.ngroup;
    // groups
    std::vector<int> gof(S.nsite, -1);
    for (int g = 0; g < S.ngroup; g++) {
      MM_REQ(S.gns[g] >= 1 && S.gns[g] <= 3, "groups must have 1..3 sites on the device");
      MM_REQ(S.gfirst[g] >= 0 && S.gfirst[g] + S.gns[g] <= S.nsite, "group out of range");
      for (int k = 0; k < S.gns[g]; k++) {
        MM_REQ(gof[S.gfirst[g] + k] < 0, "groups overlap");
        gof[S.gfirst[g] + k] = g;
      }
    }
    for (int i = 0; i < S.nsite; i++) MM_REQ(gof[i] >= 0, "site not in any group");
    std::vector<int> stype;
    P.lj0 = (int)ljt.size();
    MM_TRY(lj_types(S, stype, ljt, P.ntype));
    // exclusions: intra-group only on the device
    std::vector<int> gx(S.ngroup, 0);
    for (size_t k = 0; 2 * k + 1 < S.excl.size(); k++) {
      const int a = S.excl[2 * k], b = S.excl[2 * k + 1];
      MM_REQ(a >= 0 && a < S.nsite && b >= 0 && b < S.nsite, "exclusion out of range");
      if (a == b) continue;
      MM_REQ(gof[a] == gof[b], "exclusions between different groups are not supported on the device");
      const int g = gof[a];
      int ka = a - S.gfirst[g], kb = b - S.gfirst[g];
      if (ka > kb) std::swap(ka, kb);
      gx[g] |= 1 << (ka * MAXS + kb);
    }
    for (int g = 0; g < S.ngroup; g++) {
      gfirst.push_back(site0 + S.gfirst[g]);
      gns.push_back(S.gns[g]);
      gsys.push_back(s);
      gexcl.push_back(gx[g]);
      ginfo.push_back(group_info(S, stype, g));
    }
    for (int i = 0; i < S.nsite; i++) {
      q.push_back(S.q[i]);
      mass.push_back(S.mass.empty() ? 1.0 : S.mass[i]);
    }
    // neutralising background (eleclj.cpp:970-975)
    P.bg_phi = background_phi(S, S.eta, P.volume);
    // neighbor list
    P.ncell = nl_ncell(S.ngroup, e.cell_groups);
    P.lc = S.boxl * (1.0 / P.ncell);
    P.cell0 = cell0;
    cell0 += P.ncell * P.ncell * P.ncell;
    auto st = nl_stencil(P.ncell, S.boxl, S.rc, S.skin);
    P.sten0 = (int)sten.size();
    P.nsten = (int)st.size();
    sten.insert(sten.end(), st.begin(), st.end());
    {
      const double dens = S.ngroup / P.volume;
      const double est = dens * 4.0 / 3.0 * M_PI * pow(S.rc + S.skin, 3);
      int cap = (int)(1.6 * est) + 48;
      cap = std::min(cap, S.ngroup);
      maxnb = std::max(maxnb, cap);
    }
    // erfc table
    build_table(P, S.rc, S.eta, tab);
    MM_REQ(P.ntabA + P.ntabB <= 1536, "erfc table too large for the box/cutoff");
    // the pair kernel selects the segment as the smaller of the two index maps
    MM_REQ(P.tabAinv > P.tabBinv, "erfc table segments out of order");
    // lattice tensors
    double latv[9] = {S.boxl, 0, 0, 0, S.boxl, 0, 0, 0, S.boxl};
    if (S.has_latv) {
      for (int a = 0; a < 9; a++) latv[a] = S.latv[a];
      P.volume = fabs(-latv[6] * latv[4] * latv[2] + latv[3] * latv[7] * latv[2] +
                      latv[6] * latv[1] * latv[5] - latv[0] * latv[7] * latv[5] -
                      latv[3] * latv[1] * latv[8] + latv[0] * latv[4] * latv[8]);
    }
    double inv[9];
    tensor_inverse(inv, latv);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) P.RL[3 * a + b] = 2 * M_PI * inv[3 * b + a];
    P.kmode = S.kmode;
    if (S.kmode == 0) {
      // kspace.c:35-58 and 96-103
      const double nsqmax = S.kc * S.kc + 1e-8;
      const int nm = (int)floor(S.kc);
      const double c = 0.25 / (S.eta * S.eta), fourpi_V = 4 * M_PI / P.volume;
      P.k0 = k0;
      int nk = 0;
      for (int nx = 0; nx <= nm; nx++)
        for (int ny = (nx == 0) ? 0 : -nm; ny <= nm; ny++)
          for (int nz = (nx == 0 && ny == 0) ? 1 : -nm; nz <= nm; nz++)
            if (nx * nx + ny * ny + nz * nz <= nsqmax) {
              const double kx = P.RL[0] * nx + P.RL[3] * ny + P.RL[6] * nz;
              const double ky = P.RL[1] * nx + P.RL[4] * ny + P.RL[7] * nz;
              const double kz = P.RL[2] * nx + P.RL[5] * ny + P.RL[8] * nz;
              const double t = kx * kx + ky * ky + kz * kz;
              kvec.push_back(kx); kvec.push_back(ky); kvec.push_back(kz);
              kvec.push_back(fourpi_V * exp(-t * c) / t);
              nk++;
            }
      P.nk = nk;
      k0 += nk;
    } else if (S.kmode == 1) {
      MM_REQ(S.ik == 0 || S.ik == 1, "ik must be 0 (real-space differentiation) or 1 (-i k on the mesh)");
      MM_REQ(S.order >= 1 && S.order <= 7, "assignment order must be 1..7");
      MM_REQ(radix_ok(S.nx) && radix_ok(S.ny) && radix_ok(S.nz),
             "mesh dimensions must be <= 128 with prime factors 2,3,5,7");
      MM_REQ(S.nx >= 8 && S.ny >= 8 && S.nz >= 8, "mesh dimensions must be >= 8");
      MM_REQ(S.nalias >= 0 && S.nalias <= 7, "alias sum limit must be 0..7");
      P.nx = S.nx; P.ny = S.ny; P.nz = S.nz; P.order = S.order;
      P.ik = S.ik;
      for (int a = 0; a < 9; a++) P.T[a] = inv[a];
      P.T[0] *= S.nx; P.T[3] *= S.nx; P.T[6] *= S.nx;
      P.T[1] *= S.ny; P.T[4] *= S.ny; P.T[7] *= S.ny;
      P.T[2] *= S.nz; P.T[5] *= S.nz; P.T[8] *= S.nz;
      const int nzh = S.nz / 2 + 1;
      const int ncomp = S.ik ? 4 : 1;   // phi [, Ex, Ey, Ez]
      mesh0 = (mesh0 + 3) & ~3L;   // 16-byte aligned: an ik system's points are float4
      P.mesh0 = mesh0; mesh0 += ncomp * (long)S.nx * S.ny * S.nz;
      P.cmesh0 = cmesh0; cmesh0 += ncomp * (long)S.nx * S.ny * nzh;
      P.green0 = green0;
      green0 += (long)S.nx * S.ny * nzh;
      const double inv_4eta2 = 1 / sq(2 * S.eta), inv_v = 1.0 / P.volume;
      std::vector<float> gh((size_t)S.nx * S.ny * nzh);
      for (int ix = 0; ix < S.nx; ix++) {
        const int jx = ix > S.nx / 2 ? ix - S.nx : ix;
        for (int iy = 0; iy < S.ny; iy++) {
          const int jy = iy > S.ny / 2 ? iy - S.ny : iy;
          for (int iz = 0; iz < nzh; iz++)
            gh[((size_t)iy * nzh + iz) * S.nx + ix] =
                (float)green_at(jx, jy, iz, S.nx, S.ny, S.nz, P.RL, inv_4eta2, inv_v, S.nalias, S.order,
                                S.ik, nullptr);
        }
      }
      green.insert(green.end(), gh.begin(), gh.end());
      const int nn[3] = {S.nx, S.ny, S.nz};
      P.twid0 = (int)twid.size();
      for (int d = 0; d < 3; d++)
        for (int t = 0; t < 128; t++) {
          const double a = -2 * M_PI * (t % nn[d]) / nn[d];
          twid.push_back(make_float2((float)cos(a), (float)sin(a)));
        }
      MM_REQ(choose_spread_blocks(S.nx, S.ny, S.nz, 4096, 4096, P.bx, P.by, e.spread_block,
                                  e.spread_halo ? S.order : 0),
             "mesh dimensions give no valid spreading column blocks");
      P.nbx = S.nx / P.bx;
      P.nby = S.ny / P.by;
      P.bin0 = bin0;
      bin0 += P.nbx * P.nby;
      P.halo0 = halo0;
      if (e.spread_halo)
        halo0 += (long)P.nbx * P.nby * (P.bx + S.order - 1) * (P.by + S.order - 1) * S.nz;
    } else {
      MM_REQ(S.kmode == 2, "kmode must be 0 (Ewald), 1 (P3M) or 2 (none)");
    }
    P.dt = 0.001;
    e.hsp.push_back(P);
    e.lj_tail.push_back(S.include_lj_correction ? lj_tail(S) : 0.0);
    e.maxg = std::max(e.maxg, S.ngroup);
    e.maxs = std::max(e.maxs, S.nsite);
    site0 += S.nsite;
    group0 += S.ngroup;
  }
  e.nsite = site0;
  e.ngroup = group0;
  e.ncell_total = cell0;
  e.nk_total = k0;
  e.mesh_total = mesh0;
  e.cmesh_total = cmesh0;
  e.green_total = green0;
  e.nbin_total = bin0;
  e.halo_total = halo0;
  e.nbcap = (maxnb + 15) / 16 * 16;
  // partial-sum slots: finish blocks and x-pass blocks
  int nb = (e.maxg + 63) / 64;
  for (auto &P : e.hsp)
    if (P.kmode == 1)
      nb = std::max(nb, (P.ny * (P.nz / 2 + 1) + P3M_FFT_LINES - 1) / P3M_FFT_LINES);
  e.nblk_part = nb;
  for (auto &P : e.hsp) {
    P.water = P.ngroup > 0;
    for (int g = P.group0; g < P.group0 + P.ngroup; g++)
      if (gns[g] != 3 || (ginfo[g] & 0x030303) != WATER_MASKS) P.water = 0;
    e.any_water |= P.water;
  }
  return 0;
}

// Every system's slice of each concatenated buffer lies inside the buffer,
// and the slices are disjoint.
int engine_check_layout(const Engine &e, const HostLayout &h) {
  struct Slice { long at, n; int sys; };
  auto check = [](const char *what, std::vector<Slice> v, long total) -> int {
    std::sort(v.begin(), v.end(), [](const Slice &a, const Slice &b) { return a.at < b.at; });
    long end = 0;
    for (auto &x : v) {
      if (x.n == 0) continue;
      if (x.at < end || x.at + x.n > total || x.n < 0) {
        set_last_error(std::string("layout: ") + what + " of system " + std::to_string(x.sys) +
                       " [" + std::to_string(x.at) + ", " + std::to_string(x.at + x.n) +
                       ") overlaps another or leaves the buffer of " + std::to_string(total));
        return -1;
      }
      end = x.at + x.n;
    }
    return 0;
  };
  std::vector<Slice> site, group, cell, st, tb, lj, kv, mesh, cmesh, gr, tw, bin, halo;
  for (int s = 0; s < e.nsys; s++) {
    const SysParams &P = e.hsp[s];
    site.push_back({P.site0, P.nsite, s});
    group.push_back({P.group0, P.ngroup, s});
    cell.push_back({P.cell0, (long)P.ncell * P.ncell * P.ncell, s});
    st.push_back({P.sten0, P.nsten, s});
    tb.push_back({P.tab0, P.ntabA + P.ntabB, s});
    lj.push_back({P.lj0, (long)P.ntype * P.ntype, s});
    if (P.kmode == 0) kv.push_back({4L * P.k0, 4L * P.nk, s});
    if (P.kmode == 1) {
      const long nzh = P.nz / 2 + 1, ncomp = P.ik ? 4 : 1;
      mesh.push_back({P.mesh0, ncomp * P.nx * P.ny * P.nz, s});
      cmesh.push_back({P.cmesh0, ncomp * P.nx * P.ny * nzh, s});
      gr.push_back({P.green0, (long)P.nx * P.ny * nzh, s});
      tw.push_back({P.twid0, 3L * 128, s});
      bin.push_back({P.bin0, (long)P.nbx * P.nby, s});
      const long hpts = (long)(P.bx + P.order - 1) * (P.by + P.order - 1) * P.nz;
      if (e.spread_halo) halo.push_back({P.halo0, (long)P.nbx * P.nby * hpts, s});
      if (P.bx * P.nbx != P.nx || P.by * P.nby != P.ny || P.bx * P.by * P.nz > 4096 ||
          (e.spread_halo && hpts > SPREAD_MAXHALO)) {
        set_last_error("layout: spreading blocks of system " + std::to_string(s) +
                       " do not tile its mesh");
        return -1;
      }
    }
  }
  MM_TRY(check("sites", site, e.nsite));
  MM_TRY(check("groups", group, e.ngroup));
  MM_TRY(check("cells", cell, e.ncell_total));
  MM_TRY(check("stencil", st, (long)h.sten.size()));
  MM_TRY(check("erfc table", tb, (long)h.tab.size()));
  MM_TRY(check("LJ table", lj, (long)h.ljt.size()));
  MM_TRY(check("wave vectors", kv, (long)h.kvec.size()));
  MM_TRY(check("mesh", mesh, e.mesh_total));
  MM_TRY(check("half spectrum", cmesh, e.cmesh_total));
  MM_TRY(check("influence function", gr, (long)h.green.size()));
  MM_TRY(check("twiddles", tw, (long)h.twid.size()));
  MM_TRY(check("spreading bins", bin, e.nbin_total));
  MM_TRY(check("spreading halos", halo, e.halo_total));
  MM_REQ((long)h.kvec.size() == 4L * e.nk_total, "layout: wave vector count");
  MM_REQ((long)h.gfirst.size() == e.ngroup && (long)h.q.size() == e.nsite, "layout: site tables");
  return 0;
}

int engine_create(Engine &e, const std::vector<SystemInput> &in) {
  HostLayout h;
  MM_TRY(engine_layout(e, in, h));
  const std::vector<int> &gfirst = h.gfirst, &gns = h.gns, &gsys = h.gsys, &ginfo = h.ginfo,
                         &gexcl = h.gexcl;
  const std::vector<double> &q = h.q, &mass = h.mass;
  const std::vector<float4> &tab = h.tab;
  const std::vector<float2> &ljt = h.ljt;
  const std::vector<int4> &sten = h.sten;
  const std::vector<double> &kvec = h.kvec;
  const std::vector<float> &green = h.green;
  const std::vector<float2> &twid = h.twid;
  const int N = e.nsite, G = e.ngroup;

  MM_HIP(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
  // MOLMODEL_AMD_SERIAL_KSPACE=1 keeps the reciprocal-space branch on the
  // main stream (for timing the two arrangements against each other)
  const char *serial = getenv("MOLMODEL_AMD_SERIAL_KSPACE");
  if (!(serial && serial[0] == '1'))
    MM_HIP(hipStreamCreateWithFlags(&e.stream2, hipStreamNonBlocking));
  MM_HIP(hipEventCreateWithFlags(&e.ev_fork, hipEventDisableTiming));
  MM_HIP(hipEventCreateWithFlags(&e.ev_join, hipEventDisableTiming));
  MM_TRY(dalloc(&e.d_sp, e.nsys));
  MM_TRY(dput(e.d_sp, e.hsp));
  MM_TRY(dalloc(&e.d_gfirst, G)); MM_TRY(dput(e.d_gfirst, gfirst));
  MM_TRY(dalloc(&e.d_gns, G)); MM_TRY(dput(e.d_gns, gns));
  MM_TRY(dalloc(&e.d_gsys, G)); MM_TRY(dput(e.d_gsys, gsys));
  MM_TRY(dalloc(&e.d_gexcl, G)); MM_TRY(dput(e.d_gexcl, gexcl));
  MM_TRY(dalloc(&e.d_ginfo, G)); MM_TRY(dput(e.d_ginfo, ginfo));
  MM_TRY(dalloc(&e.d_q, N)); MM_TRY(dput(e.d_q, q));
  MM_TRY(dalloc(&e.d_mass, N)); MM_TRY(dput(e.d_mass, mass));
  MM_TRY(dalloc(&e.d_r, 3 * (size_t)N));
  MM_TRY(dalloc(&e.d_v, 3 * (size_t)N));
  MM_TRY(dalloc(&e.d_f, 3 * (size_t)N));
  MM_TRY(dalloc(&e.d_phi, (size_t)N));
  MM_TRY(dalloc(&e.d_gcell, G));
  MM_TRY(dalloc(&e.d_cell_count, e.ncell_total + 1));
  MM_TRY(dalloc(&e.d_cell_start, e.ncell_total + 1));
  MM_TRY(dalloc(&e.d_cell_fill, e.ncell_total + 1));
  {
    std::vector<int> rank_all, unrank_all, rk, ur;
    for (auto &P : e.hsp) {
      hilbert_cell_order(P.ncell, rk, ur);
      rank_all.insert(rank_all.end(), rk.begin(), rk.end());
      unrank_all.insert(unrank_all.end(), ur.begin(), ur.end());
    }
    MM_TRY(dalloc(&e.d_cell_rank, rank_all.size())); MM_TRY(dput(e.d_cell_rank, rank_all));
    MM_TRY(dalloc(&e.d_cell_unrank, unrank_all.size())); MM_TRY(dput(e.d_cell_unrank, unrank_all));
  }
  MM_TRY(dalloc(&e.d_order, G));
  {
    std::vector<int4> sl(std::max(G, 1), make_int4(0, 0, 0, 0));
    for (int g = 0; g < G; g++) sl[g] = make_int4(g, gfirst[g], gns[g], gexcl[g]);
    MM_TRY(dalloc(&e.d_gslot, sl.size()));
    MM_TRY(dput(e.d_gslot, sl));
    std::vector<int> inv(std::max(G, 1));
    for (int g = 0; g < G; g++) inv[g] = g;
    MM_TRY(dalloc(&e.d_gslot_inv, inv.size()));
    MM_TRY(dput(e.d_gslot_inv, inv));
  }
  MM_TRY(dalloc(&e.d_last, 3 * (size_t)G));
  MM_TRY(dalloc(&e.d_sten, sten.size())); MM_TRY(dput(e.d_sten, sten));
  MM_TRY(dalloc(&e.d_nbr, (size_t)G * e.nbcap));
  MM_TRY(dalloc(&e.d_nnb, G));
  MM_TRY(dalloc(&e.d_rebuild, e.nsys));
  MM_TRY(dalloc(&e.d_err, 1));
  MM_TRY(dalloc(&e.d_nlb, e.nsys));
  MM_TRY(dalloc(&e.d_gpos, G));
  MM_TRY(dalloc(&e.d_spk, (size_t)G * MAXS));
  MM_TRY(dalloc(&e.d_gmeta, G));
  MM_TRY(dalloc(&e.d_pf, (size_t)G * MAXS * 4));
  MM_TRY(dalloc(&e.d_plj, (size_t)G * MAXS * 3));
  MM_TRY(dalloc(&e.d_gacc, (size_t)G * 10));
  MM_TRY(dalloc(&e.d_tab, tab.size())); MM_TRY(dput(e.d_tab, tab));
  MM_TRY(dalloc(&e.d_ljt, ljt.size())); MM_TRY(dput(e.d_ljt, ljt));
  MM_TRY(dalloc(&e.d_kres, 4 * (size_t)N + 9 * e.nsys));
  MM_TRY(dalloc(&e.d_kvec, kvec.size())); MM_TRY(dput(e.d_kvec, kvec));
  MM_TRY(dalloc(&e.d_ksc, 2 * (size_t)e.nk_total));
  MM_TRY(dalloc(&e.d_mesh, e.mesh_total));
  MM_TRY(dalloc(&e.d_cmesh, e.cmesh_total));
  MM_TRY(dalloc(&e.d_green, green.size())); MM_TRY(dput(e.d_green, green));
  MM_TRY(dalloc(&e.d_twid, twid.size())); MM_TRY(dput(e.d_twid, twid));
  MM_TRY(dalloc(&e.d_sgrid, N));
  MM_TRY(dalloc(&e.d_sfrac, N));
  MM_TRY(dalloc(&e.d_bin_fill, e.nbin_total + 1));
  MM_TRY(dalloc(&e.d_halo, (size_t)e.halo_total));
  e.bcap = 32;
  for (auto &P : e.hsp)
    if (P.kmode == 1) e.bcap = std::max(e.bcap, p3m_bin_capacity(P.nsite, P.nbx * P.nby));
  MM_TRY(dalloc(&e.d_bgrid, (size_t)std::max(e.nbin_total, 1) * e.bcap));
  MM_TRY(dalloc(&e.d_bfrac, (size_t)std::max(e.nbin_total, 1) * e.bcap));
  MM_TRY(dalloc(&e.d_acc, (size_t)e.nsys * NACC));
  MM_TRY(dalloc(&e.d_part, (size_t)e.nsys * e.nblk_part * 16));
  MM_TRY(dalloc(&e.d_fpart, (size_t)e.nsys * e.nblk_part * 16));
  MM_TRY(dalloc(&e.d_done, 4 * (size_t)e.nsys));
  {
    // persistent pair grid: pair_slots workgroups per CU (the kernel's
    // occupancy allows 4; the default 3 leaves one slot on every CU to the
    // reciprocal-space kernels of the other stream)
    std::vector<int> ts(e.nsys + 1, 0);
    for (int s = 0; s < e.nsys; s++) ts[s + 1] = ts[s] + pair_tiles(e.hsp[s].ngroup);
    e.ntile = ts[e.nsys];
    int dev = 0, ncu = 256;
    MM_HIP(hipGetDevice(&dev));
    MM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int want = std::min(e.pair_slots * ncu, (e.ntile + 7) / 8 * 8);
    e.pair_blocks = std::max(8, want / 8 * 8);
    MM_TRY(dalloc(&e.d_tile_start, e.nsys + 1));
    MM_TRY(dput(e.d_tile_start, ts));
  }
  {
    std::vector<int> cs(e.ncell_total + 1, 0);
    cs[e.ncell_total] = G;
    MM_HIP(hipMemcpy(e.d_cell_start + e.ncell_total, &cs[e.ncell_total], sizeof(int),
                     hipMemcpyHostToDevice));
  }
  MM_HIP(hipDeviceSynchronize());
  return 0;
}

// ElecLJ::scale (eleclj.cpp:1268-1280) for one system: the box grows by f,
// eta shrinks by f (so eta L is kept), the erfc table is rebuilt for the new
// eta, the P3M influence function scales as G/f (p3m_scale, pppm.c:643-664;
// exact, since G depends on k/eta and 1/V), the Ewald wave vectors and
// coefficients and the neighbour-list cells (set_boundary_conditions,
// nbrlist.cpp:84-123) follow the new box.  Positions are the caller's.
int engine_scale_system(Engine &e, int s, double f) {
  MM_REQ(s >= 0 && s < e.nsys, "bad system index");
  MM_REQ(f > 0 && std::isfinite(f), "scale factor must be positive");
  SystemInput &S = e.in[s];
  SysParams &P = e.hsp[s];
  MM_REQ(!S.has_latv, "box scaling covers cubic boxes");
  const double L = S.boxl * f;
  MM_REQ(S.rc <= 0.49 * L, "rspace cutoff must stay <= 0.49 * box length (eleclj.cpp:164-170)");
  S.boxl = L;
  S.eta /= f;
  P.L = L;
  P.invL = 1 / L;
  P.volume = L * L * L;
  P.eta = S.eta;
  P.self_coef = S.eta * 1.12837916709551257390;
  double qsum = 0;
  for (double x : S.q) qsum += x;
  P.bg_phi = fabs(qsum) < 64.0 * 2.220446049250313e-16 ? 0.0 : -M_PI / (sq(S.eta) * P.volume) * qsum;
  P.lc = L * (1.0 / P.ncell);
  // erfc table for the new eta: same intervals (they depend on rc only)
  {
    std::vector<float4> tab;
    SysParams T = P;
    build_table(T, S.rc, S.eta, tab);
    MM_REQ(T.ntabA == P.ntabA && T.ntabB == P.ntabB, "erfc table size changed");
    P.tabAmin = T.tabAmin; P.tabMid = T.tabMid; P.tabAinv = T.tabAinv; P.tabBinv = T.tabBinv;
    P.tabBmax = T.tabBmax;
    MM_HIP(hipMemcpy(e.d_tab + P.tab0, tab.data(), sizeof(float4) * tab.size(), hipMemcpyHostToDevice));
  }
  const double latv[9] = {L, 0, 0, 0, L, 0, 0, 0, L};
  double inv[9];
  tensor_inverse(inv, latv);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) P.RL[3 * a + b] = 2 * M_PI * inv[3 * b + a];
  if (P.kmode == 0) {   // kspace.c:96-103 with the new box and eta
    std::vector<double> kv(4 * (size_t)P.nk);
    MM_HIP(hipMemcpy(kv.data(), e.d_kvec + 4L * P.k0, sizeof(double) * kv.size(), hipMemcpyDeviceToHost));
    const double c = 0.25 / (S.eta * S.eta), fourpi_V = 4 * M_PI / P.volume;
    for (int k = 0; k < P.nk; k++) {
      double *v = &kv[4 * (size_t)k];
      v[0] /= f; v[1] /= f; v[2] /= f;
      const double t = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      v[3] = fourpi_V * exp(-t * c) / t;
    }
    MM_HIP(hipMemcpy(e.d_kvec + 4L * P.k0, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice));
  } else if (P.kmode == 1) {
    for (int a = 0; a < 9; a++) P.T[a] = inv[a];
    P.T[0] *= P.nx; P.T[3] *= P.nx; P.T[6] *= P.nx;
    P.T[1] *= P.ny; P.T[4] *= P.ny; P.T[7] *= P.ny;
    P.T[2] *= P.nz; P.T[5] *= P.nz; P.T[8] *= P.nz;
    const size_t ng = (size_t)P.nx * P.ny * (P.nz / 2 + 1);
    std::vector<float> g(ng);
    MM_HIP(hipMemcpy(g.data(), e.d_green + P.green0, sizeof(float) * ng, hipMemcpyDeviceToHost));
    for (auto &x : g) x = (float)((double)x / f);
    MM_HIP(hipMemcpy(e.d_green + P.green0, g.data(), sizeof(float) * ng, hipMemcpyHostToDevice));
  }
  e.lj_tail[s] = S.include_lj_correction ? lj_tail(S) : 0.0;
  // neighbour-list stencils of every system (this one may change length)
  std::vector<int4> sten;
  for (int t = 0; t < e.nsys; t++) {
    SysParams &Q = e.hsp[t];
    const SystemInput &U = e.in[t];
    auto st = nl_stencil(Q.ncell, U.boxl, U.rc, U.skin);
    Q.sten0 = (int)sten.size();
    Q.nsten = (int)st.size();
    sten.insert(sten.end(), st.begin(), st.end());
  }
  MM_HIP(hipFree(e.d_sten));
  e.d_sten = nullptr;
  MM_TRY(dalloc(&e.d_sten, sten.size()));
  MM_TRY(dput(e.d_sten, sten));
  MM_TRY(dput(e.d_sp, e.hsp));
  e.first_eval = 1;   // cells changed: rebuild every list
  return 0;
}

// New charges and LJ parameters for the sites of system s (site masks,
// groups and exclusions stay): LJ types and the type-pair table of every
// system are rebuilt (a system's type count may change), the tail
// correction and the neutralising background follow, and the next
// evaluation packs the new charges.  This is what ElecLJ::set_lambda
// leaves behind (eleclj.cpp:273-308) for a new set of site parameters.
int engine_set_params(Engine &e, int s, const double *q, const double *sig, const double *eps) {
  MM_REQ(s >= 0 && s < e.nsys, "bad system index");
  SystemInput &S = e.in[s];
  for (int i = 0; i < S.nsite; i++)
    MM_REQ(std::isfinite(q[i]) && std::isfinite(sig[i]) && std::isfinite(eps[i]) && eps[i] >= 0,
           "site parameters must be finite, epsilon >= 0");
  const SystemInput keep = S;
  S.q.assign(q, q + S.nsite);
  S.sig.assign(sig, sig + S.nsite);
  S.eps.assign(eps, eps + S.nsite);
  std::vector<float2> ljt;
  std::vector<int> ntype(e.nsys), lj0(e.nsys), stype_s;
  for (int t = 0; t < e.nsys; t++) {
    std::vector<int> st;
    lj0[t] = (int)ljt.size();
    if (int rc = lj_types(e.in[t], st, ljt, ntype[t])) {
      S = keep;
      return rc;
    }
    if (t == s) stype_s = st;
  }
  for (int t = 0; t < e.nsys; t++) {
    e.hsp[t].ntype = ntype[t];
    e.hsp[t].lj0 = lj0[t];
  }
  SysParams &P = e.hsp[s];
  std::vector<int> ginfo(S.ngroup);
  for (int g = 0; g < S.ngroup; g++) ginfo[g] = group_info(S, stype_s, g);
  MM_HIP(hipStreamSynchronize(e.stream));
  MM_HIP(hipMemcpy(e.d_ginfo + P.group0, ginfo.data(), sizeof(int) * S.ngroup, hipMemcpyHostToDevice));
  MM_HIP(hipMemcpy(e.d_q + P.site0, S.q.data(), sizeof(double) * S.nsite, hipMemcpyHostToDevice));
  MM_HIP(hipFree(e.d_ljt));
  e.d_ljt = nullptr;
  MM_TRY(dalloc(&e.d_ljt, ljt.size()));
  MM_TRY(dput(e.d_ljt, ljt));
  P.bg_phi = background_phi(S, S.eta, P.volume);
  e.lj_tail[s] = S.include_lj_correction ? lj_tail(S) : 0.0;
  if (e.d_tail)
    MM_HIP(hipMemcpy(e.d_tail, e.lj_tail.data(), sizeof(double) * e.nsys, hipMemcpyHostToDevice));
  MM_TRY(dput(e.d_sp, e.hsp));
  e.need_pack = 1;
  engine_drop_graph(e);   // the table may have moved
  return 0;
}

void engine_drop_graph(Engine &e) {
  if (e.gexec) { hipGraphExecDestroy(e.gexec); e.gexec = nullptr; }
  if (e.graph) { hipGraphDestroy(e.graph); e.graph = nullptr; }
  if (e.gexecK) { hipGraphExecDestroy(e.gexecK); e.gexecK = nullptr; }
  if (e.graphK) { hipGraphDestroy(e.graphK); e.graphK = nullptr; }
  e.graphK_steps = 0;
}

int engine_grow_bins(Engine &e) {
  MM_HIP(hipStreamSynchronize(e.stream));
  if (e.stream2) MM_HIP(hipStreamSynchronize(e.stream2));
  int need = 0, zero = 0;
  MM_HIP(hipMemcpy(&need, e.d_bin_fill + e.nbin_total, sizeof(int), hipMemcpyDeviceToHost));
  MM_HIP(hipMemcpy(e.d_bin_fill + e.nbin_total, &zero, sizeof(int), hipMemcpyHostToDevice));
  // a bin never holds more sites than its system has
  int most = 32;
  for (auto &P : e.hsp)
    if (P.kmode == 1) most = std::max(most, P.nsite);
  most = (most + 31) / 32 * 32;
  long want = std::max(2L * e.bcap, (long)need + need / 4 + 32);
  want = std::min<long>((want + 31) / 32 * 32, most);
  MM_REQ(want > e.bcap, "P3M column bins cannot grow further");
  const size_t nb = (size_t)std::max(e.nbin_total, 1);
  MM_HIP(hipFree(e.d_bgrid));
  MM_HIP(hipFree(e.d_bfrac));
  e.d_bgrid = nullptr;
  e.d_bfrac = nullptr;
  e.bcap = (int)want;
  MM_TRY(dalloc(&e.d_bgrid, nb * e.bcap));
  MM_TRY(dalloc(&e.d_bfrac, nb * e.bcap));
  engine_drop_graph(e);
  return 0;
}

void engine_destroy(Engine &e) {
  engine_drop_graph(e);
  void *ptrs[] = {e.d_sp, e.d_gfirst, e.d_gns, e.d_gsys, e.d_gexcl, e.d_ginfo, e.d_q, e.d_mass,
                  e.d_r, e.d_v, e.d_f, e.d_r0, e.d_gcell, e.d_cell_count, e.d_cell_start,
                  e.d_cell_rank, e.d_cell_unrank,
                  e.d_cell_fill, e.d_order, e.d_gslot, e.d_gslot_inv, e.d_last, e.d_sten, e.d_nbr,
                  e.d_nnb, e.d_rebuild,
                  e.d_err, e.d_gpos, e.d_spk, e.d_gmeta, e.d_pf, e.d_plj, e.d_gacc, e.d_tab,
                  e.d_ljt, e.d_kres, e.d_kvec, e.d_ksc, e.d_mesh, e.d_cmesh, e.d_green, e.d_twid,
                  e.d_sgrid, e.d_sfrac, e.d_bin_fill, e.d_halo,
                  e.d_bgrid, e.d_bfrac, e.d_acc, e.d_part, e.d_fpart, e.d_done, e.d_phi, e.d_scratch,
                  e.d_nlb, e.d_ladder, e.d_kT, e.d_tail, e.d_rtmp, e.d_mom, e.d_perm,
                  e.d_remd_stats, e.d_tile_start, e.d_snap, e.d_snap_perm, e.d_snap_stats,
                  e.d_adjR, e.d_chi, e.d_adjpart, e.d_adjksc, e.d_adjkres, e.d_adjacc,
                  e.d_adjsrc, e.d_adjq};
  for (void *p : ptrs)
    if (p) hipFree(p);
  if (e.stream) hipStreamDestroy(e.stream);
  if (e.stream2) hipStreamDestroy(e.stream2);
  if (e.ev_fork) hipEventDestroy(e.ev_fork);
  if (e.ev_join) hipEventDestroy(e.ev_join);
  e = Engine();
}

}  // namespace mmamd
