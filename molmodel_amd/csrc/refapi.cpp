// refapi.cpp -- reference-named entry points (include/molmodel_amd_ref.h).
//
// PairBinary (pbin.c) and the Chebyshev table builder (cheby.c,
// rspace.c:25-37) are host code, as in the reference.  RSpaceEwald,
// kspace_field/kspace_phi and p3m_field run on the GPU through the same
// kernels as the engine.  Unsupported inputs abort with a message, the way
// the reference's asserts and insist() do.
#include "engine.hpp"
#include "../../include/molmodel_amd_ref.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace mmamd;

static void die_msg(const char *where, const char *msg) {
  fprintf(stderr, "molmodel_amd: %s: %s\n", where, msg);
  fflush(stderr);
  abort();
}

#define HIPCHK(x)                                                         \
  do {                                                                    \
    hipError_t e_ = (x);                                                  \
    if (e_ != hipSuccess) die_msg(__func__, hipGetErrorString(e_));       \
  } while (0)

// ------------------------------------------------------------ PairBinary
struct PairBinary {
  int n;
  int *nn, *cap;
  int **t;
};

extern "C" {

struct PairBinary *PairBinary_new(int n) {
  PairBinary *p = (PairBinary *)malloc(sizeof(PairBinary));
  p->n = n;
  p->nn = (int *)calloc(n > 0 ? n : 1, sizeof(int));
  p->cap = (int *)calloc(n > 0 ? n : 1, sizeof(int));
  p->t = (int **)calloc(n > 0 ? n : 1, sizeof(int *));
  return p;
}

void PairBinary_delete(struct PairBinary *p) {
  if (!p) return;
  for (int i = 0; i < p->n; i++) free(p->t[i]);
  free(p->t); free(p->nn); free(p->cap); free(p);
}

void PairBinary_quick_add(struct PairBinary *p, int i, int j) {   // pbin.c:53-61
  if (p->nn[i] == p->cap[i]) {
    p->cap[i] = p->cap[i] ? 2 * p->cap[i] : 16;
    p->t[i] = (int *)realloc(p->t[i], sizeof(int) * p->cap[i]);
  }
  p->t[i][p->nn[i]++] = j;
}

void PairBinary_add(struct PairBinary *p, int i, int j) {          // pbin.c:41-50
  canonical_pair(i, j);
  PairBinary_quick_add(p, i, j);
}

void PairBinary_init(struct PairBinary *p) {                        // pbin.c:73-90
  for (int i = 0; i < p->n; i++) {
    int *r = p->t[i];
    const int n = p->nn[i];
    if (n < 2) continue;
    std::sort(r, r + n);
    p->nn[i] = (int)(std::unique(r, r + n) - r);
  }
}

void PairBinary_clear(struct PairBinary *p) {
  for (int i = 0; i < p->n; i++) p->nn[i] = 0;
}

const int *PairBinary_number_of_elements(const struct PairBinary *p) { return p->nn; }
int *const *PairBinary_elements(const struct PairBinary *p) { return p->t; }
const int *PairBinary_start(const struct PairBinary *p, int i) { return p->t[i]; }
const int *PairBinary_end(const struct PairBinary *p, int i) { return p->t[i] + p->nn[i]; }

int PairBinary_exists(const struct PairBinary *p, int i, int j) {   // pbin.c:119-137
  canonical_pair(i, j);
  const int *r = p->t[i];
  return r ? (int)std::binary_search(r, r + p->nn[i], j) : 0;
}

void PairBinary_show(const struct PairBinary *p) {
  for (int i = 0; i < p->n; i++) {
    printf("%5d: ", i);
    for (int k = 0; k < p->nn[i]; k++) printf("%d ", p->t[i][k]);
    printf("\n");
  }
  fflush(stdout);
}

// ------------------------------------------------------------ Chebyshev
void Chebyshev_init(struct Chebyshev *t, double min, double max, double (*f)(double)) {
  if (!(max > min)) die_msg("Chebyshev_init", "max must exceed min");
  const double dx = (max - min) / NCHEBY;
  t->min = min;
  t->max = max;
  t->inv_dx = 1.0 / dx;
  for (int i = 0; i < NCHEBY; i++) {
    const double a = i * dx + min, b = (i + 1) * dx + min;
    double *c = t->coeff[i];
    cheb_fit(a, b, f, c);
    c[4] = -2 * c[1] / (b - a);         // -2 d/dx of the cubic (cheby.c:30-32)
    c[5] = -2 * 2 * c[2] / (b - a);
    c[6] = -2 * 3 * c[3] / (b - a);
  }
}

static void cheb_eval(const struct Chebyshev *t, double x, double *f, double *n2df) {
  x -= t->min;
  x *= t->inv_dx;
  const unsigned i = (unsigned)x;
  const double *c = t->coeff[i];
  x -= i;
  *f = ((c[3] * x + c[2]) * x + c[1]) * x + c[0];
  *n2df = (c[6] * x + c[5]) * x + c[4];
}

void Chebyshev_test(const struct Chebyshev *t, double (*f)(double), double *maxerr, double *xerr) {
  *maxerr = *xerr = 0;
  for (int i = 0; i < NCHEBY - 1; i++) {     // cheby.c:47-62
    const double x = (i + drand48()) / t->inv_dx + t->min;
    double a, da;
    cheb_eval(t, x, &a, &da);
    a = fabs(a - f(x));
    if (a > *maxerr) { *xerr = x; *maxerr = a; }
  }
}

static thread_local double cheb_eta_;
static double erfc_r(double r2) {
  const double r = sqrt(r2);
  return erfc(cheb_eta_ * r) / r;
}

struct Chebyshev *RSpaceChebyshev_new(double rhi, double eta, double *maxerr, double *xerr) {
  if (!(rhi > 0) || !(eta > 0)) die_msg("RSpaceChebyshev_new", "rhi and eta must be positive");
  struct Chebyshev *t = (struct Chebyshev *)malloc(sizeof(struct Chebyshev));
  const double r2max = 1.2 * (rhi * rhi);               // rspace.c:30-33
  const double r2min = r2max > 4 ? 4 : 0.9 * r2max;
  cheb_eta_ = eta;
  Chebyshev_init(t, r2min, r2max, erfc_r);
  if (maxerr && xerr) Chebyshev_test(t, erfc_r, maxerr, xerr);
  return t;
}

}  // extern "C"

// ------------------------------------------------------------ shared helpers
static void to_soa(const double *aos, int n, std::vector<double> &soa) {
  soa.resize(3 * (size_t)n);
  for (int i = 0; i < n; i++)
    for (int c = 0; c < 3; c++) soa[(size_t)c * n + i] = aos[3 * (size_t)i + c];
}

// ------------------------------------------------------------ RSpaceEwald
// the checks RSpaceEwald and RSpaceEwaldPhi share (rspace.c:409-411 and
// the inputs the device path does not cover)
static void rspace_checks(const char *who, int nsite, int ngroup, enum bcondtype_t bctype,
                          const tensor_t *latv, const struct PairBinary *group_excluded,
                          double rlo2, double rhi2, const double *rscreen,
                          const struct PairBinary *one_four) {
  if (bctype == non_periodic) die_msg(who, "bctype != non_periodic (rspace.c:409)");
  if (!(rhi2 >= rlo2) || !(rlo2 > 0)) die_msg(who, "need rhi2 >= rlo2 > 0 (rspace.c:410-411)");
  if (bctype != cubic || latv->yy != latv->xx || latv->zz != latv->xx || latv->xy != 0 ||
      latv->xz != 0 || latv->yx != 0 || latv->yz != 0 || latv->zx != 0 || latv->zy != 0)
    die_msg(who, "the device path covers cubic boxes only");
  for (int i = 0; rscreen && i < nsite; i++)
    if (rscreen[i] != 0) die_msg(who, "screening radii are not supported on the device");
  if (one_four)
    for (int i = 0; i < nsite; i++)
      if (PairBinary_number_of_elements(one_four)[i])
        die_msg(who, "1-4 scaled pairs are not supported on the device");
  if (group_excluded)
    for (int g = 0; g < ngroup; g++)
      if (PairBinary_number_of_elements(group_excluded)[g])
        die_msg(who, "exclusions between groups are not supported on the device");
}

// Real-space sums over the caller's neighbour list on the device: a
// transient single-system engine fed the caller's rows (as full rows),
// the pair kernel in field mode and the excluded-pair loop.  out: 7 per
// site (phi, evec, lj_force) and the LJ energy at out[7 nsite].
static void rspace_on_device(const char *who, int nsite, int ngroup, const int (*group)[2],
                             const cart_t *r, const tensor_t *latv, int *const *neighbors,
                             const int *number_of_neighbors, const struct PairBinary *excluded,
                             double rlo2, double rhi2, const int *site_mask, const double *q,
                             const double *sig, const double *eps, int geometric, double eta,
                             int min_image, std::vector<double> &out) {
  SystemInput S;
  S.nsite = nsite;
  S.ngroup = ngroup;
  for (int g = 0; g < ngroup; g++) {
    S.gfirst.push_back(group[g][0]);
    S.gns.push_back(group[g][1] - group[g][0]);
  }
  S.q.assign(q, q + nsite);
  S.sig.assign(sig, sig + nsite);
  S.eps.assign(eps, eps + nsite);
  S.mask.assign(site_mask, site_mask + nsite);
  if (excluded)
    for (int i = 0; i < nsite; i++)
      for (const int *j = PairBinary_start(excluded, i); j < PairBinary_end(excluded, i); j++) {
        S.excl.push_back(i);
        S.excl.push_back(*j);
      }
  S.boxl = latv->xx;
  S.rc = sqrt(rhi2);
  S.width = S.rc - sqrt(rlo2);
  S.skin = 1e-3;
  S.eta = eta;
  S.kmode = 2;
  S.kc = 0;
  S.nx = S.ny = S.nz = 0;
  S.order = 1;
  S.nalias = 0;
  S.ik = 0;
  S.include_lj_correction = 0;
  S.geometric_sigma = geometric;
  Engine e;
  if (engine_create(e, {S})) die_msg(who, last_error());
  e.external_nl = 1;
  // caller's neighbour rows as full rows in the caller's group order
  std::vector<std::vector<int>> rows(ngroup);
  for (int g = 0; g < ngroup; g++)
    for (int k = 0; k < number_of_neighbors[g]; k++) {
      const int h = neighbors[g][k];
      rows[g].push_back(h);
      rows[h].push_back(g);
    }
  size_t cap = 16;
  for (auto &rr : rows) cap = std::max(cap, rr.size());
  cap = (cap + 15) / 16 * 16;
  std::vector<int> nbr((size_t)ngroup * cap, 0), nnb(ngroup), order(ngroup);
  for (int g = 0; g < ngroup; g++) {
    order[g] = g;
    nnb[g] = (int)rows[g].size();
    std::copy(rows[g].begin(), rows[g].end(), nbr.begin() + (size_t)g * cap);
  }
  HIPCHK(hipFree(e.d_nbr));
  e.nbcap = (int)cap;
  HIPCHK(hipMalloc((void **)&e.d_nbr, sizeof(int) * nbr.size()));
  HIPCHK(hipMemcpy(e.d_nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e.d_nnb, nnb.data(), sizeof(int) * ngroup, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e.d_order, order.data(), sizeof(int) * ngroup, hipMemcpyHostToDevice));
  std::vector<double> soa;
  to_soa((const double *)r, nsite, soa);
  HIPCHK(hipMemcpy(e.d_r, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  out.assign(7 * (size_t)nsite + 1, 0.0);
  double *d_out = nullptr;
  HIPCHK(hipMalloc((void **)&d_out, sizeof(double) * out.size()));
  HIPCHK(hipMemset(d_out, 0, sizeof(double) * out.size()));
  if (launch_pack(e) || launch_pair(e, 1, 0) ||
      launch_field_out(e, d_out, d_out + 7 * (size_t)nsite, min_image))
    die_msg(who, last_error());
  HIPCHK(hipStreamSynchronize(e.stream));
  HIPCHK(hipMemcpy(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipFree(d_out));
  engine_destroy(e);
}

extern "C" void RSpaceEwald(int nsite, int ngroup, const int (*group)[2], const cart_t *r,
                            enum bcondtype_t bctype, const tensor_t *latv,
                            int *const *neighbors, const int *number_of_neighbors,
                            const struct PairBinary *group_excluded,
                            const struct PairBinary *excluded, double rlo2, double rhi2,
                            const int *site_mask, const double *q, double *phi, cart_t *evec,
                            double eta, const struct Chebyshev *rspace_chebyshev,
                            const double *rscreen, const double *sig, const double *eps,
                            int geometric_combining_for_sigma, double *lj_ener, cart_t *lj_force,
                            const struct PairBinary *one_four, double elec_1_4_scale_factor,
                            double lj_1_4_scale_factor) {
  (void)rspace_chebyshev;
  (void)elec_1_4_scale_factor;
  (void)lj_1_4_scale_factor;
  rspace_checks("RSpaceEwald", nsite, ngroup, bctype, latv, group_excluded, rlo2, rhi2, rscreen,
                one_four);
  std::vector<double> out;
  rspace_on_device("RSpaceEwald", nsite, ngroup, group, r, latv, neighbors, number_of_neighbors,
                   excluded, rlo2, rhi2, site_mask, q, sig, eps, geometric_combining_for_sigma,
                   eta, 1, out);
  for (int i = 0; i < nsite; i++) {
    phi[i] += out[7 * (size_t)i];
    evec[i].x += out[7 * (size_t)i + 1];
    evec[i].y += out[7 * (size_t)i + 2];
    evec[i].z += out[7 * (size_t)i + 3];
    lj_force[i].x += out[7 * (size_t)i + 4];
    lj_force[i].y += out[7 * (size_t)i + 5];
    lj_force[i].z += out[7 * (size_t)i + 6];
  }
  *lj_ener += out[7 * (size_t)nsite];
}

// phi only (rspace.c:649-782): the smoothed pair potentials of RSpaceEwald
// and the excluded-pair correction without the minimum image (:747-751)
extern "C" void RSpaceEwaldPhi(int nsite, int ngroup, const int (*group)[2], const cart_t *r,
                               enum bcondtype_t bctype, const tensor_t *latv,
                               int *const *neighbors, const int *number_of_neighbors,
                               const struct PairBinary *group_excluded,
                               const struct PairBinary *excluded, double rlo2, double rhi2,
                               const int *site_mask, const double *q, double *phi, double eta,
                               const struct Chebyshev *rspace_chebyshev, const double *rscreen,
                               const struct PairBinary *one_four, double elec_1_4_scale_factor) {
  (void)rspace_chebyshev;
  (void)elec_1_4_scale_factor;
  rspace_checks("RSpaceEwaldPhi", nsite, ngroup, bctype, latv, group_excluded, rlo2, rhi2,
                rscreen, one_four);
  std::vector<int> qmask(nsite);
  for (int i = 0; i < nsite; i++) qmask[i] = site_mask[i] & SITE_HAS_Q;   // no LJ
  const std::vector<double> zero(nsite, 0.0);
  std::vector<double> out;
  rspace_on_device("RSpaceEwaldPhi", nsite, ngroup, group, r, latv, neighbors,
                   number_of_neighbors, excluded, rlo2, rhi2, qmask.data(), q, zero.data(),
                   zero.data(), 0, eta, 0, out);
  for (int i = 0; i < nsite; i++) phi[i] += out[7 * (size_t)i];
}

// potential at grid points from all sites (rspace.c:892-967)
extern "C" void RSpaceEwaldPhiAtGridpoints(int nsite, int ngrid, const cart_t *rsite,
                                           const cart_t *rgrid, enum bcondtype_t bctype,
                                           const tensor_t *latv, double rlo2, double rhi2,
                                           const int *site_mask, const double *q, double *phi,
                                           double eta, const struct Chebyshev *rspace_chebyshev,
                                           const double *rscreen) {
  (void)rspace_chebyshev;
  const char *who = "RSpaceEwaldPhiAtGridpoints";
  rspace_checks(who, nsite, 0, bctype, latv, nullptr, rlo2, rhi2, rscreen, nullptr);
  if (ngrid <= 0 || nsite <= 0) return;
  // a one-site-per-group system supplies the table and box parameters
  SystemInput S;
  S.nsite = nsite;
  S.ngroup = nsite;
  for (int i = 0; i < nsite; i++) {
    S.gfirst.push_back(i);
    S.gns.push_back(1);
  }
  S.q.assign(nsite, 0.0);
  S.sig.assign(nsite, 0.0);
  S.eps.assign(nsite, 0.0);
  S.mask.assign(nsite, 0);
  S.boxl = latv->xx;
  S.rc = sqrt(rhi2);
  S.width = S.rc - sqrt(rlo2);
  S.skin = 1e-3;
  S.eta = eta;
  S.kmode = 2;
  S.kc = 0;
  S.nx = S.ny = S.nz = 0;
  S.order = 1;
  S.nalias = 0;
  S.ik = 0;
  S.include_lj_correction = 0;
  S.geometric_sigma = 0;
  Engine e;
  if (engine_create(e, {S})) die_msg(who, last_error());
  std::vector<double4> sites(nsite);
  for (int i = 0; i < nsite; i++)
    sites[i] = make_double4(rsite[i].x, rsite[i].y, rsite[i].z, site_mask[i] ? q[i] : 0.0);
  double4 *d_sites = nullptr;
  double *d_pts = nullptr, *d_phi = nullptr;
  HIPCHK(hipMalloc((void **)&d_sites, sizeof(double4) * nsite));
  HIPCHK(hipMalloc((void **)&d_pts, sizeof(double) * 3 * (size_t)ngrid));
  HIPCHK(hipMalloc((void **)&d_phi, sizeof(double) * ngrid));
  HIPCHK(hipMemcpy(d_sites, sites.data(), sizeof(double4) * nsite, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_pts, rgrid, sizeof(double) * 3 * (size_t)ngrid, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(d_phi, 0, sizeof(double) * ngrid));
  if (launch_phi_points(e, d_sites, nsite, d_pts, ngrid, d_phi)) die_msg(who, last_error());
  HIPCHK(hipStreamSynchronize(e.stream));
  std::vector<double> out(ngrid);
  HIPCHK(hipMemcpy(out.data(), d_phi, sizeof(double) * ngrid, hipMemcpyDeviceToHost));
  hipFree(d_sites);
  hipFree(d_pts);
  hipFree(d_phi);
  engine_destroy(e);
  for (int j = 0; j < ngrid; j++) phi[j] += out[j];
}

// ------------------------------------------------------------ non-periodic
// RSpaceNonPeriodic / RSpaceNonPeriodicPhi / RSpaceNonPeriodicPhiAtGridpoints
// (rspace.c:560-640, 784-900) on the device in FP64: all pairs of sites in
// different groups, no cutoff.  Exclusions between groups and 1-4 scaled
// pairs abort as in the periodic routines (pairs inside a group are never
// evaluated here, as in the reference).
namespace {

struct DevBuf {
  void *p = nullptr;
  explicit DevBuf(size_t bytes) { HIPCHK(hipMalloc(&p, bytes ? bytes : 8)); }
  ~DevBuf() { hipFree(p); }
  template <class T> T *as() { return static_cast<T *>(p); }
};

void nonperiodic_checks(const char *who, int nsite, int ngroup,
                        const struct PairBinary *group_excluded,
                        const struct PairBinary *one_four) {
  if (nsite < 0 || ngroup < 0) die_msg(who, "nsite and ngroup must be >= 0");
  if (group_excluded)
    for (int g = 0; g < ngroup; g++)
      if (PairBinary_number_of_elements(group_excluded)[g])
        die_msg(who, "exclusions between groups are not supported on the device");
  if (one_four)
    for (int i = 0; i < nsite; i++)
      if (PairBinary_number_of_elements(one_four)[i])
        die_msg(who, "1-4 scaled pairs are not supported on the device");
}

// 7 doubles per site (phi, evec, lj_force) and the LJ energy
void nonperiodic_on_device(const char *who, int mode, int nsite, int ngroup,
                           const int (*group)[2], const cart_t *r, const int *site_mask,
                           const double *q, const double *rscreen, const double *sig,
                           const double *eps, int geometric, std::vector<double> &out,
                           double &ulj) {
  std::vector<double4> site(nsite);
  std::vector<int4> info(nsite, make_int4(-1, 0, 0, 0));
  std::vector<double2> lj(nsite, make_double2(0, 0));
  std::vector<double> rs(nsite, 0.0);
  for (int g = 0; g < ngroup; g++)
    for (int i = group[g][0]; i < group[g][1]; i++) {
      if (i < 0 || i >= nsite) die_msg(who, "group out of range");
      info[i] = make_int4(g, site_mask[i], i == group[g][0] ? 1 : 0, 0);
    }
  for (int i = 0; i < nsite; i++) {
    if (info[i].x < 0) die_msg(who, "site not in any group");
    site[i] = make_double4(r[i].x, r[i].y, r[i].z, q[i]);
    if (sig && eps) lj[i] = make_double2(sig[i], eps[i]);
    if (rscreen) rs[i] = rscreen[i];
  }
  const int nblk = (nsite + 255) / 256;
  DevBuf ds(sizeof(double4) * nsite), di(sizeof(int4) * nsite), dl(sizeof(double2) * nsite),
      dr(sizeof(double) * nsite), dout(sizeof(double) * 7 * nsite), du(sizeof(double) * nblk);
  HIPCHK(hipMemcpy(ds.p, site.data(), sizeof(double4) * nsite, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(di.p, info.data(), sizeof(int4) * nsite, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dl.p, lj.data(), sizeof(double2) * nsite, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dr.p, rs.data(), sizeof(double) * nsite, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(dout.p, 0, sizeof(double) * 7 * nsite));
  if (launch_nonperiodic(mode, ds.as<double4>(), di.as<int4>(), dl.as<double2>(), dr.as<double>(),
                         nsite, geometric, dout.as<double>(), du.as<double>(), nullptr))
    die_msg(who, last_error());
  HIPCHK(hipDeviceSynchronize());
  out.resize(7 * (size_t)nsite);
  HIPCHK(hipMemcpy(out.data(), dout.p, sizeof(double) * out.size(), hipMemcpyDeviceToHost));
  ulj = 0;
  if (mode == 0) {
    std::vector<double> up(nblk);
    HIPCHK(hipMemcpy(up.data(), du.p, sizeof(double) * nblk, hipMemcpyDeviceToHost));
    for (double x : up) ulj += x;
  }
}

}  // namespace

extern "C" void RSpaceNonPeriodic(int nsite, int ngroup, const int (*group)[2], const cart_t *r,
                                  const struct PairBinary *group_excluded,
                                  const struct PairBinary *excluded, const int *site_mask,
                                  const double *q, double *phi, cart_t *evec,
                                  const double *rscreen, const double *sig, const double *eps,
                                  int geometric_combining_for_sigma, double *lj_ener,
                                  cart_t *lj_force, const struct PairBinary *one_four,
                                  double elec_1_4_scale_factor, double lj_1_4_scale_factor) {
  (void)excluded;   // consulted only for group-excluded pairs, which abort here
  (void)elec_1_4_scale_factor;
  (void)lj_1_4_scale_factor;
  const char *who = "RSpaceNonPeriodic";
  nonperiodic_checks(who, nsite, ngroup, group_excluded, one_four);
  if (nsite == 0) return;
  std::vector<double> out;
  double u = 0;
  nonperiodic_on_device(who, 0, nsite, ngroup, group, r, site_mask, q, rscreen, sig, eps,
                        geometric_combining_for_sigma, out, u);
  for (int i = 0; i < nsite; i++) {
    const double *o = &out[7 * (size_t)i];
    phi[i] += o[0];
    evec[i].x += o[1]; evec[i].y += o[2]; evec[i].z += o[3];
    lj_force[i].x += o[4]; lj_force[i].y += o[5]; lj_force[i].z += o[6];
  }
  *lj_ener += u;
}

extern "C" void RSpaceNonPeriodicPhi(int nsite, int ngroup, const int (*group)[2], const cart_t *r,
                                     const struct PairBinary *group_excluded,
                                     const struct PairBinary *excluded, const int *site_mask,
                                     const double *q, double *phi, const double *rscreen,
                                     const struct PairBinary *one_four,
                                     double elec_1_4_scale_factor) {
  (void)excluded;
  (void)elec_1_4_scale_factor;
  const char *who = "RSpaceNonPeriodicPhi";
  nonperiodic_checks(who, nsite, ngroup, group_excluded, one_four);
  if (nsite == 0) return;
  std::vector<double> out;
  double u = 0;
  nonperiodic_on_device(who, 1, nsite, ngroup, group, r, site_mask, q, rscreen, nullptr, nullptr,
                        0, out, u);
  for (int i = 0; i < nsite; i++) phi[i] += out[7 * (size_t)i];
}

extern "C" void RSpaceNonPeriodicPhiAtGridpoints(int nsite, int ngrid, const cart_t *rsite,
                                                 const cart_t *rgrid, const int *site_mask,
                                                 const double *q, double *phi,
                                                 const double *rscreen) {
  const char *who = "RSpaceNonPeriodicPhiAtGridpoints";
  if (nsite < 0 || ngrid < 0) die_msg(who, "nsite and ngrid must be >= 0");
  if (nsite == 0 || ngrid == 0) return;
  std::vector<double4> site(nsite);
  std::vector<double> rs(nsite, 0.0);
  for (int i = 0; i < nsite; i++) {
    site[i] = make_double4(rsite[i].x, rsite[i].y, rsite[i].z, site_mask[i] ? q[i] : 0.0);
    if (rscreen) rs[i] = rscreen[i];
  }
  DevBuf ds(sizeof(double4) * nsite), dr(sizeof(double) * nsite), dp(sizeof(double) * 3 * ngrid),
      dphi(sizeof(double) * ngrid);
  HIPCHK(hipMemcpy(ds.p, site.data(), sizeof(double4) * nsite, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dr.p, rs.data(), sizeof(double) * nsite, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dp.p, rgrid, sizeof(double) * 3 * (size_t)ngrid, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(dphi.p, 0, sizeof(double) * ngrid));
  if (launch_nonperiodic_points(ds.as<double4>(), dr.as<double>(), nsite, dp.as<double>(), ngrid,
                                dphi.as<double>(), nullptr))
    die_msg(who, last_error());
  HIPCHK(hipDeviceSynchronize());
  std::vector<double> out(ngrid);
  HIPCHK(hipMemcpy(out.data(), dphi.p, sizeof(double) * ngrid, hipMemcpyDeviceToHost));
  for (int j = 0; j < ngrid; j++) phi[j] += out[j];
}

// ------------------------------------------------------------ k space
struct kspace {
  double kc;
  int nsite, nk, nmax;
  std::vector<int> ik;   // 3 per vector, reference order (kspace.c:35-58)
  Engine e;
  int ready = 0;
};

static int ik_cmp_sq(const int *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

extern "C" struct kspace *kspace_new(double cutoff, int nsite) {
  kspace *k = new kspace();
  k->kc = cutoff;
  k->nsite = nsite;
  const double nsqmax = cutoff * cutoff + 1e-8;
  const int nmax = (int)floor(cutoff);
  k->nmax = nmax;
  for (int nx = 0; nx <= nmax; nx++)
    for (int ny = (nx == 0) ? 0 : -nmax; ny <= nmax; ny++)
      for (int nz = (nx == 0 && ny == 0) ? 1 : -nmax; nz <= nmax; nz++)
        if (nx * nx + ny * ny + nz * nz <= nsqmax) {
          k->ik.push_back(nx); k->ik.push_back(ny); k->ik.push_back(nz);
        }
  k->nk = (int)k->ik.size() / 3;
  // sorted by |n|^2, largest first (kspace.c:23-33, 58)
  std::vector<int> idx(k->nk);
  for (int i = 0; i < k->nk; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    return ik_cmp_sq(&k->ik[3 * a]) > ik_cmp_sq(&k->ik[3 * b]);
  });
  std::vector<int> s(k->ik.size());
  for (int i = 0; i < k->nk; i++)
    for (int c = 0; c < 3; c++) s[3 * i + c] = k->ik[3 * idx[i] + c];
  k->ik.swap(s);
  return k;
}

extern "C" void kspace_delete(struct kspace *k) {
  if (!k) return;
  if (k->ready) engine_destroy(k->e);
  delete k;
}

extern "C" int kspace_number_of_wavevectors(const struct kspace *k) { return k->nk; }

extern "C" void kspace_show(const struct kspace *k) {
  printf("nsite: %d  nk: %d  nmax: %d\n", k->nsite, k->nk, k->nmax);
  for (int i = 0; i < k->nk; i++) printf("ikvec %d: %d %d %d\n", i, k->ik[3 * i], k->ik[3 * i + 1], k->ik[3 * i + 2]);
  fflush(stdout);
}

static void kspace_run(struct kspace *k, const cart_t *r, const double *q, const tensor_t *L,
                       double volume, double eta, double *phi, cart_t *evec) {
  Engine &e = k->e;
  const int N = k->nsite;
  if (!k->ready) {
    e.nsys = 1;
    e.nsite = N;
    e.hsp.resize(1);
    HIPCHK(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc((void **)&e.d_sp, sizeof(SysParams)));
    HIPCHK(hipMalloc((void **)&e.d_r, sizeof(double) * 3 * N));
    HIPCHK(hipMalloc((void **)&e.d_q, sizeof(double) * N));
    HIPCHK(hipMalloc((void **)&e.d_kvec, sizeof(double) * 4 * k->nk));
    HIPCHK(hipMalloc((void **)&e.d_ksc, sizeof(double) * 2 * k->nk));
    HIPCHK(hipMalloc((void **)&e.d_kres, sizeof(double) * (4 * (size_t)N + 9)));
    HIPCHK(hipMalloc((void **)&e.d_acc, sizeof(double) * NACC));
    k->ready = 1;
  }
  SysParams &P = e.hsp[0];
  memset(&P, 0, sizeof P);
  P.kmode = 0;
  P.nk = k->nk;
  P.nsite = N;
  P.eta = eta;
  P.volume = volume;
  // kspace.c:96-103 with the caller's reciprocal lattice
  std::vector<double> kv(4 * (size_t)k->nk);
  const double c = 0.25 / (eta * eta), fourpi_V = 4 * M_PI / volume;
  for (int i = 0; i < k->nk; i++) {
    const int nx = k->ik[3 * i], ny = k->ik[3 * i + 1], nz = k->ik[3 * i + 2];
    const double kx = L->xx * nx + L->xy * ny + L->xz * nz;
    const double ky = L->yx * nx + L->yy * ny + L->yz * nz;
    const double kz = L->zx * nx + L->zy * ny + L->zz * nz;
    const double t = kx * kx + ky * ky + kz * kz;
    kv[4 * i] = kx; kv[4 * i + 1] = ky; kv[4 * i + 2] = kz;
    kv[4 * i + 3] = fourpi_V * exp(-t * c) / t;
  }
  std::vector<double> soa;
  to_soa((const double *)r, N, soa);
  HIPCHK(hipMemcpy(e.d_sp, &P, sizeof(SysParams), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e.d_kvec, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e.d_r, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e.d_q, q, sizeof(double) * N, hipMemcpyHostToDevice));
  if (launch_kspace(e, 0)) die_msg("kspace_field", last_error());
  HIPCHK(hipStreamSynchronize(e.stream));
  std::vector<double> out(4 * (size_t)N);
  HIPCHK(hipMemcpy(out.data(), e.d_kres, sizeof(double) * out.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < N; i++) {
    phi[i] += out[4 * (size_t)i];
    if (evec) {
      evec[i].x += out[4 * (size_t)i + 1];
      evec[i].y += out[4 * (size_t)i + 2];
      evec[i].z += out[4 * (size_t)i + 3];
    }
  }
}

extern "C" void kspace_field(struct kspace *k, const cart_t *r, const double *q,
                             const tensor_t *latv, double volume, double eta, double *phi,
                             cart_t *evec) {
  kspace_run(k, r, q, latv, volume, eta, phi, evec);
}

extern "C" void kspace_phi(struct kspace *k, const cart_t *r, const double *q,
                           const tensor_t *latv, double volume, double eta, double *phi) {
  kspace_run(k, r, q, latv, volume, eta, phi, nullptr);
}

// ------------------------------------------------------------ P3M
struct p3m {
  int nsite, nalias, ik, order;
  double grid_spacing;
  double latv[9], eta = 0;
  int ngrid[3];
  std::vector<double> green;   // half spectrum [ix][iy][iz], FP64
  Engine e;
  int have_green = 0;
};

int mmamd::multiple_of_2357_near(double x) {   // pppm.c:277-292
  int best = 1;
  double bestdiff = fabs(x - 1);
  for (int p2 = 1; p2 <= x; p2 *= 2)
    for (int p23 = p2; p23 <= x; p23 *= 3)
      for (int p235 = p23; p235 <= x; p235 *= 5)
        for (int p2357 = p235; p2357 <= x; p2357 *= 7) {
          const double diff = fabs(x - p2357);
          if (diff < bestdiff) { bestdiff = diff; best = p2357; }
        }
  return best;
}

static void p3m_upload_params(p3m *p) {
  SysParams &P = p->e.hsp[0];
  double inv[9];
  tensor_inverse(inv, p->latv);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) P.RL[3 * a + b] = 2 * M_PI * inv[3 * b + a];
  for (int a = 0; a < 9; a++) P.T[a] = inv[a];
  P.T[0] *= P.nx; P.T[3] *= P.nx; P.T[6] *= P.nx;
  P.T[1] *= P.ny; P.T[4] *= P.ny; P.T[7] *= P.ny;
  P.T[2] *= P.nz; P.T[5] *= P.nz; P.T[8] *= P.nz;
  P.eta = p->eta > 0 ? p->eta : 1.0;
  HIPCHK(hipMemcpy(p->e.d_sp, &P, sizeof(SysParams), hipMemcpyHostToDevice));
  if (p->have_green) {
    const int nx = P.nx, ny = P.ny, nzh = P.nz / 2 + 1;
    std::vector<float> g((size_t)nx * ny * nzh);
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++)
        for (int iz = 0; iz < nzh; iz++)
          g[((size_t)iy * nzh + iz) * nx + ix] = (float)p->green[((size_t)ix * ny + iy) * nzh + iz];
    HIPCHK(hipMemcpy(p->e.d_green, g.data(), sizeof(float) * g.size(), hipMemcpyHostToDevice));
  }
}

extern "C" p3m_t p3m_new(int nsite, const double ptv[9], double grid_spacing, int nalias,
                         int ik_differentiate, int assignment_order, int ngrid[3]) {
  if (nsite < 0) die_msg("p3m_new", "nsite >= 0 (pppm.c:407)");
  if (assignment_order <= 1 || assignment_order > 7) die_msg("p3m_new", "1 < assignment_order <= 7 (pppm.c:408-409)");
  if (nalias <= 0) die_msg("p3m_new", "nalias > 0 (pppm.c:411)");
  if (!(grid_spacing > 0)) die_msg("p3m_new", "grid_spacing > 0 (pppm.c:412)");
  p3m *p = new p3m();
  p->nsite = nsite;
  p->nalias = nalias;
  p->ik = ik_differentiate;
  p->order = assignment_order;
  p->grid_spacing = grid_spacing;
  memcpy(p->latv, ptv, sizeof(p->latv));
  for (int d = 0; d < 3; d++) {   // pppm.c:424-431, columns are lattice vectors
    const double a = sqrt(ptv[3 * d] * ptv[3 * d] + ptv[3 * d + 1] * ptv[3 * d + 1] +
                          ptv[3 * d + 2] * ptv[3 * d + 2]);
    p->ngrid[d] = multiple_of_2357_near(floor(a / grid_spacing + 0.5));
    ngrid[d] = p->ngrid[d];
    if (p->ngrid[d] > 128) die_msg("p3m_new", "mesh dimensions above 128 are not supported on the device");
  }
  Engine &e = p->e;
  const int N = std::max(nsite, 1);
  e.nsys = 1;
  e.nsite = N;
  e.hsp.resize(1);
  SysParams &P = e.hsp[0];
  memset(&P, 0, sizeof P);
  P.kmode = 1;
  P.nsite = nsite;
  P.nx = p->ngrid[0]; P.ny = p->ngrid[1]; P.nz = p->ngrid[2];
  P.order = assignment_order;
  P.ik = ik_differentiate ? 1 : 0;   // field on the mesh: three more meshes (pppm.c:444-455)
  if (P.nx < P.order || P.ny < P.order || P.nz < P.order)
    die_msg("p3m_new", "mesh dimensions below the assignment order are not supported on the device");
  const int nzh = P.nz / 2 + 1;
  if (!choose_spread_blocks(P.nx, P.ny, P.nz, 4096, 4096, P.bx, P.by))
    die_msg("p3m_new", "mesh dimensions give no valid spreading column blocks");
  P.nbx = P.nx / P.bx;
  P.nby = P.ny / P.by;
  e.nbin_total = P.nbx * P.nby;   // d_bin_fill[nbin_total] records an overfull bin's need
  e.nblk_part = std::max(1, (P.ny * nzh + P3M_FFT_LINES - 1) / P3M_FFT_LINES);
  const size_t M = (size_t)P.nx * P.ny * P.nz, Mc = (size_t)P.nx * P.ny * nzh;
  const size_t ncomp = P.ik ? 4 : 1;
  HIPCHK(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
  HIPCHK(hipMalloc((void **)&e.d_sp, sizeof(SysParams)));
  HIPCHK(hipMalloc((void **)&e.d_r, sizeof(double) * 3 * N));
  HIPCHK(hipMalloc((void **)&e.d_q, sizeof(double) * N));
  HIPCHK(hipMalloc((void **)&e.d_kres, sizeof(double) * (4 * (size_t)N + 9)));
  HIPCHK(hipMalloc((void **)&e.d_acc, sizeof(double) * NACC));
  HIPCHK(hipMalloc((void **)&e.d_mesh, sizeof(float) * M * ncomp));
  HIPCHK(hipMalloc((void **)&e.d_cmesh, sizeof(float2) * Mc * ncomp));
  HIPCHK(hipMalloc((void **)&e.d_green, sizeof(float) * Mc));
  HIPCHK(hipMemset(e.d_green, 0, sizeof(float) * Mc));
  HIPCHK(hipMalloc((void **)&e.d_sgrid, sizeof(int4) * N));
  HIPCHK(hipMalloc((void **)&e.d_sfrac, sizeof(float4) * N));
  const int nb = P.nbx * P.nby + 1;
  HIPCHK(hipMalloc((void **)&e.d_bin_fill, sizeof(int) * nb));
  HIPCHK(hipMemset(e.d_bin_fill, 0, sizeof(int) * nb));
  e.bcap = p3m_bin_capacity(N, P.nbx * P.nby);
  HIPCHK(hipMalloc((void **)&e.d_bgrid, sizeof(int4) * (size_t)(nb - 1) * e.bcap));
  HIPCHK(hipMalloc((void **)&e.d_bfrac, sizeof(float4) * (size_t)(nb - 1) * e.bcap));
  HIPCHK(hipMalloc((void **)&e.d_err, sizeof(int)));   // the binning kernel flags overfull bins
  HIPCHK(hipMemset(e.d_err, 0, sizeof(int)));
  HIPCHK(hipMalloc((void **)&e.d_part, sizeof(double) * 16 * e.nblk_part));
  HIPCHK(hipMalloc((void **)&e.d_done, sizeof(int) * 4));
  HIPCHK(hipMemset(e.d_done, 0, sizeof(int) * 4));
  std::vector<float2> tw;
  for (int d = 0; d < 3; d++)
    for (int t = 0; t < 128; t++) {
      const int n = p->ngrid[d];
      const double a = -2 * M_PI * (t % n) / n;
      tw.push_back(make_float2((float)cos(a), (float)sin(a)));
    }
  HIPCHK(hipMalloc((void **)&e.d_twid, sizeof(float2) * tw.size()));
  HIPCHK(hipMemcpy(e.d_twid, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
  p3m_upload_params(p);
  return p;
}

// pppm.c:313-393: the influence function of an ngrid mesh over the cell
// latv (columns are lattice vectors) at screening eta, into *green when
// given (half spectrum [ix][iy][iz]), and the RMS force error estimate.
double mmamd::p3m_influence(const double latv[9], const int ngrid[3], int nalias, int order, int ik,
                            double eta, std::vector<double> *green) {
  const int nx = ngrid[0], ny = ngrid[1], nz = ngrid[2], nzh = nz / 2 + 1;
  double inv[9], rl[9];
  tensor_inverse(inv, latv);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) rl[3 * a + b] = 2 * M_PI * inv[3 * b + a];
  const double *L = latv;
  const double det = -L[6] * L[4] * L[2] + L[3] * L[7] * L[2] + L[6] * L[1] * L[5] -
                     L[0] * L[7] * L[5] - L[3] * L[1] * L[8] + L[0] * L[4] * L[8];
  const double inv_v = 1.0 / fabs(det), inv_4eta2 = 1 / ((2 * eta) * (2 * eta));
  if (green) green->assign((size_t)nx * ny * nzh, 0.0);
  double c = 0;
  for (int ix = 0; ix < nx; ix++) {
    const int jx = ix > nx / 2 ? ix - nx : ix;
    for (int iy = 0; iy < ny; iy++) {
      const int jy = iy > ny / 2 ? iy - ny : iy;
      for (int iz = 0; iz < nzh; iz++) {
        const double g = green_at(jx, jy, iz, nx, ny, nz, rl, inv_4eta2, inv_v, nalias, order, ik, &c);
        if (green) (*green)[((size_t)ix * ny + iy) * nzh + iz] = g;
      }
    }
  }
  return pow(inv_v, 1.0 / 3.0) * sqrt(2 * fabs(c));
}

extern "C" void p3m_init(p3m_t p, double eta, double *chi) {   // pppm.c:313-393
  *chi = p3m_influence(p->latv, p->ngrid, p->nalias, p->order, p->ik, eta, &p->green);
  p->eta = eta;
  p->have_green = 1;
  p3m_upload_params(p);
}

extern "C" void p3m_field(p3m_t p, const double *r, const double *q, double *phi, double *evec) {
  if (!p->have_green) die_msg("p3m_field", "p3m_init or p3m_optimize first");
  Engine &e = p->e;
  const int N = p->nsite;
  if (N == 0) return;
  std::vector<double> soa;
  to_soa(r, N, soa);
  HIPCHK(hipMemcpy(e.d_r, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e.d_q, q, sizeof(double) * N, hipMemcpyHostToDevice));
  // sites far denser than the box average (a slab, a cluster) overflow the
  // column bins: grow them to what the binning recorded and evaluate again
  for (int attempt = 0;; attempt++) {
    if (launch_p3m(e, 0)) die_msg("p3m_field", last_error());
    HIPCHK(hipStreamSynchronize(e.stream));
    int flags = 0;
    HIPCHK(hipMemcpy(&flags, e.d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (!(flags & ERR_BIN_OVERFLOW)) break;
    HIPCHK(hipMemset(e.d_err, 0, sizeof(int)));
    if (attempt >= 8 || engine_grow_bins(e)) die_msg("p3m_field", "mesh column bins cannot grow further");
  }
  std::vector<double> out(4 * (size_t)N);
  HIPCHK(hipMemcpy(out.data(), e.d_kres, sizeof(double) * out.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < N; i++) {
    phi[i] += out[4 * (size_t)i];
    if (evec)
      for (int c = 0; c < 3; c++) evec[3 * (size_t)i + c] += out[4 * (size_t)i + 1 + c];
  }
}

extern "C" void p3m_scale(p3m_t p, double s) {   // pppm.c:643-664
  for (int a = 0; a < 9; a++) p->latv[a] *= s;
  for (auto &g : p->green) g /= s;
  p3m_upload_params(p);
}

// pppm.c:294-311.  The reference memcpy's n values' worth of bytes over the
// arrays' row-pointer tables (its kvec/green_hat are pointer-of-pointer
// arrays); what it means is a copy with the same influence function, which
// is what this does.
extern "C" p3m_t p3m_copy(const p3m_t t, int nsite) {
  int ng[3];
  p3m_t p = p3m_new(nsite, t->latv, t->grid_spacing, t->nalias, t->ik, t->order, ng);
  p->green = t->green;
  p->have_green = t->have_green;
  p->eta = t->eta;
  p3m_upload_params(p);
  return p;
}

extern "C" void p3m_delete(p3m_t p) {
  if (!p) return;
  Engine &e = p->e;
  void *ptrs[] = {e.d_sp, e.d_r, e.d_q, e.d_kres, e.d_acc, e.d_mesh, e.d_cmesh, e.d_green,
                  e.d_sgrid, e.d_sfrac, e.d_bin_fill, e.d_bgrid, e.d_bfrac,
                  e.d_part, e.d_done, e.d_twid, e.d_err};
  for (void *q : ptrs)
    if (q) hipFree(q);
  if (e.stream) hipStreamDestroy(e.stream);
  e = Engine();
  delete p;
}

// pppm.c:667-716: minimise chi_rspace^2 + chi_p3m^2 over eta
void mmamd::p3m_optimize_eta(const double latv[9], const int ngrid[3], int nalias, int order, int ik,
                             double grid_spacing, double rc, double *eta, double *chi_rspace,
                             double *chi_p3m) {
  const double *L = latv;
  const double v = fabs(-L[6] * L[4] * L[2] + L[3] * L[7] * L[2] + L[6] * L[1] * L[5] -
                        L[0] * L[7] * L[5] - L[3] * L[1] * L[8] + L[0] * L[4] * L[8]);
  double cr = 0, ck = 0;
  auto f = [&](double x) {   // pppm.c:672-680
    if (x < 0) return 1e8;
    cr = 2 / sqrt(rc / pow(v, 1.0 / 3.0)) * exp(-(rc * x) * (rc * x));
    ck = p3m_influence(latv, ngrid, nalias, order, ik, x, nullptr);
    return cr * cr + ck * ck;
  };
  // one-dimensional Nelder-Mead: start 0.1/h, step 0.1, size tolerance 1e-4
  double x0 = 0.1 / grid_spacing, x1 = x0 + 0.1;
  double f0 = f(x0), f1 = f(x1);
  for (int it = 1; it <= 100; it++) {
    if (f1 < f0) { std::swap(x0, x1); std::swap(f0, f1); }
    if (fabs(x1 - x0) < 1e-4) break;
    const double xr = x0 + (x0 - x1), fr = f(xr);
    if (fr < f0) {
      const double xe = x0 + 2 * (x0 - x1), fe = f(xe);
      if (fe < fr) { x1 = xe; f1 = fe; } else { x1 = xr; f1 = fr; }
    } else if (fr < f1) {
      x1 = xr; f1 = fr;
    } else {
      const double xc = x0 + 0.5 * (x1 - x0), fc = f(xc);
      if (fc < f1) { x1 = xc; f1 = fc; } else { x1 = x0 + 0.5 * (x1 - x0); f1 = f(x1); }
    }
  }
  if (f1 < f0) { std::swap(x0, x1); std::swap(f0, f1); }
  f(x0);   // both estimates at the returned eta
  *eta = x0;
  *chi_rspace = cr;
  *chi_p3m = ck;
}

extern "C" void p3m_optimize(p3m_t p, double rc, double *eta, double *chi_rspace, double *chi_p3m) {
  p3m_optimize_eta(p->latv, p->ngrid, p->nalias, p->order, p->ik, p->grid_spacing, rc, eta,
                   chi_rspace, chi_p3m);
  p3m_init(p, *eta, chi_p3m);   // leave the influence function at the returned eta
}
