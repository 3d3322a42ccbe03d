/*
 * ref_harness.cpp -- test-only entry points into the reference's own
 * nonbonded sources, compiled from where they lie under the reference tree
 * by oracle/Makefile into oracle/_ref/libref.so.  Nothing here is part of
 * the product; tests/golden/make_golden.py and the oracle tests call it to
 * pin the FP64 restatement in oracle/mm_oracle.c.
 *
 * Reference functions called:
 *   NeighborList (nbrlist.cpp:9-251) with BoundaryConditions::new_cubic
 *   PairBinary_* (pbin.c:15-137)
 *   RSpaceChebyshev_new (rspace.c:25-37), RSpaceEwald (rspace.c:382-558)
 *   kspace_new / kspace_field (kspace.c:35-174)
 *   RSpaceNonPeriodic, RSpaceNonPeriodicPhi, RSpaceNonPeriodicPhiAtGridpoints
 *   (rspace.c:560-640, 784-890)
 */
#include "nbrlist.hpp"
#include "bcond.hpp"
#include "cvec.hpp"
#include "pbin.h"
#include "rspace.h"
#include "kspace.h"
#include "cheby.h"
#include <cstdlib>
#include <cstring>

extern "C" {

/* Build a periodic cubic neighbor list over n points and write its rows as
 * (row, col) pairs in row order, each row in append order. */
long ref_nl_pairs(int n, const double *xyz, double boxl, double rc, double skin,
                  int *out, long maxp)
{
  BoundaryConditions *bc = BoundaryConditions::new_cubic(boxl);
  NeighborList nl(n, rc, skin);
  nl.set_boundary_conditions(*bc);
  CVec r(n);
  for (int i = 0; i < n; i++) r[i] = Cartesian(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  nl.coordinates_changed(r);
  long np = 0;
  for (int i = 0; i < n; i++)
    for (const int *j = nl.start(i); j < nl.end(i); j++) {
      if (np < maxp) { out[2 * np] = i; out[2 * np + 1] = *j; }
      np++;
    }
  delete bc;
  return np;
}

/* Build, then move to xyz2 and report whether coordinates_changed rebuilt
 * (the pair list of the second call differs from a forced rebuild only if it
 * did not).  Returns the pair count after the second call. */
long ref_nl_pairs_two_step(int n, const double *xyz, const double *xyz2, double boxl,
                           double rc, double skin, int *out, long maxp)
{
  BoundaryConditions *bc = BoundaryConditions::new_cubic(boxl);
  NeighborList nl(n, rc, skin);
  nl.set_boundary_conditions(*bc);
  CVec r(n);
  for (int i = 0; i < n; i++) r[i] = Cartesian(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  nl.coordinates_changed(r);
  for (int i = 0; i < n; i++) r[i] = Cartesian(xyz2[3 * i], xyz2[3 * i + 1], xyz2[3 * i + 2]);
  nl.coordinates_changed(r);
  long np = 0;
  for (int i = 0; i < n; i++)
    for (const int *j = nl.start(i); j < nl.end(i); j++) {
      if (np < maxp) { out[2 * np] = i; out[2 * np + 1] = *j; }
      np++;
    }
  delete bc;
  return np;
}

/* RSpaceChebyshev_new without the random self-test: writes min, max,
 * inv_dx followed by 512*7 coefficients. */
void ref_cheby(double rhi, double eta, double *out)
{
  struct Chebyshev *t = RSpaceChebyshev_new(rhi, eta, 0, 0);
  out[0] = t->min; out[1] = t->max; out[2] = t->inv_dx;
  memcpy(out + 3, t->coeff, sizeof(double) * NCHEBY * 7);
  free(t);
}

static struct PairBinary *pb_from(int n, int np, const int *pairs)
{
  struct PairBinary *p = PairBinary_new(n);
  for (int k = 0; k < np; k++) PairBinary_add(p, pairs[2 * k], pairs[2 * k + 1]);
  PairBinary_init(p);
  return p;
}

/* One RSpaceEwald call for a cubic box with the neighbor list built by the
 * reference's NeighborList over the groups' first sites.  excl: site pairs
 * (1-2 and 1-3); group_excluded is derived as eleclj.cpp:748-770 does.
 * Outputs are accumulated into zeroed caller arrays. */
void ref_rspace(int nsite, int ngroup, const int *group, const double *r, double boxl,
                double rc, double smoothing_width, double skin, int nexcl, const int *excl,
                const int *site_mask, const double *q, double eta, const double *sig,
                const double *eps, double *phi, double *evec, double *lj_ener, double *lj_force)
{
  BoundaryConditions *bc = BoundaryConditions::new_cubic(boxl);
  NeighborList nl(ngroup, rc, skin);
  nl.set_boundary_conditions(*bc);
  CVec gr(ngroup);
  for (int g = 0; g < ngroup; g++) {
    const int s = group[2 * g];
    gr[g] = Cartesian(r[3 * s], r[3 * s + 1], r[3 * s + 2]);
  }
  nl.coordinates_changed(gr);
  struct PairBinary *excluded = pb_from(nsite, nexcl, excl);
  int *gfor = new int[nsite > 0 ? nsite : 1];
  for (int g = 0; g < ngroup; g++)
    for (int s = group[2 * g]; s < group[2 * g + 1]; s++) gfor[s] = g;
  struct PairBinary *gexcl = PairBinary_new(ngroup);
  for (int i = 0; i < nsite; i++) {
    const int *k = PairBinary_elements(excluded)[i];
    const int *kend = k + PairBinary_number_of_elements(excluded)[i];
    for (; k < kend; k++)
      if (gfor[i] != gfor[*k]) PairBinary_add(gexcl, gfor[i], gfor[*k]);
  }
  PairBinary_init(gexcl);
  struct Chebyshev *cheb = RSpaceChebyshev_new(rc, eta, 0, 0);
  double *rscreen = new double[nsite > 0 ? nsite : 1]();
  tensor_t latv = { boxl, 0, 0, 0, boxl, 0, 0, 0, boxl };
  RSpaceEwald(nsite, ngroup, (const int(*)[2])group, (const cart_t *)r, cubic, &latv,
              nl.neighbors(), nl.number_of_neighbors(), gexcl, excluded,
              (rc - smoothing_width) * (rc - smoothing_width), rc * rc, site_mask, q, phi,
              (cart_t *)evec, eta, cheb, rscreen, sig, eps, 0, lj_ener, (cart_t *)lj_force,
              0, 1, 1);
  free(cheb);
  delete[] rscreen;
  delete[] gfor;
  PairBinary_delete(excluded);
  PairBinary_delete(gexcl);
  delete bc;
}

/* RSpaceEwaldPhiAtGridpoints as ElecLJ calls it (eleclj.cpp:1317-1321) */
void ref_phi_at_gridpoints(int nsite, int ngrid, const double *rsite, const double *rgrid,
                           double boxl, double rc, double smoothing_width, const int *site_mask,
                           const double *q, double eta, double *phi)
{
  struct Chebyshev *cheb = RSpaceChebyshev_new(rc, eta, 0, 0);
  double *rscreen = new double[nsite > 0 ? nsite : 1]();
  tensor_t latv = { boxl, 0, 0, 0, boxl, 0, 0, 0, boxl };
  RSpaceEwaldPhiAtGridpoints(nsite, ngrid, (const cart_t *)rsite, (const cart_t *)rgrid, cubic,
                             &latv, (rc - smoothing_width) * (rc - smoothing_width), rc * rc,
                             site_mask, q, phi, eta, cheb, rscreen);
  free(cheb);
  delete[] rscreen;
}

/* RSpaceEwaldPhi on the reference's own neighbour list (rspace.c:649-782) */
void ref_rspace_phi(int nsite, int ngroup, const int *group, const double *r, double boxl,
                    double rc, double smoothing_width, double skin, int nexcl, const int *excl,
                    const int *site_mask, const double *q, double eta, double *phi)
{
  BoundaryConditions *bc = BoundaryConditions::new_cubic(boxl);
  NeighborList nl(ngroup, rc, skin);
  nl.set_boundary_conditions(*bc);
  CVec gr(ngroup);
  for (int g = 0; g < ngroup; g++) {
    const int s = group[2 * g];
    gr[g] = Cartesian(r[3 * s], r[3 * s + 1], r[3 * s + 2]);
  }
  nl.coordinates_changed(gr);
  struct PairBinary *excluded = pb_from(nsite, nexcl, excl);
  struct PairBinary *gexcl = PairBinary_new(ngroup);
  PairBinary_init(gexcl);
  struct Chebyshev *cheb = RSpaceChebyshev_new(rc, eta, 0, 0);
  double *rscreen = new double[nsite > 0 ? nsite : 1]();
  tensor_t latv = { boxl, 0, 0, 0, boxl, 0, 0, 0, boxl };
  RSpaceEwaldPhi(nsite, ngroup, (const int(*)[2])group, (const cart_t *)r, cubic, &latv,
                 nl.neighbors(), nl.number_of_neighbors(), gexcl, excluded,
                 (rc - smoothing_width) * (rc - smoothing_width), rc * rc, site_mask, q, phi,
                 eta, cheb, rscreen, 0, 1);
  free(cheb);
  delete[] rscreen;
  PairBinary_delete(excluded);
  PairBinary_delete(gexcl);
  delete bc;
}

/* kspace_new + kspace_field for a cubic box (recip = 2 pi / L on the
 * diagonal, as BoundaryConditions::reciprocal_lattice_vectors gives). */
void ref_kspace(double kc, int nsite, const double *r, const double *q, const double *recip,
                double volume, double eta, double *phi, double *evec)
{
  struct kspace *k = kspace_new(kc, nsite);
  kspace_field(k, (const cart_t *)r, q, (const tensor_t *)recip, volume, eta, phi, (cart_t *)evec);
  kspace_delete(k);
}

int ref_kspace_count(double kc)
{
  struct kspace *k = kspace_new(kc, 1);
  const int n = kspace_number_of_wavevectors(k);
  kspace_delete(k);
  return n;
}

/* RSpaceNonPeriodic (rspace.c:560-640), RSpaceNonPeriodicPhi (:784-856) and
 * RSpaceNonPeriodicPhiAtGridpoints (:858-890) for a cluster: exclusions are
 * the given site pairs (inside groups), no group exclusions, no 1-4 list;
 * rscreen per site.  out: phi, evec (3), lj_force (3) per site; phi_only
 * per site; grid_phi per point; lj_energy. */
void ref_nonperiodic(int nsite, int ngroup, const int *group, const double *r, int nexcl,
                     const int *excl, const int *site_mask, const double *q,
                     const double *rscreen, const double *sig, const double *eps, int geometric,
                     int ngrid, const double *grid, double *phi, double *evec,
                     double *lj_force, double *lj_energy, double *phi_only, double *grid_phi)
{
  struct PairBinary *excluded = pb_from(nsite, nexcl, excl);
  struct PairBinary *gexcl = PairBinary_new(ngroup);
  PairBinary_init(gexcl);
  RSpaceNonPeriodic(nsite, ngroup, (const int(*)[2])group, (const cart_t *)r, gexcl, excluded,
                    site_mask, q, phi, (cart_t *)evec, rscreen, sig, eps, geometric, lj_energy,
                    (cart_t *)lj_force, 0, 1, 1);
  RSpaceNonPeriodicPhi(nsite, ngroup, (const int(*)[2])group, (const cart_t *)r, gexcl, excluded,
                       site_mask, q, phi_only, rscreen, 0, 1);
  RSpaceNonPeriodicPhiAtGridpoints(nsite, ngrid, (const cart_t *)r, (const cart_t *)grid,
                                   site_mask, q, grid_phi, rscreen);
  PairBinary_delete(excluded);
  PairBinary_delete(gexcl);
}

/* PairBinary known-answer sequence of pbin.c:176-213: rows (rowptr/cols)
 * and the 10x10 exists matrix. */
void ref_pbin_kat(const int *adds, int nadd, int n, int *rowptr, int *cols, int *exists)
{
  struct PairBinary *p = PairBinary_new(n);
  for (int k = 0; k < nadd; k++) PairBinary_add(p, adds[2 * k], adds[2 * k + 1]);
  PairBinary_init(p);
  int o = 0;
  for (int i = 0; i < n; i++) {
    rowptr[i] = o;
    for (const int *j = PairBinary_start(p, i); j < PairBinary_end(p, i); j++) cols[o++] = *j;
  }
  rowptr[n] = o;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) exists[i * n + j] = PairBinary_exists(p, i, j);
  PairBinary_delete(p);
}

}  /* extern "C" */
