/*
 * mm_oracle.c -- FP64 CPU oracle (test infrastructure only; see mm_oracle.h).
 *
 * Built with -O2 -ffp-contract=off so that every product and sum rounds the
 * way the reference's x86-64 build (no FMA) rounds it; the neighbor-list
 * distance tests depend on that for bit-exact pair sets.
 */
#define _GNU_SOURCE
#include "mm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_2_SQRTPI
#define M_2_SQRTPI 1.12837916709551257390
#endif

static double sq(double x) { return x * x; }

/* fns.h:48-58 -- non-negative modulus */
static int orc_mod(int n, int m)
{
  if (n >= 0) return n % m;
  int k = -n % m;
  return k > 0 ? m - k : 0;
}

/* fns.c:8-29 -- is_almost_zero: |x| < 64 * 2^-52 */
static int almost_zero(double x) { return fabs(x) < 64.0 * 2.220446049250313e-16; }

/* ------------------------------------------------------------------ */
/* PairBinary semantics                                                */
/* ------------------------------------------------------------------ */

/* pbin.c:6 MAKE_CANONICAL */
void orc_pbin_canonical(int *i, int *j)
{
  int a = *i, b = *j;
  if (((a + b) % 2) ^ (a < b)) { *i = b; *j = a; }
}

static int cmp_int(const void *a, const void *b)
{
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* pbin.c:41-50 (add) + 73-90 (init: per-row sort, dedup) */
int orc_pbin_build(int n, int npair, const int *pairs, int *rowptr, int *cols)
{
  int *cnt = calloc((size_t)n + 1, sizeof(int));
  int k;
  for (k = 0; k < npair; k++) {
    int i = pairs[2 * k], j = pairs[2 * k + 1];
    orc_pbin_canonical(&i, &j);
    cnt[i]++;
  }
  rowptr[0] = 0;
  for (k = 0; k < n; k++) rowptr[k + 1] = rowptr[k] + cnt[k];
  memset(cnt, 0, sizeof(int) * (size_t)n);
  for (k = 0; k < npair; k++) {
    int i = pairs[2 * k], j = pairs[2 * k + 1];
    orc_pbin_canonical(&i, &j);
    cols[rowptr[i] + cnt[i]++] = j;
  }
  /* sort + dedup, compacting in place */
  int out = 0;
  for (k = 0; k < n; k++) {
    int b = rowptr[k], e = rowptr[k + 1];
    qsort(cols + b, (size_t)(e - b), sizeof(int), cmp_int);
    int start = out, t;
    for (t = b; t < e; t++)
      if (t == b || cols[t] != cols[t - 1]) cols[out++] = cols[t];
    rowptr[k] = start;
  }
  rowptr[n] = out;
  free(cnt);
  return out;
}

/* pbin.c:119-137 */
int orc_pbin_exists(const int *rowptr, const int *cols, int i, int j)
{
  orc_pbin_canonical(&i, &j);
  int lo = rowptr[i], hi = rowptr[i + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (j < cols[mid]) hi = mid - 1;
    else if (j > cols[mid]) lo = mid + 1;
    else return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Neighbor list                                                       */
/* ------------------------------------------------------------------ */

/* nbrlist.cpp:7,11 -- ncell = floor(cbrt(n/2 + 3)), integer n/2 */
int orc_nl_ncell(int ngroup)
{
  return (int)floor(cbrt((double)(ngroup / 2 + 3)));
}

/* Cell box of a cubic box: Cubic::scale(1.0/ncell) (bcond.cpp:38, nbrlist.cpp:97) */
static double cell_length(int ncell, double boxl) { return boxl * (1.0 / ncell); }

/* nbrlist.cpp:99-102 with Cubic::max_diameter (bcond.cpp:37) and
 * Triclinic::min_diameter of the diagonal cell tensor (bcond.cpp:146-152,
 * cart.h:73-78 det). */
int orc_nl_nmax(int ncell, double boxl, double rc, double skin)
{
  const double lc = cell_length(ncell, boxl);
  const double maxd = sqrt(3.0) * lc;
  /* det of diag(lc): only the xx*yy*zz term is non-zero; the zero terms
   * added before it do not change the value */
  const double vol = fabs(lc * lc * lc);
  /* |col_i x col_j| = sqrt((lc*lc)^2) = lc*lc */
  const double a = sqrt(0.0 + 0.0 + sq(lc * lc));
  const double mind = vol / a;
  int nmax = (int)ceil((rc + skin + maxd) / mind);
  int cap = ncell / 2 + 1;
  return nmax < cap ? nmax : cap;
}

/* nbrlist.cpp:79-82 */
static int cell_index(int ix, int iy, int iz, int n)
{
  return (orc_mod(ix, n) * n + orc_mod(iy, n)) * n + orc_mod(iz, n);
}

/* nbrlist.cpp:105-120 -- half-space stencil, every offset whose cell-lattice
 * vector maps inside c2 (always true: the vector is a cell-lattice vector and
 * maps to ~0), pairs of distinct cells, then PairBinary_init. */
int orc_nl_cell_pairs(int ncell, double boxl, double rc, double skin,
                      int *rowptr, int *cols, int maxc)
{
  const double lc = cell_length(ncell, boxl);
  const double inv = 1 / lc;
  const double c2 = sq(rc + skin + sqrt(3.0) * lc);
  const int nmax = orc_nl_nmax(ncell, boxl, rc, skin);
  const int ncell3 = ncell * ncell * ncell;
  long cap = 1024, np = 0;
  int *pairs = malloc(sizeof(int) * 2 * cap);
  int jx, jy, jz;
  for (jx = 0; jx <= nmax; jx++)
    for (jy = (jx == 0 ? 0 : -nmax); jy <= nmax; jy++)
      for (jz = (jx == 0 && jy == 0 ? 1 : -nmax); jz <= nmax; jz++) {
        /* Tensor(lc) * Cartesian(jx,jy,jz), right_operate (cart.h:39-48) */
        double x = lc * jx + 0.0 * jy + 0.0 * jz;
        double y = 0.0 * jx + lc * jy + 0.0 * jz;
        double z = 0.0 * jx + 0.0 * jy + lc * jz;
        x -= lc * floor(x * inv + 0.5);
        y -= lc * floor(y * inv + 0.5);
        z -= lc * floor(z * inv + 0.5);
        if (!(x * x + y * y + z * z < c2)) continue;
        int ix, iy, iz;
        for (ix = 0; ix < ncell; ix++)
          for (iy = 0; iy < ncell; iy++)
            for (iz = 0; iz < ncell; iz++) {
              int ia = cell_index(ix, iy, iz, ncell);
              int ib = cell_index(ix + jx, iy + jy, iz + jz, ncell);
              if (ia == ib) continue;
              if (np == cap) { cap *= 2; pairs = realloc(pairs, sizeof(int) * 2 * cap); }
              pairs[2 * np] = ia; pairs[2 * np + 1] = ib; np++;
            }
      }
  int *tmp = malloc(sizeof(int) * (np > 0 ? np : 1));
  int nnz = orc_pbin_build(ncell3, (int)np, pairs, rowptr, tmp);
  free(pairs);
  if (nnz > maxc) { free(tmp); return -1; }
  memcpy(cols, tmp, sizeof(int) * (size_t)nnz);
  free(tmp);
  return nnz;
}

/* nbrlist.cpp:222-251 (with bcond.hpp:41-60 / pbc.h:8-14 / pbc.h:25-31) */
long orc_nl_build(int ngroup, const double *gr, double boxl, double rc, double skin,
                  int *pairs, long maxp)
{
  const int ncell = orc_nl_ncell(ngroup);
  const int ncell3 = ncell * ncell * ncell;
  const double lc = cell_length(ncell, boxl);
  const double c2 = sq(rc + skin);
  int *rowptr = malloc(sizeof(int) * ((size_t)ncell3 + 1));
  int maxc = ncell3 * 1400 + 16;
  int *cols = malloc(sizeof(int) * (size_t)maxc);
  if (orc_nl_cell_pairs(ncell, boxl, rc, skin, rowptr, cols, maxc) < 0) {
    free(rowptr); free(cols); return -1;
  }
  int *head = malloc(sizeof(int) * (size_t)ncell3);
  int *next = malloc(sizeof(int) * (size_t)(ngroup > 0 ? ngroup : 1));
  int i;
  for (i = 0; i < ncell3; i++) head[i] = -1;
  for (i = 0; i < ngroup; i++) {
    /* Cubic::nearest_lattice_point of the cell box: (int) floor(x/lc + 0.5) */
    int ix = (int)floor(gr[3 * i] / lc + 0.5);
    int iy = (int)floor(gr[3 * i + 1] / lc + 0.5);
    int iz = (int)floor(gr[3 * i + 2] / lc + 0.5);
    int c = cell_index(ix, iy, iz, ncell);
    next[i] = head[c];
    head[c] = i;
  }
  const double invl = 1 / boxl;
  long np = 0;
  for (i = 0; i < ncell3; i++) {
    int k;
    for (k = head[i]; k != -1; k = next[k]) {
      int l, m;
      for (l = next[k]; l != -1; l = next[l]) {
        int a = k, b = l;
        orc_pbin_canonical(&a, &b);
        if (np < maxp) { pairs[2 * np] = a; pairs[2 * np + 1] = b; }
        np++;
      }
      for (m = rowptr[i]; m < rowptr[i + 1]; m++)
        for (l = head[cols[m]]; l != -1; l = next[l]) {
          double x = gr[3 * k] - gr[3 * l];
          double y = gr[3 * k + 1] - gr[3 * l + 1];
          double z = gr[3 * k + 2] - gr[3 * l + 2];
          x -= boxl * floor(x * invl + 0.5);
          y -= boxl * floor(y * invl + 0.5);
          z -= boxl * floor(z * invl + 0.5);
          if (x * x + y * y + z * z < c2) {
            int a = k, b = l;
            orc_pbin_canonical(&a, &b);
            if (np < maxp) { pairs[2 * np] = a; pairs[2 * np + 1] = b; }
            np++;
          }
        }
    }
  }
  free(head); free(next); free(rowptr); free(cols);
  return np;
}

/* nbrlist.cpp:64-76 */
int orc_nl_needs_rebuild(int n, const double *r, const double *last, double skin)
{
  double largest = 0, second = 0;
  int i;
  for (i = 0; i < n; i++) {
    double x = r[3 * i] - last[3 * i], y = r[3 * i + 1] - last[3 * i + 1],
           z = r[3 * i + 2] - last[3 * i + 2];
    double d = sqrt(x * x + y * y + z * z);
    if (d > largest) { second = largest; largest = d; }
    else if (d > second) second = d;
  }
  return second + largest > skin;
}

/* ------------------------------------------------------------------ */
/* Chebyshev table                                                     */
/* ------------------------------------------------------------------ */

static double cheb_eta;
static double erfc_over_r(double r2)            /* rspace.c:19-23 */
{
  const double r = sqrt(r2);
  return erfc(cheb_eta * r) / r;
}

/* cheby.c:7-33: 50-node Chebyshev fit of degree 3 on [a,b], re-expressed as a
 * power series in the local coordinate x in [0,1), plus the series of
 * -2 d/d(r^2) of the same cubic. */
static void cheb_fit(double a, double b, double *coeff, double (*f)(double))
{
  double g[50], c[4];
  const double fac = 2.0 / 50, bma = 0.5 * (b - a), bpa = 0.5 * (b + a);
  int k, j;
  for (k = 0; k < 50; k++) g[k] = f(cos(M_PI * (k + 0.5) / 50) * bma + bpa);
  for (j = 0; j < 4; j++) {
    double s = 0;
    for (k = 0; k < 50; k++) s += g[k] * cos(M_PI * j * (k + 0.5) / 50);
    c[j] = fac * s;
  }
  /* T0..T3 of t = 2x-1 expanded in powers of x */
  coeff[0] = c[0] / 2 - c[1] + c[2] - c[3];
  coeff[1] = 2 * c[1] - 8 * c[2] + 18 * c[3];
  coeff[2] = 8 * c[2] - 48 * c[3];
  coeff[3] = 32 * c[3];
  coeff[4] = -2 * coeff[1] / (b - a);
  coeff[5] = -2 * 2 * coeff[2] / (b - a);
  coeff[6] = -2 * 3 * coeff[3] / (b - a);
}

/* rspace.c:25-37 + cheby.c:35-45 */
void orc_cheby_rspace(orc_cheby_t *t, double rhi, double eta)
{
  const double r2max = 1.2 * sq(rhi);
  const double r2min = r2max > 4 ? 4 : 0.9 * r2max;
  const double dx = (r2max - r2min) / ORC_NCHEBY;
  int i;
  cheb_eta = eta;
  t->min = r2min;
  t->max = r2max;
  t->inv_dx = 1.0 / dx;
  for (i = 0; i < ORC_NCHEBY; i++)
    cheb_fit(i * dx + r2min, (i + 1) * dx + r2min, t->coeff[i], erfc_over_r);
}

/* cheby.h:25-37 */
void orc_cheby_eval(const orc_cheby_t *t, double x, double *f, double *n2df)
{
  x -= t->min;
  x *= t->inv_dx;
  unsigned i = (unsigned)x;
  const double *c = t->coeff[i];
  x -= i;
  *f = ((c[3] * x + c[2]) * x + c[1]) * x + c[0];
  *n2df = (c[6] * x + c[5]) * x + c[4];
}

/* ------------------------------------------------------------------ */
/* Real space                                                          */
/* ------------------------------------------------------------------ */

/* rspace.c:52-58 */
static void switching(double r2, double rlo2, double rhi2, double *f, double *_2df)
{
  const double a = 1.0 / (rhi2 - rlo2);
  const double z = (r2 - rlo2) * a;
  *f = ((-6 * z + 15) * z - 10) * z * z * z + 1;
  *_2df = ((60 * z - 120) * z + 60) * z * z * a;
}

typedef struct {
  double *phi, *evec, *lj_ener, *lj_force, *vir;
  const double *q, *sig, *eps, *rscreen;
  const orc_cheby_t *cheb;
  double eta;
  int geometric_sigma;
} pairctx;

static void vir_add(double *vir, double x, double y, double z, double fx, double fy, double fz)
{
  if (!vir) return;
  vir[0] += x * fx; vir[1] += x * fy; vir[2] += x * fz;
  vir[3] += y * fx; vir[4] += y * fy; vir[5] += y * fz;
  vir[6] += z * fx; vir[7] += z * fy; vir[8] += z * fz;
}

/* rspace.c:104-156 (ewald) */
static void pair_ewald(pairctx *c, int i, int j, double x, double y, double z, double r2,
                       int smoothing, double smooth, double *tmpsmooth)
{
  const double qi = c->q[i], qj = c->q[j];
  double b0, b1;
  if (c->cheb && c->cheb->min < r2 && r2 < c->cheb->max) {
    orc_cheby_eval(c->cheb, r2, &b0, &b1);
  } else {
    const double rm = sqrt(r2), etar = c->eta * rm;
    const double ex = exp(-sq(etar));
    b0 = erfc(etar) / rm;
    b1 = (b0 + c->eta * M_2_SQRTPI * ex) / r2;
  }
  const double rs = c->rscreen ? c->rscreen[i] + c->rscreen[j] : 0;
  if (r2 < rs * rs) {                          /* rspace.c:61-64, 129-136 */
    const double invr = 1.0 / sqrt(r2);
    double bb1 = 1.0 / (rs * rs * rs);
    double f = -0.5 * r2 * bb1 + 1.5 / rs;
    b0 += f - invr;
    b1 += bb1 - invr * invr * invr;
  }
  if (smoothing) {
    *tmpsmooth += qi * qj * b0;
    b0 *= smooth;
    b1 *= smooth;
  }
  const double qib1 = qi * b1, qjb1 = qj * b1;
  c->phi[i] += qj * b0;
  c->evec[3 * i] += qjb1 * x; c->evec[3 * i + 1] += qjb1 * y; c->evec[3 * i + 2] += qjb1 * z;
  c->phi[j] += qi * b0;
  c->evec[3 * j] -= qib1 * x; c->evec[3 * j + 1] -= qib1 * y; c->evec[3 * j + 2] -= qib1 * z;
  const double fq = qi * qjb1;
  vir_add(c->vir, x, y, z, fq * x, fq * y, fq * z);
}

/* rspace.c:258-297 (lj) with scale factor support */
static void pair_lj(pairctx *c, int i, int j, double x, double y, double z, double r2,
                    int smoothing, double smooth, double *tmpsmooth, int scaling, double sf)
{
  const double s2 = c->geometric_sigma ? c->sig[i] * c->sig[j] : sq(0.5 * (c->sig[i] + c->sig[j]));
  const double s6 = s2 * s2 * s2;
  const double b = 4 * sqrt(c->eps[i] * c->eps[j]) * s6;
  const double a = b * s6;
  const double invr2 = 1.0 / r2, invr6 = invr2 * invr2 * invr2;
  const double binvr6 = b * invr6, ainvr12 = a * invr6 * invr6;
  double du = 12 * ainvr12 * invr2 - 6 * binvr6 * invr2;
  double u = ainvr12 - binvr6;
  if (smoothing) { *tmpsmooth += u; u *= smooth; du *= smooth; }
  if (scaling) { u *= sf; du *= sf; }
  *c->lj_ener += u;
  double cx = du * x, cy = du * y, cz = du * z;
  c->lj_force[3 * i] += cx; c->lj_force[3 * j] -= cx;
  c->lj_force[3 * i + 1] += cy; c->lj_force[3 * j + 1] -= cy;
  c->lj_force[3 * i + 2] += cz; c->lj_force[3 * j + 2] -= cz;
  vir_add(c->vir, x, y, z, cx, cy, cz);
}

/* rspace.c:66-102 (coulomb_non_periodic), used by the 1-4 loop */
static void pair_coulomb(pairctx *c, int i, int j, double x, double y, double z, double r2,
                         int scaling, double sf)
{
  const double qi = c->q[i], qj = c->q[j];
  double b0, b1;
  const double rs = c->rscreen ? c->rscreen[i] + c->rscreen[j] : 0;
  if (r2 < rs * rs) {
    b1 = 1.0 / (rs * rs * rs);
    b0 = -0.5 * r2 * b1 + 1.5 / rs;
  } else {
    const double invr = 1.0 / sqrt(r2);
    b0 = invr;
    b1 = invr * invr * invr;
  }
  if (scaling) { b0 *= sf; b1 *= sf; }
  const double qib1 = qi * b1, qjb1 = qj * b1;
  c->phi[i] += qj * b0;
  c->evec[3 * i] += qjb1 * x; c->evec[3 * i + 1] += qjb1 * y; c->evec[3 * i + 2] += qjb1 * z;
  c->phi[j] += qi * b0;
  c->evec[3 * j] -= qib1 * x; c->evec[3 * j + 1] -= qib1 * y; c->evec[3 * j + 2] -= qib1 * z;
  const double fq = qi * qjb1;
  vir_add(c->vir, x, y, z, fq * x, fq * y, fq * z);
}

static void interact_ewald(pairctx *c, int i, int j, double x, double y, double z, double r2,
                           int mask, int smoothing, double smooth, double *tmpsmooth)
{
  if (mask & ORC_SITE_HAS_Q) pair_ewald(c, i, j, x, y, z, r2, smoothing, smooth, tmpsmooth);
  if (mask & ORC_SITE_HAS_LJ) pair_lj(c, i, j, x, y, z, r2, smoothing, smooth, tmpsmooth, 0, 1);
}

/* rspace.c:382-558 for cubic boundary conditions */
static void rspace_impl(int nsite, int ngroup, const int *group, const double *r,
                        double boxl, long nnl, const int *nl,
                        int nexcl, const int *excl, double rlo2, double rhi2,
                        const int *site_mask, const double *q, double eta,
                        const orc_cheby_t *cheb, const double *rscreen,
                        const double *sig, const double *eps, int geometric_sigma,
                        int n14, const int *one_four, double s14e, double s14lj,
                        double *phi, double *evec, double *lj_ener, double *lj_force,
                        double *virial, int excl_min_image)
{
  const double negb = -boxl, invb = 1 / boxl;
  pairctx c = { phi, evec, lj_ener, lj_force, virial, q, sig, eps, rscreen, cheb, eta,
                geometric_sigma };
  /* excluded and group_excluded PairBinary (eleclj.cpp:712-724, 748-770) */
  int *erow = malloc(sizeof(int) * ((size_t)nsite + 1));
  int *ecol = malloc(sizeof(int) * (size_t)(nexcl > 0 ? nexcl : 1));
  orc_pbin_build(nsite, nexcl, excl, erow, ecol);
  int *gfor = malloc(sizeof(int) * (size_t)(nsite > 0 ? nsite : 1));
  long g, k;
  for (g = 0; g < ngroup; g++) {
    int s;
    for (s = group[2 * g]; s < group[2 * g + 1]; s++) gfor[s] = (int)g;
  }
  int *gp = malloc(sizeof(int) * 2 * (size_t)(nexcl > 0 ? nexcl : 1));
  int ngp = 0;
  for (k = 0; k < nsite; k++) {
    int t;
    for (t = erow[k]; t < erow[k + 1]; t++) {
      int ig = gfor[k], jg = gfor[ecol[t]];
      if (ig != jg) { gp[2 * ngp] = ig; gp[2 * ngp + 1] = jg; ngp++; }
    }
  }
  int *grow = malloc(sizeof(int) * ((size_t)ngroup + 1));
  int *gcol = malloc(sizeof(int) * (size_t)(ngp > 0 ? ngp : 1));
  orc_pbin_build(ngroup, ngp, gp, grow, gcol);
  free(gp);

  for (k = 0; k < nnl; k++) {
    const int ig = nl[2 * k], jg = nl[2 * k + 1];
    const int istart = group[2 * ig], iend = group[2 * ig + 1];
    const int jstart = group[2 * jg], jend = group[2 * jg + 1];
    const int gex = orc_pbin_exists(grow, gcol, ig, jg);
    double x = r[3 * istart] - r[3 * jstart];
    double y = r[3 * istart + 1] - r[3 * jstart + 1];
    double z = r[3 * istart + 2] - r[3 * jstart + 2];
    /* pbc.h:16-23 */
    const double dx = negb * floor(x * invb + 0.5);
    const double dy = negb * floor(y * invb + 0.5);
    const double dz = negb * floor(z * invb + 0.5);
    double smooth = 1, dsmooth = 0, tmpsmooth = 0, dsmx = 0, dsmy = 0, dsmz = 0;
    int smoothing = 0;
    x += dx; y += dy; z += dz;
    double r2 = x * x + y * y + z * z;
    if (r2 > rhi2) continue;
    if (r2 > rlo2) {
      smoothing = 1;
      switching(r2, rlo2, rhi2, &smooth, &dsmooth);
      dsmx = dsmooth * x; dsmy = dsmooth * y; dsmz = dsmooth * z;
    }
    if (!gex || !orc_pbin_exists(erow, ecol, istart, jstart))
      interact_ewald(&c, istart, jstart, x, y, z, r2, site_mask[istart] & site_mask[jstart],
                     smoothing, smooth, &tmpsmooth);
    int i, j;
    for (i = istart; i < iend; i++) {
      const double xi = r[3 * i] + dx, yi = r[3 * i + 1] + dy, zi = r[3 * i + 2] + dz;
      const int imask = site_mask[i];
      for (j = (i == istart ? jstart + 1 : jstart); j < jend; j++) {
        const int mask = imask & site_mask[j];
        if (!mask) continue;
        if (gex && orc_pbin_exists(erow, ecol, i, j)) continue;
        x = xi - r[3 * j]; y = yi - r[3 * j + 1]; z = zi - r[3 * j + 2];
        r2 = x * x + y * y + z * z;
        interact_ewald(&c, i, j, x, y, z, r2, mask, smoothing, smooth, &tmpsmooth);
      }
    }
    if (smoothing) {                            /* rspace.c:479-489 */
      dsmx *= tmpsmooth; dsmy *= tmpsmooth; dsmz *= tmpsmooth;
      lj_force[3 * istart] += dsmx; lj_force[3 * istart + 1] += dsmy; lj_force[3 * istart + 2] += dsmz;
      lj_force[3 * jstart] -= dsmx; lj_force[3 * jstart + 1] -= dsmy; lj_force[3 * jstart + 2] -= dsmz;
      /* force on istart is dsm, displacement of the first sites is
       * (x,y,z)/... : dsm = dsmooth*tmpsmooth*d  */
      if (virial) {
        double d0x = r[3 * istart] + dx - r[3 * jstart];
        double d0y = r[3 * istart + 1] + dy - r[3 * jstart + 1];
        double d0z = r[3 * istart + 2] + dz - r[3 * jstart + 2];
        vir_add(virial, d0x, d0y, d0z, dsmx, dsmy, dsmz);
      }
    }
  }
  /* Excluded: rspace.c:492-521 */
  int i;
  for (i = 0; i < nsite; i++) {
    if (!(site_mask[i] & ORC_SITE_HAS_Q)) continue;
    int t;
    for (t = erow[i]; t < erow[i + 1]; t++) {
      const int j = ecol[t];
      if (!(site_mask[j] & ORC_SITE_HAS_Q)) continue;
      double x = r[3 * i] - r[3 * j], y = r[3 * i + 1] - r[3 * j + 1], z = r[3 * i + 2] - r[3 * j + 2];
      if (excl_min_image) {   /* RSpaceEwaldPhi omits it (rspace.c:747-751) */
        x += negb * floor(x * invb + 0.5);
        y += negb * floor(y * invb + 0.5);
        z += negb * floor(z * invb + 0.5);
      }
      const double r2 = x * x + y * y + z * z;
      const double rm = sqrt(r2), etar = eta * rm;
      const double ex = exp(-sq(etar));
      const double b0 = -erf(etar) / rm;
      const double b1 = (b0 + eta * M_2_SQRTPI * ex) / r2;
      const double qi = q[i], qj = q[j];
      const double qib1 = qi * b1, qjb1 = qj * b1;
      phi[i] += qj * b0;
      evec[3 * i] += qjb1 * x; evec[3 * i + 1] += qjb1 * y; evec[3 * i + 2] += qjb1 * z;
      phi[j] += qi * b0;
      evec[3 * j] -= qib1 * x; evec[3 * j + 1] -= qib1 * y; evec[3 * j + 2] -= qib1 * z;
      const double fq = qi * qjb1;
      vir_add(virial, x, y, z, fq * x, fq * y, fq * z);
    }
  }
  /* One-four: rspace.c:522-557 */
  if (n14 > 0) {
    int *orow = malloc(sizeof(int) * ((size_t)nsite + 1));
    int *ocol = malloc(sizeof(int) * (size_t)n14);
    orc_pbin_build(nsite, n14, one_four, orow, ocol);
    const int se = fabs(s14e - 1) > 1e-10, sl = fabs(s14lj - 1) > 1e-10;
    for (i = 0; i < nsite; i++) {
      int t;
      for (t = orow[i]; t < orow[i + 1]; t++) {
        const int j = ocol[t];
        double x = r[3 * i] - r[3 * j], y = r[3 * i + 1] - r[3 * j + 1], z = r[3 * i + 2] - r[3 * j + 2];
        x += negb * floor(x * invb + 0.5);
        y += negb * floor(y * invb + 0.5);
        z += negb * floor(z * invb + 0.5);
        const double r2 = x * x + y * y + z * z;
        const int mask = site_mask[i] & site_mask[j];
        if (mask & ORC_SITE_HAS_Q) pair_coulomb(&c, i, j, x, y, z, r2, se, s14e - 1);
        if (mask & ORC_SITE_HAS_LJ) pair_lj(&c, i, j, x, y, z, r2, 0, 1, NULL, sl, s14lj - 1);
      }
    }
    free(orow); free(ocol);
  }
  free(erow); free(ecol); free(gfor); free(grow); free(gcol);
}

void orc_rspace_ewald(int nsite, int ngroup, const int *group, const double *r,
                      double boxl, long nnl, const int *nl,
                      int nexcl, const int *excl, double rlo2, double rhi2,
                      const int *site_mask, const double *q, double eta,
                      const orc_cheby_t *cheb, const double *rscreen,
                      const double *sig, const double *eps, int geometric_sigma,
                      int n14, const int *one_four, double s14e, double s14lj,
                      double *phi, double *evec, double *lj_ener, double *lj_force,
                      double *virial)
{
  rspace_impl(nsite, ngroup, group, r, boxl, nnl, nl, nexcl, excl, rlo2, rhi2, site_mask, q,
              eta, cheb, rscreen, sig, eps, geometric_sigma, n14, one_four, s14e, s14lj, phi,
              evec, lj_ener, lj_force, virial, 1);
}

/* RSpaceEwaldPhi (rspace.c:649-782): the same smoothed pair potentials with
 * charges only, and the excluded-pair correction without the minimum image.
 * One-four pairs are not restated (the device path rejects them). */
void orc_rspace_ewald_phi(int nsite, int ngroup, const int *group, const double *r,
                          double boxl, long nnl, const int *nl, int nexcl, const int *excl,
                          double rlo2, double rhi2, const int *site_mask, const double *q,
                          double eta, const orc_cheby_t *cheb, const double *rscreen,
                          double *phi)
{
  int *qmask = malloc(sizeof(int) * (size_t)(nsite > 0 ? nsite : 1));
  double *ev = calloc(3 * (size_t)(nsite > 0 ? nsite : 1), sizeof(double));
  double *ljf = calloc(3 * (size_t)(nsite > 0 ? nsite : 1), sizeof(double));
  double *zero = calloc((size_t)(nsite > 0 ? nsite : 1), sizeof(double));
  double lje = 0;
  int i;
  for (i = 0; i < nsite; i++) qmask[i] = site_mask[i] & ORC_SITE_HAS_Q;
  rspace_impl(nsite, ngroup, group, r, boxl, nnl, nl, nexcl, excl, rlo2, rhi2, qmask, q, eta,
              cheb, rscreen, zero, zero, 0, 0, NULL, 1, 1, phi, ev, &lje, ljf, NULL, 0);
  free(qmask); free(ev); free(ljf); free(zero);
}

/* RSpaceEwaldPhiAtGridpoints (rspace.c:892-967): every site with any mask
 * bit against every grid point, minimum image, cutoff and switching on the
 * site-grid distance, the table branch inside (min, max), screening. */
void orc_rspace_phi_at_gridpoints(int nsite, int ngrid, const double *rsite,
                                  const double *rgrid, double boxl, double rlo2, double rhi2,
                                  const int *site_mask, const double *q, double eta,
                                  const orc_cheby_t *cheb, const double *rscreen, double *phi)
{
  const double negb = -boxl, invb = 1 / boxl;
  int i, j;
  for (i = 0; i < nsite; i++) {
    if (!site_mask[i]) continue;
    const double xi = rsite[3 * i], yi = rsite[3 * i + 1], zi = rsite[3 * i + 2];
    const double qi = q[i];
    const double rscr = rscreen ? rscreen[i] : 0, rscr2 = rscr * rscr;
    for (j = 0; j < ngrid; j++) {
      double x = xi - rgrid[3 * j], y = yi - rgrid[3 * j + 1], z = zi - rgrid[3 * j + 2];
      x += negb * floor(x * invb + 0.5);
      y += negb * floor(y * invb + 0.5);
      z += negb * floor(z * invb + 0.5);
      const double r2 = x * x + y * y + z * z;
      if (r2 > rhi2) continue;
      double smooth = 1, dsmooth = 0;
      const int smoothing = r2 > rlo2;
      if (smoothing) switching(r2, rlo2, rhi2, &smooth, &dsmooth);
      double b0, b1;
      if (cheb && cheb->min < r2 && r2 < cheb->max) {
        orc_cheby_eval(cheb, r2, &b0, &b1);
      } else {
        const double rm = sqrt(r2);
        b0 = erfc(eta * rm) / rm;
      }
      if (r2 < rscr2) {   /* screening_function (rspace.c:61-64) */
        const double invr = 1.0 / sqrt(r2);
        const double b1s = 1.0 / (rscr * rscr * rscr);
        b0 += (-0.5 * r2 * b1s + 1.5 / rscr) - invr;
      }
      if (smoothing) b0 *= smooth;
      phi[j] += qi * b0;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Non-periodic real space                                             */
/* ------------------------------------------------------------------ */

/* rspace.c:61-64 (screening inside rs), else 1/r and 1/r^3 */
static void np_coulomb(double r2, double rs, double *b0, double *b1)
{
  if (r2 < rs * rs) {
    *b1 = 1.0 / (rs * rs * rs);
    *b0 = -0.5 * r2 * (*b1) + 1.5 / rs;
  } else {
    const double invr = 1.0 / sqrt(r2);
    *b0 = invr;
    *b1 = invr * invr * invr;
  }
}

/* one site pair of RSpaceNonPeriodic: interact() (rspace.c:299-321) with
 * coulomb_non_periodic (:66-102) and lj (:258-297), no smoothing/scaling */
static void np_pair(int i, int j, const double *r, const int *site_mask, const double *q,
                    const double *rscreen, const double *sig, const double *eps,
                    int geometric, double *phi, double *evec, double *lj_ener, double *lj_force)
{
  const double x = r[3 * i] - r[3 * j], y = r[3 * i + 1] - r[3 * j + 1], z = r[3 * i + 2] - r[3 * j + 2];
  const double r2 = x * x + y * y + z * z;
  const int mask = site_mask[i] & site_mask[j];
  if (mask & ORC_SITE_HAS_Q) {
    double b0, b1;
    np_coulomb(r2, rscreen[i] + rscreen[j], &b0, &b1);
    const double qib1 = q[i] * b1, qjb1 = q[j] * b1;
    phi[i] += q[j] * b0;
    evec[3 * i] += qjb1 * x; evec[3 * i + 1] += qjb1 * y; evec[3 * i + 2] += qjb1 * z;
    phi[j] += q[i] * b0;
    evec[3 * j] -= qib1 * x; evec[3 * j + 1] -= qib1 * y; evec[3 * j + 2] -= qib1 * z;
  }
  if (mask & ORC_SITE_HAS_LJ) {
    const double s2 = geometric ? sig[i] * sig[j] : sq(0.5 * (sig[i] + sig[j]));
    const double s6 = s2 * s2 * s2;
    const double b = 4 * sqrt(eps[i] * eps[j]) * s6, a = b * s6;
    const double invr2 = 1.0 / r2, invr6 = invr2 * invr2 * invr2;
    const double binvr6 = b * invr6, ainvr12 = a * invr6 * invr6;
    const double du = 12 * ainvr12 * invr2 - 6 * binvr6 * invr2;
    *lj_ener += ainvr12 - binvr6;
    lj_force[3 * i] += du * x; lj_force[3 * j] -= du * x;
    lj_force[3 * i + 1] += du * y; lj_force[3 * j + 1] -= du * y;
    lj_force[3 * i + 2] += du * z; lj_force[3 * j + 2] -= du * z;
  }
}

/* RSpaceNonPeriodic (rspace.c:560-640) without group exclusions or a 1-4
 * list: every site pair of groups ig < jg, first sites first */
void orc_rspace_nonperiodic(int nsite, int ngroup, const int *group, const double *r,
                            const int *site_mask, const double *q, const double *rscreen,
                            const double *sig, const double *eps, int geometric, double *phi,
                            double *evec, double *lj_ener, double *lj_force)
{
  int ig, jg, i, j;
  (void)nsite;
  for (ig = 0; ig < ngroup; ig++)
    for (jg = ig + 1; jg < ngroup; jg++) {
      const int is = group[2 * ig], ie = group[2 * ig + 1], js = group[2 * jg], je = group[2 * jg + 1];
      np_pair(is, js, r, site_mask, q, rscreen, sig, eps, geometric, phi, evec, lj_ener, lj_force);
      for (i = is; i < ie; i++)
        for (j = (i == is ? js + 1 : js); j < je; j++) {
          if (!(site_mask[i] & site_mask[j])) continue;
          np_pair(i, j, r, site_mask, q, rscreen, sig, eps, geometric, phi, evec, lj_ener, lj_force);
        }
    }
}

/* coulomb_non_periodic_phi (rspace.c:158-180) without smoothing/scaling */
static void np_phi_pair(int i, int j, const double *r, const double *q, const double *rscreen,
                        double *phi)
{
  const double x = r[3 * i] - r[3 * j], y = r[3 * i + 1] - r[3 * j + 1], z = r[3 * i + 2] - r[3 * j + 2];
  double b0, b1;
  np_coulomb(x * x + y * y + z * z, rscreen[i] + rscreen[j], &b0, &b1);
  phi[i] += q[j] * b0;
  phi[j] += q[i] * b0;
}

/* RSpaceNonPeriodicPhi (rspace.c:784-856): the first sites of two groups
 * always interact; other pairs when their masks share any bit */
void orc_rspace_nonperiodic_phi(int nsite, int ngroup, const int *group, const double *r,
                                const int *site_mask, const double *q, const double *rscreen,
                                double *phi)
{
  int ig, jg, i, j;
  (void)nsite;
  for (ig = 0; ig < ngroup; ig++)
    for (jg = ig + 1; jg < ngroup; jg++) {
      const int is = group[2 * ig], ie = group[2 * ig + 1], js = group[2 * jg], je = group[2 * jg + 1];
      np_phi_pair(is, js, r, q, rscreen, phi);
      for (i = is; i < ie; i++)
        for (j = (i == is ? js + 1 : js); j < je; j++) {
          if (!(site_mask[i] & site_mask[j])) continue;
          np_phi_pair(i, j, r, q, rscreen, phi);
        }
    }
}

/* RSpaceNonPeriodicPhiAtGridpoints (rspace.c:858-890) */
void orc_rspace_nonperiodic_phi_at_gridpoints(int nsite, int ngrid, const double *rsite,
                                              const double *rgrid, const int *site_mask,
                                              const double *q, const double *rscreen, double *phi)
{
  int i, j;
  for (i = 0; i < nsite; i++) {
    if (!site_mask[i]) continue;
    for (j = 0; j < ngrid; j++) {
      const double x = rsite[3 * i] - rgrid[3 * j], y = rsite[3 * i + 1] - rgrid[3 * j + 1],
                   z = rsite[3 * i + 2] - rgrid[3 * j + 2];
      double b0, b1;
      np_coulomb(x * x + y * y + z * z, rscreen[i], &b0, &b1);
      phi[j] += q[i] * b0;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Plain Ewald k-space                                                 */
/* ------------------------------------------------------------------ */

typedef struct { int x, y, z; } ivec3;
static int isq3(const ivec3 *i) { return i->x * i->x + i->y * i->y + i->z * i->z; }
static int ivec_cmp(const void *a, const void *b)    /* kspace.c:23-33 */
{
  const int i2 = isq3((const ivec3 *)a), j2 = isq3((const ivec3 *)b);
  return (i2 < j2) - (j2 < i2);
}

/* kspace.c:35-58 */
static int kspace_enumerate(double kc, ivec3 *out)
{
  const double nsqmax = kc * kc + 1e-8;
  const int nmax = (int)floor(kc);
  int nk = 0, nx, ny, nz;
  for (nx = 0; nx <= nmax; nx++)
    for (ny = (nx == 0) ? 0 : -nmax; ny <= nmax; ny++)
      for (nz = (nx == 0 && ny == 0) ? 1 : -nmax; nz <= nmax; nz++)
        if (nx * nx + ny * ny + nz * nz <= nsqmax) {
          if (out) { out[nk].x = nx; out[nk].y = ny; out[nk].z = nz; }
          nk++;
        }
  if (out) qsort(out, (size_t)nk, sizeof(ivec3), ivec_cmp);
  return nk;
}

int orc_kspace_count(double kc) { return kspace_enumerate(kc, NULL); }
void orc_kspace_vectors(double kc, int *ikvec) { kspace_enumerate(kc, (ivec3 *)ikvec); }

/* kspace.c:87-174 (+ energy and virial of the same sum) */
void orc_kspace_field(double kc, int nsite, const double *r, const double *q,
                      const double *L, double volume, double eta,
                      double *phi, double *evec, double *virial, double *energy)
{
  /* L in tensor_t layout: xx,yx,zx,xy,yy,zy,xz,yz,zz */
  const double Lxx = L[0], Lyx = L[1], Lzx = L[2], Lxy = L[3], Lyy = L[4], Lzy = L[5],
               Lxz = L[6], Lyz = L[7], Lzz = L[8];
  const int nk = orc_kspace_count(kc);
  const int nmax = (int)floor(kc);
  ivec3 *ik = malloc(sizeof(ivec3) * (size_t)nk);
  kspace_enumerate(kc, ik);
  double *kv = malloc(sizeof(double) * 3 * (size_t)nk), *ak = malloc(sizeof(double) * (size_t)nk);
  double *skr = malloc(sizeof(double) * (size_t)nk * (size_t)nsite);
  double *ckr = malloc(sizeof(double) * (size_t)nk * (size_t)nsite);
  double *sQ = calloc((size_t)nk, sizeof(double)), *cQ = calloc((size_t)nk, sizeof(double));
  double *ct = malloc(sizeof(double) * 3 * (size_t)(2 * nmax + 1));
  double *st = malloc(sizeof(double) * 3 * (size_t)(2 * nmax + 1));
  const double c = 0.25 / (eta * eta), fourpi_V = 4 * M_PI / volume;
  int i, k;
  for (k = 0; k < nk; k++) {
    kv[3 * k] = Lxx * ik[k].x + Lxy * ik[k].y + Lxz * ik[k].z;
    kv[3 * k + 1] = Lyx * ik[k].x + Lyy * ik[k].y + Lyz * ik[k].z;
    kv[3 * k + 2] = Lzx * ik[k].x + Lzy * ik[k].y + Lzz * ik[k].z;
    const double t = kv[3 * k] * kv[3 * k] + kv[3 * k + 1] * kv[3 * k + 1] + kv[3 * k + 2] * kv[3 * k + 2];
    ak[k] = fourpi_V * exp(-t * c) / t;
  }
#define CT(m, d) ct[3 * ((m) + nmax) + (d)]
#define ST(m, d) st[3 * ((m) + nmax) + (d)]
  for (i = 0; i < nsite; i++) {
    const double rx = r[3 * i], ry = r[3 * i + 1], rz = r[3 * i + 2];
    double t3[3], ck[3], sk[3];
    t3[0] = Lxx * rx + Lyx * ry + Lzx * rz;
    t3[1] = Lxy * rx + Lyy * ry + Lzy * rz;
    t3[2] = Lxz * rx + Lyz * ry + Lzz * rz;
    int d, m;
    for (d = 0; d < 3; d++) { ck[d] = cos(t3[d]); sk[d] = sin(t3[d]); CT(0, d) = 1; ST(0, d) = 0; }
    for (m = 1; m <= nmax; m++)
      for (d = 0; d < 3; d++) {
        CT(m, d) = ck[d] * CT(m - 1, d) - sk[d] * ST(m - 1, d);
        ST(m, d) = sk[d] * CT(m - 1, d) + ck[d] * ST(m - 1, d);
        CT(-m, d) = CT(m, d);
        ST(-m, d) = -ST(m, d);
      }
    for (k = 0; k < nk; k++) {
      const double cxi = CT(ik[k].x, 0), sxi = ST(ik[k].x, 0);
      const double cyi = CT(ik[k].y, 1), syi = ST(ik[k].y, 1);
      const double czi = CT(ik[k].z, 2), szi = ST(ik[k].z, 2);
      const double ccss = cxi * cyi - sxi * syi, sccs = sxi * cyi + cxi * syi;
      ckr[(size_t)i * nk + k] = ccss * czi - sccs * szi;
      skr[(size_t)i * nk + k] = sccs * czi + ccss * szi;
    }
  }
#undef CT
#undef ST
  for (i = 0; i < nsite; i++)
    for (k = 0; k < nk; k++) {
      sQ[k] += q[i] * skr[(size_t)i * nk + k];
      cQ[k] += q[i] * ckr[(size_t)i * nk + k];
    }
  for (i = 0; i < nsite; i++) {
    double gx = 0, gy = 0, gz = 0, p = 0;
    for (k = 0; k < nk; k++) {
      const double two_a = 2 * ak[k], s = skr[(size_t)i * nk + k], cc = ckr[(size_t)i * nk + k];
      const double t = two_a * (cQ[k] * s - sQ[k] * cc);
      gx += t * kv[3 * k]; gy += t * kv[3 * k + 1]; gz += t * kv[3 * k + 2];
      p += two_a * (sQ[k] * s + cQ[k] * cc);
    }
    phi[i] += p;
    evec[3 * i] += gx; evec[3 * i + 1] += gy; evec[3 * i + 2] += gz;
  }
  if (virial || energy) {
    double e = 0;
    for (k = 0; k < nk; k++) {
      const double rho2 = sQ[k] * sQ[k] + cQ[k] * cQ[k];
      const double ek = ak[k] * rho2;
      const double k2 = kv[3 * k] * kv[3 * k] + kv[3 * k + 1] * kv[3 * k + 1] + kv[3 * k + 2] * kv[3 * k + 2];
      const double fac = 2 * (1 / k2 + c);
      e += ek;
      if (virial) {
        int a, b;
        for (a = 0; a < 3; a++)
          for (b = 0; b < 3; b++)
            virial[3 * a + b] += ek * ((a == b) - fac * kv[3 * k + a] * kv[3 * k + b]);
      }
    }
    if (energy) *energy += e;
  }
  free(ik); free(kv); free(ak); free(skr); free(ckr); free(sQ); free(cQ); free(ct); free(st);
}

/* ------------------------------------------------------------------ */
/* P3M                                                                 */
/* ------------------------------------------------------------------ */

/* pppm.c:277-292 */
static int multiple_of_2357_near(double x)
{
  int p2, p23, p235, p2357, best = 1;
  double bestdiff = fabs(x - 1);
  for (p2 = 1; p2 <= x; p2 *= 2)
    for (p23 = p2; p23 <= x; p23 *= 3)
      for (p235 = p23; p235 <= x; p235 *= 5)
        for (p2357 = p235; p2357 <= x; p2357 *= 7) {
          const double diff = fabs(x - p2357);
          if (diff < bestdiff) { bestdiff = diff; best = p2357; }
        }
  return best;
}

/* pppm.c:424-431 */
int orc_p3m_grid_size(double length, double spacing)
{
  return multiple_of_2357_near(floor(length / spacing + 0.5));
}

static double sincpi(double x)                 /* pppm.c:56-59 */
{
  const double y = M_PI * x;
  return fabs(y) < 1e-16 ? 1 - sq(y) / 6 : sin(y) / y;
}

static void tensor_inverse(double *a, const double *b)   /* cart.h:73-92 */
{
  /* layout xx,yx,zx,xy,yy,zy,xz,yz,zz */
  const double xx = b[0], yx = b[1], zx = b[2], xy = b[3], yy = b[4], zy = b[5],
               xz = b[6], yz = b[7], zz = b[8];
  const double det = -xz * yy * zx + xy * yz * zx + xz * yx * zy - xx * yz * zy - xy * yx * zz + xx * yy * zz;
  const double id = 1.0 / det;
  a[0] = id * (-yz * zy + yy * zz);   /* xx */
  a[3] = id * (xz * zy - xy * zz);    /* xy */
  a[6] = id * (-xz * yy + xy * yz);   /* xz */
  a[1] = id * (yz * zx - yx * zz);    /* yx */
  a[4] = id * (-xz * zx + xx * zz);   /* yy */
  a[7] = id * (xz * yx - xx * yz);    /* yz */
  a[2] = id * (-yy * zx + yx * zy);   /* zx */
  a[5] = id * (xy * zx - xx * zy);    /* zy */
  a[8] = id * (-xy * yx + xx * yy);   /* zz */
}

static double tensor_det(const double *b)
{
  const double xx = b[0], yx = b[1], zx = b[2], xy = b[3], yy = b[4], zy = b[5],
               xz = b[6], yz = b[7], zz = b[8];
  return -xz * yy * zx + xy * yz * zx + xz * yx * zy - xx * yz * zy - xy * yx * zz + xx * yy * zz;
}

/* k = T * j with T in tensor_t layout (cart.h:13-18 multiply) */
static void tmul(const double *t, double jx, double jy, double jz, double *k)
{
  k[0] = t[0] * jx + t[3] * jy + t[6] * jz;
  k[1] = t[1] * jx + t[4] * jy + t[7] * jz;
  k[2] = t[2] * jx + t[5] * jy + t[8] * jz;
}

/* Influence function value at signed wave index (jx,jy,jz).  Follows
 * pppm.c:331-388: alias sum |p| <= nalias of g(k_p) U_p^2 with U the
 * order-p assignment function transform. */
static double green_at(int jx, int jy, int jz, int nx, int ny, int nz, const double *rl,
                       double inv_4eta2, double inv_v, int nalias, int order, int ik,
                       double *chi_acc)
{
  const double fourpi = 4 * M_PI;
  double kvec[3];
  tmul(rl, jx, jy, jz, kvec);
  if (jx == 0 && jy == 0 && jz == 0) return 0;
  const double kvec2 = kvec[0] * kvec[0] + kvec[1] * kvec[1] + kvec[2] * kvec[2];
  double ux[16], uy[16], uz[16];
  int qq, px, py, pz;
  for (qq = -nalias; qq <= nalias; qq++) {
    ux[qq + nalias] = pow(sincpi((double)jx / nx + qq), 2 * order);
    uy[qq + nalias] = pow(sincpi((double)jy / ny + qq), 2 * order);
    uz[qq + nalias] = pow(sincpi((double)jz / nz + qq), 2 * order);
  }
  double numer = 0, d1 = 0, d2 = 0, chi = 0;
  for (px = -nalias; px <= nalias; px++) {
    const double pxx = jx + px * nx;
    for (py = -nalias; py <= nalias; py++) {
      const double pyy = jy + py * ny;
      for (pz = -nalias; pz <= nalias; pz++) {
        const double pzz = jz + pz * nz;
        if (fabs(pxx) < 1e-16 && fabs(pyy) < 1e-16 && fabs(pzz) < 1e-16) continue;
        double kp[3];
        tmul(rl, pxx, pyy, pzz, kp);
        const double kp2 = kp[0] * kp[0] + kp[1] * kp[1] + kp[2] * kp[2];
        const double u2 = ux[px + nalias] * uy[py + nalias] * uz[pz + nalias];
        const double gg = fourpi / kp2 * exp(-kp2 * inv_4eta2);
        if (ik) {
          numer += gg * u2 * (kvec[0] * kp[0] + kvec[1] * kp[1] + kvec[2] * kp[2]);
          d1 += u2;
        } else {
          numer += gg * kp2 * u2;
          d1 += u2;
          d2 += u2 * kp2;
        }
        chi += gg * gg * kp2;
      }
    }
  }
  const double denom = ik ? sq(d1) * kvec2 : d1 * d2;
  if (chi_acc) *chi_acc += chi - sq(numer) / denom;
  return inv_v * numer / denom;
}

/* pppm.c:313-393 */
double orc_p3m_green(int nx, int ny, int nz, const double *latv, double eta,
                     int nalias, int order, int ik, double *green_hat)
{
  const double inv_4eta2 = 1 / sq(2 * eta);
  const double inv_v = 1.0 / fabs(tensor_det(latv));
  double inv[9], rl[9];
  int a, b, ix, iy, iz;
  tensor_inverse(inv, latv);
  /* reciprocal = 2 pi * transpose(inverse) */
  for (a = 0; a < 3; a++)
    for (b = 0; b < 3; b++) rl[3 * a + b] = 2 * M_PI * inv[3 * b + a];
  double chi = 0;
  const int nzh = nz / 2 + 1;
  for (ix = 0; ix < nx; ix++) {
    const int jx = ix > nx / 2 ? ix - nx : ix;
    for (iy = 0; iy < ny; iy++) {
      const int jy = iy > ny / 2 ? iy - ny : iy;
      for (iz = 0; iz < nzh; iz++)
        green_hat[((size_t)ix * ny + iy) * nzh + iz] =
            green_at(jx, jy, iz, nx, ny, nz, rl, inv_4eta2, inv_v, nalias, order, ik, &chi);
    }
  }
  return pow(inv_v, 1.0 / 3.0) * sqrt(2 * fabs(chi));
}

/* Cardinal B-spline M_p on [0,p] by the Cox-de Boor recursion.  The
 * reference's order-p assignment weights (pppm.c:69-172) are the values of
 * W_p(x) = M_p(x + p/2) at the p grid points around the particle, and its
 * "neg_weight_deriv" (pppm.c:174-275) is -dW/ds. */
void orc_p3m_weights(double s, int order, double *w, double *ndw)
{
  /* u = s + 1/2 in [0,1); w[k] = M_p(u + p - 1 - k) */
  const double u = s + 0.5;
  double m[8];
  int p, k;
  /* m[k] holds M_p(u + k) for k = 0..p-1 */
  m[0] = 1;
  for (k = 1; k < 8; k++) m[k] = 0;
  double prev[8];
  for (p = 2; p <= order; p++) {
    if (p == order && ndw) {
      /* derivative: M_p'(x) = M_{p-1}(x) - M_{p-1}(x-1) */
      for (k = 0; k < order; k++) {
        const double a = k < p - 1 ? m[k] : 0;
        const double b = k >= 1 ? m[k - 1] : 0;
        ndw[order - 1 - k] = -(a - b);
      }
    }
    for (k = 0; k < 8; k++) prev[k] = m[k];
    for (k = 0; k < p; k++) {
      const double x = u + k;
      const double a = k < p - 1 ? prev[k] : 0;     /* M_{p-1}(x) */
      const double b = k >= 1 ? prev[k - 1] : 0;    /* M_{p-1}(x-1) */
      m[k] = (x * a + (p - x) * b) / (p - 1);
    }
  }
  if (order == 1 && ndw) ndw[0] = 0;
  for (k = 0; k < order; k++) w[order - 1 - k] = m[k];
}

/* ---- FP64 mixed-radix complex FFT (test-only) ---- */
static void fft1d_rec(int n, int stride, const double *ire, const double *iim, double *ore,
                      double *oim, int sign)
{
  if (n <= 1) { ore[0] = ire[0]; oim[0] = iim[0]; return; }
  int p = 2;
  while (p * p <= n && n % p) p++;
  if (n % p) p = n;               /* prime: direct DFT */
  const int m = n / p;
  int r, k, q;
  /* sub-transforms of the p decimated sequences */
  for (r = 0; r < p; r++) fft1d_rec(m, stride * p, ire + (size_t)r * stride, iim + (size_t)r * stride,
                                   ore + (size_t)r * m, oim + (size_t)r * m, sign);
  double *tre = malloc(sizeof(double) * (size_t)n), *tim = malloc(sizeof(double) * (size_t)n);
  for (k = 0; k < m; k++)
    for (q = 0; q < p; q++) {
      const int kk = k + q * m;
      double sr = 0, si = 0;
      for (r = 0; r < p; r++) {
        const double ang = sign * 2 * M_PI * (double)((long)r * kk % n) / n;
        const double c = cos(ang), s = sin(ang);
        const double xr = ore[r * m + k], xi = oim[r * m + k];
        sr += xr * c - xi * s;
        si += xr * s + xi * c;
      }
      tre[kk] = sr; tim[kk] = si;
    }
  memcpy(ore, tre, sizeof(double) * (size_t)n);
  memcpy(oim, tim, sizeof(double) * (size_t)n);
  free(tre); free(tim);
}

void orc_fft3d(int nx, int ny, int nz, double *re, double *im, int sign)
{
  const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
  double *a = malloc(sizeof(double) * (size_t)nmax * 4);
  double *ai = a + nmax, *o = a + 2 * nmax, *oi = a + 3 * nmax;
  int i, j, k;
  /* z lines */
  for (i = 0; i < nx; i++)
    for (j = 0; j < ny; j++) {
      double *pr = re + ((size_t)i * ny + j) * nz, *pi = im + ((size_t)i * ny + j) * nz;
      fft1d_rec(nz, 1, pr, pi, o, oi, sign);
      memcpy(pr, o, sizeof(double) * (size_t)nz);
      memcpy(pi, oi, sizeof(double) * (size_t)nz);
    }
  /* y lines */
  for (i = 0; i < nx; i++)
    for (k = 0; k < nz; k++) {
      for (j = 0; j < ny; j++) {
        a[j] = re[((size_t)i * ny + j) * nz + k];
        ai[j] = im[((size_t)i * ny + j) * nz + k];
      }
      fft1d_rec(ny, 1, a, ai, o, oi, sign);
      for (j = 0; j < ny; j++) {
        re[((size_t)i * ny + j) * nz + k] = o[j];
        im[((size_t)i * ny + j) * nz + k] = oi[j];
      }
    }
  /* x lines */
  for (j = 0; j < ny; j++)
    for (k = 0; k < nz; k++) {
      for (i = 0; i < nx; i++) {
        a[i] = re[((size_t)i * ny + j) * nz + k];
        ai[i] = im[((size_t)i * ny + j) * nz + k];
      }
      fft1d_rec(nx, 1, a, ai, o, oi, sign);
      for (i = 0; i < nx; i++) {
        re[((size_t)i * ny + j) * nz + k] = o[i];
        im[((size_t)i * ny + j) * nz + k] = oi[i];
      }
    }
  free(a);
}

/* pppm.c:498-641 (assign, transform, convolve, back-transform, gather) */
void orc_p3m_field(int nsite, const double *r, const double *q,
                   int nx, int ny, int nz, const double *latv, double eta,
                   int order, int ik, const double *green_hat,
                   double *phi, double *evec, double *virial, double *energy)
{
  const size_t M = (size_t)nx * ny * nz;
  const int nzh = nz / 2 + 1;
  double inv[9], T[9];
  tensor_inverse(inv, latv);
  int i, a;
  /* invlatv rows scaled by n (pppm.c:523-525): row x is xx, xy, xz */
  for (a = 0; a < 9; a++) T[a] = inv[a];
  T[0] *= nx; T[3] *= nx; T[6] *= nx;
  T[1] *= ny; T[4] *= ny; T[7] *= ny;
  T[2] *= nz; T[5] *= nz; T[8] *= nz;
  const int odd = order % 2;
  const int istart = -((order + 1) / 2 - 1);
  double *mre = calloc(M, sizeof(double)), *mim = calloc(M, sizeof(double));
  double *W = malloc(sizeof(double) * 3 * (size_t)order * (size_t)nsite);
  double *D = malloc(sizeof(double) * 3 * (size_t)order * (size_t)nsite);
  int *I = malloc(sizeof(int) * 3 * (size_t)order * (size_t)nsite);
  for (i = 0; i < nsite; i++) {
    double s[3];
    tmul(T, r[3 * i], r[3 * i + 1], r[3 * i + 2], s);
    int j[3], d, p;
    for (d = 0; d < 3; d++) {
      if (odd) { j[d] = (int)floor(s[d] + 0.5); s[d] -= j[d]; }
      else { j[d] = (int)floor(s[d]); s[d] -= j[d] + 0.5; }
      orc_p3m_weights(s[d], order, W + ((size_t)i * 3 + d) * order, D + ((size_t)i * 3 + d) * order);
      const int n = d == 0 ? nx : (d == 1 ? ny : nz);
      for (p = 0; p < order; p++) I[((size_t)i * 3 + d) * order + p] = orc_mod(j[d] + istart + p, n);
    }
    const double *wx = W + ((size_t)i * 3) * order, *wy = wx + order, *wz = wy + order;
    const int *ix = I + ((size_t)i * 3) * order, *iy = ix + order, *iz = iy + order;
    int jx, jy, jz;
    for (jx = 0; jx < order; jx++)
      for (jy = 0; jy < order; jy++) {
        const double qw = wx[jx] * wy[jy] * q[i];
        for (jz = 0; jz < order; jz++)
          mre[((size_t)ix[jx] * ny + iy[jy]) * nz + iz[jz]] += qw * wz[jz];
      }
  }
  orc_fft3d(nx, ny, nz, mre, mim, -1);
  /* energy / virial from the mesh structure factor (full spectrum) */
  double rl[9];
  {
    int b;
    for (a = 0; a < 3; a++)
      for (b = 0; b < 3; b++) rl[3 * a + b] = 2 * M_PI * inv[3 * b + a];
  }
  const double c4 = 0.25 / (eta * eta);
  double e = 0;
  double *ex = NULL, *ey = NULL, *ez = NULL, *exi = NULL, *eyi = NULL, *ezi = NULL;
  if (ik && evec) {
    ex = calloc(M, sizeof(double)); ey = calloc(M, sizeof(double)); ez = calloc(M, sizeof(double));
    exi = calloc(M, sizeof(double)); eyi = calloc(M, sizeof(double)); ezi = calloc(M, sizeof(double));
  }
  int ixx, iyy, izz;
  for (ixx = 0; ixx < nx; ixx++)
    for (iyy = 0; iyy < ny; iyy++)
      for (izz = 0; izz < nz; izz++) {
        const size_t idx = ((size_t)ixx * ny + iyy) * nz + izz;
        /* Hermitian extension of the half-spectrum table */
        int hx = ixx, hy = iyy, hz = izz;
        if (izz >= nzh) { hx = (nx - ixx) % nx; hy = (ny - iyy) % ny; hz = nz - izz; }
        const double g = green_hat[((size_t)hx * ny + hy) * nzh + hz];
        const double rr = mre[idx], ii = mim[idx];
        if (virial || energy) {
          const int jx = ixx > nx / 2 ? ixx - nx : ixx, jy = iyy > ny / 2 ? iyy - ny : iyy,
                    jz = izz > nz / 2 ? izz - nz : izz;
          if (jx || jy || jz) {
            double k[3];
            tmul(rl, jx, jy, jz, k);
            const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
            const double ek = 0.5 * g * (rr * rr + ii * ii);
            e += ek;
            if (virial) {
              const double fac = 2 * (1 / k2 + c4);
              int b;
              for (a = 0; a < 3; a++)
                for (b = 0; b < 3; b++) virial[3 * a + b] += ek * ((a == b) - fac * k[a] * k[b]);
            }
          }
        }
        mre[idx] = rr * g;
        mim[idx] = ii * g;
        if (ik && evec) {
          /* pppm.c:577-594: E_hat = -i k phi_hat, k from the signed index
           * of the half-spectrum point it mirrors */
          const int sx = hx > nx / 2 ? hx - nx : hx, sy = hy > ny / 2 ? hy - ny : hy;
          double k[3];
          tmul(rl, sx, sy, hz, k);
          const double sgn = (izz >= nzh) ? -1.0 : 1.0;
          ex[idx] = sgn * k[0] * mim[idx]; exi[idx] = -sgn * k[0] * mre[idx];
          ey[idx] = sgn * k[1] * mim[idx]; eyi[idx] = -sgn * k[1] * mre[idx];
          ez[idx] = sgn * k[2] * mim[idx]; ezi[idx] = -sgn * k[2] * mre[idx];
        }
      }
  if (energy) *energy += e;
  orc_fft3d(nx, ny, nz, mre, mim, +1);
  if (ik && evec) {
    orc_fft3d(nx, ny, nz, ex, exi, +1);
    orc_fft3d(nx, ny, nz, ey, eyi, +1);
    orc_fft3d(nx, ny, nz, ez, ezi, +1);
  }
  for (i = 0; i < nsite; i++) {
    const double *wx = W + ((size_t)i * 3) * order, *wy = wx + order, *wz = wy + order;
    const double *dx = D + ((size_t)i * 3) * order, *dy = dx + order, *dz = dy + order;
    const int *ix = I + ((size_t)i * 3) * order, *iy = ix + order, *iz = iy + order;
    int jx, jy, jz;
    for (jx = 0; jx < order; jx++) {
      const double Wx = wx[jx];
      for (jy = 0; jy < order; jy++) {
        const double Wy = wy[jy], WxWy = Wx * Wy;
        for (jz = 0; jz < order; jz++) {
          const double Wz = wz[jz], w = WxWy * Wz;
          const size_t idx = ((size_t)ix[jx] * ny + iy[jy]) * nz + iz[jz];
          const double p = mre[idx];
          phi[i] += w * p;
          if (evec) {
            if (ik) {
              evec[3 * i] += w * ex[idx];
              evec[3 * i + 1] += w * ey[idx];
              evec[3 * i + 2] += w * ez[idx];
            } else {
              const double e0 = dx[jx] * Wy * Wz * p, e1 = Wx * dy[jy] * Wz * p, e2 = WxWy * dz[jz] * p;
              /* left_operate (cart.h:28-36): e -> e T */
              evec[3 * i] += e0 * T[0] + e1 * T[1] + e2 * T[2];
              evec[3 * i + 1] += e0 * T[3] + e1 * T[4] + e2 * T[5];
              evec[3 * i + 2] += e0 * T[6] + e1 * T[7] + e2 * T[8];
            }
          }
        }
      }
    }
  }
  free(mre); free(mim); free(W); free(D); free(I);
  if (ex) { free(ex); free(ey); free(ez); free(exi); free(eyi); free(ezi); }
}

/* ------------------------------------------------------------------ */
/* ElecLJ terms                                                        */
/* ------------------------------------------------------------------ */

/* eleclj.cpp:961-976 */
void orc_self_phi(int nsite, const double *q, double eta, double volume, double *phi)
{
  const double a = eta * M_2_SQRTPI;
  double qsum = 0;
  int i;
  for (i = 0; i < nsite; i++) { qsum += q[i]; phi[i] -= a * q[i]; }
  if (!almost_zero(qsum)) {
    const double b = -M_PI / (sq(eta) * volume) * qsum;
    for (i = 0; i < nsite; i++) phi[i] += b;
  }
}

/* eleclj.cpp:279-308 (totA, totB over distinct (sigma, epsilon) keys,
 * rounded through "%12.6g") and 1247-1266 (tail for the smoothed cutoff) */
double orc_lj_tail(int nsite, const int *site_mask, const double *sig, const double *eps,
                   int geometric_sigma, double rc, double width, double volume)
{
  double *ks = malloc(sizeof(double) * (size_t)(nsite > 0 ? nsite : 1));
  double *ke = malloc(sizeof(double) * (size_t)(nsite > 0 ? nsite : 1));
  int *kn = malloc(sizeof(int) * (size_t)(nsite > 0 ? nsite : 1));
  char **keys = malloc(sizeof(char *) * (size_t)(nsite > 0 ? nsite : 1));
  int nk = 0, i, j;
  for (i = 0; i < nsite; i++) {
    if (!(site_mask[i] & ORC_SITE_HAS_LJ)) continue;
    char key[128];
    snprintf(key, sizeof key, "%12.6g %12.6g", sig[i], eps[i]);
    for (j = 0; j < nk; j++)
      if (!strcmp(keys[j], key)) break;
    if (j == nk) {
      keys[nk] = strdup(key);
      sscanf(key, "%lf%lf", &ks[nk], &ke[nk]);
      kn[nk] = 0;
      nk++;
    }
    kn[j]++;
  }
  double totA = 0, totB = 0;
  for (i = 0; i < nk; i++)
    for (j = 0; j < nk; j++) {
      const double s6 = pow(geometric_sigma ? sqrt(ks[i] * ks[j]) : 0.5 * (ks[i] + ks[j]), 6);
      const double b = 4 * sqrt(ke[i] * ke[j]) * kn[i] * kn[j] * s6;
      totB -= b;
      totA += b * s6;
    }
  for (i = 0; i < nk; i++) free(keys[i]);
  free(keys); free(ks); free(ke); free(kn);
  const double rhi = rc, rlo = rhi - width;
  const double rlo2 = rlo * rlo, rlo3 = rlo2 * rlo, rlo4 = rlo3 * rlo;
  const double rlo5 = rlo4 * rlo, rlo6 = rlo5 * rlo, rlo7 = rlo6 * rlo, rlo9 = rlo7 * rlo2;
  const double rhi2 = rhi * rhi, rhi3 = rhi2 * rhi, rhi4 = rhi3 * rhi, rhi5 = rhi4 * rhi;
  const double rhi6 = rhi5 * rhi, rhi7 = rhi6 * rhi;
  return 2 * M_PI / volume *
         (totA / (9 * rlo9) + totB / (3 * rlo3) -
          (rhi - rlo) / (63 * rhi3 * rlo9 * pow(rhi + rlo, 5)) *
              (3 * totB * rhi3 * rlo6 * (7 * rhi4 + 42 * rhi3 * rlo + 112 * rhi2 * rlo2 + 150 * rhi * rlo3 + 25 * rlo4) +
               totA * (7 * rhi7 + 42 * rhi6 * rlo + 112 * rhi5 * rlo2 + 182 * rhi4 * rlo3 + 217 * rhi3 * rlo4 +
                       224 * rhi2 * rlo5 + 192 * rhi * rlo6 + 32 * rlo7)));
}

/* units.h:23-29, 91-94, 166-169, 201-204 */
double orc_e_to_charge_unit(double e) { return e * sqrt(0.5291772108 * 4.35974417 * 6.0221415 * 100 / 4.184); }
double orc_g_mol_to_mass_unit(double m) { return m / (100 * 4.184); }
double orc_K_to_kcal_mol(double t) { return t * (1.3806505 * 6.0221415 * 1e-3 / 4.184); }

/* Rows of a neighbor list visited as RSpaceEwald visits them: rows in
 * ascending order, each row in append order (rspace.c:413-418). */
static int *nl_rows(int ngroup, long np, const int *pairs)
{
  int *cnt = calloc((size_t)ngroup + 1, sizeof(int));
  int *nl = malloc(sizeof(int) * 2 * (size_t)(np > 0 ? np : 1));
  long k;
  int g;
  for (k = 0; k < np; k++) cnt[pairs[2 * k] + 1]++;
  for (g = 0; g < ngroup; g++) cnt[g + 1] += cnt[g];
  for (k = 0; k < np; k++) {
    const int row = pairs[2 * k];
    const int pos = cnt[row]++;
    nl[2 * pos] = row;
    nl[2 * pos + 1] = pairs[2 * k + 1];
  }
  free(cnt);
  return nl;
}

static int *build_nl(const orc_params_t *p, int ngroup, const int *group, const double *r,
                     long *np_out)
{
  double *gr = malloc(sizeof(double) * 3 * (size_t)ngroup);
  int g;
  for (g = 0; g < ngroup; g++) {
    gr[3 * g] = r[3 * group[2 * g]];
    gr[3 * g + 1] = r[3 * group[2 * g] + 1];
    gr[3 * g + 2] = r[3 * group[2 * g] + 2];
  }
  long np = orc_nl_build(ngroup, gr, p->boxl, p->rc, p->skin, NULL, 0);
  if (np < 0) { free(gr); return NULL; }
  int *pairs = malloc(sizeof(int) * 2 * (size_t)(np > 0 ? np : 1));
  orc_nl_build(ngroup, gr, p->boxl, p->rc, p->skin, pairs, np);
  int *nl = nl_rows(ngroup, np, pairs);
  free(pairs);
  free(gr);
  *np_out = np;
  return nl;
}

/* eleclj.cpp:785-807 + 912-959 for a cubic box with a given neighbor list */
static int evaluate_nl(const orc_params_t *p, int nsite, int ngroup, const int *group,
                       const double *r, const double *q, const double *sig, const double *eps,
                       const int *site_mask, int nexcl, const int *excl, long np, const int *nl,
                       double *f, double *energies, double *virial)
{
  const double L = p->boxl, V = L * L * L;
  int i;
  double *phi = calloc((size_t)nsite, sizeof(double));
  double *ev = calloc(3 * (size_t)nsite, sizeof(double));
  double *ljf = calloc(3 * (size_t)nsite, sizeof(double));
  double lje = 0;
  orc_cheby_t *cheb = malloc(sizeof(orc_cheby_t));
  orc_cheby_rspace(cheb, p->rc, p->eta);
  if (virial) memset(virial, 0, 9 * sizeof(double));
  orc_rspace_ewald(nsite, ngroup, group, r, L, np, nl, nexcl, excl,
                   sq(p->rc - p->smoothing_width), sq(p->rc), site_mask, q, p->eta, cheb,
                   NULL, sig, eps, p->geometric_sigma, 0, NULL, 1, 1, phi, ev, &lje, ljf, virial);
  double ek = 0;
  double latv[9] = { L, 0, 0, 0, L, 0, 0, 0, L };
  if (p->kmode == 0) {
    double inv[9], rl[9];
    int a, b;
    tensor_inverse(inv, latv);
    for (a = 0; a < 3; a++)
      for (b = 0; b < 3; b++) rl[3 * a + b] = 2 * M_PI * inv[3 * b + a];
    orc_kspace_field(p->kc, nsite, r, q, rl, V, p->eta, phi, ev, virial, &ek);
  } else {
    const size_t nh = (size_t)p->nx * p->ny * (p->nz / 2 + 1);
    double *gh = malloc(sizeof(double) * nh);
    orc_p3m_green(p->nx, p->ny, p->nz, latv, p->eta, p->nalias, p->order, p->ik, gh);
    orc_p3m_field(nsite, r, q, p->nx, p->ny, p->nz, latv, p->eta, p->order, p->ik, gh, phi, ev,
                  virial, &ek);
    free(gh);
  }
  orc_self_phi(nsite, q, p->eta, V, phi);
  double tail = 0;
  if (p->include_lj_correction)
    tail = orc_lj_tail(nsite, site_mask, sig, eps, p->geometric_sigma, p->rc, p->smoothing_width, V);
  double ue = 0;
  for (i = 0; i < nsite; i++) ue += q[i] * phi[i];
  ue /= 2;
  for (i = 0; i < nsite; i++) {
    f[3 * i] = q[i] * ev[3 * i] + ljf[3 * i];
    f[3 * i + 1] = q[i] * ev[3 * i + 1] + ljf[3 * i + 1];
    f[3 * i + 2] = q[i] * ev[3 * i + 2] + ljf[3 * i + 2];
  }
  if (virial) { virial[0] += tail; virial[4] += tail; virial[8] += tail; }
  energies[0] = ue;
  energies[1] = lje + tail;
  energies[2] = tail;
  energies[3] = ue + lje + tail;
  free(phi); free(ev); free(ljf); free(cheb);
  return 0;
}

/* eleclj.cpp:785-807 with a freshly built neighbor list */
int orc_evaluate(const orc_params_t *p, int nsite, int ngroup, const int *group,
                 const double *r, const double *q, const double *sig, const double *eps,
                 const int *site_mask, int nexcl, const int *excl,
                 double *f, double *energies, double *virial)
{
  long np = 0;
  int *nl = build_nl(p, ngroup, group, r, &np);
  if (!nl) return -1;
  int rc = evaluate_nl(p, nsite, ngroup, group, r, q, sig, eps, site_mask, nexcl, excl, np, nl,
                       f, energies, virial);
  free(nl);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Rigid water dynamics                                                */
/* ------------------------------------------------------------------ */

/* geograd.cpp:59-103 (theta, AngleGradient) */
static double angle_grad(const double *r1, const double *r2, const double *r3,
                         double *g1, double *g2, double *g3)
{
  double a[3], b[3];
  int d;
  for (d = 0; d < 3; d++) { a[d] = r1[d] - r2[d]; b[d] = r3[d] - r2[d]; }
  const double a2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
  const double b2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
  if (almost_zero(a2) || almost_zero(b2)) {
    for (d = 0; d < 3; d++) g1[d] = g2[d] = g3[d] = 0;
    return 0;
  }
  const double ma2 = 1.0 / a2, mb2 = 1.0 / b2, ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  const double ma = sqrt(ma2), mb = sqrt(mb2);
  const double ma3 = ma * ma2, mb3 = mb * mb2;
  const double t0 = ma * mb, t3 = -ma3 * mb * ab, t4 = -ab * ma * mb3;
  for (d = 0; d < 3; d++) {
    g1[d] = b[d] * t0 + a[d] * t3;
    g3[d] = a[d] * t0 + b[d] * t4;
    g2[d] = -g1[d] - g3[d];
  }
  double ct = ab * t0, dt;
  if (ct > 1) ct = 1;
  else if (ct < -1) ct = -1;
  const double s2t = 1 - ct * ct;
  double t;
  if (s2t < 64.0 * 2.220446049250313e-16) { dt = 0; t = ct > 0 ? 0 : M_PI; }
  else { dt = -1 / sqrt(s2t); t = acos(ct); }
  for (d = 0; d < 3; d++) { g1[d] *= dt; g2[d] *= dt; g3[d] *= dt; }
  return t;
}

/* constraint values and gradients for one water (atoms O=0,H1=1,H2=2):
 * c0 = length O-H1, c1 = length O-H2 (constraint.cpp:30-66, value
 * 0.5(r^2 - l^2), gradient +-r), c2 = angle H1-O-H2 (constraint.cpp:68-112) */
static void water_cnstr(const double *x, double bond, double t0, double *val, double (*grad)[9])
{
  int d, c;
  for (c = 0; c < 3; c++)
    for (d = 0; d < 9; d++) grad[c][d] = 0;
  double r01[3], r02[3];
  for (d = 0; d < 3; d++) { r01[d] = x[d] - x[3 + d]; r02[d] = x[d] - x[6 + d]; }
  val[0] = 0.5 * ((r01[0] * r01[0] + r01[1] * r01[1] + r01[2] * r01[2]) - bond * bond);
  val[1] = 0.5 * ((r02[0] * r02[0] + r02[1] * r02[1] + r02[2] * r02[2]) - bond * bond);
  for (d = 0; d < 3; d++) {
    grad[0][d] = r01[d]; grad[0][3 + d] = -r01[d];
    grad[1][d] = r02[d]; grad[1][6 + d] = -r02[d];
  }
  double g1[3], g2[3], g3[3];
  const double t = angle_grad(x + 3, x, x + 6, g1, g2, g3);
  val[2] = t - t0;
  for (d = 0; d < 3; d++) { grad[2][3 + d] = g1[d]; grad[2][d] = g2[d]; grad[2][6 + d] = g3[d]; }
}

/* Gaussian elimination with partial pivoting (the role of LAPACK dgesv in
 * linalg.cpp:29-37) */
static int solve3(double s[3][3], double *x)
{
  int c, r, k;
  for (c = 0; c < 3; c++) {
    int piv = c;
    for (r = c + 1; r < 3; r++)
      if (fabs(s[r][c]) > fabs(s[piv][c])) piv = r;
    if (s[piv][c] == 0) return -1;
    if (piv != c) {
      for (k = 0; k < 3; k++) { double t = s[c][k]; s[c][k] = s[piv][k]; s[piv][k] = t; }
      double t = x[c]; x[c] = x[piv]; x[piv] = t;
    }
    for (r = c + 1; r < 3; r++) {
      const double m = s[r][c] / s[c][c];
      for (k = c; k < 3; k++) s[r][k] -= m * s[c][k];
      x[r] -= m * x[c];
    }
  }
  for (c = 2; c >= 0; c--) {
    double t = x[c];
    for (k = c + 1; k < 3; k++) t -= s[c][k] * x[k];
    x[c] = t / s[c][c];
  }
  return 0;
}

/* constraint.cpp:295-335 for each water */
int orc_constrain_positions(int nwater, double *r, double *v, const double *r0,
                            const double *mass, double dt, double bond, double angle_deg,
                            double tol, int maxit)
{
  const double t0 = angle_deg * (M_PI / 180);
  int w, fail = 0;
  for (w = 0; w < nwater; w++) {
    double *x = r + 9 * w, *vv = v ? v + 9 * w : NULL;
    const double *m = mass + 3 * w;
    double val0[3], g0[3][9], sci0[3][9];
    int c, d, it;
    water_cnstr(r0 + 9 * w, bond, t0, val0, g0);
    for (c = 0; c < 3; c++)
      for (d = 0; d < 9; d++) sci0[c][d] = g0[c][d] / m[d / 3];
    int ok = 0;
    for (it = 0; it < maxit; it++) {
      double val[3], g[3][9], xs[3], s[3][3];
      water_cnstr(x, bond, t0, val, g);
      int sat = 1;
      for (c = 0; c < 3; c++) { xs[c] = -val[c]; if (fabs(xs[c]) > tol) sat = 0; }
      if (sat) { ok = 1; break; }
      int cb;
      for (c = 0; c < 3; c++)
        for (cb = 0; cb < 3; cb++) {
          double t = 0;
          for (d = 0; d < 9; d++) t += g[c][d] * sci0[cb][d];
          s[c][cb] = t;
        }
      if (solve3(s, xs)) break;
      const double invdt = dt > 0 ? 1 / dt : 0;
      for (c = 0; c < 3; c++)
        for (d = 0; d < 9; d++) {
          const double dx = xs[c] * sci0[c][d];
          x[d] += dx;
          if (vv) vv[d] += dx * invdt;
        }
    }
    if (!ok) fail++;
  }
  return fail;
}

/* constraint.cpp:336-358 */
void orc_constrain_velocities(int nwater, const double *r, double *v, const double *mass,
                              double bond, double angle_deg)
{
  const double t0 = angle_deg * (M_PI / 180);
  int w;
  for (w = 0; w < nwater; w++) {
    const double *x = r + 9 * w;
    double *vv = v + 9 * w;
    const double *m = mass + 3 * w;
    double val[3], g[3][9], sci0[3][9], xs[3], s[3][3];
    int c, cb, d;
    water_cnstr(x, bond, t0, val, g);
    for (c = 0; c < 3; c++) {
      xs[c] = 0;
      for (d = 0; d < 9; d++) { sci0[c][d] = g[c][d] / m[d / 3]; xs[c] -= g[c][d] * vv[d]; }
    }
    for (c = 0; c < 3; c++)
      for (cb = c; cb < 3; cb++) {
        double t = 0;
        for (d = 0; d < 9; d++) t += g[c][d] * sci0[cb][d];
        s[c][cb] = s[cb][c] = t;
      }
    if (solve3(s, xs)) continue;
    for (c = 0; c < 3; c++)
      for (d = 0; d < 9; d++) vv[d] += xs[c] * sci0[c][d];
  }
}

/* mdyn.cpp:120-160: velocity Verlet with position and velocity constraints;
 * the neighbor list is kept and rebuilt when the skin test of
 * NeighborList::coordinates_changed fires (nbrlist.cpp:51-77).
 * All groups must be waters (3 sites, O first).  energies_out receives, per
 * step, {potential, kinetic}. */
int orc_md_run(const orc_params_t *p, int nsite, int ngroup, const int *group,
               double *r, double *v, const double *mass, const double *q,
               const double *sig, const double *eps, const int *site_mask,
               int nexcl, const int *excl, double dt, int nsteps,
               double bond, double angle_deg, double tol, double *energies_out)
{
  const int nw = ngroup;
  double *f = malloc(sizeof(double) * 3 * (size_t)nsite);
  double *r0 = malloc(sizeof(double) * 3 * (size_t)nsite);
  double *gr = malloc(sizeof(double) * 3 * (size_t)ngroup);
  double *last = malloc(sizeof(double) * 3 * (size_t)ngroup);
  double en[4];
  int i, s, g;
  long np = 0;
  int *nl = NULL;
#define FIRST_SITES(dst)                                  \
  for (g = 0; g < ngroup; g++) {                          \
    dst[3 * g] = r[3 * group[2 * g]];                     \
    dst[3 * g + 1] = r[3 * group[2 * g] + 1];             \
    dst[3 * g + 2] = r[3 * group[2 * g] + 2];             \
  }
  memcpy(r0, r, sizeof(double) * 3 * (size_t)nsite);
  orc_constrain_positions(nw, r, v, r0, mass, 0, bond, angle_deg, tol, 10000);
  memcpy(r0, r, sizeof(double) * 3 * (size_t)nsite);
  nl = build_nl(p, ngroup, group, r, &np);
  FIRST_SITES(last);
  evaluate_nl(p, nsite, ngroup, group, r, q, sig, eps, site_mask, nexcl, excl, np, nl, f, en, NULL);
  for (s = 0; s < nsteps; s++) {
    for (i = 0; i < 3 * nsite; i++) v[i] += f[i] * (dt / 2 / mass[i / 3]);
    for (i = 0; i < 3 * nsite; i++) r[i] += v[i] * dt;
    if (orc_constrain_positions(nw, r, v, r0, mass, dt, bond, angle_deg, tol, 10000)) {
      free(f); free(r0); free(gr); free(last); free(nl); return -s;
    }
    memcpy(r0, r, sizeof(double) * 3 * (size_t)nsite);
    FIRST_SITES(gr);
    if (orc_nl_needs_rebuild(ngroup, gr, last, p->skin)) {
      free(nl);
      nl = build_nl(p, ngroup, group, r, &np);
      memcpy(last, gr, sizeof(double) * 3 * (size_t)ngroup);
    }
    evaluate_nl(p, nsite, ngroup, group, r, q, sig, eps, site_mask, nexcl, excl, np, nl, f, en, NULL);
    for (i = 0; i < 3 * nsite; i++) v[i] += f[i] * (dt / 2 / mass[i / 3]);
    orc_constrain_velocities(nw, r, v, mass, bond, angle_deg);
    if (energies_out) {
      double ke = 0;
      for (i = 0; i < 3 * nsite; i++) ke += mass[i / 3] * v[i] * v[i];
      energies_out[2 * s] = en[3];
      energies_out[2 * s + 1] = ke / 2;
    }
  }
#undef FIRST_SITES
  free(f); free(r0); free(gr); free(last); free(nl);
  return nsteps;
}
