/*
 * mm_oracle.h -- FP64 CPU oracle for molmodel's nonbonded force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path in
 * molmodel_amd/ is compared against.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.  The product never links it
 * and never falls back to it.
 *
 * It restates, in plain C99 (FP64, no FMA contraction), the algorithms of:
 *   nbrlist.cpp   cell neighbor list (ncell, nmax, cell pairs, rebuild, skin test)
 *   pbin.c        PairBinary canonical rows, sort/dedup, exists
 *   pbc.h/bcond.cpp  cubic minimum image, nearest lattice point
 *   cheby.c       512-piece cubic Chebyshev table of erfc(eta r)/r in r^2
 *   rspace.c      RSpaceEwald group-pair loop (+ excluded and 1-4 loops)
 *   kspace.c      plain Ewald reciprocal sum
 *   pppm.c        P3M mesh sizing, optimal influence function, field
 *   eleclj.cpp    self / background terms, LJ tail correction, totals
 *   mdyn.cpp / msys.cpp / constraint.cpp   velocity Verlet with SHAKE/RATTLE
 *                                          style Newton constraint iteration
 * Each function cites the reference file:line it follows.  What pins each
 * function (golden vectors from the compiled reference, closed forms,
 * invariants) is listed in the test files under tests/.
 *
 * Parity status: every function is pinned by at least one check in tests/
 * (reference golden vectors, closed forms or brute force) except the ones
 * whose header comment says "parity unpinned".
 */
#ifndef MM_ORACLE_H
#define MM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NCHEBY 512
#define ORC_SITE_HAS_Q 1
#define ORC_SITE_HAS_LJ 2

/* ---- PairBinary semantics (pbin.c:6, 41-50, 73-90, 119-137) ---- */
/* Canonical placement: even i+j -> row max(i,j); odd -> row min(i,j). */
void orc_pbin_canonical(int *i, int *j);
/* Build canonical sorted/deduplicated CSR from an (i,j) list.
 * rowptr has n+1 entries; cols has capacity >= npair. Returns nnz. */
int orc_pbin_build(int n, int npair, const int *pairs, int *rowptr, int *cols);
int orc_pbin_exists(const int *rowptr, const int *cols, int i, int j);

/* ---- Neighbor list (nbrlist.cpp) ---- */
int orc_nl_ncell(int ngroup);                                 /* nbrlist.cpp:11 */
int orc_nl_nmax(int ncell, double boxl, double rc, double skin); /* nbrlist.cpp:96-102 */
/* Cell-pair rows after PairBinary_init (nbrlist.cpp:105-120).  rowptr has
 * ncell^3+1 entries, cols capacity maxc.  Returns nnz or -1 on overflow. */
int orc_nl_cell_pairs(int ncell, double boxl, double rc, double skin,
                      int *rowptr, int *cols, int maxc);
/* Periodic rebuild (nbrlist.cpp:222-251).  Writes (row, col) pairs in the
 * reference's append order; returns the count (may exceed maxp, in which
 * case only maxp pairs are written). */
long orc_nl_build(int ngroup, const double *gr, double boxl, double rc, double skin,
                  int *pairs, long maxp);
/* Skin test (nbrlist.cpp:64-76): 1 if largest + second largest
 * displacement > skin. */
int orc_nl_needs_rebuild(int n, const double *r, const double *last, double skin);

/* ---- Chebyshev table (cheby.c:7-45, rspace.c:19-37) ---- */
typedef struct { double min, max, inv_dx, coeff[ORC_NCHEBY][7]; } orc_cheby_t;
void orc_cheby_rspace(orc_cheby_t *t, double rhi, double eta);
void orc_cheby_eval(const orc_cheby_t *t, double x, double *f, double *n2df);

/* ---- Real space (rspace.c:382-558) ----
 * group: 2*ngroup ints [start,end).  r: 3*nsite.  nl: 2*nnl group pairs in
 * the order they are to be visited (row, col).  excl / one_four: site pairs.
 * Outputs are accumulated (+=).  virial (9, row-major xx,xy,xz,yx..) may be
 * NULL; it receives sum over interactions of d (x) f with f the force on the
 * first member of each pair and d its displacement from the second. */
void orc_rspace_ewald(int nsite, int ngroup, const int *group, const double *r,
                      double boxl, long nnl, const int *nl,
                      int nexcl, const int *excl, double rlo2, double rhi2,
                      const int *site_mask, const double *q, double eta,
                      const orc_cheby_t *cheb, const double *rscreen,
                      const double *sig, const double *eps, int geometric_sigma,
                      int n14, const int *one_four, double s14e, double s14lj,
                      double *phi, double *evec, double *lj_ener, double *lj_force,
                      double *virial);

/* RSpaceEwaldPhi (rspace.c:649-782): phi only; excluded pairs without the
 * minimum image (:747-751).  Accumulated (+=). */
void orc_rspace_ewald_phi(int nsite, int ngroup, const int *group, const double *r,
                          double boxl, long nnl, const int *nl, int nexcl, const int *excl,
                          double rlo2, double rhi2, const int *site_mask, const double *q,
                          double eta, const orc_cheby_t *cheb, const double *rscreen,
                          double *phi);

/* RSpaceEwaldPhiAtGridpoints (rspace.c:892-967): phi at ngrid points
 * (AoS rgrid) from all sites with any mask bit.  Accumulated (+=). */
/* ---- Non-periodic real space (rspace.c:560-640, 784-890) ----
 * no group exclusions, no 1-4 list; rscreen per site (zeros: no screening).
 * Accumulated (+=). */
void orc_rspace_nonperiodic(int nsite, int ngroup, const int *group, const double *r,
                            const int *site_mask, const double *q, const double *rscreen,
                            const double *sig, const double *eps, int geometric, double *phi,
                            double *evec, double *lj_ener, double *lj_force);
void orc_rspace_nonperiodic_phi(int nsite, int ngroup, const int *group, const double *r,
                                const int *site_mask, const double *q, const double *rscreen,
                                double *phi);
void orc_rspace_nonperiodic_phi_at_gridpoints(int nsite, int ngrid, const double *rsite,
                                              const double *rgrid, const int *site_mask,
                                              const double *q, const double *rscreen, double *phi);

void orc_rspace_phi_at_gridpoints(int nsite, int ngrid, const double *rsite,
                                  const double *rgrid, double boxl, double rlo2, double rhi2,
                                  const int *site_mask, const double *q, double eta,
                                  const orc_cheby_t *cheb, const double *rscreen, double *phi);

/* ---- Plain Ewald k-space (kspace.c:35-174) ---- */
int orc_kspace_count(double kc);
/* integer vectors in the reference order (sorted by |n|^2 descending) */
void orc_kspace_vectors(double kc, int *ikvec);
void orc_kspace_field(double kc, int nsite, const double *r, const double *q,
                      const double *recip /*tensor_t layout, 9*/, double volume,
                      double eta, double *phi, double *evec, double *virial,
                      double *energy);

/* ---- P3M (pppm.c) ---- */
int orc_p3m_grid_size(double length, double spacing);         /* pppm.c:277-292, 429-431 */
/* Influence function over the half spectrum nx*ny*(nz/2+1) (pppm.c:313-393).
 * latv: tensor_t layout (columns are lattice vectors).  Returns chi. */
double orc_p3m_green(int nx, int ny, int nz, const double *latv, double eta,
                     int nalias, int order, int ik, double *green_hat);
/* Assignment weights and negative derivatives for one coordinate
 * (pppm.c:69-275), evaluated as cardinal B-splines. */
void orc_p3m_weights(double s, int order, double *w, double *ndw);
/* Field (pppm.c:498-641): phi, evec accumulated.  Uses the project's FP64
 * mixed-radix FFT.  virial / energy optional (NULL). */
void orc_p3m_field(int nsite, const double *r, const double *q,
                   int nx, int ny, int nz, const double *latv, double eta,
                   int order, int ik, const double *green_hat,
                   double *phi, double *evec, double *virial, double *energy);
/* In-place complex 3-D FFT, sign -1 forward, +1 backward, unnormalised. */
void orc_fft3d(int nx, int ny, int nz, double *re, double *im, int sign);

/* ---- ElecLJ orchestration (eleclj.cpp) ---- */
void orc_self_phi(int nsite, const double *q, double eta, double volume, double *phi); /* :961-976 */
/* LJ tail (eleclj.cpp:273-308, 1247-1266): returns energy correction. */
double orc_lj_tail(int nsite, const int *site_mask, const double *sig, const double *eps,
                   int geometric_sigma, double rc, double width, double volume);
double orc_e_to_charge_unit(double e);                         /* units.h:91-94 */
double orc_g_mol_to_mass_unit(double m);                       /* units.h:166-169 */
double orc_K_to_kcal_mol(double t);                            /* units.h:201-204 */

/* Full force evaluation for a cubic periodic system (eleclj.cpp:785-807,
 * 912-959).  kmode: 0 = Ewald (kc), 1 = P3M (nx,ny,nz,order,nalias).
 * NL is rebuilt from scratch.  Outputs: f (3N, overwritten), energies[4] =
 * {elec, lj (incl. tail), tail, total}, virial[9] (may be NULL). */
typedef struct {
  double boxl, rc, smoothing_width, skin, eta;
  int kmode;
  double kc;
  int nx, ny, nz, order, nalias, ik;
  int include_lj_correction, geometric_sigma;
} orc_params_t;
int orc_evaluate(const orc_params_t *p, int nsite, int ngroup, const int *group,
                 const double *r, const double *q, const double *sig, const double *eps,
                 const int *site_mask, int nexcl, const int *excl,
                 double *f, double *energies, double *virial);

/* ---- Rigid-water MD (mdyn.cpp:142-160, msys.cpp:1275-1285,
 *      constraint.cpp:295-358) ----
 * Constraints per water (O,H1,H2): two length constraints O-H and one angle
 * constraint H-O-H, exactly as the reference builds them.  Returns the
 * number of steps completed (or -step on constraint failure). */
int orc_md_run(const orc_params_t *p, int nsite, int ngroup, const int *group,
               double *r, double *v, const double *mass, const double *q,
               const double *sig, const double *eps, const int *site_mask,
               int nexcl, const int *excl, double dt, int nsteps,
               double bond, double angle_deg, double tol, double *energies_out);
int orc_constrain_positions(int nwater, double *r, double *v, const double *r0,
                            const double *mass, double dt, double bond, double angle_deg,
                            double tol, int maxit);
void orc_constrain_velocities(int nwater, const double *r, double *v, const double *mass,
                              double bond, double angle_deg);

#ifdef __cplusplus
}
#endif
#endif
