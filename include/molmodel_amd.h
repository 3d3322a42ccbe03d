/*
 * molmodel_amd.h -- C ABI of the MI355X nonbonded force path.
 *
 * Two layers:
 *  1. Device-resident engine (mm_*): a batch of periodic cubic systems whose
 *     positions, velocities and forces stay in HBM.  One call evaluates the
 *     full ElecLJ force (neighbor list, real-space Ewald + LJ, plain Ewald
 *     k-space or P3M, self/background terms) or runs velocity-Verlet MD with
 *     rigid-water constraints.  This is what replaces
 *     ElecLJ::add_to_energy_and_forces (eleclj.cpp:785-807) and the MD loop
 *     body of MDyn::run (mdyn.cpp:142-160).
 *  2. Reference-named drop-ins (RSpaceEwald, kspace_*, p3m_*, PairBinary_*,
 *     RSpaceChebyshev_new) with the reference's signatures and accumulate
 *     semantics (rspace.h, kspace.h, pppm.h, pbin.h, cheby.h), so existing
 *     callers relink against this library unchanged.  Declared in
 *     molmodel_amd_ref.h.
 *
 * Units are the reference's: A, kcal/mol, ps; charges in sqrt(kcal A/mol)
 * (units.h:91-94); masses in (kcal/mol) ps^2/A^2 (units.h:166-169).
 * Arrays of vectors are AoS double[3*n] in the caller's site order.
 * Every function returns 0 on success or a negative MM_ERR_* code; the text
 * of the last error is available from mm_last_error().  The reference has
 * no return codes (it asserts or calls die(), util.c:11-26); the conditions
 * that are fatal there are errors here.
 * Thread safety: one engine per host thread at a time; each engine owns one
 * HIP stream.
 */
#ifndef MOLMODEL_AMD_H
#define MOLMODEL_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define MM_OK 0
#define MM_ERR_ARG -1        /* invalid argument / unsupported input */
#define MM_ERR_HIP -2        /* HIP runtime failure */
#define MM_ERR_DEVICE -3     /* device-side error flag raised (see mm_errors) */

/* device error flags reported by mm_errors().  An overflow (neighbour row,
 * P3M column bin) is not an error for the caller: the evaluation, MD block
 * or exchange interval in which it happened runs again from its start with
 * larger buffers. */
#define MM_FLAG_NL_OVERFLOW 1
#define MM_FLAG_BIN_OVERFLOW 16
#define MM_FLAG_CONSTRAINT 4
#define MM_FLAG_NAN 8          /* non-finite energy, or a non-finite / out-of-range position */
#define MM_FLAG_CELL 32        /* positions collapsed: a neighbour-list cell held > 512 groups */

#define MM_SITE_HAS_Q 1      /* rspace.h:7 */
#define MM_SITE_HAS_LJ 2     /* rspace.h:8 */

/* One periodic cubic system.  Mirrors the inputs ElecLJ::init and
 * ElecLJ::init_sites derive (eleclj.cpp:140-265, 643-771):
 * groups = heavy atom first, then its terminal atoms, contiguous sites;
 * excluded pairs = 1-2 and 1-3 site pairs (must lie inside one group). */
typedef struct {
  int nsite, ngroup;
  const int *group;          /* 2*ngroup: [start, end) site ranges */
  const double *q;           /* nsite charges, sqrt(kcal A/mol) */
  const double *sigma;       /* nsite LJ sigma (A) */
  const double *epsilon;     /* nsite LJ epsilon (kcal/mol) */
  const int *site_mask;      /* nsite: MM_SITE_HAS_Q | MM_SITE_HAS_LJ */
  const double *mass;        /* nsite masses (may be NULL: no dynamics) */
  int nexcl;
  const int *excl;           /* 2*nexcl excluded site pairs */
  double boxl;               /* cubic box length */
  double rc;                 /* rspace_cutoff */
  double smoothing_width;    /* smoothing_width (reference default 1 A) */
  double skin;               /* neighbor_list_width */
  double eta;                /* ewald_screening */
  int kmode;                 /* 0: plain Ewald (kc), 1: P3M */
  double kc;                 /* kspace_cutoff */
  int mesh[3];               /* P3M mesh */
  int order;                 /* p3m_assignment_order (1..7) */
  int nalias;                /* p3m_alias_sum_limit */
  int ik;                    /* p3m_ik_differentiate: field as -i k phi on the mesh */
  int include_lj_correction;
  int geometric_sigma;       /* geometric_combining_for_sigma */
} mm_system_desc;

/* ElecLJ's defaults, for callers that fill a descriptor from ElecLJ's io
 * parameters.
 * mm_desc_init: zeroes *d and sets the constructor's defaults
 *   (eleclj.cpp:13-37): smoothing_width 1, kc 5, order 3, nalias 2,
 *   include_lj_correction 1, Ewald (kmode 0), and the values ElecLJ::init
 *   treats as unset: rc 0, skin 0, eta -1.
 * mm_desc_resolve: replaces the unset values as ElecLJ::init does for a
 *   periodic box (eleclj.cpp:164-201, 229-230), with d->boxl set:
 *   rc 0 -> 0.45 boxl (rc > 0.49 boxl is an error);
 *   skin 0 -> 0.2 rc, or 0.499 boxl - rc when rc + 0.2 rc >= 0.5 boxl;
 *   P3M (kmode 1) with mesh[0..2] not all positive -> the mesh of
 *     p3m_new at p3m_grid_spacing, or at boxl / cbrt(density * boxl^3)
 *     when p3m_grid_density > 0 (then the spacing must be the default 0.5);
 *   eta < 0 -> Ewald: sqrt(pi kc / (rc boxl)); P3M: p3m_optimize's minimum
 *     of the real-space and mesh RMS error estimates.
 * Host arithmetic only: both run without a GPU. */
void mm_desc_init(mm_system_desc *d);
int mm_desc_resolve(mm_system_desc *d, double p3m_grid_spacing, double p3m_grid_density);

/* Sites, groups and excluded pairs of ElecLJ::init_sites
 * (eleclj.cpp:643-771) for molecules whose sites are their atoms (no M,
 * sp2 or sp3 virtual sites).  Bonds in CSR form: the neighbours of atom a
 * are nbr[nbr_start[a] .. nbr_start[a+1]), in the order of Atom::neighbors,
 * each bond listed from both atoms.  Outputs:
 *   atom_of_site[natom]  the atom at each site.  Atoms are visited in
 *       order; a non-terminal (or isolated) atom opens a group and takes its
 *       terminal neighbours in neighbour order; two terminal atoms bonded to
 *       each other form one group, opened by the lower index;
 *   group[2*natom]       [start, end) site ranges, *ngroup of them;
 *   excl[2*maxexcl]      the 1-2 and 1-3 site pairs (i < j), sorted, each
 *       once; *nexcl is their number (entries past maxexcl are not written);
 *   *ncross              how many of them join two groups (the engine
 *       requires none: mm_engine_create rejects such pairs).
 * Host code only: runs without a GPU. */
int mm_derive_sites(int natom, const int *nbr_start, const int *nbr, int *atom_of_site,
                    int *group, int *ngroup, int *excl, long maxexcl, long *nexcl,
                    long *ncross);

/* The buffer layout mm_engine_create would give this batch, checked on the
 * host (no device needed): every system's slice of each concatenated device
 * buffer (sites, groups, cells, stencils, erfc and LJ tables, wave vectors,
 * meshes, influence functions, twiddles, spreading bins) lies inside the
 * buffer and the slices are disjoint.  0, or MM_ERR_ARG with the first
 * violation (or the input error engine creation would report). */
int mm_check_layout(int nsys, const mm_system_desc *sys);

/* ---- msim input and restart formats (host code: no GPU needed) ---- */

/* free memory the library returned (mm_msim_preprocess) */
void mm_free(void *p);

/* msim's input preprocessor (PreprocessedStream, prepro.cpp:23-318).
 * args as msim's command line after the program name: "-name value" defines
 * a variable, any other entry is a file read in turn (and sets $base to its
 * name without the extension); $pwd and $pid are predefined, and
 * $MSIM_DIR/param is searched as msim does.  Per line: '#' starts a
 * comment; $name is replaced (defined variables, then the environment);
 * !define name value, !include file or <file, !path dir (the directory is
 * taken without surrounding whitespace), !foreach v { values } ... !end,
 * !repeat n ... !end, !ignore ... !end.  !parallel_define (MPI only) and
 * reading stdin are not provided.  *out: the processed text (free with
 * mm_free). */
int mm_msim_preprocess(int nargs, const char *const *args, char **out);

/* ElecLJ's io parameters (eleclj.hpp:310-327) read from msim text, starting
 * from the constructor's defaults and table entries (eleclj.cpp:12-40).
 * deck: a whole program ("{ ... potential { eleclj { ... } } ... }", every
 * eleclj block read in turn, a bare "eleclj" meaning the defaults) or one
 * block "{ field value ... }".  Unknown fields and malformed values are
 * errors, as in the reference's generated reader (io:219-312). */
typedef struct mm_msim_eleclj mm_msim_eleclj;
int mm_msim_eleclj_read(const char *deck, mm_msim_eleclj **out);
void mm_msim_eleclj_free(mm_msim_eleclj *p);
/* a scalar field (numbers, integers, Booleans as 0/1; for the virtual-site,
 * bond-charge-increment and 1-3 tables: whether any entry was given) */
int mm_msim_eleclj_get(const mm_msim_eleclj *p, const char *field, double *value);
/* an entry of fixed_charge, screening_radius, sigma or epsilon: 0, or 1 if
 * the table has no such key */
int mm_msim_eleclj_table(const mm_msim_eleclj *p, const char *table, const char *key,
                         double *value);
/* the descriptor's parameters from the block (d->boxl set by the caller):
 * rspace_cutoff, smoothing_width, neighbor_list_width, ewald_screening,
 * kspace_cutoff, use_p3m, p3m_assignment_order, p3m_alias_sum_limit,
 * include_lj_correction, geometric_combining_for_sigma, then
 * mm_desc_resolve with p3m_grid_spacing / p3m_grid_density.  Parameters
 * the device path does not evaluate (virtual sites, bond charge
 * increments, 1-3 interactions, 1-4 scaling, a term switched off) are
 * errors; p3m_ik_differentiate sets ik.  Sites and groups are left to the caller. */
int mm_msim_eleclj_desc(const mm_msim_eleclj *p, mm_system_desc *d);
/* per-atom charge (sqrt(kcal A/mol)), sigma, epsilon and site mask from the
 * atoms' types as ElecLJ::types_changed derives them (eleclj.cpp:440-500):
 * fixed_charge[type] unless almost zero -> MM_SITE_HAS_Q; sigma and
 * epsilon of general_type[type] (or the type) when both exist and epsilon
 * is not almost zero -> MM_SITE_HAS_LJ.  A charged type with a screening
 * radius is an error (the device evaluates unscreened charges). */
int mm_msim_site_params(const mm_msim_eleclj *p, int natom, const char *const *types, double *q,
                        double *sigma, double *epsilon, int *mask);

/* xyz coordinate and restart files (msys.cpp:823-868, 916-932).
 * mm_xyz_read with natom <= 0 returns the file's atom count (> 0).  Else
 * natom must match; per atom "symbol x y z [type [vx vy vz]]": symbols and
 * types (may be NULL) receive 16-character slots (type "*" when absent),
 * r[3*natom] and v[3*natom] (may be NULL) the vectors, *has_velocities
 * whether every line held them; bc (bc_len characters, may be NULL) the
 * boundary conditions of line 2 as "cubic L" / "orthorhombic a b c" /
 * ..., or "" when line 2 gives none.
 * mm_xyz_write writes the same format: "natom", "boundary_conditions bc
 * [  comment]", then "symbol  x y z type  vx vy vz" per atom (types may be
 * NULL or hold "" for "*", v NULL for zero velocities), numbers with
 * `precision` significant digits (0: 6, the reference's stream default). */
int mm_xyz_read(const char *fname, int natom, char *symbols, char *types, double *r, double *v,
                int *has_velocities, char *bc, int bc_len);
int mm_xyz_write(const char *fname, int natom, const char *symbols, const char *types,
                 const double *r, const double *v, const char *bc, const char *comment,
                 int precision);

typedef struct mm_engine mm_engine;

/* energy slots returned per system by mm_get_energies */
#define MM_E_ELEC 0      /* 1/2 sum q phi */
#define MM_E_LJ 1        /* pair LJ + tail correction */
#define MM_E_TAIL 2      /* LJ tail correction alone */
#define MM_E_TOTAL 3     /* ELEC + LJ */
#define MM_E_KINETIC 4   /* kinetic energy (after mm_md_run) */
#define MM_E_KSPACE 5    /* reciprocal-space energy of the k sum (diagnostic) */
#define MM_NE 8

/* compute flags */
#define MM_VIRIAL 1      /* also form the 3x3 virial sum r_ij (x) f_ij */

const char *mm_last_error(void);
int mm_device_count(void);

int mm_engine_create(int nsys, const mm_system_desc *sys, mm_engine **out);

/* How a batch is laid onto the device (no reference counterpart: the
 * reference has one CPU loop).  None of these changes what is computed:
 *   cell_groups   mean groups per neighbour-list cell; 2 is the reference's
 *                 cell count (nbrlist.cpp:11), for which the pair set equals
 *                 the reference's bit for bit; other sizes give the same set
 *                 except pairs whose FP64 distance rounds across rc + skin
 *                 (outside the cutoff: no force changes), and forces to
 *                 FP64 summation order;
 *   pair_slots    pair-kernel workgroups per CU (1..4; with fewer than 4,
 *                 the rest of each CU stays free for the reciprocal-space
 *                 kernels beside it); forces bit-identical;
 *   spread_block  P3M spreading block side in mesh columns (the smallest
 *                 factor of the mesh side >= it); mesh bit-identical;
 *   graph_steps   MD steps captured per graph (1..64); trajectories
 *                 bit-identical;
 *   spread_halo   P3M spreading by owner (1: each block spreads the sites
 *                 of its own columns whole into a halo that the z pass
 *                 sums; pays where columns hold few sites) or by pulling
 *                 (0: each block gathers the sites whose stencils reach its
 *                 columns); mesh bit-identical.
 * mm_shape_default: 2, 3, 2, 10, 0 (MOLMODEL_AMD_CELL_GROUPS,
 * MOLMODEL_AMD_PAIR_SLOTS, MOLMODEL_AMD_SPREAD_B, MOLMODEL_AMD_GRAPH_STEPS
 * and MOLMODEL_AMD_SPREAD_HALO override them); mm_engine_create uses it. */
typedef struct {
  int cell_groups, pair_slots, spread_block, graph_steps, spread_halo;
} mm_shape;
void mm_shape_default(mm_shape *s);
int mm_engine_create_shaped(int nsys, const mm_system_desc *sys, const mm_shape *shape,
                            mm_engine **out);

/* The autotuner: builds a fresh engine per candidate shape from the
 * caller's batch (positions r, velocities v or NULL), runs MD with
 * mm_md_setup(dt, bond, angle_deg, tol), and times `steps`-step blocks
 * (graph replay; the best of two after a warm-up); one choice at a time
 * from the default shape: pair_slots {2, 3, 4}, spread_halo {0, 1} and
 * spread_block {2, 4, 8} (when a system uses P3M), cell_groups {1, 2, 4},
 * graph_steps {5, 10, 20}.  *best:
 * the fastest; trials[0..max_trials) the shapes timed and their ms per
 * step (-1 for a shape that could not run), *ntrials their number.  The
 * caller's own engines are not touched. */
typedef struct {
  mm_shape shape;
  double ms_per_step;
} mm_shape_trial;
int mm_tune_shape(int nsys, const mm_system_desc *sys, const double *r, const double *v,
                  double dt, double bond, double angle_deg, double tol, int steps,
                  mm_shape *best, mm_shape_trial *trials, int max_trials, int *ntrials);
int mm_engine_destroy(mm_engine *e);
/* totals over all systems of the batch */
int mm_engine_sizes(const mm_engine *e, int *nsite, int *ngroup);

/* positions / velocities, AoS double[3*nsite_total] (systems concatenated) */
int mm_set_positions(mm_engine *e, const double *r);
int mm_get_positions(mm_engine *e, double *r);
int mm_set_velocities(mm_engine *e, const double *v);
int mm_get_velocities(mm_engine *e, double *v);

/* one full force evaluation at the current positions:
 * ElecLJ::add_to_energy_and_forces (eleclj.cpp:785-807) */
int mm_compute(mm_engine *e, int flags);
int mm_get_forces(mm_engine *e, double *f);               /* AoS 3*nsite_total */
int mm_get_potential(mm_engine *e, double *phi);          /* nsite_total */
int mm_get_energies(mm_engine *e, double *en);            /* MM_NE per system */
int mm_get_virial(mm_engine *e, double *w);               /* 9 per system, row-major */
int mm_errors(mm_engine *e, int *flags);                  /* reads and clears */
/* raises MM_FLAG_NL_OVERFLOW and/or MM_FLAG_BIN_OVERFLOW on the device as a
 * kernel would, so that callers can exercise the grow-and-run-again path
 * (no reference counterpart; other flags are refused) */
int mm_raise_flags(mm_engine *e, int flags);

/* RSpaceEwald's outputs alone (rspace.c:382-558): phi, evec, lj_force (AoS)
 * and lj_energy per system, including the excluded-pair correction loop;
 * accumulated (+=) into the caller's arrays like the reference. */
int mm_rspace_field(mm_engine *e, double *phi, double *evec, double *lj_force,
                    double *lj_energy);
/* reciprocal-space phi/evec alone (kspace_field or p3m_field), += */
int mm_kspace_field(mm_engine *e, double *phi, double *evec);

/* neighbor list of system `sys` as built for the current positions:
 * canonical (row, col) group pairs (pbin.c:6) in that system's group
 * numbering; returns the pair count (writes at most maxp pairs). */
long mm_nl_pairs(mm_engine *e, int sys, int *pairs, long maxp);
long long mm_nl_builds(mm_engine *e);   /* neighbour-list builds so far, all systems */

/* MD with velocity Verlet and rigid-water constraints (mdyn.cpp:120-160):
 * setup constrains the initial positions (constraint tolerance tol) and
 * computes the initial forces; run advances nsteps of dt ps.  use_graph
 * replays each step as one captured hipGraph. */
int mm_md_setup(mm_engine *e, double dt, double bond, double angle_deg, double tol);
int mm_md_run(mm_engine *e, int nsteps, int use_graph);

/* Potential::scale(s) for one system as ElecLJ does it (eleclj.cpp:1268-1280,
 * called by the volume move, mdyn.cpp:535, 558): box length times s, eta
 * divided by s, P3M influence function divided by s (p3m_scale,
 * pppm.c:643-664), wave vectors, erfc table, tail correction and
 * neighbour-list cells for the new box.  Positions are not touched: the
 * caller scales molecule centres and sets them (MSys::scale_centroids). */
int mm_scale_system(mm_engine *e, int sys, double s);

/* NeighborList (nbrlist.hpp:8-40) on the device for n points in a cubic box
 * of side boxl: NeighborList(n, cutoff, width) + set_boundary_conditions
 * (nbrlist.cpp:9-17, 84-123).  update = coordinates_changed(r)
 * (nbrlist.cpp:51-77): r is AoS 3n; returns 1 if the list was rebuilt, 0 if
 * the skin test kept it.  rows: the pair set in PairBinary layout (each pair
 * in the row of its pbin.c:6 owner), row_start[n+1] offsets into cols, at
 * most cap columns written; returns the pair count.  The same set as the
 * reference's for the same input; order within a row is the device's. */
typedef struct mm_nlist mm_nlist;
int mm_nlist_create(int n, double cutoff, double width, double boxl, mm_nlist **out);
int mm_nlist_update(mm_nlist *nl, const double *r);
long mm_nlist_rows(mm_nlist *nl, int *row_start, int *cols, long cap);
int mm_nlist_destroy(mm_nlist *nl);

/* Replica exchange among the engine's systems (mdyn.cpp:317-351, 476-497).
 * ladder: system indices by ascending temperature, ladders separated by -1
 * (systems on one ladder must be copies of one system: same sites, groups,
 * box and k-space); kT: nsys values in kcal/mol.  Requires mm_md_setup.
 * run: nintervals x (steps MD steps (0: none), then one exchange attempt
 * per neighbouring pair from the hottest down with positions swapped,
 * then, if reinit_velocities, new Maxwell-Boltzmann velocities at each
 * system's kT, then forces recomputed).  stats: tries and accepts per system (as the lower
 * member of a pair), 2 per system; config: which starting configuration
 * each system now holds. */
int mm_remd_setup(mm_engine *e, int nladder, const int *ladder, const double *kT,
                  unsigned long long seed);
int mm_remd_run(mm_engine *e, int nintervals, int steps, int reinit_velocities, int use_graph);
/* The reference's schedule (MDyn::run, mdyn.cpp:142-186, do_every at
 * mdyn.cpp:49-52): nsteps MD steps; after each, with probability
 * 1/reinit_interval new velocities (the constant-temperature move,
 * reinit_velocities_interval) and, independently, with probability
 * 1/exchange_interval an exchange attempt (replica_exchange_interval);
 * an interval of 0 disables its move.  One schedule for the whole batch,
 * drawn from counter-based random numbers of the setup's seed and the step
 * count since mm_remd_setup; the redraw follows the exchange when both
 * fall on one step.  nexchange, nreinit (may be NULL): the moves made. */
int mm_remd_run_stochastic(mm_engine *e, int nsteps, int exchange_interval, int reinit_interval,
                           int use_graph, long long *nexchange, long long *nreinit);
int mm_remd_stats(mm_engine *e, long long *tries_accepts, int *config);

/* Reciprocal-space parameters for one cubic system chosen by this library's
 * tuner (no reference counterpart: the reference takes them from the user
 * or optimises eta for a fixed mesh, pppm.c:667-716).  eta from the
 * real-space error estimate at tol/sqrt(2) (eleclj.cpp:134-138); then the
 * cheapest of Ewald (smallest kc within tol/sqrt(2), eleclj.cpp:128-132)
 * and P3M (per order, the smallest 2357-smooth cubic mesh whose chi,
 * pppm.c:313-393, is within tol/sqrt(2)) under a device cost model.
 * Host-only: needs no GPU. */
typedef struct {
  int kmode;            /* 0 Ewald, 1 P3M */
  double eta, kc;
  int mesh[3], order;
  double err_rspace, err_kspace;   /* the estimates at the choice */
  double cost;                     /* modelled device time, microseconds */
} mm_kspace_choice;
int mm_tune_kspace(double boxl, int nsite, double rc, double tol, int allow_ewald,
                   mm_kspace_choice *out);

/* Gradients of system `sys`'s potential energy and of the force-matching
 * loss L = sum_i |F_i - f_ref_i|^2 with respect to per-type parameters
 * (no reference counterpart; the nearest is ElecLJ::set_lambda's
 * parameter interpolation, eleclj.cpp:273-319).  Site types are the
 * distinct (q, sigma, epsilon, mask) of the system's sites, numbered in
 * order of first appearance among the sites; the per-type parameters are
 * (q [internal charge units], sigma, epsilon).  One forward evaluation at
 * the current positions, then the adjoint kernels: one pass over the
 * neighbour rows that recomputes every pair term and accumulates, per site,
 * R . dF/dq (R = 2 (F - f_ref)) and, per LJ type pair, the derivatives of E
 * and L with respect to a = 4 e s^12 and b = 4 e s^6; the reciprocal-space
 * adjoint (Ewald adjoint structure factors, or one P3M spread of the
 * adjoint source with the forward's convolution and gather, or for P3M
 * with ik differentiation three forward ik passes); sigma and
 * epsilon through the mixing rule on the host.  Pair arithmetic FP32,
 * accumulation FP64.  Outputs (any may be NULL except ntype): ntype;
 * site_type[nsite]; type_params[3*ntype]; dE, dL [3*ntype] ordered
 * (dq, dsigma, depsilon) per type; loss; ms = device time of the adjoint
 * passes.  f_ref may be NULL (energy gradients only).  Leaves the engine's
 * parameters and forces as they were (the next evaluation recomputes the
 * reciprocal-space buffers the adjoint reused). */
int mm_param_gradients(mm_engine *e, int sys, const double *f_ref, int maxtypes, int *ntype,
                       int *site_type, double *type_params, double *dE, double *dL,
                       double *loss, double *ms);
/* The same gradients from a basis of forward evaluations (forces are
 * bilinear in the type charges and linear in each LJ type pair's a, b):
 * indicator charges and unit LJ entries, reassembled on the host.  An
 * independent check of mm_param_gradients; basis_err = relative RMS
 * difference between the reassembled and the evaluated forces. */
int mm_param_gradients_basis(mm_engine *e, int sys, const double *f_ref, int maxtypes,
                             int *ntype, int *site_type, double *type_params, double *dE,
                             double *dL, double *loss, double *basis_err);

/* New per-site charges (internal units), sigma and epsilon for system sys
 * (site masks, groups and exclusions stay): what ElecLJ::set_lambda leaves
 * behind for new site parameters (eleclj.cpp:273-308) -- LJ types and
 * tables, the tail correction over the "%12.6g" keys and the neutralising
 * background follow.  At most 8 distinct (sigma, epsilon) per system. */
int mm_set_site_params(mm_engine *e, int sys, const double *q, const double *sigma,
                       const double *epsilon);
/* Free-energy perturbation (Potential::types_changed(pert) + set_lambda,
 * eleclj.cpp:267-308, msimprog.cpp:80-114): set_perturbation records the
 * system's current parameters as state A and the given ones as state B;
 * set_lambda(l) then sets q = (1 - l) qA + l qB for every site and sigma,
 * epsilon likewise for sites with LJ.  Systems of a batch hold their own
 * lambda, so FEP windows run as one launch. */
int mm_set_perturbation(mm_engine *e, int sys, const double *qB, const double *sigmaB,
                        const double *epsilonB);
int mm_set_lambda(mm_engine *e, int sys, double lambda);

/* interaction counts of the current neighbour list, 3 per system:
 * group pairs with first-site r^2 <= rc^2 (unordered), charged site pairs
 * and LJ site pairs among them */
int mm_count_pairs(mm_engine *e, long long *counts);

/* average device time of one launch (HIP events on the engine stream) of
 * the real-space pair kernel (MM_KERNEL_PAIR) or of the whole P3M sequence
 * (MM_KERNEL_P3M), over reps back-to-back launches at the current state */
#define MM_KERNEL_PAIR 0
#define MM_KERNEL_P3M 1
int mm_time_kernel(mm_engine *e, int which, int reps, double *ms_per_launch);

/* device-to-device copy of `bytes` between two fresh buffers (a reference
 * kernel of known traffic for calibrating hardware byte counters) */
int mm_copy_probe(long bytes, double *ms);

/* stream the engine runs on (hipStream_t) and device pointers of the FP64
 * SoA state (x[0..N), y[N..2N), z[2N..3N)) for device-resident callers */
void *mm_stream(mm_engine *e);
double *mm_device_positions(mm_engine *e);
double *mm_device_forces(mm_engine *e);
int mm_sync(mm_engine *e);

#ifdef __cplusplus
}
#endif
#endif
