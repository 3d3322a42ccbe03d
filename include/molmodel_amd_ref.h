/*
 * molmodel_amd_ref.h -- drop-in replacements for molmodel's nonbonded C ABI.
 *
 * Same names, signatures, layouts and accumulate (+=) semantics as the
 * reference headers they replace, so a caller such as ElecLJ
 * (eleclj.cpp:912-959) links against libmolmodel_amd.so instead of the
 * reference objects.  The arithmetic runs on the GPU (FP32 pair and mesh
 * arithmetic with FP64 accumulation; FP64 k-space); PairBinary and the
 * Chebyshev table builder stay host-side, as in the reference.
 *
 * Conditions the reference asserts on abort here too, with a message;
 * inputs the device path does not cover (non-cubic real space, screening
 * radii, 1-4 scaling, exclusions between groups) also abort with a
 * message rather than being computed any other way.
 */
#ifndef MOLMODEL_AMD_REF_H
#define MOLMODEL_AMD_REF_H

#ifdef __cplusplus
extern "C" {
#endif

/* cart.h:8-11 */
typedef struct { double re, im; } complex_t;
typedef struct { double x, y, z; } cart_t;
typedef struct { int x, y, z; } icart_t;
typedef struct { double xx, yx, zx, xy, yy, zy, xz, yz, zz; } tensor_t;

/* bcondtype.h:8 */
enum bcondtype_t { non_periodic, cubic, orthorhombic, bcc, fcc, triclinic };

#define SITE_HAS_Q 1   /* rspace.h:7 */
#define SITE_HAS_LJ 2  /* rspace.h:8 */

/* ---- pbin.h:10-21 ---- */
struct PairBinary;
struct PairBinary *PairBinary_new(int n);
void PairBinary_delete(struct PairBinary *);
void PairBinary_add(struct PairBinary *, int i, int j);
void PairBinary_quick_add(struct PairBinary *, int i, int j);
void PairBinary_init(struct PairBinary *);
void PairBinary_clear(struct PairBinary *);
const int *PairBinary_number_of_elements(const struct PairBinary *);
int *const *PairBinary_elements(const struct PairBinary *);
const int *PairBinary_start(const struct PairBinary *, int i);
const int *PairBinary_end(const struct PairBinary *, int i);
int PairBinary_exists(const struct PairBinary *, int i, int j);
void PairBinary_show(const struct PairBinary *);

/* ---- cheby.h:16-37 ---- */
#define NCHEBY 512
struct Chebyshev { double min, max, inv_dx, coeff[NCHEBY][7]; };
void Chebyshev_init(struct Chebyshev *t, double min, double max, double (*f)(double));
void Chebyshev_test(const struct Chebyshev *t, double (*f)(double), double *maxerr, double *xerr);

/* ---- rspace.h:19-34 ---- */
struct Chebyshev *RSpaceChebyshev_new(double rhi, double eta, double *maxerr, double *xerr);
/* The neighbor rows given are used as they are (no rebuild); the table
 * argument is accepted for signature compatibility, the device builds its
 * own FP32 table from sqrt(rhi2) and eta. */
void RSpaceEwald(int nsite, int ngroup, const int (*group)[2], const cart_t *r,
                 enum bcondtype_t bctype, const tensor_t *latv,
                 int *const *neighbors, const int *number_of_neighbors,
                 const struct PairBinary *group_excluded,
                 const struct PairBinary *excluded,
                 double rlo2, double rhi2,
                 const int *site_mask, const double *q,
                 double *phi, cart_t *evec,
                 double eta, const struct Chebyshev *rspace_chebyshev,
                 const double *rscreen,
                 const double *sig, const double *eps,
                 int geometric_combining_for_sigma,
                 double *lj_ener, cart_t *lj_force,
                 const struct PairBinary *one_four,
                 double elec_1_4_scale_factor,
                 double lj_1_4_scale_factor);
/* rspace.h:52-62: phi only, same pair sums; the excluded-pair correction
 * without the minimum image, as rspace.c:747-751 does. */
void RSpaceEwaldPhi(int nsite, int ngroup, const int (*group)[2], const cart_t *r,
                    enum bcondtype_t bctype, const tensor_t *latv,
                    int *const *neighbors, const int *number_of_neighbors,
                    const struct PairBinary *group_excluded,
                    const struct PairBinary *excluded,
                    double rlo2, double rhi2,
                    const int *site_mask, const double *q,
                    double *phi,
                    double eta, const struct Chebyshev *rspace_chebyshev,
                    const double *rscreen,
                    const struct PairBinary *one_four,
                    double elec_1_4_scale_factor);

/* rspace.h:36-48, 74-86: non-periodic real space (clusters): every site pair
 * of two different groups, Coulomb 1/r (screened inside rscreen_i +
 * rscreen_j) and LJ, no cutoff; FP64 on the device.  rscreen may be NULL
 * (no screening).  Exclusions between groups and 1-4 scaled pairs abort. */
void RSpaceNonPeriodic(int nsite, int ngroup, const int (*group)[2], const cart_t *r,
                       const struct PairBinary *group_excluded,
                       const struct PairBinary *excluded,
                       const int *site_mask, const double *q,
                       double *phi, cart_t *evec,
                       const double *rscreen,
                       const double *sig, const double *eps,
                       int geometric_combining_for_sigma,
                       double *lj_ener, cart_t *lj_force,
                       const struct PairBinary *one_four,
                       double elec_1_4_scale_factor,
                       double lj_1_4_scale_factor);
void RSpaceNonPeriodicPhi(int nsite, int ngroup, const int (*group)[2], const cart_t *r,
                          const struct PairBinary *group_excluded,
                          const struct PairBinary *excluded,
                          const int *site_mask, const double *q,
                          double *phi, const double *rscreen,
                          const struct PairBinary *one_four,
                          double elec_1_4_scale_factor);
void RSpaceNonPeriodicPhiAtGridpoints(int nsite, int ngrid, const cart_t *rsite,
                                      const cart_t *rgrid,
                                      const int *site_mask,
                                      const double *q,
                                      double *phi, const double *rscreen);

/* rspace.h:64-72: potential at ngrid points from every site with a mask
 * bit (the real-space part of ElecLJ::electrostatic_potential_at_gridpoints,
 * eleclj.cpp:1295-1341; the caller adds kspace_phi over points + sites). */
void RSpaceEwaldPhiAtGridpoints(int nsite, int ngrid, const cart_t *rsite,
                                const cart_t *rgrid, enum bcondtype_t bctype,
                                const tensor_t *latv,
                                double rlo2, double rhi2,
                                const int *site_mask,
                                const double *q,
                                double *phi,
                                double eta, const struct Chebyshev *rspace_chebyshev,
                                const double *rscreen);

/* ---- kspace.h:10-18 ---- */
struct kspace;
struct kspace *kspace_new(double cutoff, int nsite);
void kspace_delete(struct kspace *);
void kspace_show(const struct kspace *);
void kspace_field(struct kspace *, const cart_t *r, const double *q,
                  const tensor_t *latv, double volume, double eta, double *phi, cart_t *evec);
void kspace_phi(struct kspace *, const cart_t *r, const double *q,
                const tensor_t *latv, double volume, double eta, double *phi);
int kspace_number_of_wavevectors(const struct kspace *);

/* kspace_field without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same sums and accumulate (+=) semantics as kspace_field with
 * kspace_new(cutoff, nsite), but r (AoS), q, phi and evec (AoS, may be NULL)
 * are DEVICE pointers, latv (the reciprocal lattice, tensor_t layout) is a
 * host pointer, and the work is queued on `stream` (a hipStream_t; NULL is
 * the null stream) without waiting for it.  Wave vectors are summed in
 * lexicographic order.  All other device memory the kernels use is one
 * file-scope array of 3.9 MB per device, held until the process ends, so
 * at most 65536 sites and 4096 wave vectors (cutoff <= 12); calls that
 * overlap on one device share that array and race.  Returns 0; -1 with
 * nothing queued for sizes beyond those limits or other bad arguments;
 * -2 on a HIP error (text from mm_last_error()). */
int kspace_field_noalloc(double cutoff, int nsite, const cart_t *r, const double *q,
                         const tensor_t *latv, double volume, double eta, double *phi,
                         cart_t *evec, void *stream);

/* ---- pppm.h:8-64 ---- */
/* ik_differentiate 1: the field is formed as -i k phi on the mesh and
 * interpolated with the assignment weights (three more meshes and inverse
 * transforms, pppm.c:577-626); 0: analytical differentiation of the
 * weights (pppm.c:603-640). */
typedef struct p3m *p3m_t;
p3m_t p3m_new(int nsite, const double primitive_translation_vectors[9], double grid_spacing,
              int nalias, int ik_differentiate, int assignment_order, int ngrid[3]);
p3m_t p3m_copy(const p3m_t, int nsite);
void p3m_init(p3m_t, double eta, double *chi);
/* Minimises chi_rspace^2 + chi_p3m^2 over eta (pppm.c:667-716).  The
 * reference uses GSL's Nelder-Mead simplex, which is not available here;
 * this uses its own one-dimensional Nelder-Mead with the same start
 * (0.1/grid_spacing, step 0.1), tolerance 1e-4 and 100 iterations. */
void p3m_optimize(p3m_t, double rc, double *eta, double *chi_rspace, double *chi_p3m);
void p3m_field(p3m_t, const double *r, const double *q, double *phi, double *evec);
void p3m_scale(p3m_t, double s);
void p3m_delete(p3m_t);

#ifdef __cplusplus
}
#endif
#endif
