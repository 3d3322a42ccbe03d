// capi.cpp -- the extern "C" engine API declared in include/molmodel_amd.h.
#include "../../include/molmodel_amd.h"
#include "handle.hpp"
#include <cfloat>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <vector>

using namespace mmamd;

#define CHECK_ENGINE(p)                       \
  do {                                        \
    if (!(p)) {                               \
      set_last_error("null engine");          \
      return MM_ERR_ARG;                      \
    }                                         \
  } while (0)

#define TRY(x)            \
  do {                    \
    int rc_ = (x);        \
    if (rc_) return rc_;  \
  } while (0)

extern "C" {

const char *mm_last_error(void) { return last_error(); }

int mm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// the engine's inputs from the C descriptors
static int desc_inputs(int nsys, const mm_system_desc *sys, std::vector<SystemInput> &in) {
  in.assign(nsys, SystemInput());
  for (int s = 0; s < nsys; s++) {
    const mm_system_desc &d = sys[s];
    SystemInput &S = in[s];
    if (d.nsite <= 0 || d.ngroup <= 0 || !d.group || !d.q || !d.sigma || !d.epsilon ||
        !d.site_mask) {
      set_last_error("mm_engine_create: missing arrays");
      return MM_ERR_ARG;
    }
    S.nsite = d.nsite;
    S.ngroup = d.ngroup;
    S.gfirst.resize(d.ngroup);
    S.gns.resize(d.ngroup);
    for (int g = 0; g < d.ngroup; g++) {
      S.gfirst[g] = d.group[2 * g];
      S.gns[g] = d.group[2 * g + 1] - d.group[2 * g];
    }
    S.q.assign(d.q, d.q + d.nsite);
    S.sig.assign(d.sigma, d.sigma + d.nsite);
    S.eps.assign(d.epsilon, d.epsilon + d.nsite);
    S.mask.assign(d.site_mask, d.site_mask + d.nsite);
    if (d.mass) S.mass.assign(d.mass, d.mass + d.nsite);
    if (d.nexcl > 0 && d.excl) S.excl.assign(d.excl, d.excl + 2 * d.nexcl);
    S.boxl = d.boxl;
    S.rc = d.rc;
    S.width = d.smoothing_width;
    S.skin = d.skin;
    S.eta = d.eta;
    S.kmode = d.kmode;
    S.kc = d.kc;
    S.nx = d.mesh[0];
    S.ny = d.mesh[1];
    S.nz = d.mesh[2];
    S.order = d.order;
    S.nalias = d.nalias;
    S.ik = d.ik;
    S.include_lj_correction = d.include_lj_correction;
    S.geometric_sigma = d.geometric_sigma;
  }
  return MM_OK;
}

// The default shape: two groups per cell (the reference's cell count),
// three pair workgroups per CU, 2 x 2 spreading column blocks, ten MD steps
// per graph; MOLMODEL_AMD_PAIR_SLOTS, MOLMODEL_AMD_SPREAD_B and
// MOLMODEL_AMD_GRAPH_STEPS override the last three.
void mm_shape_default(mm_shape *s) {
  if (!s) return;
  s->cell_groups = 2;
  s->pair_slots = 3;
  s->spread_block = 2;
  s->graph_steps = 10;
  s->spread_halo = 0;
  if (const char *v = getenv("MOLMODEL_AMD_CELL_GROUPS")) s->cell_groups = atoi(v);
  if (const char *v = getenv("MOLMODEL_AMD_PAIR_SLOTS")) s->pair_slots = atoi(v);
  if (const char *v = getenv("MOLMODEL_AMD_SPREAD_HALO")) s->spread_halo = atoi(v);
  if (const char *v = getenv("MOLMODEL_AMD_SPREAD_B")) s->spread_block = atoi(v);
  if (const char *v = getenv("MOLMODEL_AMD_GRAPH_STEPS")) s->graph_steps = atoi(v);
}

static int check_shape(const mm_shape &s) {
  if (s.cell_groups < 1 || s.cell_groups > 8 || s.pair_slots < 1 || s.pair_slots > 4 ||
      s.spread_block < 1 || s.spread_block > 8 || s.graph_steps < 1 || s.graph_steps > 64 ||
      s.spread_halo < 0 || s.spread_halo > 1) {
    set_last_error("mm_shape: cell_groups 1..8, pair_slots 1..4, spread_block 1..8, "
                   "graph_steps 1..64, spread_halo 0 or 1");
    return MM_ERR_ARG;
  }
  return MM_OK;
}

int mm_engine_create(int nsys, const mm_system_desc *sys, mm_engine **out) {
  mm_shape s;
  mm_shape_default(&s);
  return mm_engine_create_shaped(nsys, sys, &s, out);
}

int mm_engine_create_shaped(int nsys, const mm_system_desc *sys, const mm_shape *shape,
                            mm_engine **out) {
  if (!out || nsys <= 0 || !sys || !shape) {
    set_last_error("mm_engine_create: bad arguments");
    return MM_ERR_ARG;
  }
  *out = nullptr;
  TRY(check_shape(*shape));
  if (mm_device_count() <= 0) {
    set_last_error("mm_engine_create: no HIP device");
    return MM_ERR_HIP;
  }
  std::vector<SystemInput> in;
  TRY(desc_inputs(nsys, sys, in));
  mm_engine *m = new mm_engine();
  m->e.cell_groups = shape->cell_groups;
  m->e.pair_slots = shape->pair_slots;
  m->e.spread_block = shape->spread_block;
  m->e.graph_steps = shape->graph_steps;
  m->e.spread_halo = shape->spread_halo;
  int rc = engine_create(m->e, in);
  if (rc) {
    engine_destroy(m->e);
    delete m;
    return rc;
  }
  *out = m;
  return MM_OK;
}

int mm_check_layout(int nsys, const mm_system_desc *sys) {
  if (nsys <= 0 || !sys) {
    set_last_error("mm_check_layout: bad arguments");
    return MM_ERR_ARG;
  }
  std::vector<SystemInput> in;
  TRY(desc_inputs(nsys, sys, in));
  Engine e;
  mm_shape s;
  mm_shape_default(&s);
  TRY(check_shape(s));
  e.cell_groups = s.cell_groups;
  e.spread_block = s.spread_block;
  e.spread_halo = s.spread_halo;
  HostLayout h;
  if (engine_layout(e, in, h)) return MM_ERR_ARG;
  if (engine_check_layout(e, h)) return MM_ERR_ARG;
  return MM_OK;
}

// ElecLJ::ElecLJ (eleclj.cpp:13-37): the io parameters' defaults; the
// "unset" values are the ones ElecLJ::init replaces (rc 0, skin 0, eta -1)
void mm_desc_init(mm_system_desc *d) {
  if (!d) return;
  memset(d, 0, sizeof *d);
  d->rc = 0;
  d->smoothing_width = 1;
  d->skin = 0;
  d->kc = 5;
  d->eta = -1;
  d->kmode = 0;
  d->order = 3;
  d->nalias = 2;
  d->ik = 0;
  d->include_lj_correction = 1;
  d->geometric_sigma = 0;
}

// ElecLJ::init (eleclj.cpp:164-201, 229-230) for a cubic box
int mm_desc_resolve(mm_system_desc *d, double p3m_grid_spacing, double p3m_grid_density) {
  if (!d || !(d->boxl > 0)) {
    set_last_error("mm_desc_resolve: a positive box length is needed");
    return MM_ERR_ARG;
  }
  const double L = d->boxl, small = 64 * DBL_EPSILON;   // fns.c:22-29
  if (fabs(d->rc) < small) d->rc = 0.45 * L;
  if (d->rc > 0.49 * L) {
    set_last_error("mm_desc_resolve: rspace_cutoff must be <= 0.49 * minimum diameter of box "
                   "(eleclj.cpp:168-170)");
    return MM_ERR_ARG;
  }
  if (!(d->rc > 0)) {
    set_last_error("mm_desc_resolve: rspace_cutoff must be positive");
    return MM_ERR_ARG;
  }
  if (fabs(d->skin) < small) {
    d->skin = 0.2 * d->rc;
    if (d->rc + d->skin >= 0.5 * L) d->skin = 0.499 * L - d->rc;
  }
  if (d->smoothing_width >= d->rc) {
    set_last_error("mm_desc_resolve: smoothing_width must be less than the real-space cutoff "
                   "(eleclj.cpp:180-182)");
    return MM_ERR_ARG;
  }
  if (d->kmode == 1) {
    double h = p3m_grid_spacing;
    if (p3m_grid_density > 0) {
      if (fabs(h - 0.5) >= small) {
        set_last_error("mm_desc_resolve: only one of p3m_grid_spacing and p3m_grid_density "
                       "(eleclj.cpp:193-195)");
        return MM_ERR_ARG;
      }
      h = L / cbrt(p3m_grid_density * L * L * L);
    }
    if (!(h > 0)) {
      set_last_error("mm_desc_resolve: grid spacing must be positive (pppm.c:412)");
      return MM_ERR_ARG;
    }
    if (d->mesh[0] <= 0 || d->mesh[1] <= 0 || d->mesh[2] <= 0) {
      const int n = multiple_of_2357_near(floor(L / h + 0.5));   // pppm.c:424-431
      d->mesh[0] = d->mesh[1] = d->mesh[2] = n;
    }
    if (d->eta < 0) {
      const double latv[9] = {L, 0, 0, 0, L, 0, 0, 0, L};
      double cr = 0, ck = 0;
      p3m_optimize_eta(latv, d->mesh, d->nalias, d->order, d->ik, h, d->rc, &d->eta, &cr, &ck);
    }
  } else if (d->eta < 0) {
    d->eta = sqrt(M_PI * d->kc / (d->rc * L));   // eleclj.cpp:229-230
  }
  return MM_OK;
}

int mm_engine_destroy(mm_engine *m) {
  if (!m) return MM_OK;
  hipStreamSynchronize(m->e.stream);
  engine_destroy(m->e);
  delete m;
  return MM_OK;
}

int mm_engine_sizes(const mm_engine *m, int *nsite, int *ngroup) {
  CHECK_ENGINE(m);
  if (nsite) *nsite = m->e.nsite;
  if (ngroup) *ngroup = m->e.ngroup;
  return MM_OK;
}

static int put_aos(mm_engine *m, double *d, const double *h) {
  Engine &e = m->e;
  const int N = e.nsite;
  std::vector<double> soa(3 * (size_t)N);
  for (int i = 0; i < N; i++)
    for (int c = 0; c < 3; c++) soa[(size_t)c * N + i] = h[3 * (size_t)i + c];
  MM_HIP(hipMemcpyAsync(d, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, e.stream));
  MM_HIP(hipStreamSynchronize(e.stream));
  return 0;
}

static int get_aos(mm_engine *m, const double *d, double *h) {
  Engine &e = m->e;
  const int N = e.nsite;
  std::vector<double> soa(3 * (size_t)N);
  MM_HIP(hipMemcpyAsync(soa.data(), d, soa.size() * sizeof(double), hipMemcpyDeviceToHost, e.stream));
  MM_HIP(hipStreamSynchronize(e.stream));
  for (int i = 0; i < N; i++)
    for (int c = 0; c < 3; c++) h[3 * (size_t)i + c] = soa[(size_t)c * N + i];
  return 0;
}

int mm_set_positions(mm_engine *m, const double *r) {
  CHECK_ENGINE(m);
  if (!r) { set_last_error("null positions"); return MM_ERR_ARG; }
  TRY(put_aos(m, m->e.d_r, r));
  m->pos_dirty = 1;
  m->e.md_phase = 0;
  return MM_OK;
}
int mm_get_positions(mm_engine *m, double *r) { CHECK_ENGINE(m); return get_aos(m, m->e.d_r, r); }
int mm_set_velocities(mm_engine *m, const double *v) {
  CHECK_ENGINE(m);
  if (!v) { set_last_error("null velocities"); return MM_ERR_ARG; }
  return put_aos(m, m->e.d_v, v);
}
int mm_get_velocities(mm_engine *m, double *v) { CHECK_ENGINE(m); return get_aos(m, m->e.d_v, v); }
int mm_get_forces(mm_engine *m, double *f) { CHECK_ENGINE(m); return get_aos(m, m->e.d_f, f); }

int mm_get_potential(mm_engine *m, double *phi) {
  CHECK_ENGINE(m);
  MM_HIP(hipMemcpyAsync(phi, m->e.d_phi, sizeof(double) * m->e.nsite, hipMemcpyDeviceToHost, m->e.stream));
  MM_HIP(hipStreamSynchronize(m->e.stream));
  return MM_OK;
}

static int read_err(Engine &e, int *flags) {
  int h = 0;
  MM_HIP(hipMemcpyAsync(&h, e.d_err, sizeof(int), hipMemcpyDeviceToHost, e.stream));
  MM_HIP(hipStreamSynchronize(e.stream));
  *flags = h;
  return 0;
}

static int clear_err(Engine &e) {
  int z = 0;
  MM_HIP(hipMemcpyAsync(e.d_err, &z, sizeof(int), hipMemcpyHostToDevice, e.stream));
  MM_HIP(hipStreamSynchronize(e.stream));
  return 0;
}

int mm_errors(mm_engine *m, int *flags) {
  CHECK_ENGINE(m);
  int f = 0;
  TRY(read_err(m->e, &f));
  TRY(clear_err(m->e));
  if (flags) *flags = f;
  return MM_OK;
}

int mm_raise_flags(mm_engine *m, int flags) {
  CHECK_ENGINE(m);
  if (flags & ~(MM_FLAG_NL_OVERFLOW | MM_FLAG_BIN_OVERFLOW)) {
    set_last_error("mm_raise_flags: only the overflow flags can be raised");
    return MM_ERR_ARG;
  }
  int f = 0;
  TRY(read_err(m->e, &f));
  f |= flags;
  MM_HIP(hipMemcpyAsync(m->e.d_err, &f, sizeof(int), hipMemcpyHostToDevice, m->e.stream));
  MM_HIP(hipStreamSynchronize(m->e.stream));
  return MM_OK;
}

// grow the neighbour rows after an overflow and force a rebuild
static int grow_nbr(Engine &e) {
  MM_HIP(hipStreamSynchronize(e.stream));
  MM_HIP(hipFree(e.d_nbr));
  e.nbcap = std::min(e.nbcap * 2, ((e.maxg + 15) / 16) * 16);
  MM_HIP(hipMalloc((void **)&e.d_nbr, sizeof(int) * (size_t)e.ngroup * e.nbcap));
  e.first_eval = 1;
  // a captured step holds the old row buffer: capture again on next use
  engine_drop_graph(e);
  return 0;
}

// Grow what the flags say overflowed (neighbour rows, P3M column bins);
// returns 1 if the evaluation should run again, 0 if nothing overflowed,
// or a negative status.
static int grow_overflowed(Engine &e, int f) {
  if (!(f & (ERR_NL_OVERFLOW | ERR_BIN_OVERFLOW))) return 0;
  if (f & ERR_NL_OVERFLOW) TRY(grow_nbr(e));
  if (f & ERR_BIN_OVERFLOW) TRY(engine_grow_bins(e));
  return 1;
}

// After a non-finite or collapsed position every system's neighbour rows
// may have been left empty (the pair search reads the one error flag all
// systems share), while the skin test's reference positions were updated:
// the next evaluation rebuilds every list from scratch.
static void after_position_error(Engine &e, int f) {
  if (f & (ERR_NAN | ERR_CELL)) e.first_eval = 1;
}

// r, v (and the exchange permutation and statistics) saved before a block
// of steps, so that a block in which a neighbour row or a P3M bin
// overflowed can run again with larger buffers
static int save_state(Engine &e, int remd) {
  const size_t n3 = 3 * (size_t)e.nsite;
  if (!e.d_snap) MM_HIP(hipMalloc((void **)&e.d_snap, 2 * n3 * sizeof(double)));
  MM_HIP(hipMemcpyAsync(e.d_snap, e.d_r, n3 * sizeof(double), hipMemcpyDeviceToDevice, e.stream));
  MM_HIP(hipMemcpyAsync(e.d_snap + n3, e.d_v, n3 * sizeof(double), hipMemcpyDeviceToDevice, e.stream));
  if (remd) {
    if (!e.d_snap_perm) MM_HIP(hipMalloc((void **)&e.d_snap_perm, 3 * sizeof(int) * e.nsys));
    if (!e.d_snap_stats) MM_HIP(hipMalloc((void **)&e.d_snap_stats, 2 * sizeof(long long) * e.nsys));
    MM_HIP(hipMemcpyAsync(e.d_snap_perm, e.d_perm, 3 * sizeof(int) * e.nsys,
                          hipMemcpyDeviceToDevice, e.stream));
    MM_HIP(hipMemcpyAsync(e.d_snap_stats, e.d_remd_stats, 2 * sizeof(long long) * e.nsys,
                          hipMemcpyDeviceToDevice, e.stream));
  }
  return 0;
}

static int restore_state(Engine &e, int remd) {
  const size_t n3 = 3 * (size_t)e.nsite;
  MM_HIP(hipMemcpyAsync(e.d_r, e.d_snap, n3 * sizeof(double), hipMemcpyDeviceToDevice, e.stream));
  MM_HIP(hipMemcpyAsync(e.d_v, e.d_snap + n3, n3 * sizeof(double), hipMemcpyDeviceToDevice, e.stream));
  if (remd) {
    MM_HIP(hipMemcpyAsync(e.d_perm, e.d_snap_perm, 3 * sizeof(int) * e.nsys,
                          hipMemcpyDeviceToDevice, e.stream));
    MM_HIP(hipMemcpyAsync(e.d_remd_stats, e.d_snap_stats, 2 * sizeof(long long) * e.nsys,
                          hipMemcpyDeviceToDevice, e.stream));
  }
  e.need_pack = 1;
  e.first_eval = 1;
  e.md_phase = 0;
  return 0;
}

// nl_from_partials: the last finish left per-block partial sums (FIN_LAZY);
// energy 0: an inner MD step, whose energies nobody reads: the pair kernel
// and the P3M x pass form forces only
static int enqueue_forces(mm_engine *m, int field_mode, int want_virial,
                          int nl_from_partials = 0, int energy = 1) {
  Engine &e = m->e;
  if (e.no_rspace) {
    TRY(launch_kspace(e, want_virial));
    TRY(launch_p3m(e, want_virial));
    return 0;
  }
  // The reciprocal-space branch reads only positions and charges and writes
  // its own buffers, so it runs on the second stream beside the
  // neighbour-list and pair kernels; the join comes before anything that
  // consumes it.
  const bool fork = e.stream2 != nullptr;
  hipStream_t kst = e.stream;
  if (fork) {
    MM_HIP(hipEventRecord(e.ev_fork, e.stream));
    MM_HIP(hipStreamWaitEvent(e.stream2, e.ev_fork, 0));
    kst = e.stream2;
  }
  TRY(launch_kspace(e, want_virial, kst));
  TRY(launch_p3m(e, want_virial, kst, energy || want_virial));
  if (e.external_nl) {
    e.first_eval = 0;
  } else if (e.first_eval) {
    TRY(launch_nl_build(e, 1));
    e.first_eval = 0;
  } else {
    if (m->pos_dirty) TRY(launch_skin(e));
    TRY(launch_nl_build(e, 0, nl_from_partials));
  }
  if (e.need_pack || m->pos_dirty) TRY(launch_pack(e));
  e.need_pack = 0;
  m->pos_dirty = 0;
  TRY(launch_pair(e, field_mode, want_virial, energy || field_mode || want_virial));
  if (fork) {
    MM_HIP(hipEventRecord(e.ev_join, e.stream2));
    MM_HIP(hipStreamWaitEvent(e.stream, e.ev_join, 0));
  }
  return 0;
}

static int compute_checked(mm_engine *m, int field_mode, int want_virial) {
  Engine &e = m->e;
  for (int attempt = 0; attempt < 16; attempt++) {
    TRY(enqueue_forces(m, field_mode, want_virial));
    if (!field_mode) TRY(launch_finish_ex(e, FIN_FORCES, want_virial, e.d_phi));
    int f = 0;
    TRY(read_err(e, &f));
    if (f & (ERR_NAN | ERR_CONSTRAINT | ERR_CELL)) {   // reported once, before any retry
      TRY(clear_err(e));
      after_position_error(e, f);
      set_last_error(f & ERR_NAN ? "non-finite (or out-of-range) position or energy during force evaluation"
                     : f & ERR_CELL ? "positions collapsed: a neighbour-list cell is overfull"
                                    : "device error flag raised during force evaluation");
      return MM_ERR_DEVICE;
    }
    if (f) TRY(clear_err(e));
    const int grown = grow_overflowed(e, f);
    if (grown < 0) return grown;
    if (grown) {
      m->pos_dirty = 1;   // pack again (a retry repeats the whole evaluation)
      continue;
    }
    return 0;
  }
  set_last_error("neighbour list or P3M bin overflow persisted");
  return MM_ERR_DEVICE;
}

}  // extern "C"

int mmamd::handle_compute(mm_engine *m) { return compute_checked(m, 0, 0); }

extern "C" {

int mm_compute(mm_engine *m, int flags) {
  CHECK_ENGINE(m);
  m->want_virial_last = (flags & MM_VIRIAL) ? 1 : 0;
  TRY(compute_checked(m, 0, m->want_virial_last));
  m->e.md_phase = 0;
  return MM_OK;
}

int mm_get_energies(mm_engine *m, double *en) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  std::vector<double> acc((size_t)e.nsys * NACC);
  MM_HIP(hipMemcpyAsync(acc.data(), e.d_acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost, e.stream));
  MM_HIP(hipStreamSynchronize(e.stream));
  for (int s = 0; s < e.nsys; s++) {
    const double *a = &acc[(size_t)s * NACC];
    double *o = en + (size_t)s * MM_NE;
    o[MM_E_ELEC] = a[ACC_ELEC];
    o[MM_E_LJ] = a[ACC_LJ] + e.lj_tail[s];
    o[MM_E_TAIL] = e.lj_tail[s];
    o[MM_E_TOTAL] = o[MM_E_ELEC] + o[MM_E_LJ];
    o[MM_E_KINETIC] = a[ACC_KIN];
    o[MM_E_KSPACE] = a[ACC_KSPACE_E];
    o[6] = a[ACC_TOP1];
    o[7] = a[ACC_TOP2];
  }
  return MM_OK;
}

int mm_get_virial(mm_engine *m, double *w) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  std::vector<double> acc((size_t)e.nsys * NACC), kv(9 * (size_t)e.nsys);
  MM_HIP(hipMemcpyAsync(acc.data(), e.d_acc, acc.size() * sizeof(double), hipMemcpyDeviceToHost, e.stream));
  MM_HIP(hipMemcpyAsync(kv.data(), e.d_kres + 4L * e.nsite, kv.size() * sizeof(double),
                        hipMemcpyDeviceToHost, e.stream));
  MM_HIP(hipStreamSynchronize(e.stream));
  for (int s = 0; s < e.nsys; s++)
    for (int c = 0; c < 9; c++)
      w[9 * s + c] = acc[(size_t)s * NACC + ACC_VIR + c] + kv[9 * s + c] +
                     ((c % 4 == 0) ? e.lj_tail[s] : 0.0);
  return MM_OK;
}

static int scratch(Engine &e, size_t n) {
  if (e.scratch_n >= n) return 0;
  if (e.d_scratch) MM_HIP(hipFree(e.d_scratch));
  MM_HIP(hipMalloc((void **)&e.d_scratch, n * sizeof(double)));
  e.scratch_n = n;
  return 0;
}

int mm_rspace_field(mm_engine *m, double *phi, double *evec, double *lj_force, double *lj_energy) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  TRY(compute_checked(m, 1, 0));
  const size_t N = e.nsite;
  TRY(scratch(e, 7 * N + e.nsys));
  MM_HIP(hipMemsetAsync(e.d_scratch + 7 * N, 0, sizeof(double) * e.nsys, e.stream));
  TRY(launch_field_out(e, e.d_scratch, e.d_scratch + 7 * N));
  std::vector<double> h(7 * N + e.nsys);
  MM_HIP(hipMemcpyAsync(h.data(), e.d_scratch, h.size() * sizeof(double), hipMemcpyDeviceToHost, e.stream));
  MM_HIP(hipStreamSynchronize(e.stream));
  for (size_t i = 0; i < N; i++) {
    if (phi) phi[i] += h[7 * i];
    for (int c = 0; c < 3; c++) {
      if (evec) evec[3 * i + c] += h[7 * i + 1 + c];
      if (lj_force) lj_force[3 * i + c] += h[7 * i + 4 + c];
    }
  }
  if (lj_energy)
    for (int s = 0; s < e.nsys; s++) lj_energy[s] += h[7 * N + s];
  e.md_phase = 0;
  return MM_OK;
}

int mm_kspace_field(mm_engine *m, double *phi, double *evec) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  TRY(compute_checked(m, 0, 0));
  const size_t N = e.nsite;
  std::vector<double> h(4 * N);
  MM_HIP(hipMemcpyAsync(h.data(), e.d_kres, h.size() * sizeof(double), hipMemcpyDeviceToHost, e.stream));
  MM_HIP(hipStreamSynchronize(e.stream));
  for (size_t i = 0; i < N; i++) {
    if (phi) phi[i] += h[4 * i];
    if (evec)
      for (int c = 0; c < 3; c++) evec[3 * i + c] += h[4 * i + 1 + c];
  }
  return MM_OK;
}

long mm_nl_pairs(mm_engine *m, int sys, int *pairs, long maxp) {
  if (!m) { set_last_error("null engine"); return MM_ERR_ARG; }
  Engine &e = m->e;
  if (sys < 0 || sys >= e.nsys) { set_last_error("bad system index"); return MM_ERR_ARG; }
  if (e.first_eval || m->pos_dirty) {
    int rc = compute_checked(m, 0, 0);
    if (rc) return rc;
  }
  const SysParams &P = e.hsp[sys];
  std::vector<int> order(e.ngroup), nnb(P.ngroup), nbr((size_t)P.ngroup * e.nbcap);
  if (hipMemcpy(order.data(), e.d_order, sizeof(int) * e.ngroup, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(nnb.data(), e.d_nnb + P.group0, sizeof(int) * P.ngroup, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(nbr.data(), e.d_nbr + (size_t)P.group0 * e.nbcap, sizeof(int) * nbr.size(),
                hipMemcpyDeviceToHost) != hipSuccess) {
    set_last_error("copy of neighbour list failed");
    return MM_ERR_HIP;
  }
  long np = 0;
  for (int i = 0; i < P.ngroup; i++) {
    const int g = order[P.group0 + i] - P.group0;
    for (int t = 0; t < nnb[i]; t++) {
      const int h = order[nbr[(size_t)i * e.nbcap + t]] - P.group0;
      if (g >= h) continue;   // the full list holds both directions
      int a = g, b = h;
      canonical_pair(a, b);
      if (np < maxp && pairs) { pairs[2 * np] = a; pairs[2 * np + 1] = b; }
      np++;
    }
  }
  return np;
}

// ------------------------------------------------ standalone neighbour list
// NeighborList (nbrlist.hpp:8-40) for n points in a cubic box: a
// single-system engine whose groups are the points themselves.
struct mm_nlist {
  mm_engine *m = nullptr;
  int n = 0;
  long long builds = 0;
};

int mm_nlist_create(int n, double cutoff, double width, double boxl, mm_nlist **out) {
  if (!out || n <= 0 || !(cutoff > 0) || !(width >= 0) || !(boxl > 0)) {
    set_last_error("mm_nlist_create: bad arguments");
    return MM_ERR_ARG;
  }
  *out = nullptr;
  std::vector<int> group(2 * (size_t)n), mask(n, 0);
  std::vector<double> zero(n, 0.0);
  for (int i = 0; i < n; i++) { group[2 * i] = i; group[2 * i + 1] = i + 1; }
  mm_system_desc d;
  memset(&d, 0, sizeof d);
  d.nsite = n; d.ngroup = n; d.group = group.data();
  d.q = zero.data(); d.sigma = zero.data(); d.epsilon = zero.data(); d.site_mask = mask.data();
  d.boxl = boxl; d.rc = cutoff; d.smoothing_width = 0; d.skin = width;
  d.eta = 3.0 / cutoff; d.kmode = 2;
  mm_engine *m = nullptr;
  TRY(mm_engine_create(1, &d, &m));
  mm_nlist *nl = new mm_nlist();
  nl->m = m;
  nl->n = n;
  *out = nl;
  return MM_OK;
}

int mm_nlist_destroy(mm_nlist *nl) {
  if (!nl) return MM_OK;
  mm_engine_destroy(nl->m);
  delete nl;
  return MM_OK;
}

// coordinates_changed (nbrlist.cpp:51-77): skin test against the positions
// of the last build, rebuild when it fires (always on the first call)
int mm_nlist_update(mm_nlist *nl, const double *r) {
  if (!nl || !r) { set_last_error("mm_nlist_update: bad arguments"); return MM_ERR_ARG; }
  mm_engine *m = nl->m;
  Engine &e = m->e;
  TRY(put_aos(m, e.d_r, r));
  for (int attempt = 0; attempt < 4; attempt++) {
    if (e.first_eval) {
      TRY(launch_nl_build(e, 1));
      e.first_eval = 0;
    } else {
      TRY(launch_set_rebuild(e, 0));   // the skin test alone decides
      TRY(launch_skin(e));
      TRY(launch_nl_build(e, 0));
    }
    m->pos_dirty = 0;
    int f = 0;
    TRY(read_err(e, &f));
    if (f & (ERR_NAN | ERR_CELL)) {
      TRY(clear_err(e));
      set_last_error(f & ERR_NAN ? "mm_nlist_update: non-finite (or out-of-range) position"
                                 : "mm_nlist_update: positions collapsed, a cell is overfull");
      return MM_ERR_DEVICE;
    }
    if (f & ERR_NL_OVERFLOW) {
      TRY(clear_err(e));
      TRY(grow_nbr(e));
      continue;
    }
    int rebuilt = 0;
    MM_HIP(hipMemcpy(&rebuilt, e.d_rebuild, sizeof(int), hipMemcpyDeviceToHost));
    nl->builds += rebuilt ? 1 : 0;
    return rebuilt ? 1 : 0;
  }
  set_last_error("neighbour list overflow persisted");
  return MM_ERR_DEVICE;
}

// the pair set as PairBinary rows (pbin.c:6 owner rule): row_start[n+1]
// offsets into cols; returns the pair count (cols written up to cap)
long mm_nlist_rows(mm_nlist *nl, int *row_start, int *cols, long cap) {
  if (!nl) { set_last_error("mm_nlist_rows: bad arguments"); return MM_ERR_ARG; }
  const long np = mm_nl_pairs(nl->m, 0, nullptr, 0);
  if (np < 0) return np;
  std::vector<int> pairs(2 * (size_t)np + 2);
  mm_nl_pairs(nl->m, 0, pairs.data(), np);
  std::vector<int> cnt(nl->n + 1, 0);
  for (long k = 0; k < np; k++) cnt[pairs[2 * k] + 1]++;
  for (int i = 0; i < nl->n; i++) cnt[i + 1] += cnt[i];
  if (row_start)
    for (int i = 0; i <= nl->n; i++) row_start[i] = cnt[i];
  if (cols) {
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (long k = 0; k < np; k++) {
      const long at = fill[pairs[2 * k]]++;
      if (at < cap) cols[at] = pairs[2 * k + 1];
    }
  }
  return np;
}

int mm_scale_system(mm_engine *m, int sys, double f) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  MM_HIP(hipStreamSynchronize(e.stream));
  TRY(engine_scale_system(e, sys, f));
  // the captured step holds the old stencil buffer
  engine_drop_graph(e);
  m->pos_dirty = 1;
  e.md_phase = 0;
  return MM_OK;
}

int mm_md_setup(mm_engine *m, double dt, double bond, double angle_deg, double tol) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  if (!(dt > 0) || !(bond > 0) || !(tol > 0)) { set_last_error("bad MD parameters"); return MM_ERR_ARG; }
  for (auto &s : e.in)
    if (s.mass.empty()) { set_last_error("masses are required for dynamics"); return MM_ERR_ARG; }
  e.bond = bond;
  e.angle = angle_deg;
  e.tol = tol;
  for (auto &P : e.hsp) P.dt = dt;
  MM_HIP(hipMemcpy(e.d_sp, e.hsp.data(), sizeof(SysParams) * e.nsys, hipMemcpyHostToDevice));
  engine_drop_graph(e);
  TRY(launch_constrain_initial(e));   // mdyn.cpp:120-121
  m->pos_dirty = 1;
  TRY(compute_checked(m, 0, 0));      // mdyn.cpp:122-124
  e.md_ready = 1;
  e.md_phase = 0;
  return MM_OK;
}

// one MD step after a finish; inner steps leave partial sums (FIN_LAZY),
// which the next step's neighbour-list check reduces itself
static int enqueue_step(mm_engine *m, int mode) {
  TRY(enqueue_forces(m, 0, 0, 1, (mode & FIN_LAZY) ? 0 : 1));
  TRY(launch_finish_ex(m->e, mode, 0, m->e.d_phi));
  return 0;
}


// nsteps of velocity Verlet, queued without waiting for the device
static int md_run_async(mm_engine *m, int nsteps, int use_graph) {
  Engine &e = m->e;
  if (e.md_phase == 0) {
    // first half of a step with the current forces (mdyn.cpp:144-150)
    TRY(launch_finish_ex(e, FIN_DRIFT, 0, e.d_phi));
    e.md_phase = 1;
  }
  int full = nsteps - 1;
  if (use_graph && full > 0) {
    // the step graph: neighbour-list check/rebuild, forces, finish; blocks
    // of K steps captured as one graph save the launch gap between graphs
    // steps per captured graph (the shape's graph_steps; ten gave 549 vs
    // 538 ns/day for one on the 98k-atom P3M box)
    const int K = e.graph_steps;
    auto capture = [&](int steps, hipGraph_t *graph, hipGraphExec_t *exec) -> int {
      MM_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
      int rc = 0;
      for (int s = 0; s < steps && !rc; s++) rc = enqueue_step(m, FIN_KICK | FIN_DRIFT | FIN_LAZY);
      hipGraph_t g = nullptr;
      hipError_t ce = hipStreamEndCapture(e.stream, &g);
      if (rc) { if (g) hipGraphDestroy(g); return rc; }
      MM_HIP(ce);
      *graph = g;
      MM_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
      return 0;
    };
    int left = full;
    if (K > 1 && full >= K) {
      if (!e.gexecK || e.graphK_steps != K) {
        if (e.gexecK) { hipGraphExecDestroy(e.gexecK); e.gexecK = nullptr; }
        if (e.graphK) { hipGraphDestroy(e.graphK); e.graphK = nullptr; }
        TRY(capture(K, &e.graphK, &e.gexecK));
        e.graphK_steps = K;
      }
      for (; left >= K; left -= K) MM_HIP(hipGraphLaunch(e.gexecK, e.stream));
    }
    if (left > 0 && !e.gexec) TRY(capture(1, &e.graph, &e.gexec));
    for (; left > 0; left--) MM_HIP(hipGraphLaunch(e.gexec, e.stream));
  } else {
    for (int s = 0; s < full; s++) TRY(enqueue_step(m, FIN_KICK | FIN_DRIFT | FIN_LAZY));
  }
  // last step ends synchronised: r, v, F of the same time
  TRY(enqueue_step(m, FIN_KICK));
  e.md_phase = 0;
  return 0;
}

static int md_report(Engine &e, int f) {
  if (!f) return MM_OK;
  TRY(clear_err(e));   // reported once
  after_position_error(e, f);
  set_last_error(f & ERR_NAN ? "non-finite (or out-of-range) position or energy during MD"
                 : f & ERR_CELL ? "positions collapsed during MD: a neighbour-list cell is overfull"
                 : f & ERR_BIN_OVERFLOW ? "a P3M mesh column bin overflowed during MD"
                 : f & ERR_CONSTRAINT ? "constraint iteration did not converge"
                                      : "neighbour list overflow during MD");
  return MM_ERR_DEVICE;
}

static int md_check_errors(Engine &e) {
  int f = 0;
  TRY(read_err(e, &f));
  return md_report(e, f);
}

// An overflow alone (neighbour row or P3M bin, with no position error)
// makes the block run again from its saved start with larger buffers.
static bool replayable(int f) {
  return (f & (ERR_NL_OVERFLOW | ERR_BIN_OVERFLOW)) && !(f & (ERR_NAN | ERR_CELL | ERR_CONSTRAINT));
}

// forces of the current positions before a block starts: positions set by
// the caller, a box scaling or a recovered error leave the last
// evaluation's pair and reciprocal-space results stale, and the first half
// kick (and any captured step) must not use them
static int forces_current(mm_engine *m) {
  Engine &e = m->e;
  if (e.first_eval || m->pos_dirty) {
    TRY(compute_checked(m, 0, 0));
    e.md_phase = 0;
  }
  return 0;
}

int mm_md_run(mm_engine *m, int nsteps, int use_graph) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  if (!e.md_ready) { set_last_error("mm_md_setup first"); return MM_ERR_ARG; }
  if (nsteps <= 0) return MM_OK;
  TRY(forces_current(m));
  for (int attempt = 0;; attempt++) {
    TRY(save_state(e, 0));
    TRY(md_run_async(m, nsteps, use_graph));
    int f = 0;
    TRY(read_err(e, &f));
    if (replayable(f) && attempt < 8) {
      TRY(clear_err(e));
      const int g = grow_overflowed(e, f);
      if (g < 0) return g;
      TRY(restore_state(e, 0));
      m->pos_dirty = 1;
      TRY(forces_current(m));
      continue;
    }
    return md_report(e, f);
  }
}

int mm_remd_setup(mm_engine *m, int nladder, const int *ladder, const double *kT,
                  unsigned long long seed) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  if (nladder <= 0 || !ladder || !kT) { set_last_error("bad replica-exchange arguments"); return MM_ERR_ARG; }
  std::vector<int> seen(e.nsys, 0);
  int prev = -1;
  for (int k = 0; k < nladder; k++) {
    const int s = ladder[k];
    if (s < -1 || s >= e.nsys) { set_last_error("ladder entry out of range"); return MM_ERR_ARG; }
    if (s == -1) { prev = -1; continue; }
    if (seen[s]++) { set_last_error("a system appears twice in the ladders"); return MM_ERR_ARG; }
    if (prev >= 0) {
      const SysParams &A = e.hsp[prev], &B = e.hsp[s];
      // replicas exchange coordinates, so they must be copies of one system
      if (A.nsite != B.nsite || A.ngroup != B.ngroup || A.L != B.L || A.kmode != B.kmode) {
        set_last_error("systems on one ladder must have the same sites, groups, box and k-space");
        return MM_ERR_ARG;
      }
    }
    prev = s;
  }
  for (int s = 0; s < e.nsys; s++)
    if (!(kT[s] > 0)) { set_last_error("kT must be positive for every system"); return MM_ERR_ARG; }
  if (!e.md_ready) { set_last_error("mm_md_setup first"); return MM_ERR_ARG; }
  auto realloc_d = [](auto **p, size_t n) -> int {
    if (*p) MM_HIP(hipFree(*p));
    MM_HIP(hipMalloc((void **)p, n * sizeof(**p)));
    MM_HIP(hipMemset(*p, 0, n * sizeof(**p)));
    return 0;
  };
  TRY(realloc_d(&e.d_ladder, nladder));
  TRY(realloc_d(&e.d_kT, e.nsys));
  TRY(realloc_d(&e.d_tail, e.nsys));
  TRY(realloc_d(&e.d_perm, 3 * (size_t)e.nsys));
  TRY(realloc_d(&e.d_remd_stats, 2 * (size_t)e.nsys));
  TRY(realloc_d(&e.d_rtmp, 3 * (size_t)e.nsite));
  TRY(realloc_d(&e.d_mom, 4 * (size_t)e.nsys));
  MM_HIP(hipMemcpy(e.d_ladder, ladder, sizeof(int) * nladder, hipMemcpyHostToDevice));
  MM_HIP(hipMemcpy(e.d_kT, kT, sizeof(double) * e.nsys, hipMemcpyHostToDevice));
  MM_HIP(hipMemcpy(e.d_tail, e.lj_tail.data(), sizeof(double) * e.nsys, hipMemcpyHostToDevice));
  std::vector<int> ids(3 * e.nsys);
  for (int s = 0; s < e.nsys; s++) ids[s] = ids[e.nsys + s] = ids[2 * e.nsys + s] = s;
  MM_HIP(hipMemcpy(e.d_perm, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice));
  e.nladder = nladder;
  e.remd_seed = seed;
  e.remd_attempt = 0;
  e.remd_step = 0;
  return MM_OK;
}

// steps MD steps, then the exchange attempt (if exchange) and the velocity
// redraw (if reinit), then forces of the configurations now held; a block
// whose neighbour rows or P3M bins overflowed runs again from its start with
// the same random numbers
static int remd_block(mm_engine *m, int steps, bool exchange, bool reinit, int use_graph) {
  Engine &e = m->e;
  for (int attempt = 0;; attempt++) {
    TRY(save_state(e, 1));
    if (steps > 0) TRY(md_run_async(m, steps, use_graph));
    const long long a = e.remd_attempt;
    // end of the block's last step: the exchange attempt (mdyn.cpp:185-186),
    // then the velocity redraw (mdyn.cpp:163-167; the reference redraws
    // before exchanging, which leaves swapped rigid molecules with velocities
    // constrained for the configuration they left)
    if (exchange) TRY(launch_remd_exchange(e, a));
    if (reinit) TRY(launch_remd_velocities(e, a));
    if (exchange) {
      e.need_pack = 1;   // positions moved between systems
      // forces and energies of the configurations now held; swapped systems
      // rebuild their neighbour lists first
      TRY(enqueue_forces(m, 0, 0));
      TRY(launch_finish_ex(e, FIN_FORCES, 0, e.d_phi));
    }
    int f = 0;
    TRY(read_err(e, &f));
    if (replayable(f) && attempt < 8) {
      TRY(clear_err(e));
      const int g = grow_overflowed(e, f);
      if (g < 0) return g;
      TRY(restore_state(e, 1));
      m->pos_dirty = 1;
      TRY(forces_current(m));
      continue;
    }
    if (f) return md_report(e, f);
    break;
  }
  if (exchange || reinit) e.remd_attempt++;
  return MM_OK;
}

int mm_remd_run(mm_engine *m, int nintervals, int steps, int reinit_velocities, int use_graph) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  if (!e.d_ladder) { set_last_error("mm_remd_setup first"); return MM_ERR_ARG; }
  if (nintervals < 0 || steps < 0) { set_last_error("bad interval arguments"); return MM_ERR_ARG; }
  TRY(forces_current(m));
  for (int it = 0; it < nintervals; it++) TRY(remd_block(m, steps, true, reinit_velocities != 0, use_graph));
  return MM_OK;
}

namespace {
// splitmix64 (Steele, Lea, Flood 2014) and a uniform double in [0, 1) of
// (seed, stream, index), the construction the device's exchange uses
unsigned long long splitmix_host(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01_host(unsigned long long seed, unsigned long long stream, unsigned long long idx) {
  const unsigned long long k = splitmix_host(seed * 0x100000001B3ull + stream);
  return (double)(splitmix_host(idx ^ k) >> 11) * (1.0 / 9007199254740992.0);
}
constexpr unsigned long long STREAM_REINIT = 0x5343484544000001ull;   // schedule draws
constexpr unsigned long long STREAM_EXCHANGE = 0x5343484544000002ull;
}  // namespace

// do_every (mdyn.cpp:49-52): after each step, a move happens if
// interval * u < 1 for a fresh uniform u; the velocity redraw
// (mdyn.cpp:163-167) and the exchange attempt (mdyn.cpp:185-186) draw
// independently.  The draws are counter-based in the engine's step count,
// one schedule for all systems of the batch.
int mm_remd_run_stochastic(mm_engine *m, int nsteps, int exchange_interval, int reinit_interval,
                           int use_graph, long long *nexchange, long long *nreinit) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  if (!e.d_ladder) { set_last_error("mm_remd_setup first"); return MM_ERR_ARG; }
  if (nsteps < 0 || exchange_interval < 0 || reinit_interval < 0) {
    set_last_error("bad schedule arguments");
    return MM_ERR_ARG;
  }
  TRY(forces_current(m));
  long long nx = 0, nr = 0;
  int done = 0;
  while (done < nsteps) {
    int s = done;
    bool ex = false, rv = false;
    while (s < nsteps && !ex && !rv) {
      const unsigned long long k = (unsigned long long)(e.remd_step + s);
      rv = reinit_interval > 0 && reinit_interval * uniform01_host(e.remd_seed, STREAM_REINIT, k) < 1;
      ex = exchange_interval > 0 &&
           exchange_interval * uniform01_host(e.remd_seed, STREAM_EXCHANGE, k) < 1;
      s++;
    }
    TRY(remd_block(m, s - done, ex, rv, use_graph));
    nx += ex;
    nr += rv;
    done = s;
  }
  e.remd_step += nsteps;
  if (nexchange) *nexchange = nx;
  if (nreinit) *nreinit = nr;
  return MM_OK;
}

int mm_remd_stats(mm_engine *m, long long *tries_accepts, int *config) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  if (!e.d_ladder) { set_last_error("mm_remd_setup first"); return MM_ERR_ARG; }
  MM_HIP(hipStreamSynchronize(e.stream));
  if (tries_accepts)
    MM_HIP(hipMemcpy(tries_accepts, e.d_remd_stats, sizeof(long long) * 2 * e.nsys,
                     hipMemcpyDeviceToHost));
  if (config)
    MM_HIP(hipMemcpy(config, e.d_perm + e.nsys, sizeof(int) * e.nsys, hipMemcpyDeviceToHost));
  return MM_OK;
}

void *mm_stream(mm_engine *m) { return m ? (void *)m->e.stream : nullptr; }
double *mm_device_positions(mm_engine *m) { return m ? m->e.d_r : nullptr; }
double *mm_device_forces(mm_engine *m) { return m ? m->e.d_f : nullptr; }
int mm_sync(mm_engine *m) {
  CHECK_ENGINE(m);
  MM_HIP(hipStreamSynchronize(m->e.stream));
  return MM_OK;
}
long long mm_nl_builds(mm_engine *m) {
  if (!m) return -1;
  Engine &e = m->e;
  if (hipStreamSynchronize(e.stream) != hipSuccess) return -1;
  std::vector<unsigned long long> b(e.nsys);
  if (hipMemcpy(b.data(), e.d_nlb, sizeof(unsigned long long) * e.nsys, hipMemcpyDeviceToHost) !=
      hipSuccess)
    return -1;
  long long t = 0;
  for (auto x : b) t += (long long)x;
  return t;
}

int mm_count_pairs(mm_engine *m, long long *counts) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  if (e.first_eval || m->pos_dirty) TRY(compute_checked(m, 0, 0));
  TRY(scratch(e, 3 * (size_t)e.nsys));
  unsigned long long *d = (unsigned long long *)e.d_scratch;
  MM_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long) * 3 * e.nsys, e.stream));
  TRY(launch_count_pairs(e, d));
  MM_HIP(hipMemcpyAsync(counts, d, sizeof(long long) * 3 * e.nsys, hipMemcpyDeviceToHost, e.stream));
  MM_HIP(hipStreamSynchronize(e.stream));
  return MM_OK;
}

int mm_time_kernel(mm_engine *m, int which, int reps, double *ms_per_launch) {
  CHECK_ENGINE(m);
  Engine &e = m->e;
  if (reps <= 0 || !ms_per_launch) { set_last_error("bad arguments"); return MM_ERR_ARG; }
  if (e.first_eval || m->pos_dirty) TRY(compute_checked(m, 0, 0));
  hipEvent_t a, b;
  MM_HIP(hipEventCreate(&a));
  MM_HIP(hipEventCreate(&b));
  TRY(which == MM_KERNEL_PAIR ? launch_pair(e, 0, 0) : launch_p3m(e, 0));   // warm
  MM_HIP(hipEventRecord(a, e.stream));
  for (int i = 0; i < reps; i++) TRY(which == MM_KERNEL_PAIR ? launch_pair(e, 0, 0) : launch_p3m(e, 0));
  MM_HIP(hipEventRecord(b, e.stream));
  MM_HIP(hipEventSynchronize(b));
  float ms = 0;
  MM_HIP(hipEventElapsedTime(&ms, a, b));
  hipEventDestroy(a);
  hipEventDestroy(b);
  *ms_per_launch = ms / reps;
  return MM_OK;
}

}  // extern "C"

extern "C" int mm_copy_probe(long bytes, double *ms) {
  void *a = nullptr, *b = nullptr;
  MM_HIP(hipMalloc(&a, bytes));
  MM_HIP(hipMalloc(&b, bytes));
  MM_HIP(hipMemset(a, 1, bytes));
  hipEvent_t e0, e1;
  MM_HIP(hipEventCreate(&e0));
  MM_HIP(hipEventCreate(&e1));
  MM_HIP(hipEventRecord(e0, nullptr));
  MM_HIP(hipMemcpy(b, a, bytes, hipMemcpyDeviceToDevice));
  MM_HIP(hipEventRecord(e1, nullptr));
  MM_HIP(hipEventSynchronize(e1));
  float t = 0;
  MM_HIP(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(a);
  hipFree(b);
  return MM_OK;
}
