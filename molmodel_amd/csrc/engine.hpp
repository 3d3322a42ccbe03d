// engine.hpp -- host-side engine object: owns every device buffer of a batch
// of periodic cubic systems and sequences the kernels of one force
// evaluation / one MD step on a single HIP stream.
#pragma once
#include "mm_common.hpp"
#include <vector>
#include <string>

namespace mmamd {

#define MM_HIP(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      mmamd::set_last_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                            " at " __FILE__ ":" + std::to_string(__LINE__));   \
      return -2;                                                               \
    }                                                                          \
  } while (0)

void set_last_error(const std::string &s);
const char *last_error();

// accumulator slots per system (FP64)
enum : int {
  ACC_ELEC = 0,   // 1/2 q.phi
  ACC_LJ = 1,     // pair LJ (smoothed), without tail
  ACC_KIN = 2,    // kinetic energy
  ACC_VIR = 3,    // 9 virial components (rows x, y, z)
  ACC_KSPACE_E = 12, // reciprocal-space energy from the k-sum (diagnostic)
  ACC_TOP1 = 13,  // largest first-site displacement since the last NL build
  ACC_TOP2 = 14,  // second largest
  NACC = 16
};

struct SystemInput {       // mirrors mm_system_desc of include/molmodel_amd.h
  int nsite, ngroup;
  std::vector<int> gfirst, gns;
  std::vector<double> q, sig, eps, mass;
  std::vector<int> mask;
  std::vector<int> excl;   // pairs
  double boxl, rc, width, skin, eta;
  int kmode; double kc;
  int nx, ny, nz, order, nalias, ik;
  int include_lj_correction;
  int geometric_sigma;
  int has_latv = 0;          // P3M / k-space lattice other than the cubic box
  double latv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

struct Engine {
  int nsys = 0, nsite = 0, ngroup = 0;
  int maxg = 0;               // max groups of any system (grid x extent)
  int maxs = 0;
  std::vector<SystemInput> in;
  std::vector<SysParams> hsp;
  std::vector<double> lj_tail;   // per system (host FP64)
  hipStream_t stream = nullptr;
  // second stream for the reciprocal-space branch of a force evaluation
  // (forked from and joined back into `stream` with the two events)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

  SysParams *d_sp = nullptr;
  // topology, original order
  int *d_gfirst = nullptr, *d_gns = nullptr, *d_gsys = nullptr;
  int *d_gexcl = nullptr;      // per group: bit (k*MAXS+l) set if sites k,l excluded
  int *d_ginfo = nullptr;      // per group: packed per-site (mask | type<<2) bytes
  double *d_q = nullptr, *d_mass = nullptr;
  int *d_stype = nullptr, *d_smask = nullptr;
  // state, FP64 SoA (3 arrays of nsite each, contiguous: x..., y..., z...)
  double *d_r = nullptr, *d_v = nullptr, *d_f = nullptr, *d_r0 = nullptr;
  // neighbor list
  int ncell_total = 0;
  int *d_gcell = nullptr, *d_cell_count = nullptr, *d_cell_start = nullptr, *d_cell_fill = nullptr;
  // cells are stored in Hilbert-curve order: d_cell_rank maps a cell index
  // (ix*n + iy)*n + iz to its storage position, d_cell_unrank back (both per
  // system, relative to SysParams::cell0); groups sorted by stored cell are
  // then spatially compact in every contiguous range
  int *d_cell_rank = nullptr, *d_cell_unrank = nullptr;
  int *d_order = nullptr;     // sorted gs -> original group
  int *d_gslot_inv = nullptr; // per group: its sorted slot (inverse of d_order)
  int4 *d_gslot = nullptr;    // per sorted slot: group, first site, site count, exclusion
                              // mask (written with each neighbour-list build)
  double *d_last = nullptr;   // 3*ngroup first-site positions at the last build
  int4 *d_sten = nullptr;
  int nbcap = 0;              // neighbor slots per group
  int *d_nbr = nullptr, *d_nnb = nullptr;
  int *d_rebuild = nullptr;   // per system flag
  int *d_err = nullptr;
  long long *d_nbuilds = nullptr;
  // packed (sorted order)
  double4 *d_gpos = nullptr;
  float4 *d_spk = nullptr;    // MAXS per group: offset xyz + charge
  int2 *d_gmeta = nullptr;    // (nsites, packed masks/types)
  // pair outputs (sorted slots)
  double *d_pf = nullptr;     // 4 per slot: fx fy fz phi (force mode) or ex ey ez phi
  double *d_plj = nullptr;    // 3 per slot: LJ + smoothing force (field mode)
  double *d_gacc = nullptr;   // per sorted group: u_lj, 9 virial
  // tables
  float4 *d_tab = nullptr;
  float2 *d_ljt = nullptr;
  // k-space / P3M per-site results (original order): phi, ex, ey, ez
  double *d_kres = nullptr;
  // Ewald
  double *d_kvec = nullptr;   // 4 per k: kx ky kz a_k
  double *d_ksc = nullptr;    // 2 per k: S, C
  int nk_total = 0;
  // P3M
  long mesh_total = 0, cmesh_total = 0, green_total = 0;
  float *d_mesh = nullptr;
  float2 *d_cmesh = nullptr;
  float *d_green = nullptr;
  float2 *d_twid = nullptr;   // per P3M system (at SysParams::twid0): twiddles for nx, ny, nz (3*128)
  int4 *d_sgrid = nullptr;    // per site: jx jy jz bin
  float4 *d_sfrac = nullptr;  // per site: sx sy sz q
  int nbin_total = 0;
  long halo_total = 0;
  unsigned long long *d_halo = nullptr;   // spreading blocks' halo partials (SysParams::halo0)
  int *d_bin_fill = nullptr;   // sites per P3M column bin (cleared by the gather)
  int4 *d_bgrid = nullptr;    // bin-sorted copies of d_sgrid (w = site index) ...
  float4 *d_bfrac = nullptr;  // ... and of d_sfrac
  // reductions
  double *d_acc = nullptr;    // NACC per system
  double *d_part = nullptr;   // per-block partials (reciprocal-space and exchange kernels)
  double *d_fpart = nullptr;  // per-block partials of the finish kernel: the next
                              // step's neighbour-list check reads them while the
                              // reciprocal-space kernels write d_part
  int *d_done = nullptr;      // per-system block counters
  // pair-kernel work: tiles of 16 i-groups numbered over all systems
  int bcap = 0;               // slots per P3M column bin (bgrid/bfrac hold nbin_total * bcap)
  int any_water = 0;          // some system holds only three-site waters (SysParams::water)
  int ntile = 0;
  int pair_blocks = 0;        // persistent grid (a multiple of 8)
  // shape (mm_shape): mean groups per neighbour-list cell (2: the
  // reference's cell count), pair workgroups per CU, spreading column-block
  // side in mesh columns, MD steps per captured graph, P3M spreading by
  // owner into halos (1) or by pulling neighbouring bins' sites (0)
  int cell_groups = 2, pair_slots = 3, spread_block = 2, graph_steps = 10, spread_halo = 0;
  int *d_tile_start = nullptr;   // nsys + 1
  int nblk_finish = 0;         // finish blocks per system (max)
  int nblk_part = 0;           // partial-sum slots per system
  // MD
  double bond = 1.0, angle = 109.47, tol = 1e-8;
  int md_ready = 0;
  int first_eval = 1;
  int no_rspace = 0;         // reciprocal-space only (kspace_* / p3m_* drop-ins)
  int external_nl = 0;       // neighbour rows supplied by the caller (RSpaceEwald drop-in)
  int md_phase = 0;          // 0: r, v, F at the same step; 1: drifted half step ahead
  double *d_phi = nullptr;   // per-site total potential of the last evaluation
  double *d_scratch = nullptr; size_t scratch_n = 0;
  unsigned long long *d_nlb = nullptr;   // neighbour-list builds per system
  // the pair kernel's packed inputs are refreshed by k_finish's drift and by
  // a neighbour-list rebuild; anything else that changes positions or
  // charges sets this so the next evaluation packs explicitly
  int need_pack = 1;
  // replica exchange: ladders of system indices by ascending temperature
  // (separated by -1), kT per system, per-attempt permutation, which
  // starting configuration each system holds, tries/accepts per system
  int *d_ladder = nullptr; int nladder = 0;
  double *d_kT = nullptr, *d_tail = nullptr, *d_rtmp = nullptr, *d_mom = nullptr;
  int *d_perm = nullptr;      // 3 * nsys: perm, configuration id, scratch
  long long *d_remd_stats = nullptr;
  unsigned long long remd_seed = 0;
  long long remd_attempt = 0;
  long long remd_step = 0;   // MD steps run by the stochastic schedule
  // per system, the two parameter sets of a perturbation (ElecLJ's
  // types_changed(pert), eleclj.cpp:273-308): q, sigma, epsilon of state A
  // then of state B (6 nsite), empty when none is set
  std::vector<std::vector<double>> pert;
  // adjoint buffers (allocated by the first gradient call)
  double *d_adjR = nullptr, *d_chi = nullptr, *d_adjpart = nullptr, *d_adjksc = nullptr;
  double *d_adjkres = nullptr, *d_adjacc = nullptr;
  float4 *d_adjsrc = nullptr;
  double *d_adjq = nullptr;
  // state saved at the start of an MD block or exchange interval, restored
  // when an overflow makes the block run again (r, v: 6 N; exchange
  // permutation and statistics)
  double *d_snap = nullptr;
  int *d_snap_perm = nullptr;
  long long *d_snap_stats = nullptr;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // several consecutive steps captured as one graph (fewer graph launches)
  hipGraph_t graphK = nullptr;
  hipGraphExec_t gexecK = nullptr;
  int graphK_steps = 0;
  int graph_mode = -1;
};

// host copies of the concatenated per-system buffers engine_create uploads
struct HostLayout {
  std::vector<int> gfirst, gns, gsys, ginfo, gexcl;
  std::vector<double> q, mass;
  std::vector<float4> tab;
  std::vector<float2> ljt;
  std::vector<int4> sten;
  std::vector<double> kvec;
  std::vector<float> green;
  std::vector<float2> twid;
};
int engine_layout(Engine &e, const std::vector<SystemInput> &in, HostLayout &h);
int engine_check_layout(const Engine &e, const HostLayout &h);
int engine_create(Engine &e, const std::vector<SystemInput> &in);
// host helpers shared with the reference-named entry points
void tensor_inverse(double *a, const double *b);
double green_at(int jx, int jy, int jz, int nx, int ny, int nz, const double *rl,
                double inv_4eta2, double inv_v, int nalias, int order, int ik, double *chi_acc);
int multiple_of_2357_near(double x);
double p3m_influence(const double latv[9], const int ngrid[3], int nalias, int order, int ik,
                     double eta, std::vector<double> *green);
void p3m_optimize_eta(const double latv[9], const int ngrid[3], int nalias, int order, int ik,
                      double grid_spacing, double rc, double *eta, double *chi_rspace,
                      double *chi_p3m);
void cheb_fit(double a, double b, double (*f)(double), double *c4);
void engine_destroy(Engine &e);
void lj_tail_factors(const SystemInput &s, double *ca, double *cb);
int engine_scale_system(Engine &e, int s, double f);
// new site charges and LJ parameters for system s (ElecLJ::set_lambda's
// effect, eleclj.cpp:273-308): LJ types, tables, tail and background follow
int engine_set_params(Engine &e, int s, const double *q, const double *sig, const double *eps);
// drop the captured step graph (it holds buffer addresses that changed)
void engine_drop_graph(Engine &e);
// after ERR_BIN_OVERFLOW: grow the P3M column bins to the capacity the
// binning kernel recorded (at least twice the old one) and clear the record
int engine_grow_bins(Engine &e);

// launchers (each returns 0 or a negative status)
// from_partials: the rebuild test of each system also fires when the
// displacement partials of the last (FIN_LAZY) finish exceed the skin
int launch_nl_build(Engine &e, int force, int from_partials = 0);
int launch_pack(Engine &e);
int launch_pair(Engine &e, int field_mode, int want_virial, int energy = 1);
int pair_tiles(int ngroup);
// mesh lines per FFT workgroup (16 threads per line); the x pass writes one
// partial-sum slot per workgroup
constexpr int P3M_FFT_LINES = 16;

int launch_kspace(Engine &e, int want_virial, hipStream_t st);
// energies 0: the mesh energy (a diagnostic total) is not formed
int launch_p3m(Engine &e, int want_virial, hipStream_t st, int energies = 1);
inline int launch_kspace(Engine &e, int want_virial) { return launch_kspace(e, want_virial, e.stream); }
inline int launch_p3m(Engine &e, int want_virial) { return launch_p3m(e, want_virial, e.stream); }
int launch_finish(Engine &e, int mode);
// non-periodic real space (rspace.hip): mode 0 RSpaceNonPeriodic, 1 its
// phi-only variant; out 7 per site, upart one LJ energy per 256 sites
int launch_nonperiodic(int mode, const double4 *site, const int4 *info, const double2 *lj,
                       const double *rscreen, int n, int geometric, double *out, double *upart,
                       hipStream_t st);
int launch_nonperiodic_points(const double4 *site, const double *rscreen, int n,
                              const double *pts, int npts, double *phi, hipStream_t st);
// k-space structure factors and field of one system in buffers the caller
// provides (kspace.hip)
int launch_kspace_raw(const SysParams *sp, const double4 *kvec, int nk, int nsite,
                      const double *r, const double *q, double *ksc, double *kres, double *acc,
                      hipStream_t st);
// adjoint passes of one system (adjoint.hip, kspace.hip, pppm.hip)
int launch_kspace_field_only(Engine &e, int sys, const double *ksc, double *kres_out,
                             double *acc_out);
int launch_p3m_adjoint(Engine &e, int sys, const float4 *src, hipStream_t st);
// one forward ik P3M pass of system sys with per-site charges q (global
// site index): bin, spread, the ik convolution, gather into kres
int launch_p3m_ik_pass(Engine &e, int sys, const double *q, hipStream_t st);
int adj_part_slots();
int adj_blocks(const Engine &e, int sys);
int launch_adj_pair(Engine &e, int sys, const double *R, double *chi, double *part);
int launch_adj_ewald(Engine &e, int sys, const double *R, double *ksc_adj, double *kres_adj,
                     double *acc_scratch);
int launch_adj_p3m_src(Engine &e, int sys, const double *R, double scale, float4 *src);
int launch_adj_p3m_ik(Engine &e, int sys, const double *R, double scale, double *qsrc,
                      double *psi);
int launch_finish_ex(Engine &e, int mode, int want_virial, double *phi_out);
int launch_field_out(Engine &e, double *out, double *ulj, int min_image = 1);
int launch_skin(Engine &e);
int launch_count_pairs(Engine &e, unsigned long long *d_out);
int launch_remd_exchange(Engine &e, long long attempt);
int launch_phi_points(Engine &e, const double4 *sites, int nsite, const double *pts, int npts,
                      double *phi);
int launch_remd_velocities(Engine &e, long long attempt);
int launch_set_rebuild(Engine &e, int value = 1);
int launch_constrain_initial(Engine &e);
int launch_start_md(Engine &e);

enum : int {   // launch_finish modes
  FIN_FORCES = 0,        // forces + energies only
  FIN_KICK = 1,          // + v += F dt/2m, RATTLE velocities, kinetic energy
  FIN_DRIFT = 2,         // + v += F dt/2m, r += v dt, SHAKE positions
  FIN_LAZY = 4,          // per-block partial sums only: the per-system totals
                         // (energies, displacements) are not formed, the
                         // rebuild flags are cleared, and the next
                         // neighbour-list check reduces the displacement
                         // partials itself (launch_nl_build from_partials)
};
// finish kernel blocks and the layout of their partial sums without the
// virial: elec, lj, kin, largest and second largest displacement
constexpr int FIN_TB = 128;
constexpr int FIN_NPART = 5;
constexpr int FIN_TOP = 3;

}  // namespace mmamd
