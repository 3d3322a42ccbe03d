// shape.cpp -- measured choice of the engine's cell and tile shapes.
//
// An mm_shape fixes how a batch is laid onto the device: the neighbour-list
// cell size (mean groups per cell), the pair kernel's persistent grid
// (workgroups per CU, which also sets how many slots the reciprocal-space
// kernels keep), the P3M spreading column blocks (mesh columns per block
// side) and their way of spreading (by owner into halos, or by pulling the
// sites whose stencils reach a block), and the MD steps captured per graph.  None of them changes what is
// computed: pair-kernel tiles and spreading blocks give bit-identical
// forces, graph blocks bit-identical trajectories, and other cell sizes
// give the same pair set up to pairs whose FP64 distance lies within
// rounding of rc + skin (outside the cutoff, so no force changes) and the
// same forces to FP64 accumulation order.  mm_tune_shape times short MD
// blocks of the caller's own batch on the device for each candidate and
// keeps the fastest, one choice at a time (coordinate descent from the
// default shape).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "molmodel_amd.h"

namespace mmamd {
void set_last_error(const std::string &s);
}

namespace {

// seconds per MD step of a fresh engine with this shape: warm-up (which
// also captures the step graphs), then the best of two timed blocks;
// a negative return is the error of a shape that cannot be built or run
int time_shape(int nsys, const mm_system_desc *sys, const double *r, const double *v,
               double dt, double bond, double angle_deg, double tol, int steps,
               const mm_shape &sh, double &sec) {
  mm_engine *e = nullptr;
  int rc = mm_engine_create_shaped(nsys, sys, &sh, &e);
  if (rc) return rc;
  auto run = [&]() -> int {
    int q = mm_set_positions(e, r);
    if (!q && v) q = mm_set_velocities(e, v);
    if (!q) q = mm_md_setup(e, dt, bond, angle_deg, tol);
    if (!q) q = mm_md_run(e, std::max(2 * sh.graph_steps + 1, steps / 4), 1);
    if (q) return q;
    sec = INFINITY;
    for (int k = 0; k < 2; k++) {
      const auto t0 = std::chrono::steady_clock::now();
      q = mm_md_run(e, steps, 1);   // ends with the device's error flags read back
      const auto t1 = std::chrono::steady_clock::now();
      if (q) return q;
      sec = std::min(sec, std::chrono::duration<double>(t1 - t0).count() / steps);
    }
    return 0;
  };
  rc = run();
  mm_engine_destroy(e);
  return rc;
}

}  // namespace

extern "C" int mm_tune_shape(int nsys, const mm_system_desc *sys, const double *r,
                             const double *v, double dt, double bond, double angle_deg,
                             double tol, int steps, mm_shape *best, mm_shape_trial *trials,
                             int max_trials, int *ntrials) {
  if (nsys <= 0 || !sys || !r || !best || steps < 10 || (max_trials > 0 && !trials)) {
    mmamd::set_last_error("mm_tune_shape: bad arguments (steps >= 10)");
    return MM_ERR_ARG;
  }
  bool any_p3m = false;
  for (int s = 0; s < nsys; s++) any_p3m |= sys[s].kmode == 1;
  mm_shape cur;
  mm_shape_default(&cur);
  auto key = [](const mm_shape &s) {
    return std::make_tuple(s.cell_groups, s.pair_slots, s.spread_block, s.graph_steps,
                           s.spread_halo);
  };
  std::map<decltype(key(cur)), double> done;
  int nt = 0;
  int last_rc = 0;
  auto measure = [&](const mm_shape &s) -> double {
    auto it = done.find(key(s));
    if (it != done.end()) return it->second;
    double sec = INFINITY;
    if (last_rc == MM_ERR_HIP) return INFINITY;   // a HIP runtime error ends the tuning
    const int rc = time_shape(nsys, sys, r, v, dt, bond, angle_deg, tol, steps, s, sec);
    if (rc) { last_rc = rc; sec = INFINITY; }
    done[key(s)] = sec;
    if (nt < max_trials) {
      trials[nt].shape = s;
      trials[nt].ms_per_step = std::isfinite(sec) ? 1e3 * sec : -1.0;
    }
    nt++;
    return sec;
  };
  double tbest = measure(cur);
  // one choice at a time, in the order of their measured effect
  enum { SLOTS, SPREAD, HALO, CELLS, GRAPH };
  const std::vector<std::pair<int, std::vector<int>>> knobs = {
      {SLOTS, {2, 3, 4}}, {HALO, {0, 1}}, {SPREAD, {2, 4, 8}}, {CELLS, {1, 2, 4}},
      {GRAPH, {5, 10, 20}}};
  for (const auto &kv : knobs) {
    if ((kv.first == SPREAD || kv.first == HALO) && !any_p3m) continue;
    mm_shape keep = cur;
    for (int val : kv.second) {
      mm_shape s = cur;
      if (kv.first == SLOTS) s.pair_slots = val;
      if (kv.first == SPREAD) s.spread_block = val;
      if (kv.first == HALO) s.spread_halo = val;
      if (kv.first == CELLS) s.cell_groups = val;
      if (kv.first == GRAPH) s.graph_steps = val;
      const double t = measure(s);
      if (t < tbest) { tbest = t; keep = s; }
    }
    cur = keep;
  }
  if (ntrials) *ntrials = nt;
  if (last_rc == MM_ERR_HIP) return last_rc;
  if (!std::isfinite(tbest)) {
    if (!last_rc) mmamd::set_last_error("mm_tune_shape: no shape ran");
    return last_rc ? last_rc : MM_ERR_ARG;
  }
  *best = cur;
  return MM_OK;
}
