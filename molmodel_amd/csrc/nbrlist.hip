// nbrlist.hip -- device cell neighbor list over groups.
//
// Reproduces the pair set of NeighborList::rebuild_periodic
// (nbrlist.cpp:222-251) bit for bit:
//   * cell of a group = cell_index(floor(x/lc + 0.5), ...) of its first site,
//     FP64, lc = L*(1.0/ncell) (nbrlist.cpp:97, 228-231; pbc.h:25-31);
//   * same-cell pairs are always listed (nbrlist.cpp:243-244);
//   * pairs in distinct neighbouring cells are listed iff the FP64
//     minimum-image distance^2 < (rc+skin)^2 (nbrlist.cpp:245-248,
//     bcond.hpp:56-60, pbc.h:8-14), evaluated in the order the reference
//     evaluates it: first argument in the cell that owns the cell pair row
//     (pbin.c:6), no FMA contraction.
// Layout: groups are counting-sorted by cell (deterministic order within a
// cell), and each sorted group gets a full list (both directions) of sorted
// neighbour indices in a fixed-capacity row, for the pair kernel.
#include "engine.hpp"
#include "scan.hpp"
#include "reduce.hpp"

#pragma clang fp contract(off)

namespace mmamd {

__device__ __forceinline__ double min_image_sq(double ax, double ay, double az, double bx,
                                               double by, double bz, double L, double invL) {
#pragma clang fp contract(off)
  double x = ax - bx, y = ay - by, z = az - bz;
  x -= L * floor(x * invL + 0.5);
  y -= L * floor(y * invL + 0.5);
  z -= L * floor(z * invL + 0.5);
  return x * x + y * y + z * z;
}

// from_part: the skin test (nbrlist.cpp:64-76) on the largest and second
// largest first-site displacement since the last build, merged from the
// per-block partials of the finish kernel that moved the positions (every
// block forms the same values); a rebuild already requested stays
// requested, and block 0 records the outcome for the rest of the chain.
__global__ void __launch_bounds__(256) k_nl_assign(
    const SysParams *__restrict__ sp, const int *__restrict__ gfirst,
    const double *__restrict__ r, int nsite_total, int *__restrict__ rebuild,
    int *__restrict__ gcell, double *__restrict__ last, int *__restrict__ cell_count,
    unsigned long long *__restrict__ builds, int *__restrict__ err,
    const double *__restrict__ part, int maxblk, int from_part, const int *__restrict__ cell_rank) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  int want = rebuild[s];
  if (from_part) {
    __shared__ double w1[4], w2[4];
    __shared__ int fire;
    const int nb = (P.ngroup + FIN_TB - 1) / FIN_TB;
    double t1 = 0, t2 = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
      const double *q = part + ((long)s * maxblk + b) * FIN_NPART + FIN_TOP;
      top2_merge(t1, t2, load_coherent(q), load_coherent(q + 1));
    }
    for (int o = 32; o >= 1; o >>= 1) {
      const double b1 = __shfl_xor(t1, o), b2 = __shfl_xor(t2, o);
      top2_merge(t1, t2, b1, b2);
    }
    if ((threadIdx.x & 63) == 0) { w1[threadIdx.x >> 6] = t1; w2[threadIdx.x >> 6] = t2; }
    __syncthreads();
    if (threadIdx.x == 0) {
      double u1 = w1[0], u2 = w2[0];
      for (int k = 1; k < (int)(blockDim.x >> 6); k++) top2_merge(u1, u2, w1[k], w2[k]);
      fire = u1 + u2 > P.skin;   // nbrlist.cpp:75-76
    }
    __syncthreads();
    want |= fire;
    if (blockIdx.x == 0 && threadIdx.x == 0 && fire) rebuild[s] = 1;
  }
  if (!want) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) builds[s] += 1;
  if (i >= P.ngroup) return;
  const int g = P.group0 + i;
  const int site = gfirst[g];
  const double x = r[site], y = r[nsite_total + site], z = r[2 * nsite_total + site];
  const int n = P.ncell;
  int c = P.cell0;
  // cell indices stay far inside int range; a position beyond (or NaN) is
  // an error
  const double lim = 1e9 * P.lc;
  if (fabs(x) <= lim && fabs(y) <= lim && fabs(z) <= lim) {
    const int ix = (int)floor(x / P.lc + 0.5);
    const int iy = (int)floor(y / P.lc + 0.5);
    const int iz = (int)floor(z / P.lc + 0.5);
    // stored at the cell's position along the Hilbert curve
    c += cell_rank[P.cell0 + (mod_nonneg(ix, n) * n + mod_nonneg(iy, n)) * n + mod_nonneg(iz, n)];
  } else {
    atomicOr(err, ERR_NAN);   // reported; the group is parked in cell 0
  }
  gcell[g] = c;
  last[3 * g] = x;
  last[3 * g + 1] = y;
  last[3 * g + 2] = z;
  atomicAdd(&cell_count[c], 1);
}

// one block per system: exclusive scan of that system's cells into sorted
// group offsets; counts and fill cursors are cleared for the next build
__global__ void k_nl_scan(const SysParams *__restrict__ sp, const int *__restrict__ rebuild,
                          int *__restrict__ cell_count, int *__restrict__ cell_start,
                          int *__restrict__ cell_fill) {
  const int s = blockIdx.y;
  if (!rebuild[s]) return;
  const SysParams &P = sp[s];
  block_scan_clear(cell_count + P.cell0, cell_start + P.cell0, cell_fill + P.cell0,
                   P.ncell * P.ncell * P.ncell, P.group0);
}

__global__ void k_nl_scatter(const SysParams *__restrict__ sp, const int *__restrict__ rebuild,
                             const int *__restrict__ gcell, const int *__restrict__ cell_start,
                             int *__restrict__ cell_fill, int *__restrict__ order) {
  const int s = blockIdx.y;
  if (!rebuild[s]) return;
  const SysParams &P = sp[s];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.ngroup) return;
  const int g = P.group0 + i;
  const int c = gcell[g];
  order[cell_start[c] + atomicAdd(&cell_fill[c], 1)] = g;
}

// groups one cell may hold: far above any physical density (cells hold
// about two groups on average), low enough that the serial sort stays short
constexpr int NL_SORT_MAX = 512;

// ascending group id within each cell, so the sorted order is deterministic;
// then the first-site positions in sorted order (the pair kernel's gpos)
__global__ void k_nl_sortcell(const SysParams *__restrict__ sp, const int *__restrict__ rebuild,
                              const int *__restrict__ cell_start, int *__restrict__ order,
                              const double *__restrict__ last, double4 *__restrict__ gpos,
                              int *__restrict__ err) {
  const int s = blockIdx.y;
  if (!rebuild[s]) return;
  // after a non-finite position (its group parked in cell 0) skip the sort,
  // whose serial cost grows with the square of a cell's population; the
  // rows built from the unsorted order are discarded with the error
  if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_NAN) return;
  const SysParams &P = sp[s];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.ncell * P.ncell * P.ncell) return;
  const int b = cell_start[P.cell0 + c], e = cell_start[P.cell0 + c + 1];
  if (e - b > NL_SORT_MAX) {   // collapsed positions: reported, not sorted
    atomicOr(err, ERR_CELL);
    return;
  }
  for (int i = b + 1; i < e; i++) {
    const int v = order[i];
    int j = i - 1;
    while (j >= b && order[j] > v) { order[j + 1] = order[j]; j--; }
    order[j + 1] = v;
  }
  for (int i = b; i < e; i++) {
    const int g = order[i];
    gpos[i] = make_double4(last[3 * g], last[3 * g + 1], last[3 * g + 2], 0.0);
  }
}

// One wavefront per sorted group.  The stencil cells that can hold a pair
// (the group's own cell and those whose box lies within rc + skin of it)
// are compacted into a per-wavefront list with their counts' prefix sums;
// their candidates are then laid end to end, each lane takes one, and the
// accepted ones are compacted with a ballot.  The row lists the stencil
// cells in order and each cell's groups in sorted order.
constexpr int NL_SEG = 256;   // listed cells per wavefront before a flush

__global__ void __launch_bounds__(256) k_nl_pairs(
    const SysParams *__restrict__ sp, const int *__restrict__ rebuild,
    const int *__restrict__ order, const int *__restrict__ gcell,
    const int *__restrict__ cell_start, const int4 *__restrict__ sten, int nbcap,
    int *__restrict__ nbr, int *__restrict__ nnb, int *__restrict__ err,
    const double *__restrict__ r, int N, const double *__restrict__ q,
    const int *__restrict__ gfirst, const int *__restrict__ gns, const int *__restrict__ ginfo,
    const double4 *__restrict__ gpos, float4 *__restrict__ spk, int2 *__restrict__ gmeta,
    const int *__restrict__ gexcl, int4 *__restrict__ gslot, const int *__restrict__ cell_rank,
    const int *__restrict__ cell_unrank, int *__restrict__ gslot_inv) {
#pragma clang fp contract(off)
  __shared__ int c_start[4][NL_SEG], c_off[4][NL_SEG + 1], c_cell[4][NL_SEG];
  const int s = blockIdx.y;
  if (!rebuild[s]) return;
  const SysParams &P = sp[s];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int *cs = c_start[wv], *co = c_off[wv], *cc = c_cell[wv];
  const unsigned long long below = (1ull << lane) - 1ull;
  // one wavefront per sorted group (the loop strides only if the grid is
  // smaller); more waves in flight hide the candidate loads of a rebuild
  // better than a grid-stride over a small grid (measured, and the empty
  // launches of the steps that do not rebuild cost no more)
  const int nwave = (int)((gridDim.x * blockDim.x) >> 6);
  // after a non-finite or collapsed position the cells are meaningless (all
  // groups may share one): leave the rows empty
  const bool bad =
      (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & (ERR_NAN | ERR_CELL)) != 0;
  for (int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); wave < P.ngroup;
       wave += nwave) {
    const int gs = P.group0 + wave;
    const int g = order[gs];
    // the finish kernel's per-slot indices in the new order
    if (lane == 0) {
      gslot[gs] = make_int4(g, gfirst[g], gns[g], gexcl[g]);
      gslot_inv[g] = gs;
    }
    if (bad) {   // the error is reported; the rows stay empty
      if (lane == 0) nnb[gs] = 0;
      continue;
    }
    const int n = P.ncell;
    // A, B: cell indices (ix*n + iy)*n + iz, as the reference numbers them;
    // cells are stored at their Hilbert positions (cell_rank)
    const int A = cell_unrank[gcell[g]];
    const int ax = A / (n * n), ay = (A / n) % n, az = A % n;
    const double4 pg = gpos[gs];
    const double L = P.L, invL = 1 / P.L;
    const double c2 = P.nl_c2;
    int cnt = 0;
    int *row = nbr + (long)gs * nbcap;
    int nseg = 0, tot = 0;   // listed cells and their candidates (wave-uniform)
    if (lane == 0) co[0] = 0;
    // test the listed cells' candidates and append the accepted ones
    auto flush = [&]() {
      __builtin_amdgcn_wave_barrier();
      for (int u0 = 0; u0 < tot; u0 += 64) {
        const int u = u0 + lane;
        bool take = false;
        int hs = -1;
        if (u < tot) {
          int lo = 0, hi = nseg - 1;   // last cell with co[k] <= u
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (co[mid] <= u) lo = mid; else hi = mid - 1;
          }
          hs = cs[lo] + (u - co[lo]);
          const int B = cc[lo];
          if (hs != gs) {
            if (B == A) {
              take = true;
            } else {
              // the cell that owns the (A,B) row supplies the first argument
              int ra = A, rb = B;
              canonical_pair(ra, rb);
              const double4 ph = gpos[hs];
              const double d2 = (ra == A) ? min_image_sq(pg.x, pg.y, pg.z, ph.x, ph.y, ph.z, L, invL)
                                          : min_image_sq(ph.x, ph.y, ph.z, pg.x, pg.y, pg.z, L, invL);
              take = d2 < c2;
            }
          }
        }
        const unsigned long long bal = __ballot(take);
        if (take) {
          const int pos = cnt + __popcll(bal & below);
          if (pos < nbcap) row[pos] = hs;
        }
        cnt += __popcll(bal);
      }
      __builtin_amdgcn_wave_barrier();   // the cell list is rewritten next
      nseg = 0;
      tot = 0;
    };
    for (int t0 = 0; t0 < P.nsten; t0 += 64) {
      const int t = t0 + lane;
      int nb = 0, start = 0, B = -1;
      if (t < P.nsten) {
        const int4 o = sten[P.sten0 + t];
        const int bxx = (ax + o.x) % n, byy = (ay + o.y) % n, bzz = (az + o.z) % n;
        B = (bxx * n + byy) * n + bzz;
        const int Bs = P.cell0 + cell_rank[P.cell0 + B];
        start = cell_start[Bs];
        nb = cell_start[Bs + 1] - start;
        // a cell whose box (cell b holds x in [(b - 1/2) lc, (b + 1/2) lc)
        // mod L) lies farther than rc + skin from this group holds no pair;
        // the margin covers the rounding of the cell assignment
        if (B != A) {
          const double lc = P.lc, h = 0.5 * P.lc;
          double dx = pg.x - bxx * lc, dy = pg.y - byy * lc, dz = pg.z - bzz * lc;
          dx = fabs(dx - L * floor(dx * invL + 0.5)) - h;
          dy = fabs(dy - L * floor(dy * invL + 0.5)) - h;
          dz = fabs(dz - L * floor(dz * invL + 0.5)) - h;
          dx = dx > 0 ? dx : 0;
          dy = dy > 0 ? dy : 0;
          dz = dz > 0 ? dz : 0;
          if (dx * dx + dy * dy + dz * dz > c2 * (1 + 1e-9) + 1e-9) nb = 0;
        }
      }
      const unsigned long long keep = __ballot(nb > 0);
      const int k = __popcll(keep);
      if (nseg + k > NL_SEG) flush();
      int incl = nb;   // inclusive scan of the kept counts over the wave
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
      }
      if (nb > 0) {
        const int at = nseg + __popcll(keep & below);
        cs[at] = start;
        cc[at] = B;
        co[at + 1] = tot + incl;
      }
      tot += __shfl(incl, 63);
      nseg += k;
    }
    if (nseg > 0) flush();
    if (lane == 0) {
      if (cnt > nbcap) {
        atomicOr(err, ERR_NL_OVERFLOW);
        cnt = nbcap;
      }
      nnb[gs] = cnt;
    }
    // the pair kernel's packed site inputs in the new sorted order (as
    // k_pack; the first-site positions were written by k_nl_sortcell)
    const int s0 = gfirst[g], ns = gns[g];
    if (lane == 0) gmeta[gs] = make_int2(ns, ginfo[g]);
    if (lane < MAXS) {
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (lane < ns) {
        const int sk = s0 + lane;
        v = make_float4((float)(r[sk] - pg.x), (float)(r[N + sk] - pg.y), (float)(r[2 * N + sk] - pg.z),
                        ((ginfo[g] >> (8 * lane)) & SITE_HAS_Q) ? (float)q[sk] : 0.f);
      }
      spk[(long)gs * MAXS + lane] = v;
    }
  }
}

__global__ void k_set_flags(int *f, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = v;
}

// skin test for externally supplied positions (nbrlist.cpp:64-76): one
// block per system; a rebuild already requested stays requested
__global__ void k_skin(const SysParams *__restrict__ sp, const int *__restrict__ gfirst,
                       const double *__restrict__ r, int N, const double *__restrict__ last,
                       int *__restrict__ rebuild) {
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  double t1 = 0, t2 = 0;
  for (int i = threadIdx.x; i < P.ngroup; i += blockDim.x) {
    const int g = P.group0 + i, st = gfirst[g];
    const double dx = r[st] - last[3 * g], dy = r[N + st] - last[3 * g + 1],
                 dz = r[2 * N + st] - last[3 * g + 2];
    double d = sqrt(dx * dx + dy * dy + dz * dz);
    if (!(d <= 1e300)) d = 1e300;   // a non-finite position (or a list built from one) rebuilds
    if (d > t1) { t2 = t1; t1 = d; } else if (d > t2) t2 = d;
  }
  __shared__ double a1[16], a2[16];
  for (int o = 32; o >= 1; o >>= 1) {
    const double b1 = __shfl_xor(t1, o), b2 = __shfl_xor(t2, o);
    if (b1 > t1) { t2 = fmax(t1, b2); t1 = b1; } else if (b1 > t2) t2 = b1;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { a1[w] = t1; a2[w] = t2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double u1 = 0, u2 = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) {
      if (a1[k] > u1) { u2 = fmax(u1, a2[k]); u1 = a1[k]; } else if (a1[k] > u2) u2 = a1[k];
    }
    if (u1 + u2 > P.skin) rebuild[s] = 1;
  }
}

int launch_skin(Engine &e) {
  hipLaunchKernelGGL(k_skin, dim3(1, e.nsys), dim3(1024), 0, e.stream, e.d_sp, e.d_gfirst, e.d_r,
                     e.nsite, e.d_last, e.d_rebuild);
  MM_HIP(hipGetLastError());
  return 0;
}

int launch_set_rebuild(Engine &e, int value) {
  hipLaunchKernelGGL(k_set_flags, dim3((e.nsys + 63) / 64), dim3(64), 0, e.stream, e.d_rebuild,
                     e.nsys, value);
  MM_HIP(hipGetLastError());
  return 0;
}

int launch_nl_build(Engine &e, int force, int from_partials) {
  if (force) {
    int rc = launch_set_rebuild(e);
    if (rc) return rc;
  }
  const int tb = 256;
  dim3 gg((e.maxg + tb - 1) / tb, e.nsys);
  const double *r = e.d_r;
  hipLaunchKernelGGL(k_nl_assign, gg, dim3(tb), 0, e.stream, e.d_sp, e.d_gfirst, r, e.nsite,
                     e.d_rebuild, e.d_gcell, e.d_last, e.d_cell_count, e.d_nlb, e.d_err,
                     e.d_fpart, e.nblk_part, from_partials, e.d_cell_rank);
  hipLaunchKernelGGL(k_nl_scan, dim3(1, e.nsys), dim3(1024), 0, e.stream, e.d_sp, e.d_rebuild,
                     e.d_cell_count, e.d_cell_start, e.d_cell_fill);
  hipLaunchKernelGGL(k_nl_scatter, gg, dim3(tb), 0, e.stream, e.d_sp, e.d_rebuild, e.d_gcell,
                     e.d_cell_start, e.d_cell_fill, e.d_order);
  int maxc = 0;
  for (auto &p : e.hsp) maxc = std::max(maxc, p.ncell * p.ncell * p.ncell);
  hipLaunchKernelGGL(k_nl_sortcell, dim3((maxc + tb - 1) / tb, e.nsys), dim3(tb), 0, e.stream,
                     e.d_sp, e.d_rebuild, e.d_cell_start, e.d_order, e.d_last, e.d_gpos,
                     e.d_err);
  hipLaunchKernelGGL(k_nl_pairs, dim3((e.maxg + 3) / 4, e.nsys), dim3(256), 0,
                     e.stream, e.d_sp, e.d_rebuild, e.d_order, e.d_gcell, e.d_cell_start,
                     e.d_sten, e.nbcap, e.d_nbr, e.d_nnb, e.d_err, e.d_r, e.nsite, e.d_q,
                     e.d_gfirst, e.d_gns, e.d_ginfo, e.d_gpos, e.d_spk, e.d_gmeta, e.d_gexcl,
                     e.d_gslot, e.d_cell_rank, e.d_cell_unrank, e.d_gslot_inv);
  MM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mmamd
