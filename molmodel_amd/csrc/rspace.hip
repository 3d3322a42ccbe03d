// rspace.hip -- Ewald real-space erfc-Coulomb + Lennard-Jones pair kernel.
//
// Same interaction set and formulas as RSpaceEwald (rspace.c:382-558):
//   * per neighbour group pair, one minimum-image shift from the first sites
//     (rspace.c:423-435), FP64, and the same r^2 > rc^2 skip and
//     (rc-w)^2 < r^2 switching decision (rspace.c:436-445);
//   * every site pair of the two groups uses that shift (rspace.c:454-478);
//   * erfc(eta r)/r and -2 d/dr^2 of it from a cubic Chebyshev table in r^2
//     (cheby.c / rspace.c:19-37; built in FP64 on the host, evaluated in
//     FP32 here), exact erfc outside the table (rspace.c:116-125);
//   * Lorentz-Berthelot LJ (rspace.c:258-297) via a per-type-pair table;
//   * the switching derivative times the sum of unswitched pair energies acts
//     on the first sites (rspace.c:479-489).
// MI355X mapping: groups are cell-sorted; 16 lanes serve one i-group and
// walk its full neighbour row (both directions, so every force lands on the
// lane's own i-group and no atomics are needed); the i-group lives in
// registers; neighbours inside the cutoff are compacted into an LDS queue
// before the site-pair arithmetic (FP32) runs; each group pair's FP32
// partial sums are added into FP64 accumulators; the 16 lanes of an i-group
// are reduced with cross-lane shuffles at the end.
#include "engine.hpp"

namespace mmamd {

// ---- per-step pack: cell-sorted FP64 first sites + FP32 site offsets ----
__global__ void k_pack(const SysParams *__restrict__ sp, const int *__restrict__ order,
                       const int *__restrict__ gfirst, const int *__restrict__ gns,
                       const int *__restrict__ ginfo, const double *__restrict__ r,
                       const double *__restrict__ q, int N, double4 *__restrict__ gpos,
                       float4 *__restrict__ spk, int2 *__restrict__ gmeta) {
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.ngroup) return;
  const int gs = P.group0 + i;
  const int g = order[gs];
  const int s0 = gfirst[g], ns = gns[g];
  const double x0 = r[s0], y0 = r[N + s0], z0 = r[2 * N + s0];
  gpos[gs] = make_double4(x0, y0, z0, 0.0);
#pragma unroll
  for (int k = 0; k < MAXS; k++) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (k < ns) {
      const int sk = s0 + k;
      v = make_float4((float)(r[sk] - x0), (float)(r[N + sk] - y0), (float)(r[2 * N + sk] - z0),
                      ((ginfo[g] >> (8 * k)) & SITE_HAS_Q) ? (float)q[sk] : 0.f);
    }
    spk[(long)gs * MAXS + k] = v;
  }
  gmeta[gs] = make_int2(ns, ginfo[g]);
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 16); }


constexpr int NSI = 3;       // sites per group handled by the pair kernel
constexpr int LANES = 16;    // lanes per i-group
constexpr int PAIR_TB = 256;
constexpr int GPB = PAIR_TB / LANES;   // i-groups per block
constexpr int MAXTAB = 1536;
constexpr int MAXLJ = 64;

constexpr int QCAP = 32;      // queue entries per i-group (one batch of 16 + a fill of 16)

// The pair kernel.  Each 16-lane segment owns one i-group and alternates
// two phases over its neighbour row:
//   filter:  16 candidates at a time; FP64 minimum image and the
//            r^2 > rc^2 decision of rspace.c:423-445, the switching value
//            and derivative (rspace.c:52-58) from the FP64 distance in FP32;
//            accepted neighbours are compacted into a per-segment LDS queue
//            with a ballot (FP32 shift, switching value and derivative);
//   compute: once every segment of the wavefront holds a full batch (or has
//            no candidates left), each lane takes one queued group pair and
//            evaluates its site pairs in FP32, adding FP32 per-pair sums
//            into FP64 accumulators.
// Lanes therefore spend the expensive phase almost only on pairs inside the
// cutoff.
// erfc(eta r)/r of one charged site pair (b0) and its radial force factor
// times the smoothed charge qs (e), from the LDS table inside its range and
// from the functions outside it (rspace.c:120-125).  Both table segments
// share one index space: x_A(r2) = (r2 - Amin) / dxA and x_B(r2) = nA +
// (r2 - Mid) / dxB meet at r2 = Mid, and A is the steeper map, so the table
// index is min(x_A, x_B) and the segment is the one giving the minimum.
struct TabView {
  float amin, bmax, ainv, binv, offA, offB, m2A, m2B, eta;
  unsigned ntab;   // intervals in both segments
};

__device__ __forceinline__ void coul_pair(const TabView &T, const float4 *stab, float r2, float qs,
                                          float qsA, float qsB, float &b0, float &e) {
  if (r2 >= T.amin && r2 < T.bmax) {
    const float xa = fmaf(r2, T.ainv, T.offA), xb = fmaf(r2, T.binv, T.offB);
    const bool inA = xa < xb;
    const float xx = inA ? xa : xb;
    const int ii = (int)xx;
    const float fr = xx - (float)ii;
    const float4 c = stab[ii];
    const float p2 = fmaf(c.w, fr, c.z), p1 = fmaf(p2, fr, c.y);
    b0 = fmaf(p1, fr, c.x);
    // b1 = -2/dx * d(cubic)/dt (the derivative from the Horner
    // intermediates), times the smoothed charge (qsA = qs * -2/dxA, qsB
    // likewise)
    e = fmaf(fr, fmaf(c.w, fr, p2), p1) * (inA ? qsA : qsB);
  } else {
    const float rm = sqrtf(r2), er = T.eta * rm;
    b0 = erfcf(er) / rm;
    e = qs * (b0 + T.eta * 1.1283791670955126f * expf(-er * er)) / r2;
  }
}

// The table branch of coul_pair without the range test: the interval index
// is clamped into the table and the largest unclamped index (as unsigned, so
// that an index below the table is large too) is kept in worst; a caller
// whose worst reaches T.ntab evaluates the group pair again with coul_pair.
// Inside the range the arithmetic is coul_pair's.
__device__ __forceinline__ void coul_table(const TabView &T, const float4 *stab, float r2,
                                           float qsA, float qsB, float &b0, float &e,
                                           unsigned &worst) {
  const float xa = fmaf(r2, T.ainv, T.offA), xb = fmaf(r2, T.binv, T.offB);
  const bool inA = xa < xb;
  const float xx = inA ? xa : xb;
  int iu;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(iu) : "v"(xx));   // floor, as an integer
  const unsigned u = (unsigned)iu;
  worst = u > worst ? u : worst;
  const float fr = __builtin_amdgcn_fractf(xx);
  const float4 c = stab[u < T.ntab - 1u ? u : T.ntab - 1u];
  const float p2 = fmaf(c.w, fr, c.z), p1 = fmaf(p2, fr, c.y);
  b0 = fmaf(p1, fr, c.x);
  e = fmaf(fr, fmaf(c.w, fr, p2), p1) * (inA ? qsA : qsB);
}

// One tile: the GPB i-groups blk * GPB ... of system P.  The block's LDS
// holds that system's erfc and LJ tables.
template <bool FIELD, bool VIRIAL, bool WATER, bool ENERGY>
__device__ __forceinline__ void pair_tile(
    const SysParams &P, int blk, const float4 *stab, const float2 *slj, float4 (*qd)[QCAP],
    float2 (*qs)[QCAP], const double4 *__restrict__ gpos, const float4 *__restrict__ spk,
    const int2 *__restrict__ gmeta, const int *__restrict__ nbr, const int *__restrict__ nnb,
    int nbcap, double *__restrict__ pf, double *__restrict__ plj, double *__restrict__ gacc) {
  const int lane = threadIdx.x & (LANES - 1);
  const int segw = (threadIdx.x & 63) / LANES;     // segment within the wavefront
  const int seg = threadIdx.x / LANES;             // segment within the block
  const int i = blk * GPB + seg;
  const bool active = i < P.ngroup;
  const int gs = P.group0 + (active ? i : 0);

  const double4 pi = gpos[gs];
  const int2 mi = gmeta[gs];
  const int nsi = mi.x;
  float4 oi[NSI];
  int mki[NSI], tyi[NSI];
#pragma unroll
  for (int k = 0; k < NSI; k++) {
    oi[k] = spk[(long)gs * MAXS + k];
    const int b = (mi.y >> (8 * k)) & 0xff;
    mki[k] = k < nsi ? (b & 3) : 0;
    tyi[k] = b >> 2;
  }
  double ax[NSI], ay[NSI], az[NSI], ap[NSI];
  double lx[NSI], ly[NSI], lz[NSI];
  double ulj = 0;
  double w[9];
#pragma unroll
  for (int k = 0; k < NSI; k++) { ax[k] = ay[k] = az[k] = ap[k] = 0; lx[k] = ly[k] = lz[k] = 0; }
#pragma unroll
  for (int c = 0; c < 9; c++) w[c] = 0;

  const double L = P.L, invL = P.invL, rhi2 = P.rhi2, rlo2 = P.rlo2, inv_w2 = P.inv_w2;
  const float inv_w2f = (float)inv_w2;
  TabView T;
  T.amin = P.tabAmin; T.bmax = P.tabBmax; T.ainv = P.tabAinv; T.binv = P.tabBinv;
  T.offA = -P.tabAmin * P.tabAinv;
  T.offB = (float)P.ntabA - P.tabMid * P.tabBinv;
  T.m2A = -2.f * P.tabAinv; T.m2B = -2.f * P.tabBinv;
  T.eta = (float)P.eta;
  T.ntab = (unsigned)(P.ntabA + P.ntabB);
  const int nty = P.ntype;
  const int cnt = active ? nnb[gs] : 0;
  const int *row = nbr + (long)gs * nbcap;
  float4 *myqd = qd[seg];
  float2 *myqs = qs[seg];
  int qn = 0, t0 = 0;
  // the row entry of the lane's next candidate, loaded one filter pass ahead
  int hnext = lane < cnt ? row[lane] : 0;

  while (true) {
    // ---- filter until every segment holds a batch or has no candidates
    while (__any(qn < LANES && t0 < cnt)) {
      const bool fill = qn < LANES && t0 < cnt;
      bool take = false;
      float4 ent = make_float4(0.f, 0.f, 0.f, 1.f);
      float dsmf = 0.f;
      int hs = 0;
      if (fill && t0 + lane < cnt) {
        hs = hnext;
        const double4 pj = gpos[hs];
        if (t0 + LANES + lane < cnt) hnext = row[t0 + LANES + lane];
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        dx -= L * floor(dx * invL + 0.5);
        dy -= L * floor(dy * invL + 0.5);
        dz -= L * floor(dz * invL + 0.5);
        const double r2g = dx * dx + dy * dy + dz * dz;
        take = !(r2g > rhi2);
        // switching function (rspace.c:52-58) of z = (r^2 - rlo^2) / w2,
        // branch-free in FP32 (its value and derivative are FP32 here
        // anyway): z clamped at 0 gives S = 1, dS = 0 inside rlo
        const float z = fmaxf((float)((r2g - rlo2) * inv_w2), 0.f);
        const float z3 = z * z * z;
        const float smz = fmaf(fmaf(fmaf(-6.f, z, 15.f), z, -10.f), z3, 1.f);
        const float dsz = fmaf(fmaf(60.f, z, -120.f), z, 60.f) * z * z * inv_w2f;
        ent = make_float4((float)dx, (float)dy, (float)dz, smz);
        dsmf = dsz;
      }
      const unsigned long long bal = __ballot(take);
      const unsigned sb = (unsigned)(bal >> (segw * LANES)) & 0xffffu;
      if (take) {
        const int pos = qn + __popc(sb & ((1u << lane) - 1u));
        myqd[pos] = ent;
        myqs[pos] = make_float2(dsmf, __int_as_float(hs));
      }
      if (fill) {
        qn += __popc(sb);
        t0 += LANES;
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (!__any(qn > 0)) break;
    // ---- compute one batch per segment
    const int nb = qn < LANES ? qn : LANES;
    if (lane < nb) {
      const float4 ent = myqd[qn - nb + lane];
      const float2 e2 = myqs[qn - nb + lane];
      const int hs = __float_as_int(e2.y);
      const float smf = ent.w, dsmf = e2.x;
      const int2 mj = gmeta[hs];
      const int nsj = mj.x;
      const float Dx = ent.x, Dy = ent.y, Dz = ent.z;
      // per group pair FP32 sums (set by the path that runs)
      float fx[NSI], fy[NSI], fz[NSI], ph[NSI], gx[NSI], gy[NSI], gz[NSI];
      float tmps, uu;
      bool general = true;
      if (WATER && P.water) {
        // every group of this system is a three-site water: nine charged
        // site pairs and the first sites' LJ pair, without per-site mask
        // branches (a tile never spans two systems, so the branch is
        // uniform) and without the table's range test; a lane with a site
        // pair outside the table (closer than the table's start) takes the
        // general path below for this group pair
        unsigned worst = 0;
        const float4 o0 = spk[(long)hs * MAXS];
        {   // first sites: LJ
          const float rx = Dx - o0.x + oi[0].x, ry = Dy - o0.y + oi[0].y, rz = Dz - o0.z + oi[0].z;
          const float r2 = rx * rx + ry * ry + rz * rz;
          const float2 c = slj[tyi[0] * nty + ((mj.y >> 2) & 0x3f)];
          const float ir2 = __builtin_amdgcn_rcpf(r2);
          const float ir6 = ir2 * ir2 * ir2;
          const float bi6 = c.x * ir6, ai12 = c.y * ir6 * ir6;
          uu = ai12 - bi6;
          const float du = (12.f * ai12 - 6.f * bi6) * ir2 * smf;
          gx[0] = du * rx; gy[0] = du * ry; gz[0] = du * rz;
#pragma unroll
          for (int k = 1; k < NSI; k++) gx[k] = gy[k] = gz[k] = 0.f;
        }
        const float4 o1 = spk[(long)hs * MAXS + 1], o2 = spk[(long)hs * MAXS + 2];
#pragma unroll
        for (int l = 0; l < 3; l++) {
          const float4 oj = l == 0 ? o0 : l == 1 ? o1 : o2;
          const float px = Dx - oj.x, py = Dy - oj.y, pz = Dz - oj.z;
          const float qs = oj.w * smf, qsA = qs * T.m2A, qsB = qs * T.m2B;
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const float rx = px + oi[k].x, ry = py + oi[k].y, rz = pz + oi[k].z;
            const float r2 = rx * rx + ry * ry + rz * rz;
            float b0, e;
            coul_table(T, stab, r2, qsA, qsB, b0, e, worst);
            if (l == 0) {   // the sums start with site 0's terms
              ph[k] = oj.w * b0;   // unsmoothed here, smoothed below
              fx[k] = e * rx; fy[k] = e * ry; fz[k] = e * rz;
            } else {
              ph[k] += oj.w * b0;
              fx[k] += e * rx; fy[k] += e * ry; fz[k] += e * rz;
            }
          }
        }
        // unsmoothed pair energy over the charges, then the smoothed potentials
        tmps = oi[0].w * ph[0] + oi[1].w * ph[1] + oi[2].w * ph[2];
#pragma unroll
        for (int k = 0; k < 3; k++)
          if (ENERGY) ph[k] *= smf;
        general = worst >= T.ntab;
      }
      if (general) {
#pragma unroll
        for (int k = 0; k < NSI; k++) { fx[k] = fy[k] = fz[k] = ph[k] = 0.f; gx[k] = gy[k] = gz[k] = 0.f; }
        tmps = uu = 0.f;
#pragma unroll 1
      for (int l = 0; l < nsj; l++) {
        const float4 oj = spk[(long)hs * MAXS + l];
        const int bj = (mj.y >> (8 * l)) & 0xff;
        const int mkj = bj & 3, tyj = bj >> 2;
        const float px = Dx - oj.x, py = Dy - oj.y, pz = Dz - oj.z;
        const float qj = oj.w;
        const float qs = qj * smf;        // smoothed charge of site l
        const float qsA = qs * T.m2A, qsB = qs * T.m2B;
        float tq = 0.f;                   // sum_k q_k b0_k (unsmoothed energy / q_l)
#pragma unroll
        for (int k = 0; k < NSI; k++) {
          const int m = mki[k] & mkj;
          if (!m) continue;
          const float rx = px + oi[k].x, ry = py + oi[k].y, rz = pz + oi[k].z;
          const float r2 = rx * rx + ry * ry + rz * rz;
          if (m & SITE_HAS_Q) {
            float b0, e;
            coul_pair(T, stab, r2, qs, qsA, qsB, b0, e);
            tq += oi[k].w * b0;
            if (ENERGY) ph[k] += qs * b0;
            fx[k] += e * rx; fy[k] += e * ry; fz[k] += e * rz;
          }
          if (m & SITE_HAS_LJ) {
            const float2 c = slj[tyi[k] * nty + tyj];
            const float ir2 = __builtin_amdgcn_rcpf(r2);
            const float ir6 = ir2 * ir2 * ir2;
            const float bi6 = c.x * ir6, ai12 = c.y * ir6 * ir6;
            const float u = ai12 - bi6;
            const float du = (12.f * ai12 - 6.f * bi6) * ir2 * smf;
            uu += u;
            gx[k] += du * rx; gy[k] += du * ry; gz[k] += du * rz;
          }
        }
        tmps += qj * tq;
      }
      }
      tmps += uu;
      if (ENERGY) uu *= smf;
      // smoothing force on the first site (rspace.c:479-489)
      const float fs = dsmf * tmps;
      const float sx = fs * Dx, sy = fs * Dy, sz = fs * Dz;
      double Fx = 0, Fy = 0, Fz = 0;
#pragma unroll
      for (int k = 0; k < NSI; k++) {
        if (ENERGY) ap[k] += (double)ph[k];
        if (FIELD) {
          ax[k] += (double)fx[k]; ay[k] += (double)fy[k]; az[k] += (double)fz[k];
          lx[k] += (double)gx[k]; ly[k] += (double)gy[k]; lz[k] += (double)gz[k];
        } else {
          // the smoothing force joins the first site's FP32 sum
          const float qk = oi[k].w;
          float tx = qk * fx[k] + gx[k], ty = qk * fy[k] + gy[k], tz = qk * fz[k] + gz[k];
          if (k == 0) { tx += sx; ty += sy; tz += sz; }
          ax[k] += (double)tx; ay[k] += (double)ty; az[k] += (double)tz;
          if (VIRIAL) { Fx += tx; Fy += ty; Fz += tz; }
        }
      }
      if (FIELD) { lx[0] += (double)sx; ly[0] += (double)sy; lz[0] += (double)sz; }
      if (ENERGY) ulj += (double)uu;
      if (VIRIAL) {   // 1/2 D (x) F_gh per directed pair
        const double dx = Dx, dy = Dy, dz = Dz;
        w[0] += 0.5 * dx * Fx; w[1] += 0.5 * dx * Fy; w[2] += 0.5 * dx * Fz;
        w[3] += 0.5 * dy * Fx; w[4] += 0.5 * dy * Fy; w[5] += 0.5 * dy * Fz;
        w[6] += 0.5 * dz * Fx; w[7] += 0.5 * dz * Fy; w[8] += 0.5 * dz * Fz;
      }
    }
    qn -= nb;
    __builtin_amdgcn_wave_barrier();
  }
  // reduce the 16 lanes of each i-group
#pragma unroll
  for (int m = LANES / 2; m >= 1; m >>= 1) {
#pragma unroll
    for (int k = 0; k < NSI; k++) {
      ax[k] += shfl_xor_d(ax[k], m); ay[k] += shfl_xor_d(ay[k], m);
      az[k] += shfl_xor_d(az[k], m);
      if (ENERGY) ap[k] += shfl_xor_d(ap[k], m);
      if (FIELD) {
        lx[k] += shfl_xor_d(lx[k], m); ly[k] += shfl_xor_d(ly[k], m);
        lz[k] += shfl_xor_d(lz[k], m);
      }
    }
    if (ENERGY) ulj += shfl_xor_d(ulj, m);
    if (VIRIAL) {
#pragma unroll
      for (int c = 0; c < 9; c++) w[c] += shfl_xor_d(w[c], m);
    }
  }
  if (!active || lane != 0) return;
#pragma unroll
  for (int k = 0; k < NSI; k++) {
    double *o = pf + ((long)gs * MAXS + k) * 4;
    o[0] = ax[k]; o[1] = ay[k]; o[2] = az[k];
    if (ENERGY) o[3] = ap[k];
    if (FIELD) {
      double *l = plj + ((long)gs * MAXS + k) * 3;
      l[0] = lx[k]; l[1] = ly[k]; l[2] = lz[k];
    }
  }
  double *ga = gacc + (long)gs * 10;
  if (ENERGY) ga[0] = 0.5 * ulj;   // each unordered pair is visited from both sides
  if (VIRIAL) {
    // per-site offset term: sum_k o_k (x) F_k (o_0 = 0)
#pragma unroll
    for (int k = 1; k < NSI; k++) {
      if (k >= nsi) break;
      const double ox = oi[k].x, oy = oi[k].y, oz = oi[k].z;
      w[0] += ox * ax[k]; w[1] += ox * ay[k]; w[2] += ox * az[k];
      w[3] += oy * ax[k]; w[4] += oy * ay[k]; w[5] += oy * az[k];
      w[6] += oz * ax[k]; w[7] += oz * ay[k]; w[8] += oz * az[k];
    }
#pragma unroll
    for (int c = 0; c < 9; c++) ga[1 + c] = w[c];
  }
}

// Persistent pair kernel.  Tiles of GPB i-groups are numbered over all
// systems (system s owns tiles [tile_start[s], tile_start[s+1])) and split
// into 8 contiguous ranges, one per class of blocks with the same
// blockIdx.x mod 8.  Workgroups are dealt to the 8 XCDs round robin, so a
// class is the set of blocks on one XCD, and that XCD's L2 then holds the
// positions of one slab of the cell-sorted groups and its halo instead of
// all of them.  Block k of a class takes tiles lo + k, lo + k + nper, ...;
// the tiles in flight at any time are thus one contiguous window of the
// slab.  Waves move on to their next tile without waiting for each other
// (each wave owns its segments' queues); a block synchronises only to load
// another system's tables.  The host sizes the grid below the device's
// resident capacity, which leaves room on every CU for the reciprocal-space
// kernels running on the other stream.
template <bool FIELD, bool VIRIAL, bool WATER, bool ENERGY>
__global__ void __launch_bounds__(PAIR_TB, (FIELD || VIRIAL) ? 1 : 4) k_pair(
    const SysParams *__restrict__ sp, const int *__restrict__ tile_start, int nsys, int ntile,
    const double4 *__restrict__ gpos, const float4 *__restrict__ spk,
    const int2 *__restrict__ gmeta, const int *__restrict__ nbr, const int *__restrict__ nnb,
    int nbcap, const float4 *__restrict__ tab, const float2 *__restrict__ ljt,
    double *__restrict__ pf, double *__restrict__ plj, double *__restrict__ gacc) {
  __shared__ float4 stab[MAXTAB];
  __shared__ float2 slj[MAXLJ];
  __shared__ float4 qd[GPB][QCAP];    // Dx, Dy, Dz, switching value
  __shared__ float2 qs[GPB][QCAP];    // switching derivative, neighbour (int bits)
  const int cls = blockIdx.x & 7;
  const int nper = (int)((gridDim.x + 7 - cls) >> 3);   // blocks in this class
  const int lo = (int)((long)ntile * cls / 8), hi = (int)((long)ntile * (cls + 1) / 8);
  int cur = -1;
  for (int t = lo + (int)(blockIdx.x >> 3); t < hi; t += nper) {
    int a = 0, b = nsys - 1;   // the last system whose tiles start at or before t
    while (a < b) {
      const int m = (a + b + 1) >> 1;
      if (tile_start[m] <= t) a = m; else b = m - 1;
    }
    const SysParams &P = sp[a];
    if (a != cur) {
      __syncthreads();   // every wave is done with the previous tables
      const int ntab = P.ntabA + P.ntabB;
      for (int u = threadIdx.x; u < ntab; u += PAIR_TB) stab[u] = tab[P.tab0 + u];
      for (int u = threadIdx.x; u < P.ntype * P.ntype; u += PAIR_TB) slj[u] = ljt[P.lj0 + u];
      __syncthreads();
      cur = a;
    }
    pair_tile<FIELD, VIRIAL, WATER, ENERGY>(P, t - tile_start[a], stab, slj, qd, qs, gpos, spk, gmeta, nbr,
                             nnb, nbcap, pf, plj, gacc);
  }
}

// interaction counts of the current pair list, per system: group pairs with
// first-site r^2 <= rc^2 (each unordered pair once), charged site pairs and
// LJ site pairs among them
__global__ void k_count_pairs(const SysParams *__restrict__ sp, const double4 *__restrict__ gpos,
                              const int2 *__restrict__ gmeta, const int *__restrict__ nbr,
                              const int *__restrict__ nnb, int nbcap,
                              unsigned long long *__restrict__ out) {
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.ngroup) return;
  const int gs = P.group0 + i;
  const double4 pi = gpos[gs];
  const int2 mi = gmeta[gs];
  unsigned long long ng = 0, nq = 0, nl = 0;
  for (int t = 0; t < nnb[gs]; t++) {
    const int hs = nbr[(long)gs * nbcap + t];
    if (hs <= gs) continue;
    const double4 pj = gpos[hs];
    double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    dx -= P.L * floor(dx * P.invL + 0.5);
    dy -= P.L * floor(dy * P.invL + 0.5);
    dz -= P.L * floor(dz * P.invL + 0.5);
    if (dx * dx + dy * dy + dz * dz > P.rhi2) continue;
    ng++;
    const int2 mj = gmeta[hs];
    for (int k = 0; k < mi.x; k++)
      for (int l = 0; l < mj.x; l++) {
        const int m = ((mi.y >> (8 * k)) & 3) & ((mj.y >> (8 * l)) & 3);
        nq += (m & SITE_HAS_Q) ? 1 : 0;
        nl += (m & SITE_HAS_LJ) ? 1 : 0;
      }
  }
  atomicAdd(&out[3 * s], ng);
  atomicAdd(&out[3 * s + 1], nq);
  atomicAdd(&out[3 * s + 2], nl);
}

int launch_count_pairs(Engine &e, unsigned long long *d_out) {
  const int tb = 256;
  hipLaunchKernelGGL(k_count_pairs, dim3((e.maxg + tb - 1) / tb, e.nsys), dim3(tb), 0, e.stream,
                     e.d_sp, e.d_gpos, e.d_gmeta, e.d_nbr, e.d_nnb, e.nbcap, d_out);
  MM_HIP(hipGetLastError());
  return 0;
}

int launch_pack(Engine &e) {
  const int tb = 256;
  hipLaunchKernelGGL(k_pack, dim3((e.maxg + tb - 1) / tb, e.nsys), dim3(tb), 0, e.stream, e.d_sp,
                     e.d_order, e.d_gfirst, e.d_gns, e.d_ginfo, e.d_r, e.d_q, e.nsite, e.d_gpos,
                     e.d_spk, e.d_gmeta);
  MM_HIP(hipGetLastError());
  return 0;
}

// Real-space potential at arbitrary points (RSpaceEwaldPhiAtGridpoints,
// rspace.c:892-967): one thread per point, the sites streamed through LDS
// in tiles; minimum image, cutoff and switching on the site-point distance
// in FP64 (the reference's decision), erfc(eta r)/r from the LDS table in
// FP32, accumulation in FP64.  sites: (x, y, z, q) with q = 0 for sites
// without any mask bit.
constexpr int PHI_TB = 256;

__global__ void __launch_bounds__(PHI_TB) k_phi_points(const SysParams *__restrict__ sp,
                                                       const double4 *__restrict__ sites, int nsite,
                                                       const double *__restrict__ pts, int npts,
                                                       const float4 *__restrict__ tab,
                                                       double *__restrict__ phi) {
  __shared__ float4 stab[MAXTAB];
  __shared__ double4 ssite[PHI_TB];
  const SysParams &P = sp[0];
  const int ntab = P.ntabA + P.ntabB;
  for (int t = threadIdx.x; t < ntab; t += PHI_TB) stab[t] = tab[P.tab0 + t];
  const int j = blockIdx.x * PHI_TB + threadIdx.x;
  const bool on = j < npts;
  const double gx = on ? pts[3 * j] : 0, gy = on ? pts[3 * j + 1] : 0, gz = on ? pts[3 * j + 2] : 0;
  const double L = P.L, invL = P.invL, rhi2 = P.rhi2, rlo2 = P.rlo2;
  const float tAmin = P.tabAmin, tAinv = P.tabAinv, tBinv = P.tabBinv, tBmax = P.tabBmax;
  const float offA = -tAmin * tAinv, offB = (float)P.ntabA - P.tabMid * tBinv;
  const float eta = (float)P.eta;
  double acc = 0;
  for (int t0 = 0; t0 < nsite; t0 += PHI_TB) {
    __syncthreads();
    if (t0 + (int)threadIdx.x < nsite) ssite[threadIdx.x] = sites[t0 + threadIdx.x];
    __syncthreads();
    const int m = min(PHI_TB, nsite - t0);
    if (!on) continue;
    for (int k = 0; k < m; k++) {
      const double4 s = ssite[k];
      double x = s.x - gx, y = s.y - gy, z = s.z - gz;
      x -= L * floor(x * invL + 0.5);
      y -= L * floor(y * invL + 0.5);
      z -= L * floor(z * invL + 0.5);
      const double r2 = x * x + y * y + z * z;
      if (r2 > rhi2 || s.w == 0.0) continue;
      const float r2f = (float)r2;
      float b0;
      if (r2f >= tAmin && r2f < tBmax) {
        const float xa = fmaf(r2f, tAinv, offA), xb = fmaf(r2f, tBinv, offB);
        const float xx = xa < xb ? xa : xb;
        const int ii = (int)xx;
        const float fr = xx - (float)ii;
        const float4 c = stab[ii];
        b0 = fmaf(fmaf(fmaf(c.w, fr, c.z), fr, c.y), fr, c.x);
      } else {
        const float rm = sqrtf(r2f);
        b0 = erfcf(eta * rm) / rm;
      }
      double smooth = 1.0;
      if (r2 > rlo2) {   // rspace.c:52-58
        const double zz = (r2 - rlo2) * P.inv_w2;
        smooth = ((-6 * zz + 15) * zz - 10) * zz * zz * zz + 1;
      }
      acc += s.w * (double)b0 * smooth;
    }
  }
  if (on) phi[j] += acc;
}

int launch_phi_points(Engine &e, const double4 *sites, int nsite, const double *pts, int npts,
                      double *phi) {
  if (npts <= 0) return 0;
  hipLaunchKernelGGL(k_phi_points, dim3((npts + PHI_TB - 1) / PHI_TB), dim3(PHI_TB), 0, e.stream,
                     e.d_sp, sites, nsite, pts, npts, e.d_tab, phi);
  MM_HIP(hipGetLastError());
  return 0;
}

// ---- non-periodic real space (rspace.c:560-640, 784-900), FP64 ----------
// Every site pair of two different groups, no cutoff, no minimum image:
// Coulomb 1/r, or inside the screening radius rs = rscreen_i + rscreen_j
// the screened form b0 = 3/(2 rs) - r^2/(2 rs^3), b1 = 1/rs^3
// (rspace.c:61-64); Lennard-Jones with Lorentz-Berthelot or geometric sigma
// (rspace.c:258-297).  One thread per site (or grid point), the other
// sites streamed through LDS; each thread sums its own site's terms, so
// the two members of a pair are visited from both sides.
//   info: per site (group, mask, 1 if first site of its group, 0)
//   lj:   per site (sigma, epsilon)
//   MODE 0 (RSpaceNonPeriodic): out 7 per site (phi, evec, lj_force) and
//          per block the LJ energy of the block's sites (half of each pair);
//   MODE 1 (RSpaceNonPeriodicPhi): phi only, with that routine's pair rule:
//          the two first sites always, other pairs when their masks share
//          any bit (rspace.c:801-822).
constexpr int NP_TB = 256;

__device__ __forceinline__ void np_coulomb(double r2, double rs, double &b0, double &b1) {
  if (r2 < rs * rs) {   // screening_function, rspace.c:61-64
    b1 = 1.0 / (rs * rs * rs);
    b0 = -0.5 * r2 * b1 + 1.5 / rs;
  } else {
    const double invr = 1.0 / sqrt(r2);
    b0 = invr;
    b1 = invr * invr * invr;
  }
}

template <int MODE>
__global__ void __launch_bounds__(NP_TB) k_nonperiodic(const double4 *__restrict__ site,
                                                      const int4 *__restrict__ info,
                                                      const double2 *__restrict__ lj,
                                                      const double *__restrict__ rscreen, int n,
                                                      int geometric, double *__restrict__ out,
                                                      double *__restrict__ upart) {
  __shared__ double4 ss[NP_TB];
  __shared__ int4 si[NP_TB];
  __shared__ double2 sl[NP_TB];
  __shared__ double sr[NP_TB];
  __shared__ double red[NP_TB / 64];
  const int i = blockIdx.x * NP_TB + threadIdx.x;
  const bool on = i < n;
  const double4 a = on ? site[i] : make_double4(0, 0, 0, 0);
  const int4 ia = on ? info[i] : make_int4(-1, 0, 0, 0);
  const double2 la = on ? lj[i] : make_double2(0, 0);
  const double ra = on ? rscreen[i] : 0.0;
  double phi = 0, ex = 0, ey = 0, ez = 0, fx = 0, fy = 0, fz = 0, u = 0;
  for (int t0 = 0; t0 < n; t0 += NP_TB) {
    __syncthreads();
    const int t = t0 + threadIdx.x;
    if (t < n) {
      ss[threadIdx.x] = site[t];
      si[threadIdx.x] = info[t];
      sl[threadIdx.x] = lj[t];
      sr[threadIdx.x] = rscreen[t];
    }
    __syncthreads();
    const int m = min(NP_TB, n - t0);
    if (!on) continue;
    for (int k = 0; k < m; k++) {
      const int4 ib = si[k];
      if (ib.x == ia.x) continue;   // same group
      const double4 b = ss[k];
      const double x = a.x - b.x, y = a.y - b.y, z = a.z - b.z;
      const double r2 = x * x + y * y + z * z;
      const int mask = ia.y & ib.y;
      if (MODE == 1) {
        if (!((ia.z && ib.z) || mask)) continue;
        double b0, b1;
        np_coulomb(r2, ra + sr[k], b0, b1);
        phi += b.w * b0;
        continue;
      }
      if (mask & SITE_HAS_Q) {
        double b0, b1;
        np_coulomb(r2, ra + sr[k], b0, b1);
        phi += b.w * b0;
        ex += b.w * b1 * x;
        ey += b.w * b1 * y;
        ez += b.w * b1 * z;
      }
      if (mask & SITE_HAS_LJ) {
        const double2 lb = sl[k];
        const double s2 = geometric ? la.x * lb.x : 0.25 * (la.x + lb.x) * (la.x + lb.x);
        const double s6 = s2 * s2 * s2;
        const double bb = 4 * sqrt(la.y * lb.y) * s6, aa = bb * s6;
        const double ir2 = 1.0 / r2, ir6 = ir2 * ir2 * ir2;
        const double du = (12 * aa * ir6 * ir6 - 6 * bb * ir6) * ir2;
        u += 0.5 * (aa * ir6 * ir6 - bb * ir6);
        fx += du * x;
        fy += du * y;
        fz += du * z;
      }
    }
  }
  if (on) {
    double *o = out + 7L * i;
    o[0] = phi;
    if (MODE == 0) {
      o[1] = ex; o[2] = ey; o[3] = ez;
      o[4] = fx; o[5] = fy; o[6] = fz;
    }
  }
  if (MODE == 0) {   // the block's LJ energy, reduced in a fixed order
    for (int o = 32; o >= 1; o >>= 1) u += __shfl_xor(u, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = u;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
      for (int w = 0; w < NP_TB / 64; w++) t += red[w];
      upart[blockIdx.x] = t;
    }
  }
}

// RSpaceNonPeriodicPhiAtGridpoints (rspace.c:858-890): potential at each
// point from every site with a mask bit, screened within rscreen_i
__global__ void __launch_bounds__(NP_TB) k_nonperiodic_points(const double4 *__restrict__ site,
                                                             const double *__restrict__ rscreen,
                                                             int n, const double *__restrict__ pts,
                                                             int npts, double *__restrict__ phi) {
  __shared__ double4 ss[NP_TB];
  __shared__ double sr[NP_TB];
  const int j = blockIdx.x * NP_TB + threadIdx.x;
  const bool on = j < npts;
  const double gx = on ? pts[3 * j] : 0, gy = on ? pts[3 * j + 1] : 0, gz = on ? pts[3 * j + 2] : 0;
  double acc = 0;
  for (int t0 = 0; t0 < n; t0 += NP_TB) {
    __syncthreads();
    if (t0 + (int)threadIdx.x < n) {
      ss[threadIdx.x] = site[t0 + threadIdx.x];
      sr[threadIdx.x] = rscreen[t0 + threadIdx.x];
    }
    __syncthreads();
    const int m = min(NP_TB, n - t0);
    if (!on) continue;
    for (int k = 0; k < m; k++) {
      const double4 s = ss[k];   // w = 0 for sites without any mask bit
      const double x = s.x - gx, y = s.y - gy, z = s.z - gz;
      double b0, b1;
      np_coulomb(x * x + y * y + z * z, sr[k], b0, b1);
      acc += s.w * b0;
    }
  }
  if (on) phi[j] += acc;
}

int launch_nonperiodic(int mode, const double4 *site, const int4 *info, const double2 *lj,
                       const double *rscreen, int n, int geometric, double *out, double *upart,
                       hipStream_t st) {
  if (n <= 0) return 0;
  const dim3 grid((n + NP_TB - 1) / NP_TB);
  if (mode == 0)
    hipLaunchKernelGGL(k_nonperiodic<0>, grid, dim3(NP_TB), 0, st, site, info, lj, rscreen, n,
                       geometric, out, upart);
  else
    hipLaunchKernelGGL(k_nonperiodic<1>, grid, dim3(NP_TB), 0, st, site, info, lj, rscreen, n,
                       geometric, out, upart);
  MM_HIP(hipGetLastError());
  return 0;
}

int launch_nonperiodic_points(const double4 *site, const double *rscreen, int n,
                              const double *pts, int npts, double *phi, hipStream_t st) {
  if (n <= 0 || npts <= 0) return 0;
  hipLaunchKernelGGL(k_nonperiodic_points, dim3((npts + NP_TB - 1) / NP_TB), dim3(NP_TB), 0, st,
                     site, rscreen, n, pts, npts, phi);
  MM_HIP(hipGetLastError());
  return 0;
}

int pair_tiles(int ngroup) { return (ngroup + GPB - 1) / GPB; }

int launch_pair(Engine &e, int field_mode, int want_virial, int energy) {
  if (e.ntile == 0) return 0;
  dim3 grid(e.pair_blocks), tb(PAIR_TB);
#define PAIR_ARGS                                                                             \
  e.d_sp, e.d_tile_start, e.nsys, e.ntile, e.d_gpos, e.d_spk, e.d_gmeta, e.d_nbr, e.d_nnb,  \
      e.nbcap, e.d_tab, e.d_ljt, e.d_pf, e.d_plj, e.d_gacc
  // WATER: the kernel carries the path for systems made only of three-site
  // waters
  // ENERGY false (forces only) for steps whose energies nobody reads
#define PAIR_LAUNCH(F, V, EN)                                                                \
  if (e.any_water)                                                                           \
    hipLaunchKernelGGL((k_pair<F, V, true, EN>), grid, tb, 0, e.stream, PAIR_ARGS);          \
  else                                                                                       \
    hipLaunchKernelGGL((k_pair<F, V, false, EN>), grid, tb, 0, e.stream, PAIR_ARGS);
  if (field_mode) {
    PAIR_LAUNCH(true, false, true)
  } else if (want_virial) {
    PAIR_LAUNCH(false, true, true)
  } else if (energy) {
    PAIR_LAUNCH(false, false, true)
  } else {
    PAIR_LAUNCH(false, false, false)
  }
#undef PAIR_LAUNCH
#undef PAIR_ARGS
  MM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mmamd
