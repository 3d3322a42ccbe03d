// mm_common.hpp -- shared definitions of the MI355X nonbonded engine.
//
// Layout conventions (one engine may hold a batch of independent systems):
//   * sites and groups of all systems are concatenated; group g owns sites
//     [gfirst[g], gfirst[g] + gns[g]) with the heavy atom first, as
//     ElecLJ::init_sites builds them (eleclj.cpp:653-677);
//   * positions / velocities / forces are FP64 structure-of-arrays in the
//     caller's site order;
//   * after each neighbor-list rebuild the groups are re-ordered by cell
//     ("sorted" index gs); the pair kernel reads packed FP64 first-site
//     positions and FP32 site offsets in that order;
//   * SysParams holds every per-system constant; kernels look it up through
//     a per-group system index.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mmamd {

constexpr int MAXS = 4;          // max sites per group handled on the device
constexpr int SITE_HAS_Q = 1;    // rspace.h:7
constexpr int SITE_HAS_LJ = 2;   // rspace.h:8
// mask bytes of a three-site water (charges on all sites, LJ on the first
// only): Q|LJ, Q, Q
constexpr int WATER_MASKS = (SITE_HAS_Q | SITE_HAS_LJ) | (SITE_HAS_Q << 8) | (SITE_HAS_Q << 16);
constexpr int MAXORDER = 7;      // pppm.c:61

// error flags (bitwise OR into Engine::d_err)
enum : int {
  ERR_NL_OVERFLOW = 1,       // a group has more neighbors than the list holds
  ERR_TABLE_RANGE = 2,       // a site pair fell outside the erfc table (exact path used)
  ERR_CONSTRAINT = 4,        // SHAKE did not converge
  ERR_NAN = 8,               // non-finite energy (or force, in inner MD steps)
  ERR_BIN_OVERFLOW = 16,     // P3M column bin overflow
  ERR_CELL = 32,             // a neighbour-list cell holds more groups than its sort takes
};

struct SysParams {
  // cubic box (pbc.h:16-23 uses -L and 1/L)
  double L, invL, volume;
  double rhi2, rlo2, inv_w2;   // rc^2, (rc-w)^2, 1/(rhi2-rlo2)
  double nl_c2;                // (rc + skin)^2
  double skin;
  double eta;
  double self_coef;            // eta * 2/sqrt(pi)
  double lc;                   // neighbor-list cell length L*(1.0/ncell)
  int ncell, cell0;            // cells [cell0, cell0 + ncell^3)
  int site0, nsite, group0, ngroup;
  int sten0, nsten;            // stencil offsets [sten0, sten0+nsten)
  int tab0;                    // erfc table offset (float4 entries)
  int ntabA, ntabB;            // intervals below / above r^2 = tabMid
  float tabAmin, tabMid, tabAinv, tabBinv, tabBmax;
  int ntype, lj0;              // LJ type-pair table offset (float2 entries)
  int kmode;                   // 0 Ewald, 1 P3M
  int nk, k0;                  // Ewald k-vectors
  // P3M
  int nx, ny, nz, order;
  long mesh0, cmesh0;          // real mesh / complex half-spectrum offsets
  long green0;                 // influence function offset (layout [ky][kz][kx])
  double T[9];                 // diag(n) L^-1, tensor_t layout (pppm.c:523-525)
  double RL[9];                // reciprocal lattice 2 pi (L^-1)^T, tensor_t layout
  int bx, by, nbx, nby, bin0;  // spread column blocks (bx x by mesh columns each)
  long halo0;                  // the blocks' fixed-point halo partials: block k at
                               // halo0 + k (bx+order-1)(by+order-1) nz
  int twid0;                   // first of this system's 3 x 128 FFT twiddles
  int water;                   // every group is a three-site water (mask bytes WATER_MASKS)
  int ik;                      // P3M field by -i k on the mesh: four half spectra at
                               // cmesh0 + c nx ny (nz/2+1), c = 0..3 (phi, Ex, Ey, Ez), and
                               // 4 nx ny nz floats of real mesh from mesh0 (the density in
                               // the first nx ny nz, the transformed field interleaved as
                               // (phi, Ex, Ey, Ez) per point); mesh0 is a multiple of 4
  // dynamics
  double dt;
  double bg_phi;               // neutralising background potential (or 0)
};

// tensor_t layout helpers: T[0]=xx T[1]=yx T[2]=zx T[3]=xy T[4]=yy T[5]=zy T[6]=xz T[7]=yz T[8]=zz

__host__ __device__ inline int mod_nonneg(int n, int m) {   // fns.h:48-58
  if (n >= 0) return n % m;
  int k = -n % m;
  return k > 0 ? m - k : 0;
}

// A spreading block accumulates the sites of its bx x by mesh columns into
// the (bx + order - 1) x (by + order - 1) columns their stencils reach, all
// nz points of each, as 64-bit fixed point in LDS: at most this many points.
constexpr int SPREAD_MAXHALO = 8192;

// Mesh columns per spreading block along x and y: the smallest factor >= want
// (at most 16) of each mesh size, else the smallest factor >= 2, falling back
// to single columns where the block's points (bx by nz, 64-bit each in LDS)
// would exceed maxpts, its halo (for an order > 0) SPREAD_MAXHALO, or the
// block count maxbins.
inline bool choose_spread_blocks(int nx, int ny, int nz, int maxpts, int maxbins, int &bx,
                                 int &by, int want = 2, int order = 0) {
  auto factor = [want](int n) {
    for (int b = want; b <= 16; b++)
      if (n % b == 0) return b;
    for (int b = 2; b <= 7; b++)
      if (n % b == 0) return b;
    return 1;
  };
  const int h = order > 0 ? order - 1 : 0;
  auto fits = [&]() {
    return bx * by * nz <= maxpts && (order <= 0 || (bx + h) * (by + h) * nz <= SPREAD_MAXHALO);
  };
  bx = factor(nx);
  by = factor(ny);
  if (!fits()) bx = 1;
  if (!fits()) by = 1;
  return fits() && (nx / bx) * (ny / by) <= maxbins;
}

// slots per P3M column bin: four times the mean occupancy plus a margin,
// rounded to 32 (an overfull bin raises ERR_BIN_OVERFLOW)
inline int p3m_bin_capacity(int nsite, int nbins) {
  const int mean = (nsite + nbins - 1) / nbins;
  return (4 * mean + 64 + 31) / 32 * 32;
}

__host__ __device__ inline void canonical_pair(int &i, int &j) {   // pbin.c:6
  if (((i + j) % 2) ^ (i < j)) { int t = i; i = j; j = t; }
}

}  // namespace mmamd
