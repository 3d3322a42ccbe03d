// remd.hip -- replica exchange over the systems of one engine.
//
// Reference semantics (mdyn.cpp:317-351): neighbouring replicas in
// temperature compare
//   du = -U1/kT0 - U0/kT1 + U0/kT0 + U1/kT1
// and swap POSITIONS (velocities stay with the temperature) when du < 0 or
// a uniform random number is below exp(-du).  The MPI chain resolves the
// pairs from the hottest end down, each replica using the energy it
// computed at its last force evaluation, so one configuration can climb or
// fall more than one rung per attempt; the permutation built here from the
// top down reproduces that.  The constant-temperature move (mdyn.cpp:476-497,
// invoked at mdyn.cpp:163-167) redraws Maxwell-Boltzmann velocities at the
// system's kT, projects them onto the constraints and removes the total
// momentum.  Everything runs on the device, so a run of exchange intervals
// needs no host synchronisation.
//
// Differences from the reference, by design: with mm_remd_run, attempts and
// velocity redraws happen at fixed intervals instead of with probability
// 1/interval per step (do_every, mdyn.cpp:49-52; mm_remd_run_stochastic
// keeps those draws); the redraw follows the exchange, so swapped
// rigid molecules get velocities constrained for the configuration they
// now hold; forces are recomputed after a swap (the
// reference half-kicks once with the forces of the configuration it held
// before the swap); random numbers come from a counter-based generator
// (splitmix64 of seed, attempt, index).
#include "engine.hpp"
#include "reduce.hpp"
#include "water.hpp"

namespace mmamd {

__device__ __forceinline__ unsigned long long splitmix(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform in [0, 1) with 53 random bits
__device__ __forceinline__ double uniform01(unsigned long long seed, unsigned long long stream,
                                            unsigned long long idx) {
  const unsigned long long k = splitmix(seed * 0x100000001B3ull + stream);
  return (double)(splitmix(idx ^ k) >> 11) * (1.0 / 9007199254740992.0);
}

// One thread: every ladder, hottest pair first.  perm[s] = system whose
// positions s receives; cfg[s] tracks which starting configuration each
// system holds.  Swapped systems are flagged for a neighbour-list rebuild.
__global__ void k_remd_exchange(const int *__restrict__ ladder, int nladder,
                                const double *__restrict__ kT, const double *__restrict__ acc,
                                const double *__restrict__ tail, unsigned long long seed,
                                long long attempt, int *__restrict__ perm,
                                long long *__restrict__ stats, int *__restrict__ rebuild, int nsys) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int *cfg = perm + nsys, *tmp = perm + 2 * nsys;
  for (int s = 0; s < nsys; s++) perm[s] = s;
  int start = 0;
  while (start < nladder) {
    int end = start;
    while (end < nladder && ladder[end] >= 0) end++;
    for (int k = end - 2; k >= start; k--) {
      const int a = ladder[k], b = ladder[k + 1];
      // energies as computed at the last evaluation of each replica
      const double u0 = acc[a * NACC + ACC_ELEC] + acc[a * NACC + ACC_LJ] + tail[a];
      const double u1 = acc[b * NACC + ACC_ELEC] + acc[b * NACC + ACC_LJ] + tail[b];
      const double kT0 = kT[a], kT1 = kT[b];
      const double du = -u1 / kT0 - u0 / kT1 + u0 / kT0 + u1 / kT1;
      stats[2 * a] += 1;
      if (du < 0 || uniform01(seed, (unsigned long long)attempt, (unsigned long long)a) < exp(-du)) {
        const int t = perm[a]; perm[a] = perm[b]; perm[b] = t;
        stats[2 * a + 1] += 1;
      }
    }
    start = end + 1;
  }
  for (int s = 0; s < nsys; s++) tmp[s] = cfg[perm[s]];
  for (int s = 0; s < nsys; s++) {
    cfg[s] = tmp[s];
    if (perm[s] != s) rebuild[s] = 1;
  }
}

// positions of system s <- positions that system perm[s] held (same size)
__global__ void k_remd_gather_pos(const SysParams *__restrict__ sp, const int *__restrict__ perm,
                                  const double *__restrict__ rold, double *__restrict__ r, int N) {
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nsite) return;
  const SysParams &Q = sp[perm[s]];
  const int dst = P.site0 + i, src = Q.site0 + i;
  r[dst] = rold[src];
  r[N + dst] = rold[N + src];
  r[2 * N + dst] = rold[2 * N + src];
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// Maxwell-Boltzmann velocities at kT (mdyn.cpp:476-490), velocity
// constraints for waters (mdyn.cpp:491-492), per-system momentum sums
__global__ void __launch_bounds__(256) k_remd_velocities(
    const SysParams *__restrict__ sp, const int *__restrict__ gfirst, const int *__restrict__ gns,
    const double *__restrict__ mass, const double *__restrict__ kT, const double *__restrict__ r,
    double *__restrict__ v, int N, unsigned long long seed, long long attempt, double bond,
    double angle_rad, double *__restrict__ part, int *__restrict__ done, int maxblk,
    double *__restrict__ mom) {
  __shared__ double red[4][4];
  __shared__ int islast;
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int nblk = (P.ngroup + 255) / 256;
  if ((int)blockIdx.x >= nblk) return;
  const int i = blockIdx.x * 256 + threadIdx.x;
  double vals[4] = {0, 0, 0, 0};
  if (i < P.ngroup) {
    const int g = P.group0 + i;
    const int s0 = gfirst[g], ns = gns[g];
    double x[9], vv[9], m[3];
    for (int k = 0; k < 3; k++) {
      m[k] = 1;
      for (int d = 0; d < 3; d++) { x[3 * k + d] = 0; vv[3 * k + d] = 0; }
      if (k < ns) {
        const int sk = s0 + k;
        m[k] = mass[sk];
        const double a = sqrt(kT[s] / m[k]);
        for (int d = 0; d < 3; d++) {
          x[3 * k + d] = r[d * N + sk];
          // Box-Muller from two counter-based uniforms
          const unsigned long long idx = 3ull * (unsigned long long)sk + d;
          const double u1 = uniform01(seed, 2ull * attempt + 1, idx);
          const double u2 = uniform01(seed ^ 0x5bd1e995ull, 2ull * attempt + 1, idx);
          vv[3 * k + d] = a * sqrt(-2.0 * log1p(-u1)) * cos(2 * M_PI * u2);
        }
      }
    }
    if (ns == 3) rattle_water(x, vv, m, bond, angle_rad);
    for (int k = 0; k < 3; k++)
      if (k < ns) {
        const int sk = s0 + k;
        for (int d = 0; d < 3; d++) {
          v[d * N + sk] = vv[3 * k + d];
          vals[d] += m[k] * vv[3 * k + d];
        }
        vals[3] += m[k];
      }
  }
  if (finish_partials<4, 4, 256>(vals, part + (long)s * maxblk * 4, nblk, blockIdx.x, &done[s],
                                 red, &islast))
    for (int c = 0; c < 4; c++) mom[4 * s + c] = vals[c];
}

// remove the total momentum (MSys::rezero_linear_momentum, msys.cpp:409)
__global__ void k_remd_zero_momentum(const SysParams *__restrict__ sp,
                                     const double *__restrict__ mom, double *__restrict__ v, int N) {
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.nsite) return;
  const int si = P.site0 + i;
  const double M = mom[4 * s + 3];
  for (int d = 0; d < 3; d++) v[d * N + si] -= mom[4 * s + d] / M;
}

int launch_remd_velocities(Engine &e, long long attempt) {
  const int tb = 256;
  hipLaunchKernelGGL(k_remd_velocities, dim3((e.maxg + 255) / 256, e.nsys), dim3(256), 0, e.stream,
                     e.d_sp, e.d_gfirst, e.d_gns, e.d_mass, e.d_kT, e.d_r, e.d_v, e.nsite,
                     e.remd_seed, attempt, e.bond, e.angle * (M_PI / 180.0), e.d_part,
                     e.d_done + 2 * e.nsys, e.nblk_part, e.d_mom);
  hipLaunchKernelGGL(k_remd_zero_momentum, dim3((e.maxs + tb - 1) / tb, e.nsys), dim3(tb), 0,
                     e.stream, e.d_sp, e.d_mom, e.d_v, e.nsite);
  MM_HIP(hipGetLastError());
  return 0;
}

int launch_remd_exchange(Engine &e, long long attempt) {
  const int tb = 256;
  hipLaunchKernelGGL(k_remd_exchange, dim3(1), dim3(64), 0, e.stream, e.d_ladder, e.nladder,
                     e.d_kT, e.d_acc, e.d_tail, e.remd_seed, attempt, e.d_perm, e.d_remd_stats,
                     e.d_rebuild, e.nsys);
  const long n3 = 3L * e.nsite;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((n3 + tb - 1) / tb)), dim3(tb), 0, e.stream, e.d_r,
                     e.d_rtmp, n3);
  hipLaunchKernelGGL(k_remd_gather_pos, dim3((e.maxs + tb - 1) / tb, e.nsys), dim3(tb), 0,
                     e.stream, e.d_sp, e.d_perm, e.d_rtmp, e.d_r, e.nsite);
  MM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mmamd
