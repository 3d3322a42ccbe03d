// noalloc.hip -- kspace_field (kspace.c:87-174) without any device
// allocation, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).
//
// Everything the k-space kernels need beyond the caller's inputs and
// outputs lives in one file-scope device array (3.9 MB per device, held
// until the process ends); its device address is looked up once per device
// with that device current and used for every kernel argument.
#include "engine.hpp"
#include "../../include/molmodel_amd_ref.h"

#include <cstring>
#include <mutex>
#include <string>

namespace mmamd {

// kspace_field without allocation.  Everything the k_ks_* kernels need
// beyond the caller's inputs and outputs (system parameters, wave vectors
// and their coefficients, structure factors, SoA positions, per-site
// results, the energy slot) lives in one file-scope device array; the
// caller passes device pointers and a stream.  Each device holds its own
// copy of the array until the process ends; every call on a device uses
// that device's copy, so calls that overlap on one device race.
constexpr int NA_MAX_SITES = 65536;      // sites per call
constexpr int NA_MAX_K = 4096;           // wave vectors (|n| <= kc, half space): kc <= 12
constexpr long NA_SP = 0;                                 // SysParams (<= 1 KB)
constexpr long NA_KVEC = NA_SP + 128;                     // double4 per wave vector
constexpr long NA_KSC = NA_KVEC + 4L * NA_MAX_K;          // S_k, C_k
constexpr long NA_ACC = NA_KSC + 2L * NA_MAX_K;           // NACC energy slots
constexpr long NA_R = NA_ACC + 32;                        // x[N] y[N] z[N]
constexpr long NA_KRES = NA_R + 3L * NA_MAX_SITES;        // phi, E per site (+ 9 virial)
constexpr long NA_TOTAL = NA_KRES + 4L * NA_MAX_SITES + 32;   // 3.9 MB
static_assert(sizeof(SysParams) <= 128 * sizeof(double), "SysParams slot too small");
static_assert(NACC <= 32, "energy slot too small");

static __device__ __align__(256) double scratch[NA_TOTAL];

// the reciprocal lattice (tensor_t layout) as a kernel argument
struct Lat9 { double v[9]; };

// System parameters, then the half-space wave vectors with |n|^2 <= kc^2
// (kspace.c:41-57) in lexicographic (nx, ny, nz) order, compacted in one
// block chunk by chunk (so the order is fixed), with
// a_k = (4 pi / V) exp(-k^2 / 4 eta^2) / k^2 (kspace.c:96-103).
__global__ void __launch_bounds__(1024) k_na_setup(SysParams P, Lat9 L, double kc, double *sc) {
  __shared__ int wsum[16];
  __shared__ int base;
  if (threadIdx.x == 0) {
    *reinterpret_cast<SysParams *>(sc + NA_SP) = P;
    base = 0;
  }
  const int nm = (int)floor(kc);
  const double nsqmax = kc * kc + 1e-8;
  const int w = 2 * nm + 1;
  const long ncand = (long)(nm + 1) * w * w;
  const double c = 0.25 / (P.eta * P.eta), fourpi_V = 4 * M_PI / P.volume;
  double4 *kv = reinterpret_cast<double4 *>(sc + NA_KVEC);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  for (long c0 = 0; c0 < ncand; c0 += 1024) {
    const long cand = c0 + threadIdx.x;
    int nx = 0, ny = 0, nz = 0;
    bool ok = false;
    if (cand < ncand) {
      nx = (int)(cand / ((long)w * w));
      ny = (int)((cand / w) % w) - nm;
      nz = (int)(cand % w) - nm;
      const bool half = nx > 0 || (nx == 0 && (ny > 0 || (ny == 0 && nz > 0)));
      ok = half && (double)(nx * nx + ny * ny + nz * nz) <= nsqmax;
    }
    const unsigned long long bal = __ballot(ok);
    if (lane == 0) wsum[wv] = __popcll(bal);
    __syncthreads();
    int before = base;
    for (int k = 0; k < wv; k++) before += wsum[k];
    const int pos = before + __popcll(bal & ((1ull << lane) - 1ull));
    if (ok && pos < P.nk) {   // the host counted nk the same way; never exceeded
      const double kx = L.v[0] * nx + L.v[3] * ny + L.v[6] * nz;
      const double ky = L.v[1] * nx + L.v[4] * ny + L.v[7] * nz;
      const double kz = L.v[2] * nx + L.v[5] * ny + L.v[8] * nz;
      const double t = kx * kx + ky * ky + kz * kz;
      kv[pos] = make_double4(kx, ky, kz, fourpi_V * exp(-t * c) / t);
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int k = 0; k < 16; k++) base += wsum[k];
    __syncthreads();
  }
}

// caller's AoS positions into the SoA copy the kernels read
__global__ void k_na_soa(const double *__restrict__ r, int n, double *__restrict__ sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double *x = sc + NA_R;
  x[i] = r[3 * i];
  x[n + i] = r[3 * i + 1];
  x[2 * n + i] = r[3 * i + 2];
}

// phi += , evec += (kspace_field's accumulate semantics)
__global__ void k_na_out(const double *__restrict__ sc, int n, double *__restrict__ phi,
                         double *__restrict__ evec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *k = sc + NA_KRES + 4L * i;
  phi[i] += k[0];
  if (evec) {
    evec[3 * i] += k[1];
    evec[3 * i + 1] += k[2];
    evec[3 * i + 2] += k[3];
  }
}

namespace {
constexpr int NA_MAX_DEV = 64;
std::mutex na_mu;
double *na_ptr[NA_MAX_DEV] = {};

// the device address of this device's copy of the array, looked up once
// per device with the device current
int na_scratch(double **out) {
  int dev = 0;
  MM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEV) {
    set_last_error("kspace_field_noalloc: device index out of range");
    return -1;
  }
  std::lock_guard<std::mutex> g(na_mu);
  if (!na_ptr[dev]) {
    void *p = nullptr;
    MM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    na_ptr[dev] = static_cast<double *>(p);
  }
  *out = na_ptr[dev];
  return 0;
}

int na_count_k(double kc) {   // kspace.c:41-57
  const double nsqmax = kc * kc + 1e-8;
  const int nm = (int)floor(kc);
  int nk = 0;
  for (int nx = 0; nx <= nm; nx++)
    for (int ny = (nx == 0) ? 0 : -nm; ny <= nm; ny++)
      for (int nz = (nx == 0 && ny == 0) ? 1 : -nm; nz <= nm; nz++)
        if (nx * nx + ny * ny + nz * nz <= nsqmax) nk++;
  return nk;
}
}  // namespace

}  // namespace mmamd

extern "C" int kspace_field_noalloc(double cutoff, int nsite, const cart_t *r, const double *q,
                                    const tensor_t *latv, double volume, double eta,
                                    double *phi, cart_t *evec, void *stream) {
  using namespace mmamd;
  if (nsite < 0 || nsite > NA_MAX_SITES) {
    set_last_error("kspace_field_noalloc: nsite must be 0.." + std::to_string(NA_MAX_SITES));
    return -1;
  }
  if (!(cutoff >= 1) || cutoff > 64 || !(eta > 0) || !(volume > 0) || !latv) {
    set_last_error("kspace_field_noalloc: bad cutoff, eta, volume or lattice");
    return -1;
  }
  const int nk = na_count_k(cutoff);
  if (nk > NA_MAX_K) {
    set_last_error("kspace_field_noalloc: more than " + std::to_string(NA_MAX_K) +
                   " wave vectors (cutoff above 12)");
    return -1;
  }
  if (nsite == 0) return 0;
  if (!r || !q || !phi) {
    set_last_error("kspace_field_noalloc: null positions, charges or potential");
    return -1;
  }
  double *sc = nullptr;
  if (int rc = na_scratch(&sc)) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  SysParams P;
  memset(&P, 0, sizeof P);
  P.kmode = 0;
  P.nk = nk;
  P.k0 = 0;
  P.nsite = nsite;
  P.site0 = 0;
  P.eta = eta;
  P.volume = volume;
  Lat9 L;
  memcpy(L.v, latv, sizeof L.v);
  hipLaunchKernelGGL(k_na_setup, dim3(1), dim3(1024), 0, st, P, L, cutoff, sc);
  hipLaunchKernelGGL(k_na_soa, dim3((nsite + 255) / 256), dim3(256), 0, st, (const double *)r,
                     nsite, sc);
  const SysParams *dsp = reinterpret_cast<const SysParams *>(sc + NA_SP);
  const double4 *kv = reinterpret_cast<const double4 *>(sc + NA_KVEC);
  if (int rc = launch_kspace_raw(dsp, kv, nk, nsite, sc + NA_R, q, sc + NA_KSC, sc + NA_KRES,
                                 sc + NA_ACC, st))
    return rc;
  hipLaunchKernelGGL(k_na_out, dim3((nsite + 255) / 256), dim3(256), 0, st, sc, nsite, phi,
                     (double *)evec);
  MM_HIP(hipGetLastError());
  return 0;
}
