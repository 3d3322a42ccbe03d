// pppm.hip -- P3M on the device (pppm.c:498-641).
//
// Steps, each its own kernel on one stream:
//   bin      s = diag(n) L^-1 r in FP64, nearest/lower mesh point and the
//            fractional offset as the reference forms them (pppm.c:523-548),
//            and a copy of each site in the slots of the mesh column-block
//            bin that holds its base point (a per-block LDS rank per bin and
//            one global claim of slots per bin the block touches);
//   spread   one workgroup per column block (bx x by columns), as 64-bit
//            fixed point in LDS (exact, so independent of the order of the
//            atomics), no global atomics, one of two ways (a shape choice):
//            pulling -- the block takes the sites of its own and the
//            neighbouring bins and accumulates q w_x w_y w_z into the
//            points it owns, then writes them once; or by owner -- the
//            block takes only its own bin's sites, accumulates their whole
//            stencils into a halo of (bx + P - 1) x (by + P - 1) columns
//            and writes the halo to its own slice, which the z pass sums;
//            pulling re-reads every site for each block its stencil
//            reaches, the halos cost their traffic, so columns holding many
//            sites favour pulling and few sites the halos;
//   fft      Stockham mixed-radix (4,2,3,5,7) line FFTs in LDS: z (real to
//            half spectrum), y, then x forward, multiply by the influence
//            function (and form the mesh energy/virial), x inverse, y
//            inverse, z (half spectrum to real), unnormalised like FFTW;
//   gather   phi_i = sum w phi_m and, with real-space differentiation (the
//            reference default), E_i = (diag(n) L^-1)^T (sum -dw phi_m)
//            (pppm.c:603-640); with the field formed on the mesh (ik), the
//            x pass also forms E_a = -i k_a phi_k, the inverse passes run
//            over four components and E_i = sum w E_m (pppm.c:577-626).
// Assignment weights are the cardinal B-splines that the reference's
// polynomial tables tabulate (pppm.c:69-275), evaluated by recursion in FP32.
#include "engine.hpp"
#include "scan.hpp"
#include "reduce.hpp"

namespace mmamd {

constexpr float FIX_SCALE = 4294967296.0f;   // 2^32
constexpr double FIX_INV = 1.0 / 4294967296.0;

template <int P>
__device__ __forceinline__ void bspline_w(float s, float *w, float *ndw) {
  // w[k] = M_P(u + P - 1 - k), u = s + 1/2; ndw[k] = -dw/ds (cardinal
  // B-spline recursion, all loops unrolled for the compile-time order)
  const float u = s + 0.5f;
  float m[P + 1];
#pragma unroll
  for (int k = 0; k <= P; k++) m[k] = 0.f;
  m[0] = 1.f;
#pragma unroll
  for (int p = 2; p <= P; p++) {
    if (p == P && ndw) {
#pragma unroll
      for (int k = 0; k < P; k++) {
        const float a = k < p - 1 ? m[k] : 0.f;
        const float b = k >= 1 ? m[k - 1] : 0.f;
        ndw[P - 1 - k] = b - a;
      }
    }
    constexpr float dummy = 0.f;
    (void)dummy;
    const float inv = 1.f / (float)(p - 1);
#pragma unroll
    for (int k = p - 1; k >= 0; k--) {
      const float x = u + (float)k;
      const float a = k < p - 1 ? m[k] : 0.f;
      const float b = k >= 1 ? m[k - 1] : 0.f;
      m[k] = (x * a + ((float)p - x) * b) * inv;
    }
  }
  if (P == 1 && ndw) ndw[0] = 0.f;
#pragma unroll
  for (int k = 0; k < P; k++) w[P - 1 - k] = m[k];
}

// wrap j + off into [0, n) for |off| < n (no integer division)
__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Grid coordinates of each site (pppm.c:540-557) and its copy in the slots
// of its mesh column bin (bx x by columns).  A block ranks its sites per bin
// in LDS and claims each bin's slots with one global atomic; the order
// inside a bin varies from run to run, and the spreading's 64-bit
// fixed-point sums do not depend on it.  The gather clears the counters.
// A site beyond a bin's capacity raises ERR_BIN_OVERFLOW and records in
// *need the slot count that bin would have needed, so the host can grow the
// bins and evaluate again.
constexpr int MAXBINS = 4096;
__global__ void __launch_bounds__(256) k_p3m_bin(const SysParams *__restrict__ sp,
                                                 const double *__restrict__ r,
                                                 const double *__restrict__ q, int N,
                                                 int4 *__restrict__ sgrid,
                                                 float4 *__restrict__ sfrac,
                                                 int *__restrict__ fill, int bcap,
                                                 int4 *__restrict__ bgrid,
                                                 float4 *__restrict__ bfrac,
                                                 int *__restrict__ err, int *__restrict__ need) {
  __shared__ int cnt[MAXBINS];
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  if (P.kmode != 1) return;
  if ((int)(blockIdx.x * blockDim.x) >= P.nsite) return;
  const int nb = P.nbx * P.nby;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) cnt[t] = 0;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = i < P.nsite;
  const int si = P.site0 + (on ? i : 0);
  const double x = r[si], y = r[N + si], z = r[2 * N + si];
  const double *T = P.T;
  double sx = T[0] * x + T[3] * y + T[6] * z;
  double sy = T[1] * x + T[4] * y + T[7] * z;
  double sz = T[2] * x + T[5] * y + T[8] * z;
  int jx, jy, jz;
  if (P.order & 1) {
    jx = (int)floor(sx + 0.5); jy = (int)floor(sy + 0.5); jz = (int)floor(sz + 0.5);
    sx -= jx; sy -= jy; sz -= jz;
  } else {
    jx = (int)floor(sx); jy = (int)floor(sy); jz = (int)floor(sz);
    sx -= jx + 0.5; sy -= jy + 0.5; sz -= jz + 0.5;
  }
  jx = mod_nonneg(jx, P.nx); jy = mod_nonneg(jy, P.ny); jz = mod_nonneg(jz, P.nz);
  const int lb = (jx / P.bx) * P.nby + jy / P.by;
  const int b = P.bin0 + lb;
  const float4 f = make_float4((float)sx, (float)sy, (float)sz, (float)q[si]);
  int rank = 0;
  if (on) {
    sgrid[si] = make_int4(jx, jy, jz, b);
    sfrac[si] = f;
    rank = atomicAdd(&cnt[lb], 1);
  }
  __syncthreads();
  // one claim per bin this block touches; the base replaces the count
  for (int t = threadIdx.x; t < nb; t += blockDim.x)
    if (cnt[t]) cnt[t] = atomicAdd(&fill[P.bin0 + t], cnt[t]);
  __syncthreads();
  if (!on) return;
  const int slot = cnt[lb] + rank;
  if (slot < bcap) {
    bgrid[(long)b * bcap + slot] = make_int4(jx, jy, jz, si);
    bfrac[(long)b * bcap + slot] = f;
  } else {
    atomicOr(err, ERR_BIN_OVERFLOW);   // reported; the site is left out of the mesh
    atomicMax(need, slot + 1);
  }
}

__device__ __forceinline__ int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// LDS accumulator: bx * by columns of nz points (the host keeps the product
// within SPREAD_MAXPTS, see choose_spread_blocks)
constexpr int SPREAD_MAXPTS = 4096;

// ADJ: the transpose of the gather's field instead of the charge: site i
// contributes src_i . (dw_x w_y w_z, w_x dw_y w_z, w_x w_y dw_z) with dw the
// negated weight derivatives the gather uses and src_i = q_i (R_i T) (the
// adjoint of E_i = T^T sum_m dw phi_m for a given R_i), so that the
// convolution and an ordinary gather of the result give
// sum_i q_i R_i . dE_i/dq_j at every site j.
template <int P, bool ADJ = false>
__global__ void __launch_bounds__(256) k_p3m_spread(const SysParams *__restrict__ sp,
                                                    const int *__restrict__ fill, int bcap,
                                                    const int4 *__restrict__ bgrid,
                                                    const float4 *__restrict__ bfrac,
                                                    float *__restrict__ mesh,
                                                    const float4 *__restrict__ src = nullptr) {
  __shared__ unsigned long long acc[SPREAD_MAXPTS];
  const int s = blockIdx.y;
  const SysParams &Q = sp[s];
  if (Q.kmode != 1 || Q.order != P) return;
  const int blk = blockIdx.x;
  if (blk >= Q.nbx * Q.nby) return;
  const int bx = Q.bx, by = Q.by, nx = Q.nx, ny = Q.ny, nz = Q.nz;
  const int bi = blk / Q.nby, bj = blk % Q.nby;
  const int X0 = bi * bx, Y0 = bj * by;
  const int npts = bx * by * nz;
  for (int t = threadIdx.x; t < npts; t += blockDim.x) acc[t] = 0ull;
  __syncthreads();
  constexpr int istart = -((P + 1) / 2 - 1);
  // neighbouring column blocks whose sites reach the owned columns (at most
  // 8 per axis for P <= 7), flattened into one list of sites so that all
  // threads work whatever the bins' sizes: thread t < nbins takes bin t,
  // wave 0 scans the counts
  __shared__ int sb_start[64], sb_off[65];
  int xlo = floor_div(X0 - istart - P + 1, bx), xhi = floor_div(X0 + bx - 1 - istart, bx);
  int ylo = floor_div(Y0 - istart - P + 1, by), yhi = floor_div(Y0 + by - 1 - istart, by);
  if (xhi - xlo + 1 >= Q.nbx) { xlo = 0; xhi = Q.nbx - 1; }
  if (yhi - ylo + 1 >= Q.nby) { ylo = 0; yhi = Q.nby - 1; }
  const int ncy = yhi - ylo + 1, nbin = (xhi - xlo + 1) * ncy;   // <= 64
  if (threadIdx.x < 64) {
    int cnt = 0;
    const int t = threadIdx.x;
    if (t < nbin) {
      const int ccx = mod_nonneg(xlo + t / ncy, Q.nbx), ccy = mod_nonneg(ylo + t % ncy, Q.nby);
      const int b = Q.bin0 + ccx * Q.nby + ccy;
      sb_start[t] = b * bcap;
      cnt = min(fill[b], bcap);
    }
    int incl = cnt;   // inclusive scan over the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(incl, o);
      if (t >= o) incl += y;
    }
    sb_off[t + 1] = incl;
    if (t == 0) sb_off[0] = 0;
  }
  __syncthreads();
  const int total = sb_off[nbin];
  for (int u = threadIdx.x; u < total; u += blockDim.x) {
    int lo = 0, hi = nbin - 1;   // last bin with sb_off[bin] <= u
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sb_off[mid] <= u) lo = mid; else hi = mid - 1;
    }
    const int t = sb_start[lo] + (u - sb_off[lo]);
    const int4 g = bgrid[t];
    const float4 f = bfrac[t];
    float wx[P], wy[P], wz[P], dwx[P], dwy[P], dwz[P];
    bspline_w<P>(f.x, wx, ADJ ? dwx : nullptr);
    bspline_w<P>(f.y, wy, ADJ ? dwy : nullptr);
    bspline_w<P>(f.z, wz, ADJ ? dwz : nullptr);
    int lx[P], ly[P], iz[P];
#pragma unroll
    for (int a = 0; a < P; a++) {
      lx[a] = wrap(g.x + istart + a, nx) - X0;
      ly[a] = wrap(g.y + istart + a, ny) - Y0;
      iz[a] = wrap(g.z + istart + a, nz);
    }
    float4 sv = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ADJ) sv = src[g.w];
#pragma unroll
    for (int a = 0; a < P; a++) {
      if (lx[a] < 0 || lx[a] >= bx) continue;
      const float qa = f.w * wx[a];
#pragma unroll
      for (int bb = 0; bb < P; bb++) {
        if (ly[bb] < 0 || ly[bb] >= by) continue;
        unsigned long long *col = acc + (lx[a] * by + ly[bb]) * nz;
        if (ADJ) {
          const float al = sv.x * dwx[a] * wy[bb] + sv.y * wx[a] * dwy[bb];
          const float be = sv.z * wx[a] * wy[bb];
#pragma unroll
          for (int c = 0; c < P; c++)
            atomicAdd(col + iz[c],
                      (unsigned long long)__float2ll_rn((al * wz[c] + be * dwz[c]) * FIX_SCALE));
        } else {
          const float qw = qa * wy[bb];
#pragma unroll
          for (int c = 0; c < P; c++)
            atomicAdd(col + iz[c], (unsigned long long)__float2ll_rn(qw * wz[c] * FIX_SCALE));
        }
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < npts; t += blockDim.x) {
    const int gz = t % nz, gy = (t / nz) % by, gx = t / (nz * by);
    mesh[Q.mesh0 + ((long)(X0 + gx) * ny + (Y0 + gy)) * nz + gz] =
        (float)((double)(long long)acc[t] * FIX_INV);
  }
}

// Spreading by owner: the block of column block k takes only the sites
// binned to it (base point in its bx x by columns), so every site is read
// and weighted once, and accumulates all P^3 weights of each into an LDS
// halo of (bx + P - 1) x (by + P - 1) columns of nz points (halo column
// (hx, hy) is mesh column (X0 + istart + hx, Y0 + istart + hy) mod n) as
// 64-bit fixed point, exact and so independent of the order of the
// atomics.  The halo goes out whole to the block's own slice of `halo`;
// the z pass adds the halos that cover each mesh column (integer sums, so
// the mesh equals that of a single accumulator).  One work item per (site,
// z offset c): the P lanes of a site write P consecutive points of a
// column, in different LDS banks.
template <int P>
__global__ void __launch_bounds__(256) k_p3m_spread_halo(const SysParams *__restrict__ sp,
                                                         const int *__restrict__ fill, int bcap,
                                                         const int4 *__restrict__ bgrid,
                                                         const float4 *__restrict__ bfrac,
                                                         unsigned long long *__restrict__ halo) {
  extern __shared__ unsigned long long hacc[];
  const int s = blockIdx.y;
  const SysParams &Q = sp[s];
  if (Q.kmode != 1 || Q.order != P) return;
  const int blk = blockIdx.x;
  if (blk >= Q.nbx * Q.nby) return;
  const int nz = Q.nz, HX = Q.bx + P - 1, HY = Q.by + P - 1;
  const int npts = HX * HY * nz;
  for (int t = threadIdx.x; t < npts; t += blockDim.x) hacc[t] = 0ull;
  __syncthreads();
  const int X0 = (blk / Q.nby) * Q.bx, Y0 = (blk % Q.nby) * Q.by;
  const int b = Q.bin0 + blk;
  const int n = min(fill[b], bcap);
  constexpr int istart = -((P + 1) / 2 - 1);
  for (int u = threadIdx.x; u < n * P; u += blockDim.x) {
    const int site = u / P, c = u - site * P;
    const int4 g = bgrid[(long)b * bcap + site];
    const float4 f = bfrac[(long)b * bcap + site];
    float wx[P], wy[P], wz[P];
    bspline_w<P>(f.x, wx, nullptr);
    bspline_w<P>(f.y, wy, nullptr);
    bspline_w<P>(f.z, wz, nullptr);
    float wc = wz[0];
#pragma unroll
    for (int k = 1; k < P; k++)
      if (c == k) wc = wz[k];
    const int iz = wrap(g.z + istart + c, nz);
    unsigned long long *base = hacc + ((long)(g.x - X0) * HY + (g.y - Y0)) * nz + iz;
#pragma unroll
    for (int a = 0; a < P; a++) {
      const float qa = f.w * wx[a];
#pragma unroll
      for (int bb = 0; bb < P; bb++) {
        const float qw = qa * wy[bb];
        atomicAdd(base + ((long)a * HY + bb) * nz,
                  (unsigned long long)__float2ll_rn(qw * wc * FIX_SCALE));
      }
    }
  }
  __syncthreads();
  unsigned long long *dst = halo + Q.halo0 + (long)blk * npts;
  for (int t = threadIdx.x; t < npts; t += blockDim.x) dst[t] = hacc[t];
}

// ---------------- line FFTs in LDS ----------------
constexpr int FFT_T = 16;                 // threads per line
constexpr int FFT_LINES = P3M_FFT_LINES;  // lines per block
constexpr int FFT_TB = FFT_T * FFT_LINES; // one wave
constexpr int FFT_MAXN = 128;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// twiddle exp(sign*2*pi*i*t/n) from tw[t] = exp(-2 pi i t/n)
__device__ __forceinline__ float2 twid(const float2 *tw, int t, int n, int sign) {
  const float2 w = tw[t % n];
  return sign < 0 ? w : make_float2(w.x, -w.y);
}

// one Stockham stage of radix R over a line of length n (stride p)
template <int R>
__device__ __forceinline__ void fft_stage(const float2 *xl, float2 *yl, int n, int p,
                                          const float2 *tw, int sign, int tid) {
  const int nb = n / R;
  const int tstep = n / (p * R);
  for (int i = tid; i < nb; i += FFT_T) {
    const int k = i % p;
    float2 u[R], v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
      u[r] = xl[i + r * nb];
      if (r > 0 && k > 0) u[r] = cmul(u[r], twid(tw, r * k * tstep, n, sign));
    }
    if (R == 2) {
      v[0] = cadd(u[0], u[1]); v[1] = csub(u[0], u[1]);
    } else if (R == 4) {
      const float2 a = cadd(u[0], u[2]), b = csub(u[0], u[2]);
      const float2 c = cadd(u[1], u[3]), d = csub(u[1], u[3]);
      const float2 di = sign < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
      v[0] = cadd(a, c); v[2] = csub(a, c);
      v[1] = cadd(b, di); v[3] = csub(b, di);
    } else {
      const int ws = n / R;
#pragma unroll
      for (int r = 0; r < R; r++) {
        float2 acc = u[0];
#pragma unroll
        for (int t = 1; t < R; t++) acc = cadd(acc, cmul(u[t], twid(tw, ((r * t) % R) * ws, n, sign)));
        v[r] = acc;
      }
    }
    const int j = (i - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) yl[j + r * p] = v[r];
  }
}

// Stockham autosort FFT of FFT_LINES lines of length n held in x (line l at
// x + l*FFT_MAXN); every thread of the block must call it.  Radices are
// taken greedily (4, 2, 3, 5, 7).  Returns the buffer holding the result.
// Always inlined: as a called function (the compiler's choice for the
// mixed-radix build) it took a call and 192-416 bytes of scratch per lane in
// every kernel that uses it, and the x pass with virial and the field on
// the mesh gave wrong results in that form; inlined, the mixed-radix
// kernels fit in registers (at most 32 bytes of spill).
template <bool POW2>
__device__ __forceinline__ float2 *lds_fft_lines(float2 *x, float2 *y, int n, const float2 *tw,
                                                 int sign) {
  const int l = threadIdx.x / FFT_T, tid = threadIdx.x % FFT_T;
  int p = 1, rem = n;
  while (rem > 1) {
    float2 *xl = x + l * FFT_MAXN, *yl = y + l * FFT_MAXN;
    int R;
    if (rem % 4 == 0) { R = 4; fft_stage<4>(xl, yl, n, p, tw, sign, tid); }
    else if (rem % 2 == 0) { R = 2; fft_stage<2>(xl, yl, n, p, tw, sign, tid); }
    else if (POW2) { R = rem; }   // not reached: the host launches POW2 only for 2^k lines
    else if (rem % 3 == 0) { R = 3; fft_stage<3>(xl, yl, n, p, tw, sign, tid); }
    else if (rem % 5 == 0) { R = 5; fft_stage<5>(xl, yl, n, p, tw, sign, tid); }
    else { R = 7; fft_stage<7>(xl, yl, n, p, tw, sign, tid); }
    __syncthreads();
    float2 *t = x; x = y; y = t;
    p *= R;
    rem /= R;
  }
  return x;
}

__device__ __forceinline__ void load_tw(float2 *stw, const float2 *tw, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) stw[t] = tw[t];
}

// The z line `line` (mesh column xi = line / ny, yi = line % ny) of a P3M
// system as the sum of the spreading blocks' halo columns that cover it:
// halo column (dx, dy) of block (bi, bj) is mesh column
// (bi bx + istart + dx, bj by + istart + dy) mod n, so the covering blocks
// have dx = xi - istart (mod bx), dy likewise; up to 4 x 4 of them are
// read at once (all loads in flight), more in further rounds.  Integer
// sums, then the conversion a single accumulator would make.
__device__ __forceinline__ void halo_line(const SysParams &P,
                                          const unsigned long long *__restrict__ halo, int line,
                                          int tid, float2 *xl) {
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const int xi = line / ny, yi = line % ny;
  const int istart = -((P.order + 1) / 2 - 1);
  const int bxs = P.bx, bys = P.by, HX = bxs + P.order - 1, HY = bys + P.order - 1;
  const int hpts = HX * HY * nz;
  const unsigned long long *h = halo + P.halo0;
  constexpr int NK = FFT_MAXN / FFT_T;
  unsigned long long acc[NK];
#pragma unroll
  for (int k = 0; k < NK; k++) acc[k] = 0ull;
  const int dx0 = mod_nonneg(xi - istart, bxs), dy0 = mod_nonneg(yi - istart, bys);
  for (int dxb = dx0; dxb < HX; dxb += 4 * bxs)
    for (int dyb = dy0; dyb < HY; dyb += 4 * bys) {
      int off[16];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int dx = dxb + i * bxs;
        const int bi = mod_nonneg(xi - istart - dx, nx) / bxs;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int dy = dyb + j * bys;
          const int bj = mod_nonneg(yi - istart - dy, ny) / bys;
          off[4 * i + j] =
              (dx < HX && dy < HY) ? (bi * P.nby + bj) * hpts + (dx * HY + dy) * nz : -1;
        }
      }
#pragma unroll
      for (int k = 0; k < NK; k++) {
        const int t = tid + k * FFT_T;
        if (t < nz) {
          unsigned long long v = 0ull;
#pragma unroll
          for (int m = 0; m < 16; m++)
            if (off[m] >= 0) v += h[off[m] + t];
          acc[k] += v;
        }
      }
    }
#pragma unroll
  for (int k = 0; k < NK; k++) {
    const int t = tid + k * FFT_T;
    if (t < nz) xl[t] = make_float2((float)((double)(long long)acc[k] * FIX_INV), 0.f);
  }
}

// pass 0: z lines, real -> half spectrum; HALO: the line is the sum of the
// spreading blocks' halo columns that cover it (halo_line), else the real
// mesh.  pass 1/3: y lines fwd/inv.  pass 4: z lines, half
// spectrum -> real.  blockIdx.z: the mesh component (phi, then Ex, Ey, Ez
// of a system whose field is formed as -i k phi on the mesh; the forward
// passes run over component 0 only).
template <bool POW2, bool HALO = false>
__global__ void __launch_bounds__(FFT_TB, 4) k_fft_pass(const SysParams *__restrict__ sp,
                                                     const float2 *__restrict__ twid_all,
                                                     float *__restrict__ mesh,
                                                     float2 *__restrict__ cmesh, int pass,
                                                     const unsigned long long *__restrict__ halo) {
  __shared__ float2 bx_[FFT_LINES * FFT_MAXN], by_[FFT_LINES * FFT_MAXN];
  __shared__ float2 stw[FFT_MAXN];
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  if (P.kmode != 1) return;
  const int comp = blockIdx.z;
  if (comp > 0 && !P.ik) return;
  const int nx = P.nx, ny = P.ny, nz = P.nz, nzh = nz / 2 + 1;
  const long mesh0 = P.mesh0 + comp * (long)nx * ny * nz;
  const long cmesh0 = P.cmesh0 + comp * (long)nx * ny * nzh;
  int nlines, n, twoff;
  if (pass == 0 || pass == 4) { nlines = nx * ny; n = nz; twoff = 2 * FFT_MAXN; }
  else { nlines = nx * nzh; n = ny; twoff = FFT_MAXN; }
  if ((int)(blockIdx.x * FFT_LINES) >= nlines) return;
  load_tw(stw, twid_all + P.twid0 + twoff, n);
  const int l = threadIdx.x / FFT_T, tid = threadIdx.x % FFT_T;
  const int line = blockIdx.x * FFT_LINES + l;
  const bool on = line < nlines;
  float2 *xl = bx_ + l * FFT_MAXN;
  const int sign = (pass == 3 || pass == 4) ? 1 : -1;
  // load
  if (pass == 0) {
    if constexpr (HALO) {
      if (on) halo_line(P, halo, line, tid, xl);
    } else if (on) {
      const float *src = mesh + mesh0 + (long)line * nz;
      for (int t = tid; t < n; t += FFT_T) xl[t] = make_float2(src[t], 0.f);
    }
  } else if (pass == 4) {
    if (on) {
      const float2 *src = cmesh + cmesh0 + (long)line * nzh;
      for (int t = tid; t < nzh; t += FFT_T) xl[t] = src[t];
    }
    __syncthreads();
    if (on)
      for (int t = nzh + tid; t < n; t += FFT_T) {
        const float2 c = xl[n - t];
        xl[t] = make_float2(c.x, -c.y);
      }
  } else {
    if (on) {
      const int xi = line / nzh, kz = line % nzh;
      const float2 *src = cmesh + cmesh0 + (long)xi * ny * nzh + kz;
      for (int t = tid; t < n; t += FFT_T) xl[t] = src[(long)t * nzh];
    }
  }
  __syncthreads();
  float2 *res = lds_fft_lines<POW2>(bx_, by_, n, stw, sign);
  float2 *rl = res + l * FFT_MAXN;
  if (!on) return;
  if (pass == 0) {
    float2 *dst = cmesh + cmesh0 + (long)line * nzh;
    for (int t = tid; t < nzh; t += FFT_T) dst[t] = rl[t];
  } else if (pass == 4) {
    if (P.ik) {   // (phi, Ex, Ey, Ez) interleaved per point: the gather reads one float4
      float *dst = mesh + P.mesh0 + 4 * ((long)line * nz) + comp;
      for (int t = tid; t < n; t += FFT_T) dst[4 * t] = rl[t].x;
    } else {
      float *dst = mesh + mesh0 + (long)line * nz;
      for (int t = tid; t < n; t += FFT_T) dst[t] = rl[t].x;
    }
  } else {
    const int xi = line / nzh, kz = line % nzh;
    float2 *dst = cmesh + cmesh0 + (long)xi * ny * nzh + kz;
    for (int t = tid; t < n; t += FFT_T) dst[(long)t * nzh] = rl[t];
  }
}

// x lines: forward FFT, multiply by the influence function, inverse FFT.
// Also the mesh energy 1/2 sum G |rho_k|^2 and its virial
// W_ab = sum e_k (delta_ab - 2 k_a k_b (1/k^2 + 1/(4 eta^2))); ENER false:
// the convolution alone (inner MD steps, whose energies are not formed).
// IK: systems whose field is formed on the mesh also get E_a = -i k_a phi_k
// (k from the signed x, y index and the half-spectrum z index,
// pppm.c:577-594), each inverse-transformed along x into its own component.
template <bool VIR, bool POW2, bool ENER = true, bool IK = false>
__global__ void __launch_bounds__(FFT_TB, 4) k_fft_xgreen(const SysParams *__restrict__ sp,
                                                       const float2 *__restrict__ twid_all,
                                                       const float *__restrict__ green,
                                                       float2 *__restrict__ cmesh,
                                                       double *__restrict__ part,
                                                       int *__restrict__ done,
                                                       double *__restrict__ acc,
                                                       double *__restrict__ kvir, int want_virial,
                                                       int maxblk) {
  __shared__ float2 bx_[FFT_LINES * FFT_MAXN], by_[FFT_LINES * FFT_MAXN];
  __shared__ float2 stw[FFT_MAXN];
  constexpr int NV = VIR ? 10 : 1;   // energy [, virial]
  __shared__ double red[FFT_TB / 64][NV];
  __shared__ int last;
  (void)acc;
  const int s = blockIdx.y;
  const SysParams &P = sp[s];
  if (P.kmode != 1) return;
  const int nx = P.nx, ny = P.ny, nz = P.nz, nzh = nz / 2 + 1;
  const int nlines = ny * nzh;
  const int nblk = (nlines + FFT_LINES - 1) / FFT_LINES;
  if ((int)blockIdx.x >= nblk) return;
  load_tw(stw, twid_all + P.twid0, nx);
  const int l = threadIdx.x / FFT_T, tid = threadIdx.x % FFT_T;
  const int line = blockIdx.x * FFT_LINES + l;
  const bool on = line < nlines;
  const int ky = on ? line / nzh : 0, kz = on ? line % nzh : 0;
  float2 *xl = bx_ + l * FFT_MAXN;
  float2 *src = cmesh + P.cmesh0 + (long)ky * nzh + kz;
  const long xstride = (long)ny * nzh;
  if (on)
    for (int t = tid; t < nx; t += FFT_T) xl[t] = src[t * xstride];
  __syncthreads();
  float2 *res = lds_fft_lines<POW2>(bx_, by_, nx, stw, -1);
  float2 *oth = (res == bx_) ? by_ : bx_;
  double e = 0, w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (on) {
    float2 *rl = res + l * FFT_MAXN;
    const float *g = green + P.green0 + ((long)ky * nzh + kz) * nx;
    const double mult = (kz == 0 || (nz % 2 == 0 && kz == nz / 2)) ? 1.0 : 2.0;
    const int jy = ky > ny / 2 ? ky - ny : ky;
    const double c4 = 0.25 / (P.eta * P.eta);
    for (int t = tid; t < nx; t += FFT_T) {
      const float2 v = rl[t];
      const float gg = g[t];
      const double ek = ENER ? 0.5 * mult * (double)gg * ((double)v.x * v.x + (double)v.y * v.y) : 0.0;
      e += ek;
      if (VIR && gg != 0.f) {
        const int jx = t > nx / 2 ? t - nx : t;
        double k[3];
        for (int a = 0; a < 3; a++) k[a] = P.RL[a] * jx + P.RL[3 + a] * jy + P.RL[6 + a] * kz;
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const double f = 2 * (1 / k2 + c4);
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) w[3 * a + b] += ek * ((a == b ? 1.0 : 0.0) - f * k[a] * k[b]);
      }
      rl[t] = make_float2(v.x * gg, v.y * gg);
    }
  }
  if constexpr (IK) {
    // each thread keeps its own points of phi_k in registers and rebuilds
    // the line in LDS for every component
    constexpr int NK = FFT_MAXN / FFT_T;
    float2 keep[NK];
    float2 *rl = res + l * FFT_MAXN;
#pragma unroll
    for (int j = 0; j < NK; j++) {
      const int t = tid + j * FFT_T;
      keep[j] = (on && t < nx) ? rl[t] : make_float2(0.f, 0.f);
    }
    const int jy = ky > ny / 2 ? ky - ny : ky;
    const int ncomp = P.ik ? 4 : 1;
    const long cstride = (long)nx * ny * nzh;
    for (int c = 0; c < ncomp; c++) {
      if (c > 0) {
        const int a = c - 1;
#pragma unroll
        for (int j = 0; j < NK; j++) {
          const int t = tid + j * FFT_T;
          if (on && t < nx) {
            const int jx = t > nx / 2 ? t - nx : t;
            const float k = (float)(P.RL[a] * jx + P.RL[3 + a] * jy + P.RL[6 + a] * kz);
            rl[t] = make_float2(k * keep[j].y, -k * keep[j].x);
          }
        }
      }
      __syncthreads();
      float2 *out = lds_fft_lines<POW2>(res, oth, nx, stw, +1);
      if (on) {
        const float2 *ol = out + l * FFT_MAXN;
        for (int t = tid; t < nx; t += FFT_T) src[c * cstride + t * xstride] = ol[t];
      }
      __syncthreads();   // the next component rebuilds its line in these buffers
    }
  } else {
    __syncthreads();
    res = lds_fft_lines<POW2>(res, oth, nx, stw, +1);
    if (on) {
      float2 *rl = res + l * FFT_MAXN;
      for (int t = tid; t < nx; t += FFT_T) src[t * xstride] = rl[t];
    }
  }
  if constexpr (!ENER) return;
  // deterministic block partials, reduced by the last block
  double vals[NV];
  vals[0] = e;
  if constexpr (VIR)
    for (int c = 0; c < 9; c++) vals[1 + c] = w[c];
  if (finish_partials<NV, NV, FFT_TB>(vals, part + (long)s * maxblk * NV, nblk, blockIdx.x,
                                      &done[s], red, &last)) {
    acc[s * NACC + ACC_KSPACE_E] = vals[0];
    for (int c = 0; c < 9; c++) kvir[9 * s + c] = VIR ? vals[1 + c] : 0.0;
  }
}

// IK: systems whose field was formed on the mesh; phi and E_a are
// interpolated from their four meshes with the assignment weights
// (pppm.c:618-626), E already Cartesian
template <int P, bool IK = false>
__global__ void __launch_bounds__(256) k_p3m_gather(const SysParams *__restrict__ sp,
                                                    const int4 *__restrict__ sgrid,
                                                    const float4 *__restrict__ sfrac,
                                                    const float *__restrict__ mesh,
                                                    double *__restrict__ kres,
                                                    int *__restrict__ fill) {
  const int s = blockIdx.y;
  const SysParams &Q = sp[s];
  if (Q.kmode != 1 || Q.order != P || Q.ik != (int)IK) return;
  if (blockIdx.x == 0)   // the spreading has read the bin counters: clear them
    for (int t = threadIdx.x; t < Q.nbx * Q.nby; t += blockDim.x) fill[Q.bin0 + t] = 0;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Q.nsite) return;
  const int si = Q.site0 + i;
  const int4 g = sgrid[si];
  const float4 f = sfrac[si];
  const int nx = Q.nx, ny = Q.ny, nz = Q.nz;
  constexpr int istart = -((P + 1) / 2 - 1);
  float wx[P], wy[P], wz[P], dx[P], dy[P], dz[P];
  bspline_w<P>(f.x, wx, IK ? nullptr : dx);
  bspline_w<P>(f.y, wy, IK ? nullptr : dy);
  bspline_w<P>(f.z, wz, IK ? nullptr : dz);
  int ix[P], iy[P], iz[P];
#pragma unroll
  for (int a = 0; a < P; a++) {
    ix[a] = wrap(g.x + istart + a, nx);
    iy[a] = wrap(g.y + istart + a, ny);
    iz[a] = wrap(g.z + istart + a, nz);
  }
  const float *m0 = mesh + Q.mesh0;
  float ph = 0.f, e0 = 0.f, e1 = 0.f, e2 = 0.f;
  if constexpr (IK) {
    // the last inverse pass leaves (phi, Ex, Ey, Ez) interleaved per point
    const float4 *m4 = reinterpret_cast<const float4 *>(m0);
#pragma unroll
    for (int a = 0; a < P; a++) {
#pragma unroll
      for (int b = 0; b < P; b++) {
        const float4 *col = m4 + ((long)ix[a] * ny + iy[b]) * nz;
        const float wxy = wx[a] * wy[b];
#pragma unroll
        for (int c = 0; c < P; c++) {
          const float w = wxy * wz[c];
          const float4 v = col[iz[c]];
          ph += w * v.x;
          e0 += w * v.y;
          e1 += w * v.z;
          e2 += w * v.w;
        }
      }
    }
    kres[4 * si] = (double)ph;
    kres[4 * si + 1] = (double)e0;
    kres[4 * si + 2] = (double)e1;
    kres[4 * si + 3] = (double)e2;
    return;
  }
#pragma unroll
  for (int a = 0; a < P; a++) {
#pragma unroll
    for (int b = 0; b < P; b++) {
      const float *col = m0 + ((long)ix[a] * ny + iy[b]) * nz;
      float cp = 0.f, cz = 0.f;
#pragma unroll
      for (int c = 0; c < P; c++) {
        const float pm = col[iz[c]];
        cp += wz[c] * pm;
        cz += dz[c] * pm;
      }
      const float wxy = wx[a] * wy[b];
      ph += wxy * cp;
      e0 += dx[a] * wy[b] * cp;
      e1 += wx[a] * dy[b] * cp;
      e2 += wxy * cz;
    }
  }
  const double *T = Q.T;
  kres[4 * si] = (double)ph;
  kres[4 * si + 1] = (double)e0 * T[0] + (double)e1 * T[1] + (double)e2 * T[2];
  kres[4 * si + 2] = (double)e0 * T[3] + (double)e1 * T[4] + (double)e2 * T[5];
  kres[4 * si + 3] = (double)e0 * T[6] + (double)e1 * T[7] + (double)e2 * T[8];
}

// forward z, y passes; x pass with the influence function; inverse y, z.
// sp / nsg: the systems the grid's y extent covers (all of them, or one
// system's parameters with nsg = 1); ik: some system forms its field on the
// mesh (four components through the inverse passes)
template <bool POW2, bool IK>
static void launch_xgreen(Engine &e, int want_virial, int energies, hipStream_t st, dim3 g,
                          const SysParams *sp) {
  if (want_virial)
    hipLaunchKernelGGL((k_fft_xgreen<true, POW2, true, IK>), g, dim3(FFT_TB), 0, st, sp,
                       e.d_twid, e.d_green, e.d_cmesh, e.d_part, e.d_done + e.nsys, e.d_acc,
                       e.d_kres + 4L * e.nsite, want_virial, e.nblk_part);
  else if (energies)
    hipLaunchKernelGGL((k_fft_xgreen<false, POW2, true, IK>), g, dim3(FFT_TB), 0, st, sp,
                       e.d_twid, e.d_green, e.d_cmesh, e.d_part, e.d_done + e.nsys, e.d_acc,
                       e.d_kres + 4L * e.nsite, want_virial, e.nblk_part);
  else
    hipLaunchKernelGGL((k_fft_xgreen<false, POW2, false, IK>), g, dim3(FFT_TB), 0, st, sp,
                       e.d_twid, e.d_green, e.d_cmesh, e.d_part, e.d_done + e.nsys, e.d_acc,
                       e.d_kres + 4L * e.nsite, want_virial, e.nblk_part);
}

// halo: the spreading blocks' partials the z pass sums (null: the real
// mesh holds the charges)
template <bool POW2>
static void launch_fft(Engine &e, int want_virial, int energies, hipStream_t st, int maxx,
                       int maxy, int maxz, const SysParams *sp, int nsg, bool ik,
                       const unsigned long long *halo) {
  if (halo)
    hipLaunchKernelGGL((k_fft_pass<POW2, true>), dim3(maxz, nsg), dim3(FFT_TB), 0, st, sp,
                       e.d_twid, e.d_mesh, e.d_cmesh, 0, halo);
  else
    hipLaunchKernelGGL((k_fft_pass<POW2>), dim3(maxz, nsg), dim3(FFT_TB), 0, st, sp,
                       e.d_twid, e.d_mesh, e.d_cmesh, 0, nullptr);
  hipLaunchKernelGGL((k_fft_pass<POW2>), dim3(maxy, nsg), dim3(FFT_TB), 0, st, sp,
                     e.d_twid, e.d_mesh, e.d_cmesh, 1, nullptr);
  if (ik) launch_xgreen<POW2, true>(e, want_virial, energies, st, dim3(maxx, nsg), sp);
  else launch_xgreen<POW2, false>(e, want_virial, energies, st, dim3(maxx, nsg), sp);
  const int ncomp = ik ? 4 : 1;
  hipLaunchKernelGGL((k_fft_pass<POW2>), dim3(maxy, nsg, ncomp), dim3(FFT_TB), 0, st, sp,
                     e.d_twid, e.d_mesh, e.d_cmesh, 3, nullptr);
  hipLaunchKernelGGL((k_fft_pass<POW2>), dim3(maxz, nsg, ncomp), dim3(FFT_TB), 0, st, sp,
                     e.d_twid, e.d_mesh, e.d_cmesh, 4, nullptr);
}

int launch_p3m(Engine &e, int want_virial, hipStream_t st, int energies) {
  int maxn = 0, maxbins = 0, maxz = 0, maxy = 0, maxx = 0;
  for (auto &p : e.hsp)
    if (p.kmode == 1) {
      maxn = std::max(maxn, p.nsite);
      maxbins = std::max(maxbins, p.nbx * p.nby);
      const int nzh = p.nz / 2 + 1;
      maxz = std::max(maxz, (p.nx * p.ny + FFT_LINES - 1) / FFT_LINES);
      maxy = std::max(maxy, (p.nx * nzh + FFT_LINES - 1) / FFT_LINES);
      maxx = std::max(maxx, (p.ny * nzh + FFT_LINES - 1) / FFT_LINES);
    }
  if (maxn == 0) return 0;
  const int tb = 256;
  dim3 gs((maxn + tb - 1) / tb, e.nsys);
  hipLaunchKernelGGL(k_p3m_bin, gs, dim3(tb), 0, st, e.d_sp, e.d_r, e.d_q, e.nsite,
                     e.d_sgrid, e.d_sfrac, e.d_bin_fill, e.bcap, e.d_bgrid, e.d_bfrac, e.d_err,
                     e.d_bin_fill + e.nbin_total);
  bool used[MAXORDER + 1] = {false}, used_ad[MAXORDER + 1] = {false},
       used_ik[MAXORDER + 1] = {false};
  bool ik = false;
  for (auto &p : e.hsp)
    if (p.kmode == 1) {
      used[p.order] = true;
      (p.ik ? used_ik : used_ad)[p.order] = true;
      ik |= p.ik != 0;
    }
  // LDS of the spreading: the largest halo of the systems of each order
  size_t hbytes[MAXORDER + 1] = {0};
  for (auto &p : e.hsp)
    if (p.kmode == 1)
      hbytes[p.order] = std::max(hbytes[p.order], sizeof(unsigned long long) *
                                                      (p.bx + p.order - 1) * (p.by + p.order - 1) * p.nz);
#define SPREAD(PO)                                                                           \
  if (used[PO] && e.spread_halo)                                                             \
    hipLaunchKernelGGL(k_p3m_spread_halo<PO>, dim3(maxbins, e.nsys), dim3(256), hbytes[PO],   \
                       st, e.d_sp, e.d_bin_fill, e.bcap, e.d_bgrid, e.d_bfrac, e.d_halo);     \
  else if (used[PO])                                                                         \
    hipLaunchKernelGGL(k_p3m_spread<PO>, dim3(maxbins, e.nsys), dim3(256), 0, st, e.d_sp,    \
                       e.d_bin_fill, e.bcap, e.d_bgrid, e.d_bfrac, e.d_mesh);
  SPREAD(1) SPREAD(2) SPREAD(3) SPREAD(4) SPREAD(5) SPREAD(6) SPREAD(7)
#undef SPREAD
  const unsigned long long *halo = e.spread_halo ? e.d_halo : nullptr;
  // meshes whose sides are all powers of two take FFT kernels built with
  // the radix-4/2 stages only (fewer registers)
  bool pow2 = true;
  for (auto &p : e.hsp)
    if (p.kmode == 1)
      for (int n : {p.nx, p.ny, p.nz}) pow2 &= (n & (n - 1)) == 0;
  if (pow2)
    launch_fft<true>(e, want_virial, energies, st, maxx, maxy, maxz, e.d_sp, e.nsys, ik, halo);
  else
    launch_fft<false>(e, want_virial, energies, st, maxx, maxy, maxz, e.d_sp, e.nsys, ik, halo);
#define GATHER(PO)                                                                        \
  if (used_ad[PO])                                                                        \
    hipLaunchKernelGGL(k_p3m_gather<PO>, gs, dim3(tb), 0, st, e.d_sp, e.d_sgrid,     \
                       e.d_sfrac, e.d_mesh, e.d_kres, e.d_bin_fill);                      \
  if (used_ik[PO])                                                                        \
    hipLaunchKernelGGL((k_p3m_gather<PO, true>), gs, dim3(tb), 0, st, e.d_sp, e.d_sgrid, \
                       e.d_sfrac, e.d_mesh, e.d_kres, e.d_bin_fill);
  GATHER(1) GATHER(2) GATHER(3) GATHER(4) GATHER(5) GATHER(6) GATHER(7)
#undef GATHER
  MM_HIP(hipGetLastError());
  return 0;
}

// The mesh adjoint of one P3M system: bin its sites again (the last gather
// cleared the counters; positions are those of the last evaluation), spread
// the adjoint source src (per site, global site index), convolve with the
// influence function (no energy), and gather the potential of the result
// into kres[4 i] of the system's sites (kres[4 i + 1..3] receive its
// field, unused).
int launch_p3m_adjoint(Engine &e, int sys, const float4 *src, hipStream_t st) {
  const SysParams &P = e.hsp[sys];
  if (P.kmode != 1) return 0;
  const SysParams *sp = e.d_sp + sys;
  const int tb = 256;
  const dim3 gs((P.nsite + tb - 1) / tb, 1);
  hipLaunchKernelGGL(k_p3m_bin, gs, dim3(tb), 0, st, sp, e.d_r, e.d_q, e.nsite, e.d_sgrid,
                     e.d_sfrac, e.d_bin_fill, e.bcap, e.d_bgrid, e.d_bfrac, e.d_err,
                     e.d_bin_fill + e.nbin_total);
  const dim3 gb(P.nbx * P.nby, 1);
#define ADJ_ORDER(PO)                                                                        \
  case PO:                                                                                   \
    hipLaunchKernelGGL((k_p3m_spread<PO, true>), gb, dim3(256), 0, st, sp, e.d_bin_fill,     \
                       e.bcap, e.d_bgrid, e.d_bfrac, e.d_mesh, src);                         \
    break;
  switch (P.order) {
    ADJ_ORDER(1) ADJ_ORDER(2) ADJ_ORDER(3) ADJ_ORDER(4) ADJ_ORDER(5) ADJ_ORDER(6) ADJ_ORDER(7)
    default: set_last_error("assignment order out of range"); return -1;
  }
#undef ADJ_ORDER
  const int nzh = P.nz / 2 + 1;
  const int maxz = (P.nx * P.ny + FFT_LINES - 1) / FFT_LINES;
  const int maxy = (P.nx * nzh + FFT_LINES - 1) / FFT_LINES;
  const int maxx = (P.ny * nzh + FFT_LINES - 1) / FFT_LINES;
  bool pow2 = true;
  for (int n : {P.nx, P.ny, P.nz}) pow2 &= (n & (n - 1)) == 0;
  // the engine chose one FFT build for all its meshes; keep to it
  for (auto &p : e.hsp)
    if (p.kmode == 1)
      for (int n : {p.nx, p.ny, p.nz}) pow2 &= (n & (n - 1)) == 0;
  if (pow2) launch_fft<true>(e, 0, 0, st, maxx, maxy, maxz, sp, 1, false, nullptr);
  else launch_fft<false>(e, 0, 0, st, maxx, maxy, maxz, sp, 1, false, nullptr);
#define ADJ_GATHER(PO)                                                                      \
  case PO:                                                                                  \
    hipLaunchKernelGGL(k_p3m_gather<PO>, gs, dim3(tb), 0, st, sp, e.d_sgrid, e.d_sfrac,     \
                       e.d_mesh, e.d_kres, e.d_bin_fill);                                   \
    break;
  switch (P.order) {
    ADJ_GATHER(1) ADJ_GATHER(2) ADJ_GATHER(3) ADJ_GATHER(4) ADJ_GATHER(5) ADJ_GATHER(6)
    ADJ_GATHER(7)
    default: break;
  }
#undef ADJ_GATHER
  MM_HIP(hipGetLastError());
  return 0;
}

// One forward ik pass of one system with the charges q (global site index)
// in place of its own: bin, spread by pulling, the ik convolution (no
// energy) and the ik gather, which leaves (phi, E) of the pass in kres[4 i]
// of the system's sites.  The ik field operator is a real circulant whose
// multiplier -i k_a G(k) is odd, so its transpose is its negative; the
// adjoint of the ik force therefore needs nothing but forward passes.
int launch_p3m_ik_pass(Engine &e, int sys, const double *q, hipStream_t st) {
  const SysParams &P = e.hsp[sys];
  if (P.kmode != 1 || !P.ik) return 0;
  const SysParams *sp = e.d_sp + sys;
  const int tb = 256;
  const dim3 gs((P.nsite + tb - 1) / tb, 1);
  hipLaunchKernelGGL(k_p3m_bin, gs, dim3(tb), 0, st, sp, e.d_r, q, e.nsite, e.d_sgrid,
                     e.d_sfrac, e.d_bin_fill, e.bcap, e.d_bgrid, e.d_bfrac, e.d_err,
                     e.d_bin_fill + e.nbin_total);
  const dim3 gb(P.nbx * P.nby, 1);
#define IK_SPREAD(PO)                                                                        \
  case PO:                                                                                   \
    hipLaunchKernelGGL((k_p3m_spread<PO, false>), gb, dim3(256), 0, st, sp, e.d_bin_fill,    \
                       e.bcap, e.d_bgrid, e.d_bfrac, e.d_mesh, nullptr);                     \
    break;
  switch (P.order) {
    IK_SPREAD(1) IK_SPREAD(2) IK_SPREAD(3) IK_SPREAD(4) IK_SPREAD(5) IK_SPREAD(6) IK_SPREAD(7)
    default: set_last_error("assignment order out of range"); return -1;
  }
#undef IK_SPREAD
  const int nzh = P.nz / 2 + 1;
  const int maxz = (P.nx * P.ny + FFT_LINES - 1) / FFT_LINES;
  const int maxy = (P.nx * nzh + FFT_LINES - 1) / FFT_LINES;
  const int maxx = (P.ny * nzh + FFT_LINES - 1) / FFT_LINES;
  bool pow2 = true;
  for (auto &p : e.hsp)   // the engine's one FFT build, as in the forward
    if (p.kmode == 1)
      for (int n : {p.nx, p.ny, p.nz}) pow2 &= (n & (n - 1)) == 0;
  if (pow2) launch_fft<true>(e, 0, 0, st, maxx, maxy, maxz, sp, 1, true, nullptr);
  else launch_fft<false>(e, 0, 0, st, maxx, maxy, maxz, sp, 1, true, nullptr);
#define IK_GATHER(PO)                                                                       \
  case PO:                                                                                  \
    hipLaunchKernelGGL((k_p3m_gather<PO, true>), gs, dim3(tb), 0, st, sp, e.d_sgrid,        \
                       e.d_sfrac, e.d_mesh, e.d_kres, e.d_bin_fill);                        \
    break;
  switch (P.order) {
    IK_GATHER(1) IK_GATHER(2) IK_GATHER(3) IK_GATHER(4) IK_GATHER(5) IK_GATHER(6) IK_GATHER(7)
    default: break;
  }
#undef IK_GATHER
  MM_HIP(hipGetLastError());
  return 0;
}

}  // namespace mmamd
