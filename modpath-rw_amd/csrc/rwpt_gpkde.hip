// gpkde bandwidth optimisation (nOptimizationLoops > 0) on the device.
//
// GridProjectedKDE%ComputeDensityOptimization (GridProjectedKDE.F90:4396-5160)
// gives every active bin its own anisotropic Gaussian kernel and refines the
// kernels in a loop: support sizes from a sigma-kernel gather of the density
// (:4808-4858), curvature bandwidths (:5300-5400), the net roughness from
// second-derivative kernels scattered into curvature grids and gathered back
// with the sigma kernels (:1905-2289, :2295-2709), the optimal smoothing and
// shape (:5404-5724), a new density scatter (:4930-4955), and the
// convergence tests on the ALMISE proxy, density and smoothing changes
// (:4993-5038).
//
// The work that grows with bins x kernel footprint -- every scatter and
// gather -- runs here, one workgroup per active bin.  The kernels of
// KernelMultiGaussian.f90 are products of 1-D factors (the W kernel of
// :720-790 and the second-derivative kernels V of :794-1107), so a workgroup
// builds three 1-D factor arrays in LDS, applies the reflection fold of
// :303-390 to each (the fold of a product kernel is the product of the folded
// factors), and walks the footprint with the outer product.  The per-bin
// formulas between the passes run on the device too, one thread per active
// bin in FP64, so the loop keeps every per-bin array in HBM and the host
// only reads back the widest kernel and the tile-list length before each
// pass and three sums per loop.  Scatters sum per grid tile in a fixed bin
// order (no atomics), so the density is the same from run to run.
#include <algorithm>
#include <cmath>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "rwpt_internal.h"

namespace rwpt {
namespace {

constexpr double kSqrt2 = 1.4142135623730951;
constexpr double kSqrtPi = 1.7724538509055159;
constexpr int kRangeW = 3;   // defaultKernelRange
constexpr int kRangeSD = 4;  // defaultKernelSDRange
constexpr int kBlock = 256;

struct GkGrid {
  int nb[3];
  int active[3];
  double bin[3];
  int ndim;
};

// half-width of a kernel of bandwidth h in a dimension of bin size b:
// ceiling(matrixRange*bandwidth/binSize), KernelMultiGaussian.f90:553-559
__host__ __device__ inline int half_width(int range, double h, double b) {
  if (b == 0.0) return 0;
  return (int)ceil((double)range * h / b);
}

// erf difference of the W kernel factor at offset i >= 0
__device__ inline double w_factor(int i, double hd) {
  const double s = hd * kSqrt2;
  return erf((i + 0.5) / s) - erf((i - 0.5) / s);
}

// second-derivative factor at offset i >= 0 (without the constant)
__device__ inline double v_factor(int i, double hd) {
  const double a = i + 0.5, b = i - 0.5, d = 2.0 * (hd * hd);
  return a * exp(-(a * a) / d) - b * exp(-(b * b) / d);
}

// Reflection fold of one factor array k[0..L-1] (L = 2m+1) for a bin at
// 1-based index c in a dimension of n bins (KernelMultiGaussian.f90:303-390).
__device__ void fold(double *k, int m, int c, int n) {
  const int L = 2 * m + 1;
  if (c - m < 1) {
    const int b1 = 1 - c + m;  // entries 1..b1 fall below the grid
    for (int t = 1; t <= b1; ++t) {
      const int dst = b1 + t, src = b1 + 1 - t;
      if (dst <= L) k[dst - 1] += k[src - 1];
    }
    for (int t = 1; t <= b1 && t <= L; ++t) k[t - 1] = 0.0;
  }
  if (c + m > n) {
    const int b2 = L - (c + m - n) + 1;  // entries b2..L fall above the grid
    const int lenb = L - b2 + 1;
    // kernel(b2-lenb:b2-1) += kernel(L:b2:-1); the two ranges are disjoint
    for (int t = 0; t < lenb; ++t) {
      const int dst = b2 - lenb + t, src = L - t;
      if (dst >= 1) k[dst - 1] += k[src - 1];
    }
    for (int t = b2; t <= L; ++t) k[t - 1] = 0.0;
  }
}

// Factor arrays of one bin's kernel in LDS.
//  kind 0: W kernel with bandwidth h[d] (dimensions with h = 0 or bin = 0
//          are skipped: factor 1), normalised by the product of the full
//          sums, folded in the active dimensions;
//  kind 1: W kernel used as sigma kernel: every dimension with a bin size
//          carries the bandwidth (inactive ones contribute their centre
//          weight only, the span being one bin);
//  kind 2: second-derivative kernel along dimension `der` (V_x, V_y, V_z of
//          KernelMultiGaussian.f90:794-1107; raw erf differences across, the
//          first-difference of x*exp(-x^2/2h^2) along `der`).
struct Factors {
  int m[3];
  int lo[3], hi[3];  // 1-based grid span
  double scale;      // scalar factor of the product
};

// Per-dimension values are staged in LDS before any thread indexes them
// by its thread id: a dynamically indexed private array or by-value
// argument would live in scratch memory.
__device__ void build_factors(double *lds, int cap, const GkGrid &G, const int c[3],
                              const double h[3], int kind, int der, int range,
                              int collapse, Factors &F) {
  __shared__ double s_h[3], s_bin[3], sums[3], sumsq[3];
  __shared__ int s_c[3], s_nb[3], s_act[3], shm[3];
  const int tid = threadIdx.x;
  if (tid == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      s_h[d] = h[d];
      s_bin[d] = G.bin[d];
      s_c[d] = c[d];
      s_nb[d] = G.nb[d];
      s_act[d] = G.active[d];
    }
  }
  __syncthreads();
  if (tid < 3) {
    const int d = tid;
    double hd = 0.0;
    if (s_bin[d] != 0.0) hd = s_h[d] / s_bin[d];
    int m = 0;
    if ((kind == 1 || s_act[d] || kind == 2) && hd > 0.0 && hd < 1.0e9)
      m = half_width(range, s_h[d], s_bin[d]);
    if (kind == 0 && s_h[d] == 0.0) m = 0;
    if (2 * m + 1 > cap) m = -1;  // wider than the launch allows: bin skipped
    shm[d] = m;
  }
  __syncthreads();
  bool bad = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    F.m[d] = shm[d] < 0 ? 0 : shm[d];
    bad = bad || shm[d] < 0;
  }
  // fill
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int m = F.m[d], L = 2 * m + 1;
    const double hd = G.bin[d] != 0.0 ? h[d] / G.bin[d] : 0.0;
    const bool vd = kind == 2 && d == der;
    for (int i = tid; i < L; i += blockDim.x) {
      const int off = i - m < 0 ? m - i : i - m;
      double v = 1.0;
      if (hd > 0.0) v = vd ? v_factor(off, hd) : w_factor(off, hd);
      lds[d * cap + i] = v;
    }
  }
  __syncthreads();
  if (tid < 3) {
    const int d = tid, L = 2 * (shm[d] < 0 ? 0 : shm[d]) + 1;
    const double *a = lds + d * cap;
    double sm = 0.0, q = 0.0;
    for (int i = 0; i < L; ++i) {
      sm += a[i];
      q += a[i] * a[i];
    }
    sums[d] = sm;
    sumsq[d] = q;
  }
  __syncthreads();
  // a dimension collapsed to its centre entry keeps the normalisation over
  // its full width (the x-y transposed x-z / y-z database lookup)
  if (collapse >= 0 && collapse < 3) {
    if (tid == 0 && shm[collapse] > 0) lds[collapse * cap] = lds[collapse * cap + shm[collapse]];
    __syncthreads();
    if (tid == 0) shm[collapse] = 0;
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (d == collapse) F.m[d] = 0;
  }
  // scalar factor
  double sc = 1.0;
  if (kind == 0 || kind == 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (F.m[d] > 0 || (G.bin[d] != 0.0 && h[d] > 0.0)) sc /= sums[d];
  } else {
    double hdv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) hdv[d] = G.bin[d] != 0.0 ? h[d] / G.bin[d] : 0.0;
    const double hder = der == 0 ? hdv[0] : (der == 1 ? hdv[1] : hdv[2]);
    const double fnd = (double)G.ndim;
    const double C = -1.0 / (pow(2.0, fnd - 0.5) * kSqrtPi * pow(hder, 3.0));
    // a kernel whose squares sum to zero is zeroed (:871-875)
    sc = C * C * sumsq[0] * sumsq[1] * sumsq[2] <= 0.0 ? 0.0 : C;
    // the correction loop divides by sqrt(h/delta) in the dimensions before
    // the derivative dimension and then returns unconditionally (the
    // `if (...) matrix = 0; return;` line of :887, :993, :1097), so the
    // sqrt(3/normprod) factor and the later dimensions are never applied
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (d < der && hdv[d] > 0.0) sc = sc / sqrt(hdv[d]);
  }
  // SetupMatrix with a zero bandwidth gives a zero matrix (:431-434)
  if (h[0] == 0.0 && h[1] == 0.0 && h[2] == 0.0) sc = 0.0;
  if (bad) sc = 0.0;
  F.scale = sc;
  // fold (active dimensions) and spans
  if (tid < 3) {
    const int d = tid, m = shm[d] < 0 ? 0 : shm[d];
    if (s_act[d] && m > 0) fold(lds + d * cap, m, s_c[d], s_nb[d]);
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    F.lo[d] = max(c[d] - F.m[d], 1);
    F.hi[d] = min(c[d] + F.m[d], G.nb[d]);
  }
  __syncthreads();
}

// Scatters (density and curvature grids): grid += w(bin) * K(bin) over every
// active bin, output-stationary.  gk_factors builds each bin's three folded
// 1-D factor arrays once (one workgroup per bin) and stores them with the
// bin's span and amplitude; every grid tile then lists the bins whose span
// reaches it (sorted by tile, ascending bin within a tile) and one
// workgroup per tile sums those bins' contributions for its cells in
// registers and writes each cell once -- no atomics, and a summation order
// fixed by the bin order.
constexpr int kSpan = 9;  // per bin: lo[3], hi[3] (1-based, inclusive), base[3]

__global__ void gk_factors(const int32_t *__restrict__ bins, int nact,
                           const double *__restrict__ hv, int hstride,
                           const double *__restrict__ w, const GkGrid G, int kind,
                           int der, int range, int cap, const uint8_t *__restrict__ skip,
                           const int8_t *__restrict__ collapse, double *__restrict__ fac,
                           int32_t *__restrict__ span, double *__restrict__ amp) {
  extern __shared__ double lds[];
  const int n = blockIdx.x;
  if (n >= nact) return;
  int32_t *sp = span + (int64_t)n * kSpan;
  if (skip && skip[n]) {
    if (threadIdx.x == 0) {
      sp[0] = 1; sp[3] = 0;  // empty span
      amp[n] = 0.0;
    }
    return;
  }
  const int id = bins[n];
  const int c[3] = {id % G.nb[0] + 1, (id / G.nb[0]) % G.nb[1] + 1,
                    id / (G.nb[0] * G.nb[1]) + 1};
  double h[3];
  for (int d = 0; d < 3; ++d) h[d] = hv[(int64_t)n * hstride + (hstride == 1 ? 0 : d)];
  if (kind == 2)  // second-derivative kernels carry one bandwidth in active dims
    for (int d = 0; d < 3; ++d) h[d] = G.active[d] ? h[d] : 0.0;
  Factors F;
  build_factors(lds, cap, G, c, h, kind, der, range, collapse ? collapse[n] : -1, F);
  double *out = fac + (int64_t)n * 3 * cap;
  for (int d = 0; d < 3; ++d)
    for (int i = threadIdx.x; i <= 2 * F.m[d]; i += blockDim.x)
      out[d * cap + i] = lds[d * cap + i];
  if (threadIdx.x == 0) {
    const double a = w[n] * F.scale;
    amp[n] = a;
    for (int d = 0; d < 3; ++d) {
      sp[d] = F.lo[d];
      sp[3 + d] = F.hi[d];
      sp[6 + d] = c[d] - F.m[d];
    }
    if (a == 0.0) sp[3] = sp[0] - 1;  // adds nothing: no tiles
  }
}

struct Tiles {
  int td[3];  // tile extent per dimension
  int nt[3];  // tiles per dimension
};

__device__ __forceinline__ bool span_tiles(const int32_t *sp, const Tiles &T, int lo[3],
                                           int hi[3]) {
  for (int d = 0; d < 3; ++d) {
    if (sp[3 + d] < sp[d]) return false;
    lo[d] = (sp[d] - 1) / T.td[d];
    hi[d] = (sp[3 + d] - 1) / T.td[d];
  }
  return true;
}

__global__ void gk_tile_count(int nact, const int32_t *__restrict__ span, const Tiles T,
                              int32_t *__restrict__ count) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nact) return;
  int lo[3], hi[3];
  count[n] = span_tiles(span + (int64_t)n * kSpan, T, lo, hi)
                 ? (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
                 : 0;
}

__global__ void gk_tile_pairs(int nact, const int32_t *__restrict__ span, const Tiles T,
                              const int32_t *__restrict__ offs, int32_t *__restrict__ keys,
                              int32_t *__restrict__ vals) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nact) return;
  int lo[3], hi[3];
  if (!span_tiles(span + (int64_t)n * kSpan, T, lo, hi)) return;
  int o = offs[n];
  for (int k = lo[2]; k <= hi[2]; ++k)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int i = lo[0]; i <= hi[0]; ++i) {
        keys[o] = i + T.nt[0] * (j + T.nt[1] * k);
        vals[o] = n;
        ++o;
      }
}

__global__ void gk_fill_int(int64_t n, int32_t *__restrict__ a, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// [start, end) of every tile's run in the sorted pair list
__global__ void gk_tile_bounds(int npairs, const int32_t *__restrict__ keys,
                               int32_t *__restrict__ start, int32_t *__restrict__ end) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const int32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) start[k] = i;
  if (i == npairs - 1 || keys[i + 1] != k) end[k] = i + 1;
}

// One workgroup per tile, one cell per thread: the contributions of the
// tile's bins in ascending bin order, then one read-add-write of the cell.
// The bins are taken in chunks: each bin's amplitude and the parts of its
// three factors over the tile (zero outside the bin's span) are staged in
// LDS, so the inner loop is three LDS reads and three multiplies and adds.
constexpr int kChunk = 32;
constexpr int kTileSpan = 16;  // largest tile extent in one dimension

__global__ void __launch_bounds__(kBlock)
    gk_tile_accumulate(const Tiles T, const GkGrid G, const int32_t *__restrict__ start,
                       const int32_t *__restrict__ end, const int32_t *__restrict__ vals,
                       const int32_t *__restrict__ span, const double *__restrict__ amp,
                       const double *__restrict__ fac, int cap, double *__restrict__ grid) {
  __shared__ double s_f[kChunk][3][kTileSpan];
  __shared__ double s_a[kChunk];
  const int t = blockIdx.x;
  const int q0 = start[t], q1 = end[t];
  if (q1 <= q0) return;
  const int tx = t % T.nt[0], ty = (t / T.nt[0]) % T.nt[1], tz = t / (T.nt[0] * T.nt[1]);
  const int g0[3] = {tx * T.td[0] + 1, ty * T.td[1] + 1, tz * T.td[2] + 1};
  const int tid = threadIdx.x;
  const int li = tid % T.td[0], lj = (tid / T.td[0]) % T.td[1], lk = tid / (T.td[0] * T.td[1]);
  const int gi = g0[0] + li, gj = g0[1] + lj, gk = g0[2] + lk;
  // staging: the chunk's (bin, factor entry) pairs, bin-major, dealt to the
  // threads in turn
  const int tsum = T.td[0] + T.td[1] + T.td[2];
  const int b_start = tid / tsum, e_start = tid - (tid / tsum) * tsum;
  double acc = 0.0;
  for (int q = q0; q < q1; q += kChunk) {
    const int nb = min(kChunk, q1 - q);
    for (int b = b_start, e = e_start; b < nb;) {
      int o = e, d = 0;
      while (d < 2 && o >= T.td[d]) {
        o -= T.td[d];
        ++d;
      }
      const int n = vals[q + b];
      const int32_t *sp = span + (int64_t)n * kSpan;
      const int g = g0[d] + o;
      double f = 0.0;
      if (g >= sp[d] && g <= sp[3 + d]) f = fac[((int64_t)n * 3 + d) * cap + g - sp[6 + d]];
      s_f[b][d][o] = f;
      if (e == 0) s_a[b] = amp[n];
      e += kBlock;
      while (e >= tsum) {
        e -= tsum;
        ++b;
      }
    }
    __syncthreads();
    for (int b = 0; b < nb; ++b)
      acc += s_a[b] * (s_f[b][0][li] * s_f[b][1][lj] * s_f[b][2][lk]);
    __syncthreads();
  }
  if (gi > G.nb[0] || gj > G.nb[1] || gk > G.nb[2]) return;
  const int64_t g = (int64_t)(gi - 1) + (int64_t)G.nb[0] * ((gj - 1) + (int64_t)G.nb[1] * (gk - 1));
  grid[g] += acc;
}

// The roughness gathers read the curvature grids transformed once per cell:
// k_d = curv_d / binVol / delta_d^2 (3-D: squared, 2-D: as is), the same
// expressions the gather would otherwise form at every footprint cell.
__global__ void gk_curv_transform(int64_t ncells, int squared, double binVol, double cs0,
                                  double cs1, double cs2, double *__restrict__ g0,
                                  double *__restrict__ g1, double *__restrict__ g2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncells) return;
  double k0 = g0[i] / binVol, k1 = g1[i] / binVol;
  k0 = k0 / cs0;
  k1 = k1 / cs1;
  if (squared) {
    double k2 = g2[i] / binVol;
    k2 = k2 / cs2;
    g0[i] = k0 * k0;
    g1[i] = k1 * k1;
    g2[i] = k2 * k2;
  } else {
    g0[i] = k0;
    g1[i] = k1;
  }
}

// sigma-kernel gathers.  mode 0: sum density*K; mode 1 (3-D roughness):
// over the squared curvatures k_d^2 (gk_curv_transform) the sums of kx^2,
// ky^2, kz^2, kx^2 ky^2, kx^2 kz^2, ky^2 kz^2 (cross terms after squaring,
// GridProjectedKDE.F90:2464-2469); mode 2 (2-D): over k1, k2 the sums of
// k1^2, k2^2, k1 k2.  Threads walk the footprint x fastest with the row
// and layer carried incrementally (no division per cell).
__global__ void gk_gather(const int32_t *__restrict__ bins, int nact,
                         const double *__restrict__ hv, int hstride, int kind,
                         const GkGrid G, int mode, int cap, const double *__restrict__ g0,
                         const double *__restrict__ g1, const double *__restrict__ g2,
                         double *__restrict__ out) {
  extern __shared__ double lds[];
  __shared__ double red[6][kBlock];
  const int n = blockIdx.x;
  if (n >= nact) return;
  const int id = bins[n];
  const int c[3] = {id % G.nb[0] + 1, (id / G.nb[0]) % G.nb[1] + 1,
                    id / (G.nb[0] * G.nb[1]) + 1};
  double h[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) h[d] = hv[(int64_t)n * hstride + (hstride == 1 ? 0 : d)];
  Factors F;
  build_factors(lds, cap, G, c, h, kind, 0, kRangeW, -1, F);
  const double *ax = lds + F.lo[0] - c[0] + F.m[0];
  const double *ay = lds + cap + F.lo[1] - c[1] + F.m[1];
  const double *az = lds + 2 * cap + F.lo[2] - c[2] + F.m[2];
  const int sx = F.hi[0] - F.lo[0] + 1, sy = F.hi[1] - F.lo[1] + 1,
            sz = F.hi[2] - F.lo[2] + 1;
  const int64_t nx = G.nb[0], nxy = (int64_t)G.nb[0] * G.nb[1];
  const int64_t g00 = (int64_t)(F.lo[0] - 1) + nx * (F.lo[1] - 1) + nxy * (F.lo[2] - 1);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  // element t = i + sx (j + sy k); thread t0 = threadIdx.x, stride B
  const int B = blockDim.x;
  const int di = B % sx, dr = B / sx;
  int i = threadIdx.x % sx, j = threadIdx.x / sx, k = 0;
  while (j >= sy && k < sz) {
    j -= sy;
    ++k;
  }
  while (k < sz) {
    const int64_t g = g00 + i + nx * j + nxy * k;
    const double kv = ax[i] * ay[j] * az[k] * F.scale;
    if (mode == 0) {
      acc[0] += g0[g] * kv;
    } else if (mode == 1) {
      const double kx = g0[g], ky = g1[g], kz = g2[g];
      acc[0] += kx * kv;
      acc[1] += ky * kv;
      acc[2] += kz * kv;
      acc[3] += (kx * ky) * kv;
      acc[4] += (kx * kz) * kv;
      acc[5] += (ky * kz) * kv;
    } else {
      const double k1 = g0[g], k2 = g1[g];
      acc[0] += (k1 * k1) * kv;
      acc[1] += (k2 * k2) * kv;
      acc[2] += (k1 * k2) * kv;
    }
    i += di;
    j += dr;
    if (i >= sx) {
      i -= sx;
      ++j;
    }
    while (j >= sy && k < sz) {
      j -= sy;
      ++k;
    }
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
#pragma unroll
      for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 6; ++q) out[(int64_t)n * 6 + q] = red[q][0];
  }
}

__global__ void gk_fill(int64_t n, double *__restrict__ a, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void gk_div(int64_t n, double *__restrict__ a, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = a[i] / v;
}

__global__ void gk_extract(const int32_t *__restrict__ bins, int nact,
                           const double *__restrict__ grid,
                           double *__restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < nact) out[n] = grid[bins[n]];
}

__global__ void gk_flag_active(int64_t ncells, const double *__restrict__ counts,
                               int32_t *__restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncells) flags[i] = counts[i] != 0.0 ? 1 : 0;
}


// Grow-only device buffer `slot` of the engine, at least `bytes` large.
hipError_t gk_reserve(rwpt_engine *e, int slot, size_t bytes, void **out) {
  bytes = std::max(bytes, (size_t)256);
  if (e->gk_bytes[slot] < bytes) {
    if (e->gk_mem[slot]) (void)hipFree(e->gk_mem[slot]);
    e->gk_mem[slot] = nullptr;
    e->gk_bytes[slot] = 0;
    hipError_t err = hipMalloc(&e->gk_mem[slot], bytes);
    if (err != hipSuccess) return err;
    e->gk_bytes[slot] = bytes;
  }
  *out = e->gk_mem[slot];
  return hipSuccess;
}

enum GkSlot : int {
  kSlotBins = 0, kSlotFlags, kSlotSelect, kSlotPool, kSlotGrids, kSlotSmall,
  kSlotFactors, kSlotSpans, kSlotPairs, kSlotSort, kSlotTiles
};

// indices of the non-zero histogram bins, ascending, and their number
hipError_t compact_nonzero(rwpt_engine *e, const double *counts, int64_t ncells,
                           int32_t *bins, int32_t *dcount, int *count, hipStream_t st) {
  int32_t *flags = nullptr;
  void *tmp = nullptr;
  size_t bytes = 0;
  const hipcub::CountingInputIterator<int32_t> ids(0);
  hipError_t err = hipcub::DeviceSelect::Flagged(nullptr, bytes, ids, (const int32_t *)nullptr,
                                                 bins, dcount, (int)ncells, st);
  if (err == hipSuccess)
    err = gk_reserve(e, kSlotFlags, (size_t)ncells * 4, reinterpret_cast<void **>(&flags));
  if (err == hipSuccess) err = gk_reserve(e, kSlotSelect, bytes, &tmp);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(gk_flag_active, dim3((int)((ncells + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, st, ncells, counts, flags);
    err = hipGetLastError();
  }
  if (err == hipSuccess)
    err = hipcub::DeviceSelect::Flagged(tmp, bytes, ids, (const int32_t *)flags, bins, dcount,
                                        (int)ncells, st);
  int32_t c = 0;
  if (err == hipSuccess) err = hipMemcpyAsync(&c, dcount, 4, hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  *count = c;
  return err;
}

// Tile shape: 8 x 8 x 4 cells in 3-D, 16 x 16 over the two active
// dimensions of a 2-D grid (one cell per thread of a tile's workgroup).
Tiles make_tiles(const GkGrid &G) {
  Tiles T{};
  for (int d = 0; d < 3; ++d)
    T.td[d] = G.ndim == 3 ? (d < 2 ? 8 : 4) : (G.active[d] ? 16 : 1);
  for (int d = 0; d < 3; ++d) T.nt[d] = (G.nb[d] + T.td[d] - 1) / T.td[d];
  return T;
}

// One scatter pass (grid += w K over the active bins) as the tiled,
// output-stationary sum: per-bin factors, the (tile, bin) list sorted by
// tile, one workgroup per tile.
hipError_t scatter_pass(rwpt_engine *e, hipStream_t st, const GkGrid &G,
                        const int32_t *bins, int N, const double *hv, int hstride,
                        const double *w, int kind, int der, int range, int cap,
                        const uint8_t *skip, const int8_t *collapse, double *grid) {
  hipError_t err;
  double *fac = nullptr, *amp = nullptr;
  int32_t *span = nullptr, *cnt = nullptr, *offs = nullptr;
  const size_t nB = (size_t)N;
  if ((err = gk_reserve(e, kSlotFactors, nB * 3 * (size_t)cap * 8,
                        reinterpret_cast<void **>(&fac))) != hipSuccess)
    return err;
  void *spans = nullptr;
  if ((err = gk_reserve(e, kSlotSpans, nB * (kSpan * 4 + 8 + 8), &spans)) != hipSuccess)
    return err;
  amp = reinterpret_cast<double *>(spans);
  span = reinterpret_cast<int32_t *>(amp + nB);
  cnt = span + nB * kSpan;
  offs = cnt + nB;
  // one wavefront per bin: a bin's factors are at most 3 x cap values
  hipLaunchKernelGGL(gk_factors, dim3(N), dim3(64), cap * 3 * 8, st, bins, N, hv,
                     hstride, w, G, kind, der, range, cap, skip, collapse, fac, span, amp);
  const Tiles T = make_tiles(G);
  const unsigned bgrid = (unsigned)((N + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(gk_tile_count, dim3(bgrid), dim3(kBlock), 0, st, N, span, T, cnt);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  size_t bytes = 0;
  if ((err = hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt, offs, N, st)) != hipSuccess)
    return err;
  void *tmp = nullptr;
  if ((err = gk_reserve(e, kSlotSort, bytes, &tmp)) != hipSuccess) return err;
  if ((err = hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnt, offs, N, st)) != hipSuccess)
    return err;
  int32_t last[2] = {0, 0};
  if ((err = hipMemcpyAsync(&last[0], offs + N - 1, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (err = hipMemcpyAsync(&last[1], cnt + N - 1, 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (err = hipStreamSynchronize(st)) != hipSuccess)
    return err;
  const int P = last[0] + last[1];
  if (P == 0) return hipSuccess;
  int32_t *pairs = nullptr;
  if ((err = gk_reserve(e, kSlotPairs, (size_t)P * 16, reinterpret_cast<void **>(&pairs))) !=
      hipSuccess)
    return err;
  int32_t *keys = pairs, *vals = pairs + P, *skeys = pairs + 2 * (size_t)P,
          *svals = pairs + 3 * (size_t)P;
  hipLaunchKernelGGL(gk_tile_pairs, dim3(bgrid), dim3(kBlock), 0, st, N, span, T, offs, keys,
                     vals);
  const int ntiles = T.nt[0] * T.nt[1] * T.nt[2];
  int bits = 1;
  while ((1ll << bits) <= ntiles) ++bits;
  bytes = 0;
  if ((err = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, skeys, vals, svals, P, 0,
                                                bits, st)) != hipSuccess)
    return err;
  if ((err = gk_reserve(e, kSlotSort, bytes, &tmp)) != hipSuccess) return err;
  if ((err = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys, skeys, vals, svals, P, 0,
                                                bits, st)) != hipSuccess)
    return err;
  int32_t *tiles = nullptr;
  if ((err = gk_reserve(e, kSlotTiles, (size_t)ntiles * 8,
                        reinterpret_cast<void **>(&tiles))) != hipSuccess)
    return err;
  int32_t *tstart = tiles, *tend = tiles + ntiles;
  hipLaunchKernelGGL(gk_fill_int, dim3((unsigned)((2 * (int64_t)ntiles + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, st, 2 * (int64_t)ntiles, tiles, 0);
  hipLaunchKernelGGL(gk_tile_bounds, dim3((unsigned)((P + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, st, P, skeys, tstart, tend);
  hipLaunchKernelGGL(gk_tile_accumulate, dim3(ntiles), dim3(kBlock), 0, st, T, G, tstart, tend,
                     svals, span, amp, fac, cap, grid);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The per-bin formulas between the passes, one thread per active bin, in
// FP64.  Whole-set quantities (the widest kernel, the smallest net
// roughness, the largest scale and density, whether a bound was hit) are
// folded with atomics on the bit patterns of non-negative doubles, which
// order like the values.

struct GkC {
  int nd, dims[3], active[3];
  double bin[3], maxKS[3], maxKSD[3], minKS[3], minKSD[3];
  double maxKSv, minKSv;
  int bounded, isotropic, useDb;
  double supportC, alphaC, betaC, fnd, opq, o4, m6, sqrtEightPi, pi, nPoints;
  // kernel database: h/delta of entries 1..ndb and the log-spaced index
  const double *hod;
  int ndb;
  double dbmin, dbdelta;
};

struct GkA {
  double *hk, *scale, *shape, *sigma, *nEst, *dens, *densOld, *scaleOld, *curv,
      *rough, *net, *hw, *hs, *hsd, *s6, *part;
  int8_t *col;
  uint8_t *skip;
  // 0 widest half-width, 1 smallest positive net roughness, 2 largest
  // positive scale, 3 a smoothing was bounded, 4 largest density, 5 largest
  // previous density
  unsigned long long *u;
};

constexpr int kSumBlocks = 256;

__device__ inline void amax_pos(unsigned long long *a, double v) {
  atomicMax(a, (unsigned long long)__double_as_longlong(v));
}
__device__ inline void amin_pos(unsigned long long *a, double v) {
  atomicMin(a, (unsigned long long)__double_as_longlong(v));
}
__device__ inline double as_double(unsigned long long b) {
  return __longlong_as_double((long long)b);
}

// prComputeKernelSmoothingScale (GridProjectedKDE.F90:5164-5203)
__device__ inline double smoothing_scale_d(const GkC &K, const double *h) {
  double s = 1.0;
  int dc = K.nd;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!K.active[d]) continue;
    if (h[d] != 0.0) s = s * h[d];
    else dc = dc - 1;
  }
  return dc != 0 ? pow(s, 1.0 / K.fnd) : 0.0;
}

// database index of a bandwidth, 1-based and clamped (:5740-5750)
__device__ inline int db_index(const GkC &K, double h, double bin) {
  if (!(h > 0.0)) return 1;
  const int i = (int)floor(log(h / bin / K.dbmin) / K.dbdelta) + 1;
  return min(max(i, 1), K.ndb);
}

__device__ inline double pick3(const double v[3], int d) {
  return d == 0 ? v[0] : (d == 1 ? v[1] : v[2]);
}
__device__ inline int pick3i(const int v[3], int d) {
  return d == 0 ? v[0] : (d == 1 ? v[1] : v[2]);
}

// density kernel from the database (prSetKernelFromDatabase :5965-6003):
// bandwidths, and the dimension left at its centre entry by the x-y
// transpose of an x-z or y-z lookup (-1 none)
__device__ inline int db_w(const GkC &K, const double h[3], double out[3]) {
  int idx[3] = {1, 1, 1};
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (K.active[d]) idx[d] = db_index(K, h[d], K.bin[d]);
  if (K.nd == 3) {
#pragma unroll
    for (int d = 0; d < 3; ++d) out[d] = K.hod[idx[d] - 1] * K.bin[d];
    return -1;
  }
  const int a = K.dims[0], b = K.dims[1];
  const int ia = pick3i(idx, a), ib = pick3i(idx, b);
  const double ba = pick3(K.bin, a), bb = pick3(K.bin, b);
  double ha, hb;
  int c = -1;
  if (ia >= ib || (a == 0 && b == 1)) {
    ha = K.hod[ia - 1] * ba;
    hb = K.hod[ib - 1] * bb;
  } else {
    ha = K.hod[ib - 1] * ba;
    hb = K.hod[ia - 1] * bb;
    c = a;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) out[d] = d == a ? ha : (d == b ? hb : 0.0);
  return c;
}

// sigma kernel from the database (:6006-6043): the flat entry, the first
// two active dimensions taking the larger and the smaller index
__device__ inline void db_sigma(const GkC &K, double s, double out[3]) {
  int idx[3] = {1, 1, 1};
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (K.active[d]) idx[d] = db_index(K, s, K.bin[d]);
  const int a = K.nd == 3 ? 0 : K.dims[0], b = K.nd == 3 ? 1 : K.dims[1];
  const int ia = pick3i(idx, a), ib = pick3i(idx, b);
  const int hi = max(ia, ib), lo = min(ia, ib);
  const double va = K.hod[hi - 1] * pick3(K.bin, a), vb = K.hod[lo - 1] * pick3(K.bin, b);
  const double vz = K.nd == 3 ? K.hod[idx[2] - 1] * K.bin[2] : 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) out[d] = d == a ? va : (d == b ? vb : vz);
}

// widest half-width of a set of kernels, for the LDS size of the next pass
__device__ inline void cap_update(const GkC &K, unsigned long long *u, const double h[3],
                                  int range, bool allDims) {
  int m = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!allDims && !K.active[d]) continue;
    if (h[d] > 0.0 && isfinite(h[d])) m = max(m, half_width(range, h[d], K.bin[d]));
  }
  if (m > 0) atomicMax(u, (unsigned long long)m);
}

__global__ void gk_reset(unsigned long long *u) {
  const int t = threadIdx.x;
  if (t < 6) u[t] = t == 1 ? (unsigned long long)__double_as_longlong(INFINITY) : 0ull;
}

// initial smoothing in the active dimensions (:4109-4117, :4606-4618)
__global__ void gk_init(int N, const GkC K, GkA A, double hinit) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double h[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) h[d] = K.active[d] ? hinit : 0.0;
  const double sc = smoothing_scale_d(K, h);
  A.scale[n] = sc;
  A.scaleOld[n] = sc;
  A.sigma[n] = 2.0 * sc;  // defaultInitialSigmaFactor
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    A.hk[3 * (size_t)n + d] = h[d];
    A.shape[3 * (size_t)n + d] = (K.active[d] && sc > 0.0) ? h[d] / sc : 0.0;
  }
}

// density kernels of the current smoothing
__global__ void gk_wparams(int N, const GkC K, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double h[3], o[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) h[d] = A.hk[3 * (size_t)n + d];
  A.skip[n] = (h[0] < 0.0 || h[1] < 0.0 || h[2] < 0.0) ? 1 : 0;
  int c = -1;
  if (K.useDb) {
    c = db_w(K, h, o);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) o[d] = h[d];
  }
  A.col[n] = (int8_t)c;
#pragma unroll
  for (int d = 0; d < 3; ++d) A.hw[3 * (size_t)n + d] = o[d];
  cap_update(K, A.u, o, kRangeW, false);
}

// sigma kernels of the current support scales: brute force carries the
// scale in every dimension with a bin size, a database entry only in the
// active ones
__global__ void gk_sigparams(int N, const GkC K, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const double s = A.sigma[n];
  double o[3];
  if (K.useDb) {
    db_sigma(K, s, o);
  } else {
    o[0] = o[1] = o[2] = s;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) A.hs[3 * (size_t)n + d] = o[d];
  cap_update(K, A.u, o, kRangeW, !K.useDb);
}

// support scale from the first n-estimate (:4816-4858)
__global__ void gk_support(int N, const GkC K, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const double n1 = A.s6[6 * (size_t)n];
  double ks = pow(n1, 0.5) * pow(A.scale[n], K.opq) / pow(4.0 * A.dens[n], 0.25) * K.supportC;
  if (K.bounded) {
    const double g = 1.5 * A.sigma[n];  // defaultMaxSigmaGrowth
    if (ks > g) ks = g;
    if (K.maxKSv < ks) ks = K.maxKSv;
    if (ks < K.minKSv) ks = K.minKSv;
  }
  A.sigma[n] = ks;
}

// n-estimate, then the curvature bandwidths (:5300-5400), then the
// second-derivative kernels along each derivative dimension and whether a
// bin's curvature kernels are skipped
__global__ void gk_curv(int N, const GkC K, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const double ne = A.s6[6 * (size_t)n];
  A.nEst[n] = ne;
  const double dn = A.dens[n], sc = A.scale[n];
  double sh[3], cv[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int d = 0; d < 3; ++d) sh[d] = A.shape[3 * (size_t)n + d];
  if (dn > 0.0) {
    const double nvpb =
        pow(pow(K.sqrtEightPi * A.sigma[n], K.fnd) * pow(ne, 2.0) / dn, K.betaC);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (!K.active[d]) continue;
      double sum = 0.0;
#pragma unroll
      for (int e = 0; e < 3; ++e)
        if (K.active[e]) sum = sum + (e == d ? 5.0 : 1.0) / pow(sh[e], 2.0);
      const double term = pow((K.o4 / pow(sh[d], 4.0)) * sum, K.m6);
      double cb = K.alphaC * nvpb * term * sc;
      if (K.bounded) {
        if (cb > K.maxKSD[d]) cb = K.maxKSD[d];
        if (cb < K.minKSD[d]) cb = K.minKSD[d];
      }
      cv[d] = cb;
    }
  }
  uint8_t sk = 0;
  if (K.nd == 2)
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (cv[d] < 0.0 || cv[d] != cv[d]) sk = 1;
  A.skip[n] = sk;
  int m = 0;
#pragma unroll
  for (int der = 0; der < 3; ++der) {
    A.curv[3 * (size_t)n + der] = cv[der];
    const double cb = cv[der];
    double o[3];
    if (K.useDb) {
      // prSetKernelSD*FromDatabase (:6081-6124, :6184-6221)
      const int i = cb > 0.0 ? db_index(K, cb, K.bin[der]) : 1;
#pragma unroll
      for (int d = 0; d < 3; ++d) o[d] = K.active[d] ? K.hod[i - 1] * K.bin[d] : 0.0;
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) o[d] = K.active[d] ? cb : 0.0;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      A.hsd[9 * (size_t)n + 3 * der + d] = o[d];
      if (K.active[der] && K.active[d] && o[d] > 0.0 && isfinite(o[d]))
        m = max(m, half_width(kRangeSD, o[d], K.bin[d]));
    }
  }
  if (m > 0) atomicMax(A.u, (unsigned long long)m);
}

// net roughness from the gathered sums (:2137-2289, :2464-2709)
__global__ void gk_rough(int N, const GkC K, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const double *s = A.s6 + 6 * (size_t)n;
  double r[3] = {0.0, 0.0, 0.0};
  double net;
  if (K.nd == 3) {
    const double rXX = s[0], rYY = s[1], rZZ = s[2], rXY = s[3], rXZ = s[4], rYZ = s[5];
    r[0] = rXX;
    r[1] = rYY;
    r[2] = rZZ;
    if (!K.isotropic) {
      double v = 3.0 * pow(rXX * rYY * rZZ, 1.0 / 3.0);
      if (rYZ > 0.0) v = v + 2.0 * rYZ * pow(pow(rXX, 2.0) / rYY / rZZ, 1.0 / 6.0);
      if (rXZ > 0.0) v = v + 2.0 * rXZ * pow(pow(rYY, 2.0) / rXX / rZZ, 1.0 / 6.0);
      if (rXY > 0.0) v = v + 2.0 * rXY * pow(pow(rZZ, 2.0) / rXX / rYY, 1.0 / 6.0);
      net = v;
    } else {
      net = rXX + 2.0 * rXY + 2.0 * rXZ + rYY + 2.0 * rYZ + rZZ;
    }
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) r[d] = d == K.dims[0] ? s[0] : (d == K.dims[1] ? s[1] : 0.0);
    if (!K.isotropic) net = 2.0 * s[2] + 2.0 * sqrt(s[0] * s[1]);
    else net = 2.0 * s[2] + s[0] + s[1];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) A.rough[3 * (size_t)n + d] = r[d];
  A.net[n] = net;
  if (net > 0.0) amin_pos(&A.u[1], net);
}

// isotropy correction of the net roughness and the optimal scale
// (:5404-5540)
__global__ void gk_scale(int N, const GkC K, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double net = A.net[n];
  if (!K.isotropic) {
    double r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) r[d] = A.rough[3 * (size_t)n + d];
    const double r11 = pick3(r, K.dims[0]), r22 = pick3(r, K.dims[1]);
    double sumMain = K.nd == 2 ? r11 + r22 : r[0] + r[1] + r[2];
    if (sumMain == 0.0) sumMain = as_double(A.u[1]);
    bool iso;
    if (K.nd == 2)
      iso = (r11 / sumMain > 0.85) || (r22 / sumMain > 0.85);
    else
      iso = (r[0] / sumMain > 0.85) || (r[1] / sumMain > 0.85) || (r[2] / sumMain > 0.85);
    if (iso) net = (1.0 - 0.5) * net + 0.5 * sumMain;
    A.net[n] = net;
  }
  double sc = 0.0;
  if (net > 0.0) sc = pow(K.fnd * A.nEst[n] / (pow(4.0 * K.pi, 0.5 * K.fnd) * net), K.o4);
  A.scale[n] = sc;
  if (sc > 0.0) amax_pos(&A.u[2], sc);
}

// shape factors and the new smoothing (:5540-5724), bounded per dimension
__global__ void gk_shape(int N, const GkC K, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const double net = A.net[n];
  double sc = A.scale[n];
  if (net == 0.0) {
    sc = A.u[2] ? as_double(A.u[2]) : -INFINITY;
    A.scale[n] = sc;
  }
  double r[3], sh[3] = {1.0, 1.0, 1.0};
#pragma unroll
  for (int d = 0; d < 3; ++d) r[d] = A.rough[3 * (size_t)n + d];
  if (!K.isotropic) {
    if (K.nd == 2) {
      const int a = K.dims[0], b = K.dims[1];
      const double r11 = pick3(r, a), r22 = pick3(r, b);
      if (r11 != 0.0 && r22 != 0.0) {
        double sa = pow(net / r11, 1.0 / 8.0);
        if (K.bounded) {
          const double lo = pick3(K.minKS, a) / sc;
          if (sa < lo) sa = lo;
        }
        double sb = 1.0 / sa;
        const double ns = sqrt(sa * sb);
        if (ns == 0.0) {
          sa = sb = 1.0;
        } else {
          sa = sa / ns;
          sb = sb / ns;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) sh[d] = d == a ? sa : (d == b ? sb : 1.0);
      }
    } else if (r[0] != 0.0 && r[1] != 0.0 && r[2] != 0.0) {
      double s0 = pow(net / r[0], 1.0 / 12.0), s1 = pow(net / r[1], 1.0 / 12.0);
      if (K.bounded) {
        if (s0 < K.minKS[0] / sc) s0 = K.minKS[0] / sc;
        if (s1 < K.minKS[1] / sc) s1 = K.minKS[1] / sc;
      }
      double s2 = 1.0 / s0 / s1;
      // the reference's exponent is the single-precision literal 0.333
      const double ns = pow(s0 * s1 * s2, (double)0.333f);
      if (ns == 0.0) {
        s0 = s1 = s2 = 1.0;
      } else {
        s0 = s0 / ns;
        s1 = s1 / ns;
        s2 = s2 / ns;
      }
      sh[0] = s0;
      sh[1] = s1;
      sh[2] = s2;
    }
  }
  // a smoothing above the maximum is set to it, then one below the minimum
  // to that; any bounded smoothing recomputes every bin's scale
  bool clamped = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double h = 0.0;
    if (K.active[d]) {
      h = sh[d] * sc;
      if (K.bounded) {
        if (h > K.maxKS[d]) {
          h = K.maxKS[d];
          clamped = true;
        }
        if (h < K.minKS[d]) {
          h = K.minKS[d];
          clamped = true;
        }
      }
    }
    A.shape[3 * (size_t)n + d] = sh[d];
    A.hk[3 * (size_t)n + d] = h;
  }
  if (clamped) atomicOr(&A.u[3], 1ull);
}

__global__ void gk_rescale(int N, const GkC K, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N || A.u[3] == 0ull) return;
  A.scale[n] = smoothing_scale_d(K, A.hk + 3 * (size_t)n);
}

__global__ void gk_max_density(int N, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  amax_pos(&A.u[4], fmax(A.dens[n], 0.0));
  amax_pos(&A.u[5], fmax(A.densOld[n], 0.0));
}

// squared relative density and smoothing changes and squared ALMISE terms
// (:4962-4991), summed per block in a fixed order; then the previous
// density and scale take the new ones
__global__ void __launch_bounds__(kBlock) gk_conv(int N, const GkC K, GkA A) {
  __shared__ double red[3][kBlock];
  const double maxD = as_double(A.u[4]), maxDO = as_double(A.u[5]);
  double acc[3] = {0.0, 0.0, 0.0};
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < N; n += gridDim.x * kBlock) {
    const double dO = A.densOld[n], dN = A.dens[n];
    double em = 0.0;
    if (dO != 0.0) em = fabs(dN / maxD - dO / maxDO) / (dO / maxDO);
    const double sO = A.scaleOld[n], sN = A.scale[n];
    double es = 0.0;
    if (sO != 0.0) es = fabs(sN - sO) / sO;
    double ea = 0.0;
    if (sN != 0.0)
      ea = (A.nEst[n] / (pow(sN, K.fnd) * pow(4.0 * K.pi, 0.5 * K.fnd)) +
            0.25 * A.net[n] * pow(sN, 4.0)) /
           pow(K.nPoints, 2.0);
    acc[0] += em * em;
    acc[1] += es * es;
    acc[2] += ea * ea;
    A.densOld[n] = dN;
    A.scaleOld[n] = sN;
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 3; ++q) A.part[3 * blockIdx.x + q] = red[q][0];
}

__global__ void __launch_bounds__(kSumBlocks) gk_conv_total(GkA A, double *out) {
  __shared__ double red[3][kSumBlocks];
#pragma unroll
  for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = A.part[3 * threadIdx.x + q];
  __syncthreads();
  for (int s = kSumBlocks / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 3; ++q) out[q] = red[q][0];
}

__global__ void gk_keep_density(int N, GkA A) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) A.densOld[n] = A.dens[n];
}

}  // namespace

// Host side of one adaptive reconstruction.  counts_dev: the histogram in
// the units the reference's loop uses (weighted: mass sums / effective mass,
// point counts or per-bin Kish counts); final_weights_dev, when given, is
// scattered with the final kernels in place of the counts (effective weight
// formats 2 and 3, GridProjectedKDE.F90:4712-4760); density_dev (ncells)
// receives the density (before the effective-mass, unit-volume and scaling
// factors).  Returns the loops run in *loops_run.  Every per-bin quantity
// stays on the device: the host reads back the widest kernel before each
// scatter or gather (its LDS size) and three sums per loop for the
// convergence tests.
hipError_t launch_gpkde_opt(rwpt_engine *e, const rwpt_kde_options *opt,
                            const double *counts_dev, double hinit,
                            double nPoints, const double *final_weights_dev,
                            double *density_dev, int *loops_run) {
  hipStream_t st = e->stream;
  GkGrid G{};
  int nd = 0;
  double binVol = 1.0;
  for (int d = 0; d < 3; ++d) {
    G.nb[d] = opt->nbins[d];
    G.active[d] = opt->nbins[d] > 1;
    G.bin[d] = opt->bin_size[d];
    nd += G.active[d];
    if (G.bin[d] > 0.0) binVol = binVol * G.bin[d];
  }
  G.ndim = nd;
  if (nd < 2) return hipErrorNotSupported;
  const int64_t ncells = (int64_t)G.nb[0] * G.nb[1] * G.nb[2];
  const double fnd = (double)nd;
  const double pi = 4.0 * std::atan(1.0);
  GkC K{};
  K.nd = nd;
  for (int d = 0, di = 0; d < 3; ++d) {
    K.active[d] = G.active[d];
    K.bin[d] = G.bin[d];
    if (G.active[d]) K.dims[di++] = d;
  }
  // prInitializeModuleConstants, GridProjectedKDE.F90:1582-1611
  K.supportC = std::pow((fnd + 2.0) * std::pow(8.0 * pi, 0.5 * fnd), 0.25);
  K.alphaC = std::pow((1.0 + std::pow(2.0, 0.5 * fnd + 2.0)) /
                          (3.0 * std::pow(2.0, 4.0 / (fnd + 4.0))),
                      1.0 / (fnd + 6.0)) *
             std::pow(fnd + 2.0, 1.0 / (fnd + 4.0)) / std::pow(fnd + 4.0, 1.0 / (fnd + 6.0));
  K.betaC = 2.0 / (fnd + 4.0) / (fnd + 6.0);
  K.fnd = fnd;
  K.o4 = 1.0 / (fnd + 4.0);
  K.m6 = -1.0 / (fnd + 6.0);
  K.opq = 1.0 + 0.25 * fnd;
  K.sqrtEightPi = std::sqrt(8.0 * pi);
  K.pi = pi;
  K.nPoints = nPoints;
  K.isotropic = opt->isotropic != 0;
  // kernel size bounds (:788-858): format 0 by the domain, 1 by the given
  // h/delta range, 2 none
  K.bounded = opt->bound_kernel_size != 2;
  for (int d = 0; d < 3; ++d) {
    if (!G.active[d]) continue;
    if (opt->bound_kernel_size == 1) {
      K.maxKS[d] = K.maxKSD[d] = G.bin[d] * opt->max_h_over_delta;
      K.minKS[d] = K.minKSD[d] = G.bin[d] * opt->min_h_over_delta;
    } else {
      K.maxKS[d] = G.bin[d] * (0.9 * G.nb[d] - 1) / (double)kRangeW;
      K.maxKSD[d] = G.bin[d] * (0.9 * G.nb[d] - 1) / (double)kRangeSD;
      K.minKS[d] = 1.2 * G.bin[d] / (double)kRangeW;
      K.minKSD[d] = 1.2 * G.bin[d] / (double)kRangeSD;
    }
  }
  {
    int maxDim = -1, minDim = -1;
    for (int d = 0; d < 3; ++d) {
      if (K.maxKS[d] > 0.0 && (maxDim < 0 || K.maxKS[d] < K.maxKS[maxDim])) maxDim = d;
      if (K.minKS[d] > 0.0 && (minDim < 0 || K.minKS[d] > K.minKS[minDim])) minDim = d;
    }
    K.maxKSv = maxDim >= 0 ? K.maxKS[maxDim] : 0.0;
    K.minKSv = minDim >= 0 ? K.minKS[minDim] : 0.0;
  }
  const double tol = opt->rel_error_convergence > 0.0 ? opt->rel_error_convergence : 0.02;
  const bool skipBreak = opt->skip_error_convergence != 0;

  // active bins (ascending linear index = the reference's column-major order)
  int32_t *dbins = nullptr, *dcount = nullptr;
  hipError_t err = gk_reserve(e, kSlotBins, (size_t)ncells * 4 + 256,
                              reinterpret_cast<void **>(&dbins));
  if (err != hipSuccess) return err;
  dcount = dbins + ncells;
  int N = 0;
  err = compact_nonzero(e, counts_dev, ncells, dbins, dcount, &N, st);
  if (loops_run) *loops_run = 0;
  if (err != hipSuccess || N == 0)
    return err != hipSuccess ? err : hipMemsetAsync(density_dev, 0, ncells * 8, st);

  KernelDb db;
  K.useDb = opt->kernel_database != 0;
  if (K.useDb) {
    db.build(opt->min_h_over_delta, opt->delta_h_over_delta, opt->max_h_over_delta,
             opt->bin_size, opt->nbins);
    K.ndb = db.n;
    K.dbmin = db.min;
    K.dbdelta = db.delta;
  }

  // per-bin doubles in one allocation: hk 3, scale, shape 3, sigma, nEst,
  // dens, densOld, scaleOld, curv 3, rough 3, net, hw 3, hs 3, hsd 9, s6 6,
  // weights
  const size_t nB = (size_t)N;
  const size_t nDoubles = nB * 41 + 3 * kSumBlocks + 3 + (size_t)std::max(db.n, 1);
  double *pool = nullptr, *grids = nullptr;
  unsigned long long *du = nullptr;
  int8_t *dcol = nullptr;
  uint8_t *dskip = nullptr;
  // buffers belong to the engine and are reused by later reconstructions
  void *small = nullptr;
  if ((err = gk_reserve(e, kSlotPool, nDoubles * 8, reinterpret_cast<void **>(&pool))) !=
          hipSuccess ||
      (err = gk_reserve(e, kSlotGrids, 3 * (size_t)ncells * 8,
                        reinterpret_cast<void **>(&grids))) != hipSuccess ||
      (err = gk_reserve(e, kSlotSmall, 64 + 2 * nB, &small)) != hipSuccess)
    return err;
  du = reinterpret_cast<unsigned long long *>(small);
  dcol = reinterpret_cast<int8_t *>(du + 8);
  dskip = reinterpret_cast<uint8_t *>(dcol + nB);
  GkA A{};
  double *p = pool;
  auto take = [&](size_t k) {
    double *q = p;
    p += k;
    return q;
  };
  A.hk = take(3 * nB);
  A.scale = take(nB);
  A.shape = take(3 * nB);
  A.sigma = take(nB);
  A.nEst = take(nB);
  A.dens = take(nB);
  A.densOld = take(nB);
  A.scaleOld = take(nB);
  A.curv = take(3 * nB);
  A.rough = take(3 * nB);
  A.net = take(nB);
  A.hw = take(3 * nB);
  A.hs = take(3 * nB);
  A.hsd = take(9 * nB);
  A.s6 = take(6 * nB);
  double *wbins = take(nB);
  A.part = take(3 * kSumBlocks);
  double *dsums = take(3);
  double *dhod = take((size_t)std::max(db.n, 1));
  A.col = dcol;
  A.skip = dskip;
  A.u = du;
  K.hod = dhod;
  if (K.useDb &&
      (err = hipMemcpyAsync(dhod, db.hod.data(), db.n * 8, hipMemcpyHostToDevice, st)) !=
          hipSuccess)
    return err;
  double *cg[3] = {grids, grids + ncells, grids + 2 * ncells};

  const dim3 binGrid((N + kBlock - 1) / kBlock), blk(kBlock);
  const int cellBlocks = (int)((ncells + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(gk_extract, binGrid, blk, 0, st, dbins, N, counts_dev, wbins);

  // the widest kernel of the last parameter kernel -> LDS entries per factor
  auto read_cap = [&](int *cap) -> hipError_t {
    unsigned long long m = 0;
    hipError_t r = hipGetLastError();
    if (r == hipSuccess) r = hipMemcpyAsync(&m, du, 8, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return r;
    *cap = 2 * (int)m + 1;
    // dynamic LDS beside the 12 KiB reduction buffer of the gather
    return (size_t)*cap * 3 * 8 <= 49152 ? hipSuccess : hipErrorNotSupported;
  };
  auto reset_cap = [&]() { return hipMemsetAsync(du, 0, 8, st); };

  // density scatter with the current smoothing (:4623-4668, :4930-4955)
  auto density_pass = [&](const double *w) -> hipError_t {
    hipError_t r = reset_cap();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(gk_wparams, binGrid, blk, 0, st, N, K, A);
    int cap = 0;
    if ((r = read_cap(&cap)) != hipSuccess) return r;
    hipLaunchKernelGGL(gk_fill, dim3(cellBlocks), blk, 0, st, ncells, density_dev, 0.0);
    if ((r = scatter_pass(e, st, G, dbins, N, (const double *)A.hw, 3, w, 0, 0, kRangeW,
                          cap, (const uint8_t *)dskip,
                          K.useDb ? (const int8_t *)dcol : (const int8_t *)nullptr,
                          density_dev)) != hipSuccess)
      return r;
    hipLaunchKernelGGL(gk_div, dim3(cellBlocks), blk, 0, st, ncells, density_dev, binVol);
    hipLaunchKernelGGL(gk_extract, binGrid, blk, 0, st, dbins, N, density_dev, A.dens);
    return hipGetLastError();
  };
  // sigma gather of the density grid (the n-estimate) into s6[6n]
  const int sigKind = K.useDb ? 0 : 1;
  int sigCap = 0;
  auto nest_pass = [&]() -> hipError_t {
    hipError_t r = reset_cap();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(gk_sigparams, binGrid, blk, 0, st, N, K, A);
    if ((r = read_cap(&sigCap)) != hipSuccess) return r;
    hipLaunchKernelGGL(gk_gather, dim3(N), blk, sigCap * 3 * 8, st, dbins, N,
                       (const double *)A.hs, 3, sigKind, G, 0, sigCap,
                       (const double *)density_dev, (const double *)nullptr,
                       (const double *)nullptr, A.s6);
    return hipGetLastError();
  };

  hipLaunchKernelGGL(gk_init, binGrid, blk, 0, st, N, K, A, hinit);
  if ((err = density_pass(wbins)) != hipSuccess) return err;
  hipLaunchKernelGGL(gk_keep_density, binGrid, blk, 0, st, N, A);
  // nEstimate and the net roughness start at zero, so the first ALMISE is 0
  double almiseOld = 0.0;
  double cumsum = almiseOld, cumsumOld = cumsum;
  int loops = 0;
  const double c0 = G.bin[K.dims[0]] * G.bin[K.dims[0]];
  const double c1 = G.bin[K.dims[1]] * G.bin[K.dims[1]];
  const double c2 = G.bin[2] * G.bin[2];
  for (int m = 1; m <= opt->n_opt_loops; ++m) {
    loops = m;
    hipLaunchKernelGGL(gk_reset, dim3(1), dim3(64), 0, st, du);
    // support sizes and nEstimate (:4816-4858)
    if ((err = nest_pass()) != hipSuccess) return err;
    hipLaunchKernelGGL(gk_support, binGrid, blk, 0, st, N, K, A);
    if ((err = nest_pass()) != hipSuccess) return err;
    // curvature bandwidths and kernels (:5300-5400), curvature grids
    // (:1905-2136)
    if ((err = reset_cap()) != hipSuccess) return err;
    hipLaunchKernelGGL(gk_curv, binGrid, blk, 0, st, N, K, A);
    int sdCap = 0;
    if ((err = read_cap(&sdCap)) != hipSuccess) return err;
    for (int g = 0; g < 3; ++g)
      hipLaunchKernelGGL(gk_fill, dim3(cellBlocks), blk, 0, st, ncells, cg[g], 0.0);
    const int nder = nd == 3 ? 3 : 2;
    for (int q = 0; q < nder; ++q) {
      const int der = nd == 3 ? q : K.dims[q];
      if ((err = scatter_pass(e, st, G, dbins, N, (const double *)(A.hsd + 3 * der), 9,
                              (const double *)wbins, 2, der, kRangeSD, sdCap,
                              (const uint8_t *)dskip, (const int8_t *)nullptr, cg[q])) !=
          hipSuccess)
        return err;
    }
    // roughness: sigma gathers of the squared curvatures (:2137-2709)
    hipLaunchKernelGGL(gk_curv_transform, dim3(cellBlocks), blk, 0, st, ncells,
                       nd == 3 ? 1 : 0, binVol, c0, c1, c2, cg[0], cg[1], cg[2]);
    hipLaunchKernelGGL(gk_gather, dim3(N), blk, sigCap * 3 * 8, st, dbins, N,
                       (const double *)A.hs, 3, sigKind, G, nd == 3 ? 1 : 2, sigCap,
                       (const double *)cg[0], (const double *)cg[1], (const double *)cg[2],
                       A.s6);
    hipLaunchKernelGGL(gk_rough, binGrid, blk, 0, st, N, K, A);
    // optimal smoothing and shape (:5404-5724)
    hipLaunchKernelGGL(gk_scale, binGrid, blk, 0, st, N, K, A);
    hipLaunchKernelGGL(gk_shape, binGrid, blk, 0, st, N, K, A);
    hipLaunchKernelGGL(gk_rescale, binGrid, blk, 0, st, N, K, A);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    // new density (:4930-4960)
    if ((err = density_pass(wbins)) != hipSuccess) return err;
    // convergence (:4962-5038)
    hipLaunchKernelGGL(gk_max_density, binGrid, blk, 0, st, N, A);
    hipLaunchKernelGGL(gk_conv, dim3(kSumBlocks), blk, 0, st, N, K, A);
    hipLaunchKernelGGL(gk_conv_total, dim3(1), dim3(kSumBlocks), 0, st, A, dsums);
    double sums[3];
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if ((err = hipMemcpyAsync(sums, dsums, 24, hipMemcpyDeviceToHost, st)) != hipSuccess)
      return err;
    if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
    const double errDensity = std::sqrt(sums[0] / (double)N);
    const double errSmoothing = std::sqrt(sums[1] / (double)N);
    const double alm = std::sqrt(sums[2] / (double)N);
    cumsum = cumsum + alm;
    if (!skipBreak) {
      if (almiseOld > 0.0 && std::fabs(alm - almiseOld) / almiseOld < tol) break;
      if (errDensity < tol) break;
      if (errSmoothing < tol) break;
      if (m > 7) {
        if (std::fabs(cumsumOld / (double)(m - 1) - cumsum / (double)m) / cumsumOld /
                (double)(m - 1) <
            tol)
          break;
      }
    }
    almiseOld = alm;
    cumsumOld = cumsum;
  }
  if (final_weights_dev) {
    // last reconstruction over the mass sums (:4730-4758)
    hipLaunchKernelGGL(gk_extract, binGrid, blk, 0, st, dbins, N, final_weights_dev, wbins);
    if ((err = density_pass(wbins)) != hipSuccess) return err;
  }
  if (loops_run) *loops_run = loops;
  return hipStreamSynchronize(st);
}

void KernelDb::build(double min_hod, double delta_hod, double max_hod,
                     const double binsz[3], const int nbins[3]) {
  min = min_hod;
  // nDelta = ceiling(log10(max/min)/log10(1 + delta)) + 1 and the log-spaced
  // grid of prGenerateLogSpaceData (GridProjectedKDE.F90:3197-3207, 6467-6500)
  n = (int)std::ceil(std::log10(max_hod / min_hod) / std::log10(1.0 + delta_hod)) + 1;
  const double lo = std::log10(min_hod), hi = std::log10(max_hod);
  const double step = (hi - lo) / (double)(n - 1);
  double acc = 0.0;
  hod.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    hod[i] = std::pow(10.0, lo + acc);
    acc = acc + step;
  }
  // the lookups use log(h2/h1) in place of the given spacing (:3207)
  delta = std::log(hod[1] / hod[0]);
  ndim = 0;
  for (int d = 0; d < 3; ++d) {
    bin[d] = binsz[d];
    active[d] = nbins[d] > 1;
    if (active[d]) dims[ndim++] = d;
  }
}

int KernelDb::index(double h, int d) const {
  // :5740-5750, 1-based and clamped to the grid
  if (!(h > 0.0)) return 1;
  const int i = (int)std::floor(std::log(h / bin[d] / min) / delta) + 1;
  return std::min(std::max(i, 1), n);
}

int KernelDb::w(const double h[3], double out[3]) const {
  int idx[3] = {1, 1, 1};
  for (int d = 0; d < 3; ++d)
    if (active[d]) idx[d] = index(h[d], d);
  out[0] = out[1] = out[2] = 0.0;
  if (ndim == 3) {
    // the x-y transpose of a lookup with idx(x) < idx(y) restores the
    // bandwidths in 3-D
    for (int d = 0; d < 3; ++d) out[d] = hod[idx[d] - 1] * bin[d];
    return -1;
  }
  if (ndim == 1) {
    // the 1-D entry is the largest index over the three dimensions (:5752-5757)
    const int a = dims[0];
    out[a] = hod[std::max(idx[0], std::max(idx[1], idx[2])) - 1] * bin[a];
    return -1;
  }
  const int a = dims[0], b = dims[1];
  if (idx[a] >= idx[b] || (a == 0 && b == 1)) {
    out[a] = hod[idx[a] - 1] * bin[a];
    out[b] = hod[idx[b] - 1] * bin[b];
    return -1;
  }
  // x-z or y-z plane, transposed entry: the x-y transpose leaves the first
  // active dimension at its centre entry and the bandwidths swapped
  // (ComputeSpansBoundedTranspose, KernelMultiGaussian.f90:393-484)
  out[a] = hod[idx[b] - 1] * bin[a];
  out[b] = hod[idx[a] - 1] * bin[b];
  return a;
}

void KernelDb::sigma(double s, double out[3]) const {
  int idx[3] = {1, 1, 1};
  for (int d = 0; d < 3; ++d)
    if (active[d]) idx[d] = index(s, d);
  out[0] = out[1] = out[2] = 0.0;
  if (ndim == 1) {
    const int a = dims[0];
    out[a] = hod[std::max(idx[0], std::max(idx[1], idx[2])) - 1] * bin[a];
    return;
  }
  // the flat entry without the transpose: the first two (active)
  // dimensions take the larger and the smaller index
  const int a = ndim == 3 ? 0 : dims[0], b = ndim == 3 ? 1 : dims[1];
  const int hi = std::max(idx[a], idx[b]), lo = std::min(idx[a], idx[b]);
  out[a] = hod[hi - 1] * bin[a];
  out[b] = hod[lo - 1] * bin[b];
  if (ndim == 3) out[2] = hod[idx[2] - 1] * bin[2];
}

void KernelDb::sd(double cb, int der, double out[3]) const {
  const int i = cb > 0.0 ? index(cb, der) : 1;
  for (int d = 0; d < 3; ++d) out[d] = active[d] ? hod[i - 1] * bin[d] : 0.0;
}

}  // namespace rwpt
