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
// formulas between the passes run on the host in FP64 with the same libm
// calls as the reference.  Scatters add with atomics, so the summation order
// (and the last bits of the density) can vary from run to run.
#include <algorithm>
#include <cmath>
#include <vector>

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

// density or curvature scatter: grid += w(bin) * K(bin)
__global__ void gk_scatter(const int32_t *__restrict__ bins, int nact,
                           const double *__restrict__ hv, int hstride,
                           const double *__restrict__ w, const GkGrid G,
                           int kind, int der, int range, int cap,
                           const uint8_t *__restrict__ skip,
                           const int8_t *__restrict__ collapse,
                           double *__restrict__ grid) {
  extern __shared__ double lds[];
  const int n = blockIdx.x;
  if (n >= nact) return;
  if (skip && skip[n]) return;
  const int id = bins[n];
  const int c[3] = {id % G.nb[0] + 1, (id / G.nb[0]) % G.nb[1] + 1,
                    id / (G.nb[0] * G.nb[1]) + 1};
  double h[3];
  for (int d = 0; d < 3; ++d) h[d] = hv[(int64_t)n * hstride + (hstride == 3 ? d : 0)];
  if (kind == 2)  // second-derivative kernels carry one bandwidth in active dims
    for (int d = 0; d < 3; ++d) h[d] = G.active[d] ? h[d] : 0.0;
  Factors F;
  build_factors(lds, cap, G, c, h, kind, der, range, collapse ? collapse[n] : -1, F);
  const double *ax = lds, *ay = lds + cap, *az = lds + 2 * cap;
  const int sx = F.hi[0] - F.lo[0] + 1, sy = F.hi[1] - F.lo[1] + 1,
            sz = F.hi[2] - F.lo[2] + 1;
  const double a = w[n] * F.scale;
  const int64_t tot = (int64_t)sx * sy * sz;
  for (int64_t t = threadIdx.x; t < tot; t += blockDim.x) {
    const int i = (int)(t % sx), j = (int)((t / sx) % sy), k = (int)(t / ((int64_t)sx * sy));
    const int gi = F.lo[0] + i, gj = F.lo[1] + j, gk = F.lo[2] + k;
    const double v = ax[gi - c[0] + F.m[0]] * ay[gj - c[1] + F.m[1]] *
                     az[gk - c[2] + F.m[2]];
    const int64_t g = (int64_t)(gi - 1) + (int64_t)G.nb[0] * ((gj - 1) + (int64_t)G.nb[1] * (gk - 1));
    unsafeAtomicAdd(&grid[g], a * v);
  }
}

// sigma-kernel gathers.  mode 0: sum density*K; mode 1 (3-D roughness):
// with k_d = curv_d / binVol / delta_d^2 the sums of kx^2, ky^2, kz^2,
// kx^2 ky^2, kx^2 kz^2, ky^2 kz^2 (cross terms after squaring,
// GridProjectedKDE.F90:2464-2469); mode 2 (2-D): k1^2, k2^2, k1 k2.
__global__ void gk_gather(const int32_t *__restrict__ bins, int nact,
                         const double *__restrict__ hv, int hstride, int kind,
                         const GkGrid G,
                         int mode, int cap, const double *__restrict__ g0,
                         const double *__restrict__ g1,
                         const double *__restrict__ g2, double binVol,
                         double cs0, double cs1, double cs2,
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
  for (int d = 0; d < 3; ++d) h[d] = hv[(int64_t)n * hstride + (hstride == 3 ? d : 0)];
  Factors F;
  build_factors(lds, cap, G, c, h, kind, 0, kRangeW, -1, F);
  const double *ax = lds, *ay = lds + cap, *az = lds + 2 * cap;
  const int sx = F.hi[0] - F.lo[0] + 1, sy = F.hi[1] - F.lo[1] + 1,
            sz = F.hi[2] - F.lo[2] + 1;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int64_t tot = (int64_t)sx * sy * sz;
  for (int64_t t = threadIdx.x; t < tot; t += blockDim.x) {
    const int i = (int)(t % sx), j = (int)((t / sx) % sy), k = (int)(t / ((int64_t)sx * sy));
    const int gi = F.lo[0] + i, gj = F.lo[1] + j, gk = F.lo[2] + k;
    const double kv = ax[gi - c[0] + F.m[0]] * ay[gj - c[1] + F.m[1]] *
                      az[gk - c[2] + F.m[2]] * F.scale;
    const int64_t g = (int64_t)(gi - 1) + (int64_t)G.nb[0] * ((gj - 1) + (int64_t)G.nb[1] * (gk - 1));
    if (mode == 0) {
      acc[0] += g0[g] * kv;
    } else if (mode == 1) {
      double kx = g0[g] / binVol, ky = g1[g] / binVol, kz = g2[g] / binVol;
      kx = kx / cs0;
      ky = ky / cs1;
      kz = kz / cs2;
      kx = kx * kx;
      ky = ky * ky;
      kz = kz * kz;
      acc[0] += kx * kv;
      acc[1] += ky * kv;
      acc[2] += kz * kv;
      acc[3] += (kx * ky) * kv;
      acc[4] += (kx * kz) * kv;
      acc[5] += (ky * kz) * kv;
    } else {
      double k1 = g0[g] / binVol, k2 = g1[g] / binVol;
      k1 = k1 / cs0;
      k2 = k2 / cs1;
      acc[0] += (k1 * k1) * kv;
      acc[1] += (k2 * k2) * kv;
      acc[2] += (k1 * k2) * kv;
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

}  // namespace

// Host side of one adaptive reconstruction.  counts_dev: the histogram in
// the units the reference's loop uses (weighted: mass sums / effective mass,
// point counts or per-bin Kish counts); final_weights_dev, when given, is
// scattered with the final kernels in place of the counts (effective weight
// formats 2 and 3, GridProjectedKDE.F90:4712-4760); density_dev (ncells)
// receives the density (before the effective-mass, unit-volume and scaling
// factors).  Returns the loops run in *loops_run.
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
  const double sqrtEightPi = std::sqrt(8.0 * pi);
  // prInitializeModuleConstants, GridProjectedKDE.F90:1582-1611
  const double supportC = std::pow((fnd + 2.0) * std::pow(8.0 * pi, 0.5 * fnd), 0.25);
  const double alphaC =
      std::pow((1.0 + std::pow(2.0, 0.5 * fnd + 2.0)) /
                   (3.0 * std::pow(2.0, 4.0 / (fnd + 4.0))),
               1.0 / (fnd + 6.0)) *
      std::pow(fnd + 2.0, 1.0 / (fnd + 4.0)) / std::pow(fnd + 4.0, 1.0 / (fnd + 6.0));
  const double betaC = 2.0 / (fnd + 4.0) / (fnd + 6.0);
  const double oneOverNDimPlusFour = 1.0 / (fnd + 4.0);
  const double minusOneOverNDimPlusSix = -1.0 / (fnd + 6.0);
  const double onePlusNDimQuarter = 1.0 + 0.25 * fnd;
  // kernel size bounds (:788-858): format 0 by the domain, 1 by the given
  // h/delta range, 2 none
  double maxKS[3] = {0, 0, 0}, maxKSD[3] = {0, 0, 0}, minKS[3] = {0, 0, 0},
         minKSD[3] = {0, 0, 0};
  const bool bounded = opt->bound_kernel_size != 2;
  for (int d = 0; d < 3; ++d) {
    if (!G.active[d]) continue;
    if (opt->bound_kernel_size == 1) {
      maxKS[d] = maxKSD[d] = G.bin[d] * opt->max_h_over_delta;
      minKS[d] = minKSD[d] = G.bin[d] * opt->min_h_over_delta;
    } else {
      maxKS[d] = G.bin[d] * (0.9 * G.nb[d] - 1) / (double)kRangeW;
      maxKSD[d] = G.bin[d] * (0.9 * G.nb[d] - 1) / (double)kRangeSD;
      minKS[d] = 1.2 * G.bin[d] / (double)kRangeW;
      minKSD[d] = 1.2 * G.bin[d] / (double)kRangeSD;
    }
  }
  int maxDim = -1, minDim = -1;
  for (int d = 0; d < 3; ++d) {
    if (maxKS[d] > 0.0 && (maxDim < 0 || maxKS[d] < maxKS[maxDim])) maxDim = d;
    if (minKS[d] > 0.0 && (minDim < 0 || minKS[d] > minKS[minDim])) minDim = d;
  }
  const double maxSigmaGrowth = 1.5, isoThreshold = 0.85, isoFactor = 0.5;
  const bool isotropic = opt->isotropic != 0;
  const double tol = opt->rel_error_convergence > 0.0 ? opt->rel_error_convergence : 0.02;
  const bool skipBreak = opt->skip_error_convergence != 0;
  int dims[3], di = 0;
  for (int d = 0; d < 3; ++d)
    if (G.active[d]) dims[di++] = d;
  const int id1 = dims[0], id2 = dims[1];

  // active bins (ascending linear index = the reference's column-major order)
  std::vector<double> hcounts(ncells);
  hipError_t err = hipMemcpyAsync(hcounts.data(), counts_dev, ncells * 8,
                                  hipMemcpyDeviceToHost, st);
  if (err == hipSuccess) err = hipStreamSynchronize(st);
  if (err != hipSuccess) return err;
  std::vector<int32_t> hbins;
  for (int64_t i = 0; i < ncells; ++i)
    if (hcounts[i] != 0.0) hbins.push_back((int32_t)i);
  const int N = (int)hbins.size();
  if (loops_run) *loops_run = 0;
  if (N == 0) {
    return hipMemsetAsync(density_dev, 0, ncells * 8, st);
  }
  std::vector<double> cnt(N), rawDensity(N);
  for (int n = 0; n < N; ++n) {
    cnt[n] = hcounts[hbins[n]];
    rawDensity[n] = cnt[n] / binVol;
  }
  // device buffers
  int32_t *dbins = nullptr;
  double *dcnt = nullptr, *dh = nullptr, *dsig = nullptr, *dsd = nullptr,
         *dsums = nullptr, *darr = nullptr, *grids = nullptr;
  uint8_t *dskip = nullptr;
  int8_t *dcol = nullptr;
  auto cleanup = [&]() {
    for (void *p : {(void *)dbins, (void *)dcnt, (void *)dh, (void *)dsig,
                    (void *)dsd, (void *)dsums, (void *)darr, (void *)grids,
                    (void *)dskip, (void *)dcol})
      if (p) (void)hipFree(p);
  };
  if ((err = hipMalloc(&dbins, N * 4)) != hipSuccess) return cleanup(), err;
  if ((err = hipMalloc(&dcnt, N * 8)) != hipSuccess) return cleanup(), err;
  if ((err = hipMalloc(&dh, N * 24)) != hipSuccess) return cleanup(), err;
  if ((err = hipMalloc(&dsig, N * 24)) != hipSuccess) return cleanup(), err;
  if ((err = hipMalloc(&dsd, N * 24)) != hipSuccess) return cleanup(), err;
  if ((err = hipMalloc(&dcol, N)) != hipSuccess) return cleanup(), err;
  if ((err = hipMalloc(&dsums, N * 48)) != hipSuccess) return cleanup(), err;
  if ((err = hipMalloc(&darr, N * 8)) != hipSuccess) return cleanup(), err;
  if ((err = hipMalloc(&dskip, N)) != hipSuccess) return cleanup(), err;
  if ((err = hipMalloc(&grids, 3 * ncells * 8)) != hipSuccess) return cleanup(), err;
  double *cg[3] = {grids, grids + ncells, grids + 2 * ncells};
  err = hipMemcpyAsync(dbins, hbins.data(), N * 4, hipMemcpyHostToDevice, st);
  if (err == hipSuccess)
    err = hipMemcpyAsync(dcnt, cnt.data(), N * 8, hipMemcpyHostToDevice, st);
  if (err != hipSuccess) return cleanup(), err;

  const int gridBlocks = (int)((ncells + kBlock - 1) / kBlock);
  int maxnb = std::max(G.nb[0], std::max(G.nb[1], G.nb[2]));
  // LDS capacity per factor array: the widest kernel a launch needs
  auto cap_for = [&](const std::vector<double> &hv, int stride, int range,
                     bool allDims) {
    int m = 0;
    for (int n = 0; n < N; ++n)
      for (int d = 0; d < 3; ++d) {
        if (!allDims && !G.active[d]) continue;
        const double h = hv[(size_t)n * stride + (stride == 3 ? d : 0)];
        if (h > 0.0 && std::isfinite(h)) m = std::max(m, half_width(range, h, G.bin[d]));
      }
    (void)maxnb;
    return 2 * m + 1;
  };
  // dynamic LDS beside the 12 KiB reduction buffer of the gather
  auto lds_ok = [&](int cap) { return (size_t)cap * 3 * 8 <= 49152; };

  std::vector<double> hk(3 * (size_t)N), scale(N), shape(3 * (size_t)N), sigma(N);
  std::vector<double> nEst(N, 0.0), density(N, 0.0), curvBW(3 * (size_t)N, 0.0);
  std::vector<double> rXX(N, 0.0), rYY(N, 0.0), rZZ(N, 0.0), net(N, 0.0);
  std::vector<uint8_t> skip(N, 0);

  // prComputeKernelSmoothingScale (:5164-5203)
  auto smoothing_scale = [&](int n) {
    double s = 1.0;
    int dc = nd;
    for (int d = 0; d < 3; ++d) {
      if (!G.active[d]) continue;
      if (hk[3 * (size_t)n + d] != 0.0) s = s * hk[3 * (size_t)n + d];
      else dc = dc - 1;
    }
    return dc != 0 ? std::pow(s, 1.0 / fnd) : 0.0;
  };

  // kernel database: every lookup becomes the bandwidths of its entry
  KernelDb db;
  const bool useDb = opt->kernel_database != 0;
  if (useDb)
    db.build(opt->min_h_over_delta, opt->delta_h_over_delta, opt->max_h_over_delta,
             opt->bin_size, opt->nbins);
  std::vector<double> hw(3 * (size_t)N), hs(3 * (size_t)N), hsd(3 * (size_t)N);
  std::vector<int8_t> col(N, -1);

  // density scatter with the current smoothing (:4623-4668, :4930-4955)
  auto density_pass = [&](const double *wbins) -> hipError_t {
    for (int n = 0; n < N; ++n) {
      const double *hn = &hk[3 * (size_t)n];
      skip[n] = (hn[0] < 0.0 || hn[1] < 0.0 || hn[2] < 0.0) ? 1 : 0;
      if (useDb) col[n] = (int8_t)db.w(hn, &hw[3 * (size_t)n]);
      else for (int d = 0; d < 3; ++d) hw[3 * (size_t)n + d] = hn[d];
    }
    const int cap = cap_for(hw, 3, kRangeW, false);
    if (!lds_ok(cap)) return hipErrorNotSupported;
    hipError_t r = hipMemcpyAsync(dh, hw.data(), N * 24, hipMemcpyHostToDevice, st);
    if (r == hipSuccess) r = hipMemcpyAsync(dskip, skip.data(), N, hipMemcpyHostToDevice, st);
    if (r == hipSuccess && useDb)
      r = hipMemcpyAsync(dcol, col.data(), N, hipMemcpyHostToDevice, st);
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(gk_fill, dim3(gridBlocks), dim3(kBlock), 0, st, ncells,
                       density_dev, 0.0);
    hipLaunchKernelGGL(gk_scatter, dim3(N), dim3(kBlock), cap * 3 * 8, st, dbins, N,
                       dh, 3, wbins, G, 0, 0, kRangeW, cap, dskip,
                       useDb ? (const int8_t *)dcol : (const int8_t *)nullptr, density_dev);
    hipLaunchKernelGGL(gk_div, dim3(gridBlocks), dim3(kBlock), 0, st, ncells,
                       density_dev, binVol);
    hipLaunchKernelGGL(gk_extract, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       st, dbins, N, density_dev, darr);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    r = hipMemcpyAsync(density.data(), darr, N * 8, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    return r;
  };
  // sigma kernels of the current support scales: brute force carries the
  // scale in every dimension with a bin size (kind 1); a database entry
  // has zero bandwidth outside the active dimensions (kind 0)
  int sigCap = 0;
  auto upload_sigma = [&]() -> hipError_t {
    if (useDb) {
      for (int n = 0; n < N; ++n) db.sigma(sigma[n], &hs[3 * (size_t)n]);
      sigCap = cap_for(hs, 3, kRangeW, false);
    } else {
      for (int n = 0; n < N; ++n)
        for (int d = 0; d < 3; ++d) hs[3 * (size_t)n + d] = sigma[n];
      sigCap = cap_for(sigma, 1, kRangeW, true);
    }
    if (!lds_ok(sigCap)) return hipErrorNotSupported;
    return hipMemcpyAsync(dsig, hs.data(), N * 24, hipMemcpyHostToDevice, st);
  };
  const int sigKind = useDb ? 0 : 1;
  // sigma gather of the density grid (nEstimate)
  auto nest_pass = [&](std::vector<double> &out) -> hipError_t {
    hipError_t r = upload_sigma();
    if (r != hipSuccess) return r;
    const int cap = sigCap;
    hipLaunchKernelGGL(gk_gather, dim3(N), dim3(kBlock), cap * 3 * 8, st, dbins, N,
                       dsig, 3, sigKind, G, 0, cap, density_dev, (const double *)nullptr,
                       (const double *)nullptr, 1.0, 1.0, 1.0, 1.0, dsums);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    std::vector<double> s6(6 * (size_t)N);
    r = hipMemcpyAsync(s6.data(), dsums, N * 48, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    for (int n = 0; n < N; ++n) out[n] = s6[6 * (size_t)n];
    return r;
  };

  // initial smoothing in the active dimensions (:4109-4117, :4616)
  for (int n = 0; n < N; ++n)
    for (int d = 0; d < 3; ++d) hk[3 * (size_t)n + d] = G.active[d] ? hinit : 0.0;
  for (int n = 0; n < N; ++n) {
    scale[n] = smoothing_scale(n);
    sigma[n] = 2.0 * scale[n];  // defaultInitialSigmaFactor
    for (int d = 0; d < 3; ++d) {
      if (G.active[d]) {
        if (scale[n] > 0.0) shape[3 * (size_t)n + d] = hk[3 * (size_t)n + d] / scale[n];
      } else {
        hk[3 * (size_t)n + d] = 0.0;
        shape[3 * (size_t)n + d] = 0.0;
      }
    }
  }
  std::vector<double> scaleOld = scale;
  if ((err = density_pass(dcnt)) != hipSuccess) return cleanup(), err;

  auto almise = [&]() {
    double s2 = 0.0;
    for (int n = 0; n < N; ++n) {
      double em = 0.0;
      if (scale[n] != 0.0)
        em = (nEst[n] / (std::pow(scale[n], fnd) * std::pow(4.0 * pi, 0.5 * fnd)) +
              0.25 * net[n] * std::pow(scale[n], 4.0)) /
             std::pow(nPoints, 2.0);
      s2 += std::pow(em, 2.0);
    }
    return std::sqrt(s2 / (double)N);
  };
  double almiseOld = almise();
  double cumsum = almiseOld, cumsumOld = cumsum;
  std::vector<double> densityOld = density;
  int loops = 0;
  for (int m = 1; m <= opt->n_opt_loops; ++m) {
    loops = m;
    // support sizes and nEstimate (:4816-4858)
    std::vector<double> n1(N);
    if ((err = nest_pass(n1)) != hipSuccess) return cleanup(), err;
    for (int n = 0; n < N; ++n) {
      double ks = std::pow(n1[n], 0.5) * std::pow(scale[n], onePlusNDimQuarter) /
                  std::pow(4.0 * density[n], 0.25) * supportC;
      if (bounded) {
        if (ks > maxSigmaGrowth * sigma[n]) ks = maxSigmaGrowth * sigma[n];
        ks = std::min(ks, maxKS[maxDim]);
        if (ks < minKS[minDim]) ks = minKS[minDim];
      }
      sigma[n] = ks;
    }
    if ((err = nest_pass(nEst)) != hipSuccess) return cleanup(), err;
    // curvature bandwidths (:5300-5400)
    for (int n = 0; n < N; ++n) {
      for (int d = 0; d < 3; ++d) curvBW[3 * (size_t)n + d] = 0.0;
      if (density[n] <= 0.0) continue;
      const double nvpb = std::pow(std::pow(sqrtEightPi * sigma[n], fnd) *
                                       std::pow(nEst[n], 2.0) / density[n],
                                   betaC);
      for (int a = 0; a < nd; ++a) {
        const int did = dims[a];
        double num[3] = {1.0, 1.0, 1.0};
        num[did] = 5.0;
        double sum = 0.0;
        for (int b = 0; b < nd; ++b) {
          const int dd = dims[b];
          sum = sum + num[dd] / std::pow(shape[3 * (size_t)n + dd], 2.0);
        }
        const double term = std::pow(
            (oneOverNDimPlusFour / std::pow(shape[3 * (size_t)n + did], 4.0)) * sum,
            minusOneOverNDimPlusSix);
        double cb = alphaC * nvpb * term * scale[n];
        if (bounded) {
          if (cb > maxKSD[did]) cb = maxKSD[did];
          if (cb < minKSD[did]) cb = minKSD[did];
        }
        curvBW[3 * (size_t)n + did] = cb;
      }
    }
    // curvature grids and roughness (:1905-2709)
    for (int g = 0; g < 3; ++g)
      hipLaunchKernelGGL(gk_fill, dim3(gridBlocks), dim3(kBlock), 0, st, ncells, cg[g], 0.0);
    const int nder = nd == 3 ? 3 : 2;
    for (int q = 0; q < nder; ++q) {
      const int der = nd == 3 ? q : dims[q];
      for (int n = 0; n < N; ++n) {
        const double cb = curvBW[3 * (size_t)n + der];
        double *ho = &hsd[3 * (size_t)n];
        if (useDb) db.sd(cb, der, ho);
        else for (int d = 0; d < 3; ++d) ho[d] = G.active[d] ? cb : 0.0;
        skip[n] = 0;
        if (nd == 2) {
          for (int d = 0; d < 3; ++d) {
            const double v = curvBW[3 * (size_t)n + d];
            if (v < 0.0 || v != v) skip[n] = 1;
          }
        }
      }
      const int cap = cap_for(hsd, 3, kRangeSD, false);
      if (!lds_ok(cap)) return cleanup(), hipErrorNotSupported;
      err = hipMemcpyAsync(dsd, hsd.data(), N * 24, hipMemcpyHostToDevice, st);
      if (err == hipSuccess)
        err = hipMemcpyAsync(dskip, skip.data(), N, hipMemcpyHostToDevice, st);
      if (err != hipSuccess) return cleanup(), err;
      hipLaunchKernelGGL(gk_scatter, dim3(N), dim3(kBlock), cap * 3 * 8, st, dbins, N,
                         dsd, 3, dcnt, G, 2, der, kRangeSD, cap, dskip,
                         (const int8_t *)nullptr, cg[q]);
      if ((err = hipGetLastError()) != hipSuccess) return cleanup(), err;
    }
    {
      if ((err = upload_sigma()) != hipSuccess) return cleanup(), err;
      const int cap = sigCap;
      const double c0 = G.bin[nd == 3 ? 0 : id1] * G.bin[nd == 3 ? 0 : id1];
      const double c1 = G.bin[nd == 3 ? 1 : id2] * G.bin[nd == 3 ? 1 : id2];
      const double c2 = G.bin[2] * G.bin[2];
      hipLaunchKernelGGL(gk_gather, dim3(N), dim3(kBlock), cap * 3 * 8, st, dbins, N,
                         dsig, 3, sigKind, G, nd == 3 ? 1 : 2, cap, cg[0], cg[1], cg[2],
                         binVol, c0, c1, c2, dsums);
      if ((err = hipGetLastError()) != hipSuccess) return cleanup(), err;
      std::vector<double> s6(6 * (size_t)N);
      err = hipMemcpyAsync(s6.data(), dsums, N * 48, hipMemcpyDeviceToHost, st);
      if (err == hipSuccess) err = hipStreamSynchronize(st);
      if (err != hipSuccess) return cleanup(), err;
      for (int n = 0; n < N; ++n) {
        const double *s = &s6[6 * (size_t)n];
        rXX[n] = rYY[n] = rZZ[n] = 0.0;
        if (nd == 3) {
          rXX[n] = s[0];
          rYY[n] = s[1];
          rZZ[n] = s[2];
          const double rXY = s[3], rXZ = s[4], rYZ = s[5];
          if (!isotropic) {
            double v = 3.0 * std::pow(rXX[n] * rYY[n] * rZZ[n], 1.0 / 3.0);
            if (rYZ > 0.0)
              v = v + 2.0 * rYZ * std::pow(std::pow(rXX[n], 2.0) / rYY[n] / rZZ[n], 1.0 / 6.0);
            if (rXZ > 0.0)
              v = v + 2.0 * rXZ * std::pow(std::pow(rYY[n], 2.0) / rXX[n] / rZZ[n], 1.0 / 6.0);
            if (rXY > 0.0)
              v = v + 2.0 * rXY * std::pow(std::pow(rZZ[n], 2.0) / rXX[n] / rYY[n], 1.0 / 6.0);
            net[n] = v;
          } else {
            net[n] = rXX[n] + 2.0 * rXY + 2.0 * rXZ + rYY[n] + 2.0 * rYZ + rZZ[n];
          }
        } else {
          double *r11 = id1 == 0 ? &rXX[n] : (id1 == 1 ? &rYY[n] : &rZZ[n]);
          double *r22 = id2 == 0 ? &rXX[n] : (id2 == 1 ? &rYY[n] : &rZZ[n]);
          *r11 = s[0];
          *r22 = s[1];
          if (!isotropic) net[n] = 2.0 * s[2] + 2.0 * std::sqrt(s[0] * s[1]);
          else net[n] = 2.0 * s[2] + s[0] + s[1];
        }
      }
    }
    // optimal smoothing and shape (:5404-5724)
    {
      std::vector<double> r11v(N), r22v(N);
      for (int n = 0; n < N; ++n) {
        r11v[n] = id1 == 0 ? rXX[n] : (id1 == 1 ? rYY[n] : rZZ[n]);
        r22v[n] = id2 == 0 ? rXX[n] : (id2 == 1 ? rYY[n] : rZZ[n]);
      }
      if (!isotropic) {
        double minNet = INFINITY;
        for (int n = 0; n < N; ++n)
          if (net[n] > 0.0) minNet = std::min(minNet, net[n]);
        for (int n = 0; n < N; ++n) {
          double sumMain = nd == 2 ? r11v[n] + r22v[n] : rXX[n] + rYY[n] + rZZ[n];
          if (sumMain == 0.0) sumMain = minNet;
          bool iso;
          if (nd == 2)
            iso = (r11v[n] / sumMain > isoThreshold) || (r22v[n] / sumMain > isoThreshold);
          else
            iso = (rXX[n] / sumMain > isoThreshold) || (rYY[n] / sumMain > isoThreshold) ||
                  (rZZ[n] / sumMain > isoThreshold);
          if (iso) net[n] = (1.0 - isoFactor) * net[n] + isoFactor * sumMain;
        }
      }
      for (int n = 0; n < N; ++n) scale[n] = 0.0;
      for (int n = 0; n < N; ++n)
        if (net[n] > 0.0)
          scale[n] = std::pow(fnd * nEst[n] / (std::pow(4.0 * pi, 0.5 * fnd) * net[n]),
                              oneOverNDimPlusFour);
      double maxScale = -INFINITY;
      for (int n = 0; n < N; ++n)
        if (scale[n] > 0.0) maxScale = std::max(maxScale, scale[n]);
      for (int n = 0; n < N; ++n)
        if (net[n] == 0.0) scale[n] = maxScale;
      for (int n = 0; n < N; ++n) {
        double *sh = &shape[3 * (size_t)n];
        for (int d = 0; d < 3; ++d) {
          hk[3 * (size_t)n + d] = 0.0;
          sh[d] = 1.0;
        }
        if (isotropic) continue;
        if (nd == 2) {
          if (r11v[n] == 0.0 || r22v[n] == 0.0) continue;
          sh[id1] = std::pow(net[n] / r11v[n], 1.0 / 8.0);
          if (bounded) sh[id1] = std::max(sh[id1], minKS[id1] / scale[n]);
          sh[id2] = 1.0 / sh[id1];
          const double ns = std::sqrt(sh[id1] * sh[id2]);
          if (ns == 0.0) {
            sh[id1] = sh[id2] = 1.0;
            continue;
          }
          sh[id1] = sh[id1] / ns;
          sh[id2] = sh[id2] / ns;
        } else {
          if (rXX[n] == 0.0 || rYY[n] == 0.0 || rZZ[n] == 0.0) continue;
          sh[0] = std::pow(net[n] / rXX[n], 1.0 / 12.0);
          sh[1] = std::pow(net[n] / rYY[n], 1.0 / 12.0);
          if (bounded) {
            sh[0] = std::max(sh[0], minKS[0] / scale[n]);
            sh[1] = std::max(sh[1], minKS[1] / scale[n]);
          }
          sh[2] = 1.0 / sh[0] / sh[1];
          // the reference's exponent is the single-precision literal 0.333
          const double ns = std::pow(sh[0] * sh[1] * sh[2], (double)0.333f);
          if (ns == 0.0) {
            sh[0] = sh[1] = sh[2] = 1.0;
            continue;
          }
          sh[0] = sh[0] / ns;
          sh[1] = sh[1] / ns;
          sh[2] = sh[2] / ns;
        }
      }
      bool update = false;
      for (int d = 0; d < 3; ++d) {
        if (!G.active[d]) continue;
        for (int n = 0; n < N; ++n) {
          double &hv = hk[3 * (size_t)n + d];
          hv = shape[3 * (size_t)n + d] * scale[n];
        }
        if (!bounded) continue;
        bool anyMax = false, anyMin = false;
        for (int n = 0; n < N; ++n) anyMax |= hk[3 * (size_t)n + d] > maxKS[d];
        if (anyMax) {
          update = true;
          for (int n = 0; n < N; ++n)
            if (hk[3 * (size_t)n + d] > maxKS[d]) hk[3 * (size_t)n + d] = maxKS[d];
        }
        for (int n = 0; n < N; ++n) anyMin |= hk[3 * (size_t)n + d] < minKS[d];
        if (anyMin) {
          update = true;
          for (int n = 0; n < N; ++n)
            if (hk[3 * (size_t)n + d] < minKS[d]) hk[3 * (size_t)n + d] = minKS[d];
        }
      }
      if (update)
        for (int n = 0; n < N; ++n) scale[n] = smoothing_scale(n);
    }
    // new density (:4930-4960)
    if ((err = density_pass(dcnt)) != hipSuccess) return cleanup(), err;
    // convergence (:4962-5038)
    double maxD = -INFINITY, maxDO = -INFINITY;
    for (int n = 0; n < N; ++n) {
      maxD = std::max(maxD, density[n]);
      maxDO = std::max(maxDO, densityOld[n]);
    }
    double sDen = 0.0, sSm = 0.0;
    for (int n = 0; n < N; ++n) {
      double em = 0.0;
      if (densityOld[n] != 0.0)
        em = std::fabs(density[n] / maxD - densityOld[n] / maxDO) / (densityOld[n] / maxDO);
      sDen += std::pow(em, 2.0);
      double es = 0.0;
      if (scaleOld[n] != 0.0) es = std::fabs(scale[n] - scaleOld[n]) / scaleOld[n];
      sSm += std::pow(es, 2.0);
    }
    const double errDensity = std::sqrt(sDen / (double)N);
    const double errSmoothing = std::sqrt(sSm / (double)N);
    const double alm = almise();
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
    densityOld = density;
    scaleOld = scale;
  }
  if (final_weights_dev) {
    // last reconstruction over the mass sums (:4730-4758): the counts'
    // device copy is no longer needed and takes the masses of the bins
    hipLaunchKernelGGL(gk_extract, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       dbins, N, final_weights_dev, dcnt);
    if ((err = hipGetLastError()) != hipSuccess) return cleanup(), err;
    if ((err = density_pass(dcnt)) != hipSuccess) return cleanup(), err;
  }
  if (loops_run) *loops_run = loops;
  cleanup();
  return hipSuccess;
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
