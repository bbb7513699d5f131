// gpkde reconstruction on the device: histogram of the active particles and
// one density pass with a uniform (Silverman) bandwidth.
//
// GridProjectedKDE%ComputeDensity (GridProjectedKDE.F90:3598-4392) with
// nOptimizationLoops = 0 scatters, for every active bin, the bin count times
// the erf-integrated multi-Gaussian kernel and folds the part that falls
// outside the grid back across the face (KernelMultiGaussian.f90:303-390).
// With one bandwidth for all bins that kernel is a product of 1-D kernels
// and the fold is a half-sample symmetric reflection, so the same estimate
// is three 1-D convolutions with mirrored padding -- O(cells * taps) instead
// of O(active bins * taps^3).
//
// The histogram is exact and reproducible: unit masses are integer counts
// (atomic increments, order-free); weighted particles are sorted by bin
// (stable, so ascending particle index within a bin) and every bin sums its
// run sequentially, the order the reference's particle loop adds weights in.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "rwpt_internal.h"

namespace rwpt {
namespace {

struct KdeGrid {
  double origin[3], bin[3];
  int nb[3];
  int active[3];
  int member;             // ensemble member reconstructed, -1 all
  int64_t ncell_member;   // flow cells of one member
};

// the particle enters the reconstruction: active, and in the member asked for
__device__ __forceinline__ bool kde_takes(int32_t status, int32_t cell, const KdeGrid &G) {
  return status == 1 && (G.member < 0 || (int64_t)(cell - 1) / G.ncell_member == G.member);
}

// What the global coordinates of a particle are computed from: the DIS
// face coordinates and cell sizes, and the particle's local coordinates.
struct KdeGeom {
  int ncol, nrow;
  const double *xface, *yface, *delr, *delc, *top, *bot;
  const double *lx, *ly, *lz;
  // unstructured grids: per-cell centres and sizes (null on DIS grids)
  const double *ucx, *ucy, *udx, *udy;
};

// ConvertToModelXYZ (ModflowRectangularGrid.f90:524-549) of particle p in
// 0-based cell c0, clamped to the cell.
__device__ __forceinline__ void kde_point(const KdeGeom &Q, int c0, int64_t p,
                                          double g[3]) {
  double left, right, front, back;
  if (Q.ucx) {
    left = Q.ucx[c0] - (Q.udx[c0] / 2.0);
    right = Q.ucx[c0] + (Q.udx[c0] / 2.0);
    front = Q.ucy[c0] - (Q.udy[c0] / 2.0);
    back = Q.ucy[c0] + (Q.udy[c0] / 2.0);
  } else {
    const int lc = Q.nrow * Q.ncol, r = c0 % lc, i = r / Q.ncol, j = r % Q.ncol;
    const double cx = (Q.xface[j] + Q.xface[j + 1]) / 2.0;
    const double cy = (Q.yface[i] + Q.yface[i + 1]) / 2.0;
    left = cx - (Q.delr[j] / 2.0);
    right = cx + (Q.delr[j] / 2.0);
    front = cy - (Q.delc[i] / 2.0);
    back = cy + (Q.delc[i] / 2.0);
  }
  const double b = Q.bot[c0], t = Q.top[c0];
  const double x = Q.lx[p], y = Q.ly[p], z = Q.lz[p];
  g[0] = fmin(fmax(((1.0 - x) * left) + (x * right), left), right);
  g[1] = fmin(fmax(((1.0 - y) * front) + (y * back), front), back);
  g[2] = fmin(fmax(((1.0 - z) * b) + (z * t), b), t);
}

// The initial-condition reconstruction of MPathRW.F90:1043-1100 takes the
// particles pending release (status 0) whose release time is 0: sel holds 1
// for them (the status value the reconstruction kernels take) and 0 else.
__global__ void kde_select_initial(int64_t n, const int32_t *__restrict__ status,
                                   const double *__restrict__ t,
                                   int32_t *__restrict__ sel) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  sel[p] = (status[p] == 0 && t[p] == 0.0) ? 1 : 0;
}

// Global coordinates (ConvertToModelXYZ) and bin key of every particle;
// inactive particles and points outside the grid get the key ncells.  gx
// may be null: the coordinates are then not kept (the statistics kernel
// recomputes them).
__global__ void kde_keys(int64_t n, const int32_t *__restrict__ status,
                         const int32_t *__restrict__ cellv, const KdeGeom Q,
                         const KdeGrid G, double *__restrict__ gx,
                         double *__restrict__ gy, double *__restrict__ gz,
                         int32_t *__restrict__ keys, int32_t *__restrict__ vals) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t ncells = (int64_t)G.nb[0] * G.nb[1] * G.nb[2];
  vals[p] = (int32_t)p;
  if (!kde_takes(status[p], cellv[p], G)) {
    keys[p] = (int32_t)ncells;
    return;
  }
  double g[3];
  kde_point(Q, cellv[p] - 1, p, g);
  if (gx) {
    gx[p] = g[0];
    gy[p] = g[1];
    gz[p] = g[2];
  }
  // Histogram.F90:276-350: floor((x - origin)/bin) + 1 in every dimension
  // with a bin size (a single-bin dimension still drops points outside it)
  int64_t key = 0, stride = 1;
  for (int d = 0; d < 3; ++d) {
    int id = 0;
    if (G.bin[d] > 0.0) {
      id = (int)floor((g[d] - G.origin[d]) / G.bin[d]);
      if (id < 0 || id >= G.nb[d]) {
        keys[p] = (int32_t)ncells;
        return;
      }
    }
    key += stride * id;
    stride *= G.nb[d];
  }
  keys[p] = (int32_t)key;
}

// Per-bin sums over the sorted runs (first element of a run sums it):
// what 0 the masses, 1 the points, 2 the squared masses.
__global__ void kde_runs(int64_t n, const int32_t *__restrict__ keys,
                         const int32_t *__restrict__ vals,
                         const double *__restrict__ mass, int64_t ncells,
                         int what, double *__restrict__ counts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t k = keys[i];
  if (k >= ncells) return;
  if (i > 0 && keys[i - 1] == k) return;
  double s = 0.0;
  for (int64_t q = i; q < n && keys[q] == k; ++q) {
    const double w = mass ? mass[vals[q]] : 1.0;
    s = s + (what == 0 ? w : (what == 1 ? 1.0 : w * w));
  }
  counts[k] = s;
}

// Per-bin Kish counts (sum w)^2 / sum w^2 where sum w^2 != 0
// (Histogram.F90:476-480); sq holds the squared-mass sums on entry.
__global__ void kde_kish_counts(int64_t n, const double *__restrict__ wsum,
                                double *__restrict__ sq) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && sq[i] != 0.0) sq[i] = (wsum[i] * wsum[i]) / sq[i];
}

// Deterministic sums over the active particles: each block reduces a fixed
// contiguous chunk in a fixed tree, then one block reduces the partials in
// a fixed tree. mode 0: sum of x, y, z, w, w^2, the active count and the
// count of points inside the grid; mode 1: sum of squared deviations from
// mean[].
constexpr int kStatBlocks = 512;
constexpr int kStats = 8;
__global__ void kde_stats(int64_t n, const int32_t *__restrict__ status,
                          const int32_t *__restrict__ cellv, const KdeGrid G,
                          const int32_t *__restrict__ keys, int64_t ncells,
                          const double *__restrict__ gx,
                          const double *__restrict__ gy,
                          const double *__restrict__ gz, const KdeGeom Q,
                          const double *__restrict__ mass, int mode,
                          const double *__restrict__ mean,
                          double *__restrict__ partial) {
  __shared__ double sh[kStats][256];
  double acc[kStats] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  for (int64_t p = lo + threadIdx.x; p < hi; p += blockDim.x) {
    if (!kde_takes(status[p], cellv[p], G)) continue;
    double g[3];
    if (gx) {
      g[0] = gx[p];
      g[1] = gy[p];
      g[2] = gz[p];
    } else {
      kde_point(Q, cellv[p] - 1, p, g);
    }
    if (mode == 0) {
      const double w = mass ? mass[p] : 1.0;
      acc[0] += g[0];
      acc[1] += g[1];
      acc[2] += g[2];
      acc[3] += w;
      acc[4] += w * w;
      acc[5] += 1.0;
      acc[6] += keys[p] < ncells ? 1.0 : 0.0;
    } else {
      const double ex = g[0] - mean[0], ey = g[1] - mean[1], ez = g[2] - mean[2];
      acc[0] += ex * ex;
      acc[1] += ey * ey;
      acc[2] += ez * ez;
    }
  }
  for (int q = 0; q < kStats; ++q) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < kStats; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < kStats; ++q) partial[kStats * blockIdx.x + q] = sh[q][0];
}

// One block of kStatBlocks threads: fixed-tree sum of the block partials.
__global__ void kde_stats_final(const double *__restrict__ partial,
                                double *__restrict__ out) {
  __shared__ double sh[kStats][kStatBlocks];
  const int t = threadIdx.x;
  for (int q = 0; q < kStats; ++q) sh[q][t] = partial[kStats * t + q];
  __syncthreads();
  for (int s = kStatBlocks / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int q = 0; q < kStats; ++q) sh[q][t] += sh[q][t + s];
    __syncthreads();
  }
  if (t < kStats) out[t] = sh[t][0];
}

// Unit masses: integer bin counts by atomics (order-free, exact), then
// converted to the double histogram.
__global__ void kde_count(int64_t n, const int32_t *__restrict__ keys,
                          int64_t ncells, uint32_t *__restrict__ cnt) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t k = keys[p];
  if (k < ncells) atomicAdd(&cnt[k], 1u);
}

__global__ void kde_count_to_hist(int64_t ncells, const uint32_t *__restrict__ cnt,
                                  double *__restrict__ hist) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncells) hist[i] = (double)cnt[i];
}

// zeroes n words (a kernel, not a memset node, so that a captured graph
// replays it correctly)
__global__ void kde_zero_u32(int64_t n, uint32_t *__restrict__ a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0u;
}

__global__ void kde_scale(int64_t n, const double *__restrict__ in,
                          double *__restrict__ out, double scale) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i] * scale;
}

// density * effectiveMass (GridProjectedKDE.F90:5074-5080), * binVolume
// (unitVolume, :4292-4296), * scalingFactor (:4297-4300), in that order
__global__ void kde_finish(int64_t n, double *__restrict__ d, double eff,
                           double vol, double sf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = d[i];
  if (eff != 1.0) v = v * eff;
  if (vol != 1.0) v = v * vol;
  d[i] = v * sf;
}

// One 1-D pass of the separable kernel along `axis` with the reflection of
// KernelMultiGaussian.f90:303-390: a bin's kernel entries below the grid
// fold once onto the grid (target -1-q), then the entries above it fold
// once (target 2n-1-q), and what either fold leaves outside is dropped.
// As a gather, output p takes the source bins p+o (direct), o-1-p (folded
// at the lower face), 2n-1-p+o (at the upper face) and p-2n+o (at the lower
// then the upper face) with weight w[|o|].  For m <= n only single folds
// occur, half-sample symmetric padding (d c b a | a b c d | d c b a); wider
// kernels lose the mass folded beyond the opposite face, as the reference
// does.  Weights w[0..m] are the normalised 1-D kernel at offsets 0..m.
__global__ void kde_pass(const double *__restrict__ in, double *__restrict__ out,
                         int nx, int ny, int nz, int axis,
                         const double *__restrict__ w, int m, double scale) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)nx * ny * nz;
  if (idx >= total) return;
  const int i = (int)(idx % nx);
  const int j = (int)((idx / nx) % ny);
  const int k = (int)(idx / ((int64_t)nx * ny));
  const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  const int64_t stride = axis == 0 ? 1 : (axis == 1 ? nx : (int64_t)nx * ny);
  const int pos = axis == 0 ? i : (axis == 1 ? j : k);
  const int64_t base = idx - (int64_t)pos * stride;
  double s = 0.0;
  for (int o = -m; o <= m; ++o) {
    const double wo = w[o < 0 ? -o : o];
    const int a = pos + o, b = o - 1 - pos, c = 2 * n - 1 - pos + o, d = pos - 2 * n + o;
    if (a >= 0 && a < n) s += wo * in[base + (int64_t)a * stride];
    if (b >= 0 && b < n) s += wo * in[base + (int64_t)b * stride];
    if (c >= 0 && c < n) s += wo * in[base + (int64_t)c * stride];
    if (d >= 0 && d < n) s += wo * in[base + (int64_t)d * stride];
  }
  out[idx] = s * scale;
}

}  // namespace

// Normalised 1-D erf kernel (KernelMultiGaussian.f90
// prComputeKernelWMatrix): offsets 0..m with m = ceiling(3 h / bin); with
// centre_only the centre weight alone, normalised over the full width (a
// dimension left one bin wide by a transposed database lookup).
static void kernel_1d(double h, double bin, std::vector<double> &w,
                      bool centre_only = false) {
  const int m = (int)std::ceil(3 * h / bin);
  const double hd = h / bin, s2 = std::sqrt(2.0);
  std::vector<double> e(m + 1);
  double sum = 0.0;
  for (int i = 0; i <= m; ++i) {
    e[i] = std::erf(((double)i + 0.5) / (hd * s2)) -
           std::erf(((double)i - 0.5) / (hd * s2));
    sum += (i == 0 ? 1.0 : 2.0) * e[i];
  }
  w.assign(centre_only ? 1 : m + 1, 0.0);
  for (int i = 0; i < (int)w.size(); ++i) w[i] = e[i] / sum;
}

namespace {

KdeGrid make_grid(const rwpt_engine *e, const rwpt_kde_options *opt, int &nd) {
  KdeGrid G{};
  nd = 0;
  for (int d = 0; d < 3; ++d) {
    G.origin[d] = opt->origin[d];
    G.bin[d] = opt->bin_size[d];
    G.nb[d] = opt->nbins[d];
    G.active[d] = opt->nbins[d] > 1;
    nd += G.active[d];
  }
  G.member = e->prob.members > 1 ? opt->member : -1;
  G.ncell_member = e->prob.ncell_member;
  return G;
}

KdeGeom make_geom(const rwpt_engine *e) {
  const DeviceProblem &P = e->prob;
  const Particles &S = e->part;
  const UsgData *U = e->usg;
  return KdeGeom{P.ncol, P.nrow, e->xface, e->yface, P.delr, P.delc, model_top(P), P.bot,
                 S.lx, S.ly, S.lz, U ? U->cx : nullptr, U ? U->cy : nullptr,
                 U ? U->dx : nullptr, U ? U->dy : nullptr};
}

// The particle statistics of the reconstruction: hs = sums of x, y, z, w,
// w^2, the active count and the count inside the grid; dev2 = the sums of
// squared deviations from the mean.  Two deterministic fixed-tree passes
// and two host read-backs.
hipError_t kde_moments(rwpt_engine *e, const KdeGrid &G, const KdeGeom &Q,
                       const int32_t *sel, const double *gx, const double *gy,
                       const double *gz, const double *mass_dev, int64_t ncells,
                       double *partial, double *sums, double hs[kStats],
                       double dev2[kStats]) {
  const Particles &S = e->part;
  const int64_t n = S.n;
  hipStream_t st = e->stream;
  hipError_t err = hipSuccess;
  for (int q = 0; q < kStats; ++q) hs[q] = dev2[q] = 0.0;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(kde_stats, dim3(kStatBlocks), dim3(256), 0, st, n, sel, S.cell,
                     G, S.keys[0], ncells, gx, gy, gz, Q, mass_dev, 0,
                     (const double *)nullptr, partial);
  hipLaunchKernelGGL(kde_stats_final, dim3(1), dim3(kStatBlocks), 0, st, partial, sums);
  if ((err = hipMemcpyAsync(hs, sums, kStats * sizeof(double), hipMemcpyDeviceToHost,
                            st)) != hipSuccess)
    return err;
  if ((err = hipStreamSynchronize(st)) != hipSuccess) return err;
  const double nact = hs[5];
  if (nact <= 0) return hipSuccess;
  double mean[3];
  for (int d = 0; d < 3; ++d) mean[d] = hs[d] / nact;
  if ((err = hipMemcpyAsync(sums + kStats, mean, sizeof mean, hipMemcpyHostToDevice,
                            st)) != hipSuccess)
    return err;
  hipLaunchKernelGGL(kde_stats, dim3(kStatBlocks), dim3(256), 0, st, n, sel, S.cell,
                     G, S.keys[0], ncells, gx, gy, gz, Q, mass_dev, 1,
                     (const double *)(sums + kStats), partial);
  hipLaunchKernelGGL(kde_stats_final, dim3(1), dim3(kStatBlocks), 0, st, partial, sums);
  if ((err = hipMemcpyAsync(dev2, sums, kStats * sizeof(double), hipMemcpyDeviceToHost,
                            st)) != hipSuccess)
    return err;
  return hipStreamSynchronize(st);
}

struct KdeBandwidth {
  double h = 0.0, effMass = 1.0, nPoints = 0.0, binVolume = 1.0;
};

// Bin volume, effective mass, point count and initial smoothing of
// ComputeDensity for the statistics hs and dev2.
KdeBandwidth kde_bandwidth(const rwpt_kde_options *opt, const KdeGrid &G, int nd,
                           bool weighted, const double hs[kStats],
                           const double dev2[kStats]) {
  KdeBandwidth B;
  // bin volume and bin distance over the dimensions with a bin size
  // (Histogram.F90:121, 154-155)
  int nbd = 0;
  for (int d = 0; d < 3; ++d)
    if (G.bin[d] > 0.0) {
      B.binVolume *= G.bin[d];
      ++nbd;
    }
  const double binDistance = std::pow(B.binVolume, 1.0 / (nbd > 0 ? nbd : 1));
  // effective mass (Histogram.F90:326-343; formats 2 and 3 keep 1) and
  // point count: the reference divides the weighted counts by the
  // effective mass and multiplies the density back by it; both cancel in
  // a single pass, so the bins hold the plain sums of the masses
  const int wfmt = opt->effective_weight_format;
  const double nact = hs[5];
  B.nPoints = nact;
  if (weighted) {
    if (wfmt == 0) {
      const double neff = hs[3] * hs[3] / hs[4];
      B.effMass = hs[3] / neff;
    } else if (wfmt == 1) {
      B.effMass = hs[3] / nact;
    }
  } else {
    // unit masses: nPoints counts the points inside the grid
    B.nPoints = hs[6];
  }
  // initial smoothing: Silverman (1986) over the standard deviations
  // (GridProjectedKDE.F90:4034-4054, 4109-4117), or a factor of the bin
  // distance (:643-648)
  const double fac = opt->initial_smoothing_factor;
  if (opt->initial_smoothing == 1) {
    if (nact > 0 && nd > 0) B.h = fac * binDistance;
  } else if (nact > 0 && nd > 0) {
    double sscale = 1.0;
    for (int d = 0; d < 3; ++d)
      if (G.active[d]) sscale *= std::sqrt(dev2[d] / nact);
    sscale = std::pow(sscale, 1.0 / nd);
    B.h = sscale * std::pow(4.0 / ((nd + 2.0) * B.nPoints), 1.0 / (nd + 4.0));
    if (sscale == 0.0) B.h = 2.0 * binDistance;
    else if (fac > 0.0 && fac != 1.0) B.h = B.h * fac;
  }
  return B;
}

// The 1-D kernels of the one-pass density (the same bandwidth in every
// active dimension, or the database entry it looks up), concatenated:
// dimension d has m[d] + 1 weights from off[d].
void kde_weights(const rwpt_kde_options *opt, const KdeGrid &G, double h,
                 std::vector<double> &all, int off[3], int m[3]) {
  double hv[3] = {0, 0, 0};
  int collapse = -1;
  for (int d = 0; d < 3; ++d) hv[d] = G.active[d] ? h : 0.0;
  if (opt->kernel_database) {
    KernelDb db;
    db.build(opt->min_h_over_delta, opt->delta_h_over_delta, opt->max_h_over_delta,
             opt->bin_size, opt->nbins);
    const double hin[3] = {hv[0], hv[1], hv[2]};
    collapse = db.w(hin, hv);
  }
  all.clear();
  for (int d = 0; d < 3; ++d) {
    off[d] = m[d] = 0;
    if (!G.active[d]) continue;
    std::vector<double> w;
    kernel_1d(hv[d], G.bin[d], w, d == collapse);
    m[d] = (int)w.size() - 1;
    off[d] = (int)all.size();
    all.insert(all.end(), w.begin(), w.end());
  }
}

// The separable passes from the histogram to the density.  With tmpB null
// the passes alternate between the density and tmpA so that the last one
// ends in the density.
hipError_t kde_passes(hipStream_t st, const KdeGrid &G, const rwpt_kde_options *opt,
                      double binVolume, const double *hist, double *density,
                      double *tmpA, double *tmpB, const double *weights,
                      const int off[3], const int m[3]) {
  const int blk = 256;
  const int64_t ncells = (int64_t)G.nb[0] * G.nb[1] * G.nb[2];
  int last = -1, passes = 0;
  for (int d = 0; d < 3; ++d)
    if (G.active[d]) {
      last = d;
      ++passes;
    }
  const double *src = hist;
  int ping = 0, k = 0;
  for (int d = 0; d < 3; ++d) {
    if (!G.active[d]) continue;
    double *dst;
    if (d == last) dst = density;
    else if (tmpB) dst = ping ? tmpB : tmpA;
    else dst = ((passes - 1 - k) % 2 == 0) ? density : tmpA;
    const double sc = d != last ? 1.0
                      : (opt->unit_volume ? opt->scaling_factor
                                          : opt->scaling_factor / binVolume);
    hipLaunchKernelGGL(kde_pass, dim3((unsigned)((ncells + blk - 1) / blk)), dim3(blk),
                       0, st, src, dst, G.nb[0], G.nb[1], G.nb[2], d, weights + off[d],
                       m[d], sc);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    src = dst;
    ping ^= 1;
    ++k;
  }
  return hipSuccess;
}

void record_stats(double *stats_host, const KdeBandwidth &B, double nact) {
  if (!stats_host) return;
  stats_host[0] = B.h;
  stats_host[1] = nact;
  stats_host[2] = B.nPoints;
  stats_host[3] = B.effMass;
}

}  // namespace

hipError_t launch_kde(rwpt_engine *e, const rwpt_kde_options *opt,
                      const double *mass_dev, double *density_dev,
                      double *hist_dev, double *stats_host) {
  Particles &S = e->part;
  const int64_t n = S.n;
  int nd = 0;
  const KdeGrid G = make_grid(e, opt, nd);
  const KdeGeom Q = make_geom(e);
  const int64_t ncells = (int64_t)G.nb[0] * G.nb[1] * G.nb[2];
  hipStream_t st = e->stream;
  // scratch: gx gy gz (3n), keys/vals reuse the walk's work lists, two grid
  // buffers, stats partials, the selection of an initial-condition pass
  const bool initial = opt->initial != 0;
  const int64_t need = 3 * n + 2 * ncells + kStats * kStatBlocks + 2 * kStats +
                       (initial ? (n + 1) / 2 : 0);
  hipError_t err = hipSuccess;
  if (need > e->kde_scratch_size) {
    if (e->kde_scratch) (void)hipFree(e->kde_scratch);
    e->kde_scratch = nullptr;
    e->kde_scratch_size = 0;
    if ((err = hipMalloc(&e->kde_scratch, need * sizeof(double))) != hipSuccess)
      return err;
    e->kde_scratch_size = need;
  }
  double *gx = e->kde_scratch, *gy = gx + n, *gz = gy + n;
  double *tmpA = gz + n, *tmpB = tmpA + ncells;
  double *partial = tmpB + ncells, *sums = partial + kStats * kStatBlocks;
  // the particles taken: status 1, or the selection of the initial condition
  const int32_t *sel = S.status;
  KernelTimer &TH = e->timers[1], &TD = e->timers[2];
  (void)hipEventRecord(TH.start, st);
  const int blk = 256;
  if (n > 0 && initial) {
    int32_t *s = reinterpret_cast<int32_t *>(sums + 2 * kStats);
    hipLaunchKernelGGL(kde_select_initial, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk),
                       0, st, n, S.status, S.t, s);
    sel = s;
  }
  if (n > 0) {
    hipLaunchKernelGGL(kde_keys, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk), 0, st,
                       n, sel, S.cell, Q, G, gx, gy, gz, S.keys[0], S.vals[0]);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  }
  if (n > 0 && !mass_dev) {
    uint32_t *cnt = reinterpret_cast<uint32_t *>(tmpA);
    if ((err = hipMemsetAsync(cnt, 0, ncells * sizeof(uint32_t), st)) != hipSuccess)
      return err;
    hipLaunchKernelGGL(kde_count, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk), 0,
                       st, n, S.keys[0], ncells, cnt);
    hipLaunchKernelGGL(kde_count_to_hist, dim3((unsigned)((ncells + blk - 1) / blk)),
                       dim3(blk), 0, st, ncells, cnt, hist_dev);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  } else {
    if ((err = hipMemsetAsync(hist_dev, 0, ncells * sizeof(double), st)) != hipSuccess)
      return err;
  }
  if (n > 0 && mass_dev) {
    // weighted: stable sort by bin so that every bin adds its masses in
    // ascending particle order
    int bits = 1;
    while ((1ll << bits) <= ncells) ++bits;
    size_t bytes = 0;
    err = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, S.keys[0], S.keys[1],
                                             S.vals[0], S.vals[1], (int)n, 0, bits, st);
    if (err != hipSuccess) return err;
    if (bytes > S.sort_tmp_bytes) {
      if (S.sort_tmp) (void)hipFree(S.sort_tmp);
      S.sort_tmp = nullptr;
      S.sort_tmp_bytes = 0;
      if ((err = hipMalloc(&S.sort_tmp, bytes)) != hipSuccess) return err;
      S.sort_tmp_bytes = bytes;
    }
    bytes = S.sort_tmp_bytes;
    err = hipcub::DeviceRadixSort::SortPairs(S.sort_tmp, bytes, S.keys[0], S.keys[1],
                                             S.vals[0], S.vals[1], (int)n, 0, bits, st);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(kde_runs, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk), 0,
                       st, n, S.keys[1], S.vals[1], mass_dev, ncells, 0, hist_dev);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  }
  (void)hipEventRecord(TH.stop, st);
  // statistics for the bandwidth and the effective mass
  double hs[kStats], dev2[kStats];
  if ((err = kde_moments(e, G, Q, sel, gx, gy, gz, mass_dev, ncells, partial, sums, hs,
                         dev2)) != hipSuccess)
    return err;
  const KdeBandwidth B = kde_bandwidth(opt, G, nd, mass_dev != nullptr, hs, dev2);
  const double h = B.h, effMass = B.effMass, nPoints = B.nPoints;
  record_stats(stats_host, B, hs[5]);
  const int wfmt = opt->effective_weight_format;
  (void)hipEventRecord(TD.start, st);
  e->kde_loops = 0;
  if (opt->mode == 1 && opt->n_opt_loops > 0 && h > 0.0) {
    // bandwidth optimisation: the loop sees the counts the reference's
    // weighted histogram holds: mass sums / effective mass (formats 0, 1),
    // point counts (2) or per-bin Kish counts (3); formats 2 and 3 then
    // scatter the mass sums with the final kernels
    const unsigned gb = (unsigned)((ncells + blk - 1) / blk);
    const bool final_masses = mass_dev && (wfmt == 2 || wfmt == 3);
    if (mass_dev && final_masses) {
      if ((err = hipMemsetAsync(tmpA, 0, ncells * sizeof(double), st)) != hipSuccess)
        return err;
      if (n > 0)
        hipLaunchKernelGGL(kde_runs, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk),
                           0, st, n, S.keys[1], S.vals[1], mass_dev, ncells,
                           wfmt == 2 ? 1 : 2, tmpA);
      if (wfmt == 3)
        hipLaunchKernelGGL(kde_kish_counts, dim3(gb), dim3(blk), 0, st, ncells,
                           hist_dev, tmpA);
    } else {
      if ((err = hipMemcpyAsync(tmpA, hist_dev, ncells * sizeof(double),
                                hipMemcpyDeviceToDevice, st)) != hipSuccess)
        return err;
      if (mass_dev)
        hipLaunchKernelGGL(kde_scale, dim3(gb), dim3(blk), 0, st, ncells, tmpA, tmpA,
                           1.0 / effMass);
    }
    if ((err = hipGetLastError()) != hipSuccess) return err;
    int loops = 0;
    if ((err = launch_gpkde_opt(e, opt, tmpA, h, nPoints,
                                final_masses ? hist_dev : nullptr, density_dev,
                                &loops)) != hipSuccess)
      return err;
    e->kde_loops = loops;
    hipLaunchKernelGGL(kde_finish, dim3((unsigned)((ncells + blk - 1) / blk)),
                       dim3(blk), 0, st, ncells, density_dev, mass_dev ? effMass : 1.0,
                       opt->unit_volume ? B.binVolume : 1.0, opt->scaling_factor);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  } else if (opt->mode == 0 || h <= 0.0) {
    // histogram only: the raw density counts / binVolume
    hipLaunchKernelGGL(kde_scale, dim3((unsigned)((ncells + blk - 1) / blk)),
                       dim3(blk), 0, st, ncells, hist_dev, density_dev,
                       opt->scaling_factor / B.binVolume);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  } else {
    std::vector<double> all;
    int off[3], m[3];
    kde_weights(opt, G, h, all, off, m);
    if ((int64_t)all.size() > e->kde_weights_size) {
      if (e->kde_weights) (void)hipFree(e->kde_weights);
      e->kde_weights = nullptr;
      e->kde_weights_size = 0;
      if ((err = hipMalloc(&e->kde_weights, all.size() * sizeof(double))) != hipSuccess)
        return err;
      e->kde_weights_size = (int64_t)all.size();
    }
    if ((err = hipMemcpyAsync(e->kde_weights, all.data(), all.size() * sizeof(double),
                              hipMemcpyHostToDevice, st)) != hipSuccess)
      return err;
    if ((err = kde_passes(st, G, opt, B.binVolume, hist_dev, density_dev, tmpA, tmpB,
                          e->kde_weights, off, m)) != hipSuccess)
      return err;
  }
  (void)hipEventRecord(TD.stop, st);
  return hipStreamSynchronize(st);
}

// ---------------------------------------------------------------------------
// Allocation-free one-pass reconstruction (rwpt_kde_device).  Its scratch
// memory -- the count grid that becomes the ping-pong grid of the
// separable passes, the statistics partials and the 1-D kernel weights --
// is one file-scope device array, of which each GPU holds its own copy for
// the life of the process.  Host code never uses the array's name as an
// address: hipGetSymbolAddress gives each device's address once, at engine
// creation with that device current, and that address is what the kernels
// receive.
constexpr int64_t kNoAllocGrid = RWPT_KDE_DEVICE_MAX_CELLS;
constexpr int64_t kNoAllocTaps = RWPT_KDE_DEVICE_MAX_TAPS;
constexpr int64_t kNoAllocDoubles =
    kNoAllocGrid + kStats * kStatBlocks + 2 * kStats + kNoAllocTaps;

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t kde_device_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void *ptr = nullptr;
  hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (err != hipSuccess) return err;
  g_scratch_ptr[device] = static_cast<double *>(ptr);
  return hipSuccess;
}

hipError_t launch_kde_device(rwpt_engine *e, const rwpt_kde_options *opt,
                             double *density_dev, double *hist_dev, double *stats_host) {
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (e->device >= 0 && e->device < kMaxDevices) base = g_scratch_ptr[e->device];
  }
  if (!base) return hipErrorNotInitialized;
  Particles &S = e->part;
  const int64_t n = S.n;
  int nd = 0;
  const KdeGrid G = make_grid(e, opt, nd);
  const KdeGeom Q = make_geom(e);
  const int64_t ncells = (int64_t)G.nb[0] * G.nb[1] * G.nb[2];
  if (ncells > kNoAllocGrid) return hipErrorInvalidValue;
  double *gridA = base;
  double *partial = gridA + kNoAllocGrid, *sums = partial + kStats * kStatBlocks;
  double *weights = sums + 2 * kStats;
  hipStream_t st = e->stream;
  hipError_t err = hipSuccess;
  KernelTimer &TH = e->timers[1], &TD = e->timers[2];
  (void)hipEventRecord(TH.start, st);
  const int blk = 256;
  const unsigned gcells = (unsigned)((ncells + blk - 1) / blk);
  // unit masses: exact integer counts in the scratch grid (cleared first:
  // it holds the previous call's values)
  uint32_t *cnt = reinterpret_cast<uint32_t *>(gridA);
  hipLaunchKernelGGL(kde_zero_u32, dim3(gcells), dim3(blk), 0, st, ncells, cnt);
  if (n > 0) {
    hipLaunchKernelGGL(kde_keys, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk), 0, st,
                       n, S.status, S.cell, Q, G, (double *)nullptr, (double *)nullptr,
                       (double *)nullptr, S.keys[0], S.vals[0]);
    hipLaunchKernelGGL(kde_count, dim3((unsigned)((n + blk - 1) / blk)), dim3(blk), 0,
                       st, n, S.keys[0], ncells, cnt);
  }
  hipLaunchKernelGGL(kde_count_to_hist, dim3(gcells), dim3(blk), 0, st, ncells, cnt,
                     hist_dev);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  (void)hipEventRecord(TH.stop, st);
  double hs[kStats], dev2[kStats];
  if ((err = kde_moments(e, G, Q, S.status, nullptr, nullptr, nullptr, nullptr, ncells,
                         partial, sums, hs, dev2)) != hipSuccess)
    return err;
  const KdeBandwidth B = kde_bandwidth(opt, G, nd, false, hs, dev2);
  record_stats(stats_host, B, hs[5]);
  (void)hipEventRecord(TD.start, st);
  e->kde_loops = 0;
  if (opt->mode == 0 || B.h <= 0.0) {
    hipLaunchKernelGGL(kde_scale, dim3(gcells), dim3(blk), 0, st, ncells, hist_dev,
                       density_dev, opt->scaling_factor / B.binVolume);
    if ((err = hipGetLastError()) != hipSuccess) return err;
  } else {
    std::vector<double> all;
    int off[3], m[3];
    kde_weights(opt, G, B.h, all, off, m);
    if ((int64_t)all.size() > kNoAllocTaps) return hipErrorInvalidConfiguration;
    if ((err = hipMemcpyAsync(weights, all.data(), all.size() * sizeof(double),
                              hipMemcpyHostToDevice, st)) != hipSuccess)
      return err;
    if ((err = kde_passes(st, G, opt, B.binVolume, hist_dev, density_dev, gridA, nullptr,
                          weights, off, m)) != hipSuccess)
      return err;
  }
  (void)hipEventRecord(TD.stop, st);
  // the host vector of weights is read by the copy above before it goes
  return hipStreamSynchronize(st);
}

}  // namespace rwpt
