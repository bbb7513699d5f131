// Random-walk advance: the plain FP64 walk, the orchestration of one
// tracking interval for either arithmetic type, rng stream seeding and
// global coordinates.
//
// This translation unit is built with -ffp-contract=off: floating-point
// expressions keep the reference operand order and with the EULERIAN solver
// every result is the double the reference forms.  The FP32-local walk is
// instantiated in rwpt_walk32.hip, built with contraction.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "rwpt_walk_usg.h"

namespace rwpt {
namespace {

// Stream p = J^(offset+p) base, J the jump matrix over GF(2)^128; the k-th
// 128-column matrix in mats is J^(2^k).
__global__ void seed_streams(int64_t n, int64_t offset, uint64_t b0,
                             uint64_t b1, const uint64_t *__restrict__ mats,
                             uint64_t *__restrict__ rng) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint64_t s0 = b0, s1 = b1;
  const uint64_t idx = (uint64_t)(offset + p);
  for (int k = 0; k < 64; ++k) {
    if (!((idx >> k) & 1ull)) continue;
    const uint64_t *M = mats + (size_t)k * 256;
    uint64_t r0 = 0, r1 = 0;
    for (int b = 0; b < 128; ++b) {
      const uint64_t bit = b < 64 ? (s0 >> b) & 1ull : (s1 >> (b - 64)) & 1ull;
      const uint64_t m = 0ull - bit;
      r0 ^= M[2 * b] & m;
      r1 ^= M[2 * b + 1] & m;
    }
    s0 = r0;
    s1 = r1;
  }
  rng[2 * p] = s0;
  rng[2 * p + 1] = s1;
}

__global__ void global_xyz_kernel(int64_t n, const int32_t *__restrict__ cellv,
                                  const double *__restrict__ lxv,
                                  const double *__restrict__ lyv,
                                  const double *__restrict__ lzv, int ncol,
                                  int nrow, const double *__restrict__ xface,
                                  const double *__restrict__ yface,
                                  const double *__restrict__ delr,
                                  const double *__restrict__ delc,
                                  const double *__restrict__ top,
                                  const double *__restrict__ bot,
                                  double *__restrict__ gx,
                                  double *__restrict__ gy,
                                  double *__restrict__ gz) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int c0 = cellv[p] - 1;
  const int lc = nrow * ncol, r = c0 % lc, i = r / ncol, j = r % ncol;
  // ConvertToModelXYZ with CellX/CellY of RectangularGridDis.f90:352-361
  const double cx = (xface[j] + xface[j + 1]) / 2.0;
  const double cy = (yface[i] + yface[i + 1]) / 2.0;
  const double left = cx - (delr[j] / 2.0), right = cx + (delr[j] / 2.0);
  const double front = cy - (delc[i] / 2.0), back = cy + (delc[i] / 2.0);
  const double bt = bot[c0], tp = top[c0];
  const double x = lxv[p], y = lyv[p], z = lzv[p];
  gx[p] = fmin(fmax(((1.0 - x) * left) + (x * right), left), right);
  gy[p] = fmin(fmax(((1.0 - y) * front) + (y * back), front), back);
  gz[p] = fmin(fmax(((1.0 - z) * bt) + (z * tp), bt), tp);
}

}  // namespace

// J^(2^k) for k < kmax as 128 columns of 2 words each.
void jump_power_matrices(int kmax, std::vector<uint64_t> &out) {
  auto next = [](uint64_t s[2]) {
    const uint64_t s1 = s[0], s2 = s[1] ^ s[0];
    s[0] = ((s1 << 24) | (s1 >> 40)) ^ s2 ^ (s2 << 16);
    s[1] = (s2 << 37) | (s2 >> 27);
  };
  auto jump = [&](uint64_t s[2]) {
    const uint64_t jm[2] = {~0x206ffd6b270aab5aull, 0x170865df4b3201fcull};
    uint64_t a = 0, b = 0;
    for (int i = 0; i < 2; ++i)
      for (int bit = 0; bit < 64; ++bit) {
        if (jm[i] & (1ull << bit)) {
          a ^= s[0];
          b ^= s[1];
        }
        next(s);
      }
    s[0] = a;
    s[1] = b;
  };
  out.assign((size_t)kmax * 256, 0);
  for (int b = 0; b < 128; ++b) {
    uint64_t s[2] = {b < 64 ? 1ull << b : 0ull, b >= 64 ? 1ull << (b - 64) : 0ull};
    jump(s);
    out[2 * b] = s[0];
    out[2 * b + 1] = s[1];
  }
  auto apply = [](const uint64_t *M, const uint64_t v[2], uint64_t r[2]) {
    r[0] = r[1] = 0;
    for (int b = 0; b < 128; ++b) {
      const uint64_t bit = b < 64 ? (v[0] >> b) & 1 : (v[1] >> (b - 64)) & 1;
      if (bit) {
        r[0] ^= M[2 * b];
        r[1] ^= M[2 * b + 1];
      }
    }
  };
  for (int k = 1; k < kmax; ++k) {
    const uint64_t *P = &out[(size_t)(k - 1) * 256];
    uint64_t *Q = &out[(size_t)k * 256];
    for (int b = 0; b < 128; ++b) apply(P, &P[2 * b], &Q[2 * b]);
  }
}

hipError_t launch_seed_streams(rwpt_engine *e, int64_t n, int64_t offset) {
  if (n <= 0) return hipSuccess;
  const int block = 256;
  hipLaunchKernelGGL(seed_streams, dim3((unsigned)((n + block - 1) / block)),
                     dim3(block), 0, e->stream, n, offset, e->opt.base_seed[0],
                     e->opt.base_seed[1], e->jump_mats, e->part.rng);
  return hipGetLastError();
}

static int radix_bits(int64_t ncell) {
  int b = 1;
  while ((1ll << b) <= ncell) ++b;
  return b;
}

hipError_t ensure_sort_storage(rwpt_engine *e) {
  Particles &P = e->part;
  size_t bytes = 0;
  hipError_t err = hipcub::DeviceRadixSort::SortPairs(
      nullptr, bytes, P.keys[0], P.keys[1], P.vals[0], P.vals[1],
      (int)std::max<int64_t>(P.cap, 1), 0,
      radix_bits(e->usg ? std::max<int64_t>(e->prob.ncell, e->usg->nunit) : e->prob.ncell),
      e->stream);
  if (err != hipSuccess) return err;
  if (bytes > P.sort_tmp_bytes) {
    if (P.sort_tmp) (void)hipFree(P.sort_tmp);
    P.sort_tmp = nullptr;
    P.sort_tmp_bytes = 0;
    err = hipMalloc(&P.sort_tmp, bytes);
    if (err != hipSuccess) return err;
    P.sort_tmp_bytes = bytes;
  }
  return hipSuccess;
}

UsgWalk usg_walk_view(rwpt_engine *e) {
  UsgData &U = *e->usg;
  return UsgWalk{U.cell_unit, U.side, U.off, U.cx, U.cy, U.dx, U.dy, e->prob.ibound,
                 U.unit, U.pcell, U.ex, U.ey, U.ez, U.et, U.px, U.py, U.pz};
}

// The unstructured walk for the engine's dimensions, model and advection.
hipError_t walk_launch_usg(rwpt_engine *e, const WalkParams &W, dim3 grid, int block,
                           const int32_t *work, int64_t cnt, int budget, int32_t *nk,
                           int32_t *nv) {
  const UsgWalk G = usg_walk_view(e);
  const ParticlePtrs<double> S = particle_ptrs<double>(e->part);
  const int mode = (W.ndim - 1) * 4 + W.disp_model * 2 + (W.advection_kind - 1);
#define RWPT_WALK_USG(ND, MO, AD)                                                     \
  case (ND - 1) * 4 + MO * 2 + (AD - 1):                                              \
    hipLaunchKernelGGL((walk_kernel_usg<ND, MO, AD>), grid, dim3(block), 0, e->stream, \
                       e->prob.cells, G, work, cnt, S, W, budget, nk, nv, e->counters,  \
                       e->nwork_dev);                                                  \
    break;
  switch (mode) {
    RWPT_WALK_USG(1, 0, 1) RWPT_WALK_USG(1, 0, 2) RWPT_WALK_USG(1, 1, 1)
    RWPT_WALK_USG(1, 1, 2) RWPT_WALK_USG(2, 0, 1) RWPT_WALK_USG(2, 0, 2)
    RWPT_WALK_USG(2, 1, 1) RWPT_WALK_USG(2, 1, 2) RWPT_WALK_USG(3, 0, 1)
    RWPT_WALK_USG(3, 0, 2) RWPT_WALK_USG(3, 1, 1) RWPT_WALK_USG(3, 1, 2)
    default: return hipErrorInvalidValue;
  }
#undef RWPT_WALK_USG
  return hipGetLastError();
}

hipError_t walk_launch64(rwpt_engine *e, const WalkParams &W, dim3 grid, int block,
                         const int32_t *work, int64_t cnt, int budget, int32_t *nk,
                         int32_t *nv) {
  if (e->prob.pow2_cells)
    return launch_walk_typed<double, 0, 1>(e, W, grid, block, work, cnt, budget, nk, nv);
  return launch_walk_typed<double, 0, 0>(e, W, grid, block, work, cnt, budget, nk, nv);
}

// One tracking interval: prepare, then walk launches of at most `budget`
// iterations per particle over the cell-sorted list of unfinished
// particles, until every particle has reached t_stop or stopped.
WalkParams walk_params(const rwpt_engine *e, double t_stop) {
  WalkParams W{};
  W.advection_kind = e->opt.advection_kind;
  W.time_step_kind = e->opt.time_step_kind;
  W.ndim = 0;
  for (int d = 0; d < 3; ++d)
    if (e->opt.dimension_mask[d]) W.dims[W.ndim++] = d + 1;
  W.disp_model = e->prob.disp_model;
  W.backward = e->opt.backward_tracking;
  W.stop_zone = e->opt.stop_zone;
  W.stop_weak_sinks = e->opt.stop_at_weak_sinks;
  W.stop_weak_sources = e->opt.stop_at_weak_sources;
  W.steady_state = e->opt.steady_state;
  W.t_engine = e->t_engine;
  W.t_stop = t_stop;
  return W;
}

// Graph mode: the list written by the last launch becomes the list the
// next one reads (counters[2] -> counters[3]), by a kernel, so that a
// captured graph holds no memset node.
__global__ void rotate_list(unsigned long long *counters) {
  counters[3] = counters[2];
  counters[2] = 0;
}

// Records the start or end of a timed span of walk launches.
hipError_t timed_event(rwpt_engine *e, int64_t pair, int end, int32_t launches) {
  if ((size_t)(2 * pair + 2) > e->walk_events.size()) {
    hipEvent_t ev[2];
    hipError_t err;
    if ((err = hipEventCreate(&ev[0])) != hipSuccess) return err;
    if ((err = hipEventCreate(&ev[1])) != hipSuccess) return err;
    e->walk_events.push_back(ev[0]);
    e->walk_events.push_back(ev[1]);
  }
  if ((size_t)(pair + 1) > e->walk_event_launches.size())
    e->walk_event_launches.resize(pair + 1);
  e->walk_event_launches[pair] = launches;
  return hipEventRecord(e->walk_events[2 * pair + end], e->stream);
}

// The rest of an interval's walk as replays of a captured graph of K
// launches (K even), from the list in keys[0]/vals[0] whose length is in
// counters[2]: launch k reads list k % 2 with the length in counters[3] and
// writes list (k + 1) % 2, its grid covering every particle.  The lists are
// not sorted between launches (the results do not depend on it); the host
// reads the list length once per replay instead of once per launch.  The
// graph is captured for every advance (the kernels take t_stop by value)
// and updated into the engine's executable when its shape allows.
hipError_t advance_graph(rwpt_engine *e, const WalkParams &W, unsigned long long cnt,
                         int64_t max_launches, int64_t launches, int64_t timed) {
  Particles &P = e->part;
  const int K = e->graph_launches;
  const int budget = std::max(1, e->tune.budget);
  const int block = e->tune.block;
  const bool usg = e->usg != nullptr;
  const bool fp32 = e->opt.precision != 0;
  const dim3 grid((unsigned)((P.n + block - 1) / block));
  hipStream_t st = e->stream;
  hipError_t err = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) return err;
  e->nwork_dev = e->counters + 3;
  for (int k = 0; k < K && err == hipSuccess; ++k) {
    hipLaunchKernelGGL(rotate_list, dim3(1), dim3(1), 0, st, e->counters);
    const int a = k & 1, b = a ^ 1;
    const int32_t *work = P.vals[a];
    err = usg    ? walk_launch_usg(e, W, grid, block, work, P.n, budget, P.keys[b], P.vals[b])
          : fp32 ? walk_launch32(e, W, grid, block, work, P.n, budget, P.keys[b], P.vals[b])
                 : walk_launch64(e, W, grid, block, work, P.n, budget, P.keys[b], P.vals[b]);
  }
  e->nwork_dev = nullptr;
  hipGraph_t graph = nullptr;
  hipError_t end = hipStreamEndCapture(st, &graph);
  if (err == hipSuccess) err = end;
  if (err != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return err;
  }
  if (e->graph_exec) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(e->graph_exec, graph, &bad, &res) != hipSuccess ||
        res != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(e->graph_exec);
      e->graph_exec = nullptr;
    }
  }
  if (!e->graph_exec) err = hipGraphInstantiate(&e->graph_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (err != hipSuccess) return err;
  while (cnt > 0) {
    if (launches >= max_launches) {  // step guard reached
      const int32_t *work = P.vals[0];
      if (usg) {
        hipLaunchKernelGGL(terminate_walk_usg, dim3((unsigned)((cnt + 255) / 256)),
                           dim3(256), 0, st, usg_walk_view(e), work, (int64_t)cnt,
                           particle_ptrs<double>(P));
        err = hipGetLastError();
      } else {
        err = fp32 ? walk_terminate32(e, work, (int64_t)cnt)
                   : launch_terminate_typed<double>(e, work, (int64_t)cnt);
      }
      if (err != hipSuccess) return err;
      break;
    }
    if ((err = timed_event(e, timed, 0, K)) != hipSuccess) return err;
    if ((err = hipGraphLaunch(e->graph_exec, st)) != hipSuccess) return err;
    if ((err = timed_event(e, timed, 1, K)) != hipSuccess) return err;
    timed += 1;
    launches += K;
    err = hipMemcpyAsync(&cnt, e->counters + 2, sizeof cnt, hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    if (err != hipSuccess) return err;
  }
  e->last_launches = launches;
  e->last_timed = timed;
  return hipSuccess;
}

hipError_t launch_advance(rwpt_engine *e, double t_stop) {
  Particles &P = e->part;
  const WalkParams W = walk_params(e, t_stop);
  e->last_launches = 0;
  e->last_timed = 0;
  hipError_t err = hipMemsetAsync(e->counters, 0, 4 * sizeof(unsigned long long),
                                  e->stream);
  if (err != hipSuccess || P.n == 0) return err;
  const bool fp32 = e->opt.precision != 0;
  err = fp32 ? walk_prepare32(e, W) : launch_prepare_typed<double>(e, W);
  if (err != hipSuccess) return err;
  unsigned long long cnt = 0;
  err = hipMemcpyAsync(&cnt, e->counters + 2, sizeof cnt, hipMemcpyDeviceToHost,
                       e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  if (err != hipSuccess) return err;
  const int budget = std::max(1, e->tune.budget);
  const int block = e->tune.block;
  const int64_t max_launches = (kMaxSteps + budget - 1) / budget;
  const bool usg = e->usg != nullptr;
  const int bits = radix_bits(usg ? std::max<int64_t>(e->prob.ncell, e->usg->nunit)
                                  : e->prob.ncell);
  int64_t launches = 0;
  while (cnt > 0) {
    // graph mode: the rest of the walk once the list is short enough
    if (e->graph_launches > 0 && (e->graph_max_list <= 0 ||
                                  cnt <= (unsigned long long)e->graph_max_list))
      return advance_graph(e, W, cnt, max_launches, launches, launches);
    const int32_t *work = P.vals[0];
    if (e->tune.sort) {
      size_t bytes = P.sort_tmp_bytes;
      err = hipcub::DeviceRadixSort::SortPairs(P.sort_tmp, bytes, P.keys[0],
                                               P.keys[1], P.vals[0], P.vals[1],
                                               (int)cnt, 0, bits, e->stream);
      if (err != hipSuccess) return err;
      work = P.vals[1];
    }
    if (launches >= max_launches) {  // step guard reached
      if (usg) {
        hipLaunchKernelGGL(terminate_walk_usg, dim3((unsigned)((cnt + 255) / 256)),
                           dim3(256), 0, e->stream, usg_walk_view(e), work, (int64_t)cnt,
                           particle_ptrs<double>(P));
        err = hipGetLastError();
      } else {
        err = fp32 ? walk_terminate32(e, work, (int64_t)cnt)
                   : launch_terminate_typed<double>(e, work, (int64_t)cnt);
      }
      if (err != hipSuccess) return err;
      break;
    }
    // the next list is written into the (consumed) sort input buffers,
    // or into the second pair when the list is not sorted
    int32_t *nk = e->tune.sort ? P.keys[0] : P.keys[1];
    int32_t *nv = e->tune.sort ? P.vals[0] : P.vals[1];
    err = hipMemsetAsync(e->counters + 2, 0, sizeof(unsigned long long), e->stream);
    if (err != hipSuccess) return err;
    if ((err = timed_event(e, launches, 0, 1)) != hipSuccess) return err;
    const dim3 grid((unsigned)((cnt + block - 1) / block));
    err = usg    ? walk_launch_usg(e, W, grid, block, work, (int64_t)cnt, budget, nk, nv)
          : fp32 ? walk_launch32(e, W, grid, block, work, (int64_t)cnt, budget, nk, nv)
                 : walk_launch64(e, W, grid, block, work, (int64_t)cnt, budget, nk, nv);
    if (err != hipSuccess) return err;
    if ((err = timed_event(e, launches, 1, 1)) != hipSuccess) return err;
    launches += 1;
    err = hipMemcpyAsync(&cnt, e->counters + 2, sizeof cnt, hipMemcpyDeviceToHost,
                         e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    if (err != hipSuccess) return err;
    if (!e->tune.sort) {
      std::swap(P.keys[0], P.keys[1]);
      std::swap(P.vals[0], P.vals[1]);
    }
  }
  e->last_launches = launches;
  e->last_timed = launches;
  return hipSuccess;
}

hipError_t launch_global_xyz(rwpt_engine *e, double *gx, double *gy,
                             double *gz) {
  if (e->usg) return launch_usg_global_xyz(e, gx, gy, gz);
  DeviceProblem &P = e->prob;
  const int block = 256;
  const int64_t n = e->part.n;
  if (n > 0)
    hipLaunchKernelGGL(global_xyz_kernel, dim3((unsigned)((n + block - 1) / block)),
                       dim3(block), 0, e->stream, n, e->part.cell, e->part.lx,
                       e->part.ly, e->part.lz, P.ncol, P.nrow, e->xface, e->yface,
                       P.delr, P.delc, model_top(P), P.bot, gx, gy, gz);
  return hipGetLastError();
}
}  // namespace rwpt
