// Per-cell table build: once per flow time step, every cell gathers its
// 18-cell stencil and stores the corner quantities a random-walk sub-step
// needs.  The reference instead rebuilds the cell buffer, the 18 neighbour
// buffers and the corner variables on every cell entry of every particle
// (ParticleTrackingEngine.f90:1093-1108, TrackCell.f90:511-518); building
// them once turns that O(entries) work into O(cells).
//
// Two passes, one thread per cell:
//  1. assemble_cells: face flows with boundary flows, GetDZRW, connections
//     and mass-boundary flags (SetMassTransportDataStructured,
//     ModpathCellData.f90:1099-1487; GetDZRW :3208-3256).
//  2. corner_tables: corner discharge, corner porosity, discharge products,
//     corner dispersion and their derivative terms
//     (TrackSubCell.F90:3667-4363, 2557-2691, 3117-3303, 4432-4543).
// Floating-point expressions keep the reference's operand order and the
// file is compiled with -ffp-contract=off, so every table entry is the
// double the reference computes.
#include "rwpt_internal.h"
#include "rwpt_record.h"

namespace rwpt {
namespace {

struct GridDims {
  int ncol, nrow, nlay;
  int nlay_member;  // layers of one ensemble member (nlay for one member)
};

__device__ __forceinline__ int face_neighbor(const GridDims g, int c0, int f) {
  // c0 is 0-based; returns the 1-based neighbour or 0 (ComputeFaceConnection)
  const int lc = g.nrow * g.ncol;
  const int k = c0 / lc, r = c0 % lc, i = r / g.ncol, j = r % g.ncol;
  switch (f) {
    case 0: return j > 0 ? c0 : 0;                       // (c0+1)-1
    case 1: return j < g.ncol - 1 ? c0 + 2 : 0;
    case 2: return i < g.nrow - 1 ? c0 + 1 + g.ncol : 0;
    case 3: return i > 0 ? c0 + 1 - g.ncol : 0;
    // vertical connections stop at the bottom/top of each ensemble member
    case 4: return (k % g.nlay_member) < g.nlay_member - 1 ? c0 + 1 + lc : 0;
    default: return (k % g.nlay_member) > 0 ? c0 + 1 - lc : 0;
  }
}

__global__ void assemble_cells(const GridDims g, int64_t ncell,
                               const double *__restrict__ frf,
                               const double *__restrict__ fff,
                               const double *__restrict__ flf,
                               const double *__restrict__ bflow,
                               const int32_t *__restrict__ ibound,
                               const int32_t *__restrict__ icbound,
                               int32_t default_icbound,
                               const double *__restrict__ top,
                               const double *__restrict__ bot,
                               const double *__restrict__ head,
                               const int32_t *__restrict__ layer_type,
                               const int32_t *__restrict__ zone,
                               const double *__restrict__ source,
                               const double *__restrict__ sink,
                               double *__restrict__ q6,
                               double *__restrict__ dzrw,
                               double *__restrict__ cells) {
  const int64_t c0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c0 >= ncell) return;
  int conn[6], mb[6], bcount[6];
  double Q[6] = {0, 0, 0, 0, 0, 0};
  for (int f = 0; f < 6; ++f) {
    conn[f] = face_neighbor(g, (int)c0, f);
    if (conn[f] > 0) {
      bcount[f] = (ibound && ibound[conn[f] - 1] == 0) ? 1 : 0;
      mb[f] = icbound ? icbound[conn[f] - 1] : 0;
    } else {
      bcount[f] = 1;
      mb[f] = default_icbound;
    }
  }
  if (conn[0] > 0) Q[0] = frf[conn[0] - 1];
  Q[1] = frf[c0];
  Q[2] = -fff[c0];
  if (conn[3] > 0) Q[3] = -fff[conn[3] - 1];
  Q[4] = -flf[c0];
  if (conn[5] > 0) Q[5] = -flf[conn[5] - 1];
  double src = source ? source[c0] : 0.0;
  double snk = sink ? sink[c0] : 0.0;
  if (bflow) {
    for (int f = 0; f < 6; ++f) {
      const double bf = bflow[6 * c0 + f];
      if (bf != 0.0) {
        if (bcount[f] > 0) {
          const double flow = bf / bcount[f];
          if (conn[f] == 0 || (ibound && ibound[conn[f] - 1] == 0))
            Q[f] = (f % 2 == 0) ? flow : -flow;
        } else if (bf > 0.0) {
          src = src + bf;
        } else {
          snk = snk + bf;
        }
      }
    }
  }
  // GetDZRW and VerifyDryCell
  const double tp = top[c0], bt = bot[c0];
  const double hd = head ? head[c0] : tp;
  const int lt = layer_type ? layer_type[c0] : 0;
  double dz = tp - bt;
  const double dzc = tp - bt;
  bool dry = false, vpdry = false;
  bool anyflow = false;
  for (int f = 0; f < 6; ++f) anyflow |= (Q[f] != 0.0);
  if (lt == 1 && hd < tp) {
    dz = hd - bt;
    if ((dz <= 0.0) || ((hd - bt) / dzc < 0.01)) {
      dz = tp - bt;
      dry = !anyflow;
    }
    // VerifyDryCell: partially dry unless head <= bottom with no flow
    vpdry = !((hd - bt) <= 0.0) || anyflow;
  }
  for (int f = 0; f < 6; ++f) q6[6 * c0 + f] = Q[f];
  dzrw[c0] = dz;
  double *rec = cells + c0 * CELL_DOUBLES;
  int32_t *ic = reinterpret_cast<int32_t *>(rec + F_CONN);
  int32_t *im = reinterpret_cast<int32_t *>(rec + F_MASSBND);
  for (int f = 0; f < 6; ++f) {
    ic[f] = conn[f];
    im[f] = mb[f];
  }
  bool cell_exit = (Q[0] < 0.0) || (Q[1] > 0.0) || (Q[2] < 0.0) ||
                   (Q[3] > 0.0) || (Q[4] < 0.0) || (Q[5] > 0.0);
  int32_t *fl = reinterpret_cast<int32_t *>(rec + F_FLAGS);
  fl[0] = ibound ? ibound[c0] : 1;
  int32_t nbi = 0;
  for (int f = 0; f < 6; ++f)
    if (conn[f] > 0 && ibound && ibound[conn[f] - 1] == 0) nbi |= CF_NB_INACTIVE << f;
  fl[1] = (dry ? CF_DRY : 0) | (cell_exit ? CF_CELL_EXIT : 0) |
          (vpdry ? CF_VERIFY_PDRY : 0) | (src != 0.0 ? CF_SOURCE : 0) |
          (snk != 0.0 ? CF_SINK : 0) | nbi;
  int32_t *zn = reinterpret_cast<int32_t *>(rec + F_ZONE);
  zn[0] = zone ? zone[c0] : 0;
  zn[1] = 0;
  rec[F_SOURCE] = src;
  rec[F_SINK] = snk;
}

// Stencil slot offsets (dk, di, dj) in the order of
// pr_FillNeighborCellDataStructured (ParticleTrackingEngine.f90:2009-2126):
// W, W-front, W-back, E, E-front, E-back, front, front-bottom, front-top,
// back, back-bottom, back-top, bottom, bottom-W, bottom-E, top, top-W, top-E.
__constant__ int kSlot[18][3] = {
    {0, 0, -1}, {0, 1, -1}, {0, -1, -1}, {0, 0, 1},  {0, 1, 1},  {0, -1, 1},
    {0, 1, 0},  {1, 1, 0},  {-1, 1, 0},  {0, -1, 0}, {1, -1, 0}, {-1, -1, 0},
    {1, 0, 0},  {1, 0, -1}, {1, 0, 1},   {-1, 0, 0}, {-1, 0, -1}, {-1, 0, 1}};

__global__ void corner_tables(const GridDims g, int64_t ncell,
                              const double *__restrict__ delr,
                              const double *__restrict__ delc,
                              const double *__restrict__ porosity,
                              const double *__restrict__ retardation,
                              const double *__restrict__ disp, int64_t ndisp,
                              const double *__restrict__ q6,
                              const double *__restrict__ dzrw,
                              const TableParams tp,
                              const int32_t *__restrict__ uniform_member,
                              double *__restrict__ cells,
                              double *__restrict__ debug) {
  const int64_t c0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c0 >= ncell) return;
  const int lc = g.nrow * g.ncol;
  const int k = (int)(c0 / lc), r = (int)(c0 % lc), i = r / g.ncol,
            j = r % g.ncol;
  double *rec = cells + c0 * CELL_DOUBLES;
  // sub-cell buffer, FillMassSubCellDataBuffer (ModpathCellData.f90:2522-2684)
  const double DX = delr[j] / 1.0, DY = delc[i] / 1.0, DZ = dzrw[c0];
  const double phi = porosity[c0];
  const double R = retardation ? retardation[c0] : 1.0;
  double Qc[6];
  for (int f = 0; f < 6; ++f) Qc[f] = q6[6 * c0 + f];
  // dispersion parameters (pr_FillDispersionParameters*)
  const int64_t d = ndisp == 1 ? 0 : c0;
  const double aLH = disp[d], aTH = disp[ndisp + d];
  const double aLV = tp.disp_model == 1 ? disp[2 * ndisp + d] : 0.0;
  const double aTV = tp.disp_model == 1 ? disp[3 * ndisp + d] : 0.0;
  const double dm = disp[4 * ndisp + d];

  // stencil data (pr_FillNeighborSubCellVariables, unrefined branch)
  double nQ[18][6], nA[18][3], nV[18], nP[18];
  for (int n = 0; n < 18; ++n) {
    const int kk = k + kSlot[n][0], ii = i + kSlot[n][1], jj = j + kSlot[n][2];
    const bool in = kk >= 0 && kk < g.nlay && kk / g.nlay_member == k / g.nlay_member &&
                    ii >= 0 && ii < g.nrow && jj >= 0 && jj < g.ncol;
    if (!in) {
      for (int f = 0; f < 6; ++f) nQ[n][f] = 0.0;
      nA[n][0] = nA[n][1] = nA[n][2] = 0.0;
      nV[n] = 0.0;
      nP[n] = 0.0;
      continue;
    }
    const int64_t nc = (int64_t)kk * lc + (int64_t)ii * g.ncol + jj;
    for (int f = 0; f < 6; ++f) nQ[n][f] = q6[6 * nc + f];
    const double dx = delr[jj], dy = delc[ii], dz = dzrw[nc];
    nA[n][0] = dy * dz;
    nA[n][1] = dx * dz;
    nA[n][2] = dx * dy;
    nV[n] = dx * dy * dz;
    nP[n] = porosity[nc];
  }
  RecordDbg dbg;
  record_from_stencil(rec, tp, DX, DY, DZ, phi, R, Qc, aLH, aTH, aLV, aTV, dm,
                      uniform_member[k / g.nlay_member] != 0, nQ, nA, nV, nP, dbg);
  if (debug) {
    double *o = debug + c0 * DEBUG_FIELDS;
    record_debug(rec, dbg, o);
  }
}

// FP32-local record from the FP64 one: float fields rounded once, integer
// fields copied, full reciprocals of DX DY DZ R phi (every divisor of the
// sub-step becomes a product).
__global__ void float_records(int64_t ncell, const double *__restrict__ cells,
                              float *__restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const double *rec = cells + c * CELL_DOUBLES;
  float *o = out + c * CELL_FLOATS;
  for (int f = 0; f < F_CONN; ++f) o[f] = (float)rec[f];
  // face velocity slopes
  for (int a = 0; a < 3; ++a)
    o[G_DV + a] = (float)(rec[F_VX1 + 2 * a + 1] - rec[F_VX1 + 2 * a]);
  // dispersion factors 2 alpha / R, 2 Dm / R
  for (int k = 0; k < 5; ++k) o[G_S2ALH + k] = (float)(2.0 * rec[F_ALH + k] / rec[F_R]);
  // trilinear coefficients of the corner velocities (corner i + 2j + 4k
  // at x = i, y = j, z = k)
  for (int comp = 0; comp < 3; ++comp) {
    const double *v = rec + F_VC + 8 * comp;
    float *cc = o + G_VC + 8 * comp;
    cc[0] = (float)v[0];
    cc[1] = (float)(v[1] - v[0]);
    cc[2] = (float)(v[2] - v[0]);
    cc[3] = (float)(v[4] - v[0]);
    cc[4] = (float)((v[3] - v[2]) - (v[1] - v[0]));
    cc[5] = (float)((v[5] - v[4]) - (v[1] - v[0]));
    cc[6] = (float)((v[6] - v[4]) - (v[2] - v[0]));
    cc[7] = (float)(((v[7] - v[6]) - (v[5] - v[4])) - ((v[3] - v[2]) - (v[1] - v[0])));
  }
  // bilinear coefficients of the drift terms over phi R
  const double s = 1.0 / (rec[F_PHI] * rec[F_R]);
  for (int term = 0; term < 9; ++term) {
    const double *d = rec + F_DD + 4 * term;
    float *e = o + G_DD + 4 * term;
    e[0] = (float)(d[0] * s);
    e[1] = (float)((d[1] - d[0]) * s);
    e[2] = (float)((d[2] - d[0]) * s);
    e[3] = (float)(((d[3] - d[2]) - (d[1] - d[0])) * s);
  }
  const int32_t *ri = reinterpret_cast<const int32_t *>(rec);
  int32_t *oi = reinterpret_cast<int32_t *>(o);
  for (int k = 0; k < 6; ++k) {
    oi[G_CONN + k] = ri[2 * F_CONN + k];
    oi[G_MASSBND + k] = ri[2 * F_MASSBND + k];
  }
  oi[G_FLAGS] = ri[2 * F_FLAGS];
  oi[G_FLAGS + 1] = ri[2 * F_FLAGS + 1];
  oi[G_ZONE] = ri[2 * F_ZONE];
  o[G_SOURCE] = (float)rec[F_SOURCE];
  o[G_SINK] = (float)rec[F_SINK];
  o[G_RDX] = (float)(1.0 / rec[F_DX]);
  o[G_RDY] = (float)(1.0 / rec[F_DY]);
  o[G_RDZ] = (float)(1.0 / rec[F_DZ]);
  o[G_RR] = (float)(1.0 / rec[F_R]);
  o[G_RPHI] = (float)(1.0 / rec[F_PHI]);
  for (int f = G_RPHI + 1; f < CELL_FLOATS; ++f) o[f] = 0.0f;
}

__global__ void to_float3(int64_t n, const double *__restrict__ a,
                          const double *__restrict__ b,
                          const double *__restrict__ c, float *__restrict__ fa,
                          float *__restrict__ fb, float *__restrict__ fc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fa[i] = (float)a[i];
  fb[i] = (float)b[i];
  fc[i] = (float)c[i];
}

__global__ void to_double3(int64_t n, const float *__restrict__ fa,
                           const float *__restrict__ fb,
                           const float *__restrict__ fc, double *__restrict__ a,
                           double *__restrict__ b, double *__restrict__ c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[i] = fa[i];
  b[i] = fb[i];
  c[i] = fc[i];
}

// Clears *flag unless every cell stores the exact reciprocals of DX, DY, DZ
// and R (all powers of two).
__global__ void all_pow2(int64_t ncell, const double *__restrict__ cells,
                         unsigned long long *__restrict__ flag) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = true;
  if (c < ncell) {
    const double *rec = cells + c * CELL_DOUBLES;
    ok = rec[F_RDX] != 0.0 && rec[F_RDY] != 0.0 && rec[F_RDZ] != 0.0 &&
         rec[F_RR] != 0.0;
  }
  if (__ballot(!ok) != 0 && __lane_id() == 0) atomicAnd(flag, 0ull);
}

// a[i] * f (porosity) or a[i] + sign * h[i] over the blocks of the
// dispersion array that hold the perturbed parameter
__global__ void scaled_copy(int64_t n, const double *__restrict__ a, double f,
                            double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] * f;
}

__global__ void shifted_copy(int64_t n, const double *__restrict__ a, int64_t lo,
                             int64_t hi, double sign, const double *__restrict__ h,
                             double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = (i >= lo && i < hi) ? a[i] + sign * h[i - lo] : a[i];
}

// step of the central difference: relative to the value, absolute 1 at 0
__global__ void fd_steps(int64_t n, const double *__restrict__ a, double rel,
                         double *__restrict__ h) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) h[i] = (a && a[i] != 0.0) ? rel * fabs(a[i]) : rel;
}

// (plus - minus) / (span h) of the double fields 0..F_CONN-1 of every
// record for parameter k (span 2 for a central difference, 1 for a forward
// one), stored as out[(c * F_CONN + f) * 4 + k]; h per cell from the step
// array (one value when nh == 1)
__global__ void fd_tangent(int64_t ncell, const double *__restrict__ plus,
                           const double *__restrict__ minus,
                           const double *__restrict__ h, int64_t nh, int k, double span,
                           double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncell * F_CONN) return;
  const int64_t c = i / F_CONN, f = i % F_CONN;
  const double hc = h[nh == 1 ? 0 : c];
  out[i * 4 + k] = (plus[c * CELL_DOUBLES + f] - minus[c * CELL_DOUBLES + f]) / (span * hc);
}

}  // namespace

hipError_t build_record_tangents(rwpt_engine *e, double *out) {
  DeviceProblem &P = e->prob;
  const GridDims g{P.ncol, P.nrow, P.nlay, P.nlay_member};
  const int block = 256;
  const unsigned grid = (unsigned)((P.ncell + block - 1) / block);
  const TableParams tp{P.disp_model, e->opt.backward_tracking, e->opt.time_step_kind,
                       e->opt.time_step_params[0], e->opt.time_step_params[1]};
  // the records are rational in the porosity scale and linear in aLH, aTH
  // and Dm (for FIXED time steps): a relative step of 1e-5 leaves the
  // truncation error near 1e-10 and the rounding error near 1e-11 relative.
  // Where a record is linear in the parameter, a forward difference from
  // the engine's own records has no truncation error either and needs one
  // table build instead of two.
  const double rel = 1.0e-5;
  const bool linear = e->opt.time_step_kind == 4;
  const size_t recBytes = (size_t)P.ncell * CELL_DOUBLES * sizeof(double);
  const int64_t nd = 5 * P.ndisp;
  double *plus = nullptr, *minus = nullptr, *por = nullptr, *disp = nullptr, *h = nullptr;
  auto release = [&]() {
    for (double *q : {plus, minus, por, disp, h})
      if (q) (void)hipFree(q);
  };
  hipError_t err = hipMalloc(&plus, recBytes);
  if (err == hipSuccess) err = hipMalloc(&minus, recBytes);
  if (err == hipSuccess) err = hipMalloc(&por, (size_t)P.ncell * sizeof(double));
  if (err == hipSuccess) err = hipMalloc(&disp, (size_t)nd * sizeof(double));
  if (err == hipSuccess)
    err = hipMalloc(&h, (size_t)std::max<int64_t>(P.ncell, P.ndisp) * sizeof(double));
  if (err != hipSuccess) return release(), err;
  const unsigned dgrid = (unsigned)((nd + block - 1) / block);
  const unsigned tgrid = (unsigned)((P.ncell * F_CONN + block - 1) / block);
  for (int k = 0; k < 4; ++k) {
    // dispersion block of parameter k: aL [0, n), aT [n, 2n), Dm [4n, 5n)
    const int64_t lo = k == 0 ? 0 : (k == 1 ? P.ndisp : 4 * P.ndisp), hi = lo + P.ndisp;
    if (k < 3)
      hipLaunchKernelGGL(fd_steps, dim3((unsigned)((P.ndisp + block - 1) / block)),
                         dim3(block), 0, e->stream, P.ndisp, P.disp + lo, rel, h);
    const bool forward = linear && k < 3;
    for (int sgn = 0; sgn < (forward ? 1 : 2); ++sgn) {
      double *dst = sgn == 0 ? plus : minus;
      const double sign = sgn == 0 ? 1.0 : -1.0;
      err = hipMemcpyAsync(dst, P.cells, recBytes, hipMemcpyDeviceToDevice, e->stream);
      if (err != hipSuccess) return release(), err;
      const double *pin = P.porosity, *din = P.disp;
      if (k < 3) {
        hipLaunchKernelGGL(shifted_copy, dim3(dgrid), dim3(block), 0, e->stream, nd, P.disp,
                           lo, hi, sign, (const double *)h, disp);
        din = disp;
      } else {
        hipLaunchKernelGGL(scaled_copy, dim3(grid), dim3(block), 0, e->stream, P.ncell,
                           P.porosity, 1.0 + sign * rel, por);
        pin = por;
      }
      hipLaunchKernelGGL(corner_tables, dim3(grid), dim3(block), 0, e->stream, g, P.ncell,
                         P.delr, P.delc, pin, P.retardation, din, P.ndisp, P.q6, P.dzrw,
                         tp, P.uniform_member, dst, (double *)nullptr);
      if ((err = hipGetLastError()) != hipSuccess) return release(), err;
    }
    if (k == 3)
      hipLaunchKernelGGL(fd_steps, dim3(1), dim3(1), 0, e->stream, (int64_t)1,
                         (const double *)nullptr, rel, h);
    hipLaunchKernelGGL(fd_tangent, dim3(tgrid), dim3(block), 0, e->stream, P.ncell,
                       (const double *)plus, forward ? (const double *)P.cells : minus,
                       (const double *)h, k == 3 ? (int64_t)1 : P.ndisp, k,
                       forward ? 1.0 : 2.0, out);
    if ((err = hipGetLastError()) != hipSuccess) return release(), err;
  }
  err = hipStreamSynchronize(e->stream);
  release();
  return err;
}

hipError_t build_cell_tables32(rwpt_engine *e) {
  DeviceProblem &P = e->prob;
  const int block = 256;
  hipLaunchKernelGGL(float_records, dim3((unsigned)((P.ncell + block - 1) / block)),
                     dim3(block), 0, e->stream, P.ncell, P.cells, P.cells32);
  return hipGetLastError();
}

hipError_t launch_positions_to32(rwpt_engine *e) {
  Particles &S = e->part;
  if (S.n == 0) return hipSuccess;
  const int block = 256;
  hipLaunchKernelGGL(to_float3, dim3((unsigned)((S.n + block - 1) / block)),
                     dim3(block), 0, e->stream, S.n, S.lx, S.ly, S.lz, S.flx,
                     S.fly, S.flz);
  return hipGetLastError();
}

hipError_t launch_positions_from32(rwpt_engine *e) {
  Particles &S = e->part;
  if (S.n == 0) return hipSuccess;
  const int block = 256;
  hipLaunchKernelGGL(to_double3, dim3((unsigned)((S.n + block - 1) / block)),
                     dim3(block), 0, e->stream, S.n, S.flx, S.fly, S.flz, S.lx,
                     S.ly, S.lz);
  return hipGetLastError();
}

hipError_t build_cell_tables(rwpt_engine *e, double *debug_out) {
  DeviceProblem &P = e->prob;
  const GridDims g{P.ncol, P.nrow, P.nlay, P.nlay_member};
  const int block = 256;
  const unsigned grid = (unsigned)((P.ncell + block - 1) / block);
  hipLaunchKernelGGL(assemble_cells, dim3(grid), dim3(block), 0, e->stream, g,
                     P.ncell, P.frf, P.fff, P.flf, P.bflow, P.ibound,
                     P.icbound, P.default_icbound, P.top, P.bot, P.head,
                     P.layer_type, P.zone, P.source, P.sink, P.q6, P.dzrw,
                     P.cells);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  TableParams tp{P.disp_model, e->opt.backward_tracking,
                 e->opt.time_step_kind, e->opt.time_step_params[0],
                 e->opt.time_step_params[1]};
  hipLaunchKernelGGL(corner_tables, dim3(grid), dim3(block), 0, e->stream, g,
                     P.ncell, P.delr, P.delc, P.porosity, P.retardation,
                     P.disp, P.ndisp, P.q6, P.dzrw, tp, P.uniform_member, P.cells,
                     debug_out);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  // counters[3] is free outside an advance
  unsigned long long one = 1ull, flag = 0ull;
  err = hipMemcpyAsync(e->counters + 3, &one, sizeof one, hipMemcpyHostToDevice,
                       e->stream);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(all_pow2, dim3(grid), dim3(block), 0, e->stream, P.ncell,
                     P.cells, e->counters + 3);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  err = hipMemcpyAsync(&flag, e->counters + 3, sizeof flag, hipMemcpyDeviceToHost,
                       e->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
  P.pow2_cells = err == hipSuccess && flag == 1ull;
  return err;
}

}  // namespace rwpt
