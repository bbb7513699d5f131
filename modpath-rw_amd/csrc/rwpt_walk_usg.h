// The random walk on unstructured grids of rectangular cells (FP64):
// TrackCell's sub-cell loop (TrackCell.f90:375-669) and pr_RWPTrackPath's
// cell transfers (ParticleTrackingEngine.f90:1051-1254) over tracking units
// (an unrefined cell, or one sub-cell of a refined one), each with the
// record the structured walk reads; the time loop itself is the structured
// walk's `iterate`.
//
// Between advances a particle holds its cell and its location in the cell
// (what pr_RWPTrackPath reports); during a visit, its unit and its location
// in the unit's sub-cell.  A cell entry runs TrackCell's checks on the
// cell, finds the sub-cell (FindSubCell, :47-90) and converts the location
// (ConvertFromLocalParentCoordinate, ModpathSubCellData.f90:129-149); a
// sub-cell exit through an internal face moves to the sibling sub-cell
// with the face coordinate set (TrackCell.f90:530-611), at most 100
// sub-cell visits per cell entry; a cell exit converts the exit point to
// cell coordinates (ConvertToLocalParentCoordinate, :114-126) and then into
// the neighbour (ConvertFromNeighbor, ModflowRectangularGrid.f90:366-477:
// the face coordinate set on a one-to-one face, through global coordinates
// with a 1e-3 snap otherwise).  Reported points follow pr_RWPTrackPath's
// pathline: the cell-entry point for statuses 3-7, the last point before
// the cell entry for status 9.
#pragma once

#include "rwpt_walk_kernels.h"

namespace rwpt {
namespace {

struct UsgWalk {
  const int32_t *cell_unit, *side;
  const double *off, *cx, *cy, *dx, *dy;
  const int32_t *ibound;  // IBoundTS, null: all active
  int32_t *unit, *pcell;
  double *ex, *ey, *ez, *et, *px, *py, *pz;
};

constexpr int kMaxSubCellVisits = 100;  // TrackCell.f90:521
constexpr int kCountShift = 22;         // sub-cell visit count in the aux word

// TrackCell's checks before FindSubCell (TrackCell.f90:375-497): the
// location, IboundTS, stop zone, weak sink/source, no exit face.  0: go on.
__device__ __forceinline__ int usg_cell_checks(const double *__restrict__ rec,
                                               const WalkParams &W, double x, double y,
                                               double z) {
  if (outside(x, y, z)) return 9;
  const int32_t *fl = ints(rec) + Lay<double>::FLAGS;
  if (fl[0] == 0) return 7;
  if (W.stop_zone > 0 && ints(rec)[Lay<double>::ZONE] == W.stop_zone) return 6;
  const int32_t b = fl[1];
  const bool src = (b & CF_SOURCE) != 0, snk = (b & CF_SINK) != 0;
  if (W.backward) {
    if (W.stop_weak_sources && src) return 4;
  } else if (W.stop_weak_sinks && snk) {
    return 3;
  }
  if (!(b & CF_CELL_EXIT)) {
    if (W.steady_state || (W.backward && src) || (!W.backward && snk)) return 5;
  }
  return 0;
}

// TrackSubCell's early returns (TrackSubCell.F90:870-945) with TrackCell's
// stopIfNoSubCellExit (TrackCell.f90:441-469).  0: go on.
__device__ __forceinline__ int usg_subcell_checks(const double *__restrict__ rec,
                                                  const WalkParams &W) {
  const int32_t b = ints(rec)[Lay<double>::FLAGS + 1];
  bool stopIfNoExit = true;
  if (W.backward) {
    if (W.stop_weak_sources && !W.steady_state) stopIfNoExit = false;
  } else if (W.stop_weak_sinks && !W.steady_state) {
    stopIfNoExit = false;
  }
  if (stopIfNoExit && !(b & CF_SUBCELL_EXIT)) return 5;
  if (b & CF_DRY) return 7;
  if (b & CF_DT_NONPOS) return 9;
  return 0;
}

template <int NDIM, int MODEL, int ADV>
__global__ void __launch_bounds__(256)
    walk_kernel_usg(const double *__restrict__ cells, const UsgWalk G,
                    const int32_t *__restrict__ work, int64_t nwork,
                    ParticlePtrs<double> S, const WalkParams W, int budget,
                    int32_t *__restrict__ nkeys, int32_t *__restrict__ nvals,
                    unsigned long long *__restrict__ counters,
                    const unsigned long long *__restrict__ nwork_dev) {
  // a launch captured in a graph reads the list length from the device
  if (nwork_dev) nwork = (int64_t)*nwork_dev;
  const int64_t nblk = (nwork + blockDim.x - 1) / blockDim.x;
  if ((int64_t)blockIdx.x >= nblk) return;
  const int64_t blk = xcd_block(blockIdx.x, nblk);
  const int64_t i = blk * blockDim.x + threadIdx.x;
  bool more = false;
  int c = 0, p = 0, u = 0;
  unsigned long long steps = 0, ended_active = 0;
  if (i < nwork) {
    p = work[i];
    c = S.cell[p];
    Visit<double> V;
    V.x = S.lx[p]; V.y = S.ly[p]; V.z = S.lz[p]; V.t = S.t[p];
    const int32_t aux = S.aux[p];
    Aux A = unpack_aux(aux);
    int count = (aux >> kCountShift) & 127;
    double ex, ey, ez, et, px, py, pz;
    int pc;
    if (A.phase) {
      u = G.unit[p];
      V.ix = S.ix[p]; V.iy = S.iy[p]; V.iz = S.iz[p]; V.t0 = S.t0[p];
      ex = G.ex[p]; ey = G.ey[p]; ez = G.ez[p]; et = G.et[p];
      pc = G.pcell[p]; px = G.px[p]; py = G.py[p]; pz = G.pz[p];
    } else {
      // the advance's start location is the pathline's first point
      u = G.cell_unit[c - 1];
      pc = c; px = V.x; py = V.y; pz = V.z;
      ex = V.x; ey = V.y; ez = V.z; et = V.t;
      V.ix = V.x; V.iy = V.y; V.iz = V.z; V.t0 = V.t;
    }
    V.restarts = A.restarts;
    V.exitFace = A.exitFace;
    Stream rng{S.rng[2 * p], S.rng[2 * p + 1]};
    int status = 1;
    bool done = false;
    const double *rec = cells + (int64_t)u * CELL_DOUBLES;
    // the reported point of a stop: the cell-entry point, or the last
    // point before the entry
    auto at_entry = [&]() { V.x = ex; V.y = ey; V.z = ez; V.t = et; };
    auto at_previous = [&]() { c = pc; V.x = px; V.y = py; V.z = pz; V.t = et; };
    // a TrackSubCell call: its checks, then a new visit
    auto enter_subcell = [&]() -> bool {
      const int st = usg_subcell_checks(rec, W);
      if (st != 0) {
        status = st;
        if (st == 9) at_previous();
        else at_entry();
        return false;
      }
      A.phase = 1;
      V.ix = V.x; V.iy = V.y; V.iz = V.z; V.t0 = V.t;
      V.restarts = 0;
      V.exitFace = 0;
      return true;
    };
    // a TrackCell call on cell c at the cell location (V.x, V.y, V.z, V.t)
    auto enter_cell = [&]() -> bool {
      ex = V.x; ey = V.y; ez = V.z; et = V.t;
      const int u0 = G.cell_unit[c - 1];
      const double *rec0 = cells + (int64_t)u0 * CELL_DOUBLES;
      const int st = usg_cell_checks(rec0, W, V.x, V.y, V.z);
      if (st != 0) {
        status = st;
        if (st == 9) at_previous();
        else at_entry();
        return false;
      }
      u = u0;
      if (G.side[(int64_t)USG_SIDE * u0 + US_NSUB] == 4) {
        // FindSubCell: rows top-down, columns left-right, edges inclusive;
        // a location it cannot place (NaN) ends the visit as Undefined
        const double x = V.x, y = V.y;
        if (!(x >= 0.0 && x <= 1.0 && y >= 0.0 && y <= 1.0)) {
          status = 9;
          at_previous();
          return false;
        }
        const int row = (y >= 0.5 && y <= 1.0) ? 1 : 2;
        const int col = (x >= 0.0 && x <= 0.5) ? 1 : 2;
        u = u0 + (row - 1) * 2 + (col - 1);
        const double *of = G.off + 4 * (int64_t)u;
        V.x = (x - of[0]) / (of[1] - of[0]);
        V.y = (y - of[2]) / (of[3] - of[2]);
        V.z = (V.z - 0.0) / (1.0 - 0.0);
      }
      rec = cells + (int64_t)u * CELL_DOUBLES;
      count = 1;
      return enter_subcell();
    };
    if (!A.phase) done = !enter_cell();
    while (!done) {
      if ((int)steps >= budget) {
        more = true;
        break;
      }
      steps += 1;
      bool zero_dt = false;
      const int out = iterate<double, 0, 0, NDIM, MODEL, ADV>(rec, W, V, rng, zero_dt);
      if (out == IT_CONTINUE) continue;
      const double *of = G.off + 4 * (int64_t)u;
      if (out == IT_MAXTIME) {
        status = 1;
        const double sx = V.x, sy = V.y, sz = V.z;
        V.x = ((1.0 - sx) * of[0]) + (sx * of[1]);
        V.y = ((1.0 - sy) * of[2]) + (sy * of[3]);
        V.z = ((1.0 - sz) * 0.0) + (sz * 1.0);
        done = true;
        break;
      }
      if (out == IT_UNDEFINED) {
        status = 9;
        at_previous();
        done = true;
        break;
      }
      // IT_EXIT
      const int face = V.exitFace;
      const int conn = ints(rec)[Lay<double>::CONN + face - 1];
      const int code = G.side[(int64_t)USG_SIDE * u + US_FACE + face - 1];
      if (conn < 0) {  // internal face: the sibling sub-cell
        if (count >= kMaxSubCellVisits) {
          status = 9;
          at_previous();
          done = true;
          break;
        }
        count += 1;
        switch (face) {
          case 1: V.x = 1.0; break;
          case 2: V.x = 0.0; break;
          case 3: V.y = 1.0; break;
          default: V.y = 0.0; break;
        }
        u = code;
        rec = cells + (int64_t)u * CELL_DOUBLES;
        done = !enter_subcell();
        continue;
      }
      // the exit point in cell coordinates is the last pathline point
      const double sx = V.x, sy = V.y, sz = V.z;
      V.x = ((1.0 - sx) * of[0]) + (sx * of[1]);
      V.y = ((1.0 - sy) * of[2]) + (sy * of[3]);
      V.z = ((1.0 - sz) * 0.0) + (sz * 1.0);
      A.stalls = zero_dt ? A.stalls + 1 : 0;
      if (A.stalls > kMaxStalls) {
        status = 9;
        done = true;
        break;
      }
      if (conn <= 0 || (G.ibound && G.ibound[conn - 1] == 0)) {
        status = 2;
        done = true;
        break;
      }
      const int tf = code & 7;
      const double fx = V.x, fy = V.y, fz = V.z;
      const bool bad = tf == 0 || (tf == 1 && fx != 1.0) || (tf == 2 && fx != 0.0) ||
                       (tf == 3 && fy != 1.0) || (tf == 4 && fy != 0.0) ||
                       (tf == 5 && fz != 1.0) || (tf == 6 && fz != 0.0);
      if (bad) {  // the reference finds no location (cell 0)
        status = 9;
        done = true;
        break;
      }
      pc = c; px = fx; py = fy; pz = fz;
      double nx = fx, ny = fy, nz = fz;
      if (code & 8) {
        switch (tf) {
          case 1: nx = 0.0; break;
          case 2: nx = 1.0; break;
          case 3: ny = 0.0; break;
          case 4: ny = 1.0; break;
          case 5: nz = 0.0; break;
          default: nz = 1.0; break;
        }
      } else {
        // ConvertToModelXY of this cell, ConvertToLocalXY of the next
        double left = G.cx[c - 1] - (G.dx[c - 1] / 2.0);
        double right = G.cx[c - 1] + (G.dx[c - 1] / 2.0);
        double front = G.cy[c - 1] - (G.dy[c - 1] / 2.0);
        double back = G.cy[c - 1] + (G.dy[c - 1] / 2.0);
        double gx = ((1.0 - fx) * left) + (fx * right);
        if (gx < left) gx = left;
        if (gx > right) gx = right;
        double gy = ((1.0 - fy) * front) + (fy * back);
        if (gy < front) gy = front;
        if (gy > back) gy = back;
        left = G.cx[conn - 1] - (G.dx[conn - 1] / 2.0);
        right = G.cx[conn - 1] + (G.dx[conn - 1] / 2.0);
        front = G.cy[conn - 1] - (G.dy[conn - 1] / 2.0);
        back = G.cy[conn - 1] + (G.dy[conn - 1] / 2.0);
        nx = (gx - left) / (right - left);
        if (nx > 1.0) nx = 1.0;
        if (nx < 0.0) nx = 0.0;
        ny = (gy - front) / (back - front);
        if (ny > 1.0) ny = 1.0;
        if (ny < 0.0) ny = 0.0;
        // SnapToValue with a 1e-3 tolerance on the face coordinate
        const double tol = 1.0e-3;
        if (tf <= 2) {
          const double sv = tf == 1 ? 0.0 : 1.0;
          if (nx >= sv - (tol / 2.0) && nx <= sv + (tol / 2.0)) nx = sv;
        } else if (tf <= 4) {
          const double sv = tf == 3 ? 0.0 : 1.0;
          if (ny >= sv - (tol / 2.0) && ny <= sv + (tol / 2.0)) ny = sv;
        }
        if (tf == 5) nz = 0.0;
        if (tf == 6) nz = 1.0;
      }
      c = conn;
      V.x = nx; V.y = ny; V.z = nz;
      A.phase = 0;
      done = !enter_cell();
    }
    if (more) {
      S.cell[p] = c;
      S.lx[p] = V.x; S.ly[p] = V.y; S.lz[p] = V.z; S.t[p] = V.t;
      A.restarts = V.restarts;
      A.exitFace = V.exitFace;
      S.aux[p] = pack_aux(A) | (count << kCountShift);
      S.ix[p] = V.ix; S.iy[p] = V.iy; S.iz[p] = V.iz; S.t0[p] = V.t0;
      G.unit[p] = u;
      G.ex[p] = ex; G.ey[p] = ey; G.ez[p] = ez; G.et[p] = et;
      G.pcell[p] = pc; G.px[p] = px; G.py[p] = py; G.pz[p] = pz;
    } else {
      S.cell[p] = c;
      S.lx[p] = V.x; S.ly[p] = V.y; S.lz[p] = V.z; S.t[p] = V.t;
      S.status[p] = status;
      ended_active = status == 1 ? 1ull : 0ull;
    }
    S.rng[2 * p] = rng.a;
    S.rng[2 * p + 1] = rng.b;
    S.substeps[p] += (int64_t)steps;
  }
  append_work(more, u, p, nkeys, nvals, &counters[2]);
  wave_count(ended_active, steps, &counters[0], &counters[1]);
}

// Particles still listed after the step guard: status 9 at the last point
// before their cell entry.
__global__ void terminate_walk_usg(const UsgWalk G, const int32_t *__restrict__ work,
                                   int64_t nwork, ParticlePtrs<double> S) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwork) return;
  const int p = work[i];
  const Aux A = unpack_aux(S.aux[p]);
  if (A.phase) {
    S.cell[p] = G.pcell[p];
    S.lx[p] = G.px[p]; S.ly[p] = G.py[p]; S.lz[p] = G.pz[p]; S.t[p] = G.et[p];
  }
  S.status[p] = 9;
}

}  // namespace
}  // namespace rwpt
