// Unstructured grids of rectangular cells (MODFLOW-USG, MODFLOW 6 DISV
// smoothed quadtrees: a horizontal face meets one cell of its size, one
// twice its size or two of half its size).
//
// The reference splits a cell with a face of two sub-faces into 2 x 2
// sub-cells and tracks in the sub-cells (TrackCell.f90:500-669); at every
// cell entry it rebuilds the cell, a 2 x 18 neighbour buffer that may hold
// real cells, cells forced into 2 x 2 sub-cells and sub-cells of bigger
// cells (ParticleTrackingEngine.f90:1257-2005), and for every sub-cell the
// 18-cell stencil of its corner interpolation (TrackSubCell.F90:3856-4012).
//
// Here the tracking unit is the sub-cell (one per unrefined cell, four per
// refined one) and every unit gets, once per flow field, the same 96-double
// record the structured walk reads.  The work splits in two:
//  * host, once per grid (integers only): JaFace (ComputeFaceAssignments,
//    ModflowRectangularGrid.f90:98-253), the sub-face connections of every
//    cell, and for every unit a symbolic run of the neighbour fill that
//    records how each of its 18 stencil entries is formed: which real cell,
//    whether that cell is forced into sub-cells, which sub-cell (and
//    sub-sub-cell) of it, which of its sub-cells is read and whether its
//    sub-cell indexing is skipped;
//  * device, once per flow field (one thread per cell, then per unit): the
//    cell assembly with FLOW-JA-FACE and boundary flows
//    (SetMassTransportDataUnstructured, ModpathCellData.f90:674-1096), then
//    each unit's stencil numbers from its descriptors (sub-cell flows,
//    ComputeSubCellFlows :2340-2441, face areas and volumes :3035-3130) and
//    the record (record_from_stencil, shared with the structured build).
//
// Buffer states follow what the reference's compiled code does: neighbour
// buffers start reset (sub-face and sub-cell counts 1); a buffer filled
// from a sub-cell, an intent(out) dummy, and every local buffer start with
// all components zero, counts included (the compiler initialises such
// objects from the type's default initialisation), so that a sub-cell
// buffer not forced into sub-cells has zero sub-cell counts.  The
// reference's own reset loop (:1269-1273) skips buffers whose initialized
// flag is false, which makes its results on refined grids depend on the
// cells a thread visited before; every neighbourhood here is built from
// reset buffers.
//
// Built with -ffp-contract=off: the records are the doubles the reference
// forms.
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "rwpt_internal.h"
#include "rwpt_record.h"

namespace rwpt {

// ---------------------------------------------------------------------------
// Host: grid topology
// ---------------------------------------------------------------------------

namespace {

struct HostGrid {
  int ncell = 0;
  const int32_t *ia = nullptr, *ja = nullptr, *layer = nullptr;
  const double *cx = nullptr, *cy = nullptr, *dx = nullptr, *dy = nullptr;
  std::vector<int32_t> jaface;
  // per cell: potential count, sub-face counts and connections per face
  std::vector<int32_t> pcount, sfc, conn;  // 6, 6, 24 per cell
  std::vector<int32_t> qidx;               // 24 per cell: flow index or -1

  int off(int n) const { return ia[n - 1] - 1; }
  int cnt(int n) const { return ia[n] - ia[n - 1]; }
  double left(int n) const { return cx[n - 1] - (dx[n - 1] / 2.0); }
  double right(int n) const { return cx[n - 1] + (dx[n - 1] / 2.0); }
  double front(int n) const { return cy[n - 1] - (dy[n - 1] / 2.0); }
  double back(int n) const { return cy[n - 1] + (dy[n - 1] / 2.0); }

  // GetPotentialFaceConnectionCount (:321-343)
  int potential(int n, int face) const {
    const int o = off(n), f1 = 10000 * face, f2 = 10000 * (face + 1);
    for (int i = 1; i < cnt(n); ++i) {
      const int v = jaface[o + i];
      if (v > f1 && v < f2) return v - 100 * (int)((double)v / 100.0);
    }
    return 0;
  }
  // GetFaceConnection (:345-364)
  int face_conn(int n, int face, int sub) const {
    if (n <= 0) return 0;
    const int o = off(n), lo = 10000 * face + 100 * sub, hi = lo + 100;
    for (int i = 0; i < cnt(n); ++i) {
      const int v = jaface[o + i];
      if (v >= lo && v < hi) return ja[o + i];
    }
    return 0;
  }
  // FindConnectionIndex (:498-513), 0-based, -1 when absent
  int find_index(int n, int m) const {
    const int o = off(n);
    for (int i = 1; i < cnt(n); ++i)
      if (ja[o + i] == m) return o + i;
    return -1;
  }
  // FindConnectionPosition (:479-496): the face of n that m meets
  int connection_face(int n, int m) const {
    if (n < 1 || n > ncell || m == n) return 0;
    const int idx = find_index(n, m);
    if (idx < 0) return 0;
    const int k = 100 * (int)((double)jaface[idx] / 100.0);
    for (int f = 1; f <= 6; ++f)
      if (k > 10000 * f && k < 10000 * (f + 1)) return f;
    return 0;
  }
};

double snap_ratio(double r, double tol) {
  if (std::fabs(r - 0.5) <= tol) return 0.5;
  if (std::fabs(r - 1.0) <= tol) return 1.0;
  if (std::fabs(r - 2.0) <= tol) return 2.0;
  return r;
}

// ComputeFaceAssignments (ModflowRectangularGrid.f90:98-253) for the
// connection n -> m: face * 10000 + sub-face * 100 + sub-face count; -1 for
// a spacing a smoothed quadtree cannot have (the reference stops there).
int face_assignment(const HostGrid &G, int n, int m) {
  const double tol = (double)0.0001f;
  const int ln = G.layer[n - 1], lm = G.layer[m - 1];
  const double rx = snap_ratio(G.dx[m - 1] / G.dx[n - 1], tol);
  const double ry = snap_ratio(G.dy[m - 1] / G.dy[n - 1], tol);
  const double xn = G.cx[n - 1], yn = G.cy[n - 1], xm = G.cx[m - 1], ym = G.cy[m - 1];
  if (lm != ln) {
    int code = lm > ln ? 50000 : 60000;
    if (rx == 0.5 && ry == 0.5) {
      if (ym > yn && xm < xn) code += 104;
      else if (ym > yn && xm > xn) code += 204;
      else if (ym < yn && xm < xn) code += 304;
      else if (ym < yn && xm > xn) code += 404;
    } else if ((rx == 1.0 || rx == 2.0) && rx == ry) {
      code += 101;
    } else {
      return -1;
    }
    return code;
  }
  const bool west = std::fabs(G.left(n) - G.right(m)) <= tol;
  const bool east = std::fabs(G.right(n) - G.left(m)) <= tol;
  if (west || east) {
    int code = west ? 10000 : 20000;
    if (ry == 0.5) {
      if (ym > yn) code += 102;
      else if (ym < yn) code += 202;
    } else if (ry == 1.0 || ry == 2.0) {
      code += 101;
    } else {
      return -1;
    }
    return code;
  }
  const bool south = std::fabs(G.front(n) - G.back(m)) <= tol;
  const bool north = std::fabs(G.back(n) - G.front(m)) <= tol;
  if (south || north) {
    int code = south ? 30000 : 40000;
    if (rx <= 0.5) {
      if (xm < xn) code += 102;
      else if (xm > xn) code += 202;
    } else if (rx == 1.0 || rx == 2.0) {
      code += 101;
    } else {
      return -1;
    }
    return code;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Host: symbolic neighbour buffers
// ---------------------------------------------------------------------------

// What a buffer of the reference's neighbour fill holds: the fields its
// control flow reads, and how its numbers are formed (the real cell `base`,
// forced into sub-cells or not, then up to two sub-cell selections).
struct SBuf {
  int cell = 0;
  int sfc[6] = {1, 1, 1, 1, 1, 1};
  int conn[6][4] = {};
  int scr = 1, scc = 1;
  bool fromSub = false, isParent = false;
  int parentCell = 0, parentRow = 0, parentCol = 0, rfd = 0;
  int base = 0;
  bool force0 = false;
  int nsub = 0;
  int r[2] = {0, 0}, c[2] = {0, 0};
  bool f[2] = {false, false};

  static SBuf reset() { return SBuf(); }  // pr_Reset
  static SBuf zero() {                    // all components zero
    SBuf b;
    for (int k = 0; k < 6; ++k) b.sfc[k] = 0;
    b.scr = b.scc = 0;
    return b;
  }
  int subcount() const { return scr * scc; }
  // GetFaceConnection of the buffer (ModpathCellData.f90:2257-2283)
  int face_conn(int face, int sub) const {
    if (sub < 1 || sub > sfc[face - 1]) return 0;
    return conn[face - 1][sub - 1];
  }
};

struct Filler {
  const HostGrid &G;
  bool error = false;

  // pr_FillCellBuffer: the cell's topology and numbers; the sub-cell flags
  // keep what the buffer held (SetMassTransportDataUnstructured resets the
  // arrays only)
  void real(int n, SBuf &b) const {
    if (n <= 0) return;
    b.cell = n;
    for (int f = 0; f < 6; ++f) {
      b.sfc[f] = G.sfc[6 * (n - 1) + f];
      for (int s = 0; s < 4; ++s) b.conn[f][s] = G.conn[24 * (n - 1) + 4 * f + s];
    }
    b.scr = b.scc = 1;
    for (int f = 0; f < 6; ++f)
      if (b.sfc[f] > 1) b.scr = b.scc = 2;
    b.base = n;
    b.force0 = false;
    b.nsub = 0;
  }
  // pr_FillCellFromRealData (:1974-2002)
  void real_forced(int n, SBuf &b, bool force) const {
    if (n <= 0) return;
    real(n, b);
    if (force) {
      b.scr = b.scc = 2;
      b.force0 = true;
    }
  }
  // pr_FillCellBufferFromSubCell (:1907-1971)
  void from_sub(const SBuf &parent, int r, int c, SBuf &out, bool force) {
    SBuf o = SBuf::zero();
    if (force) o.scr = o.scc = 2;
    o.fromSub = true;
    o.parentCell = parent.cell;
    o.parentRow = r;
    o.parentCol = c;
    o.base = parent.base;
    o.force0 = parent.force0;
    o.nsub = parent.nsub;
    for (int k = 0; k < 2; ++k) {
      o.r[k] = parent.r[k];
      o.c[k] = parent.c[k];
      o.f[k] = parent.f[k];
    }
    if (o.nsub >= 2 || o.base <= 0) {
      error = true;  // three levels of sub-cells: not a smoothed quadtree
      out = o;
      return;
    }
    o.r[o.nsub] = r;
    o.c[o.nsub] = c;
    o.f[o.nsub] = force;
    o.nsub += 1;
    out = o;
  }

  // pr_FillNeighborCellsConnectionFromHorizontalFace (:1649-1903)
  void from_horizontal(const SBuf &cc, int face, SBuf *nbuf, bool force) {
    if (cc.cell <= 0) return;
    SBuf parentBuf = SBuf::zero(), auxBuf = SBuf::zero();
    const int direct = cc.face_conn(face, 1);
    const int dir = face <= 2 ? 1 : (face <= 4 ? 2 : 3);
    if (cc.sfc[face - 1] > 1) {
      if (!cc.isParent) {
        for (int m = 0; m < cc.sfc[face - 1] && m < 2; ++m) {
          const int n = cc.face_conn(face, m + 1);
          if (n > 0) real(n, nbuf[m]);
          nbuf[m].rfd = dir;
        }
      } else {
        if (dir != 2) {  // the reference stops ("inconsistency")
          error = true;
          return;
        }
        const int n = cc.face_conn(face, cc.parentCol);
        if (n > 0) real(n, nbuf[0]);
        nbuf[0].rfd = dir;
        if (force) {
          nbuf[0].scr = nbuf[0].scc = 2;
          nbuf[0].force0 = true;
        }
      }
      return;
    }
    if (direct <= 0) return;
    int orth[2], sci[2][2];
    switch (face) {
      case 1: orth[0] = 3; orth[1] = 4; sci[0][0] = 1; sci[0][1] = 2; sci[1][0] = 2; sci[1][1] = 2; break;
      case 2: orth[0] = 3; orth[1] = 4; sci[0][0] = 1; sci[0][1] = 1; sci[1][0] = 2; sci[1][1] = 1; break;
      case 3: orth[0] = 1; orth[1] = 2; sci[0][0] = 1; sci[0][1] = 2; sci[1][0] = 1; sci[1][1] = 1; break;
      default: orth[0] = 1; orth[1] = 2; sci[0][0] = 2; sci[0][1] = 2; sci[1][0] = 2; sci[1][1] = 1; break;
    }
    const int oc0 = cc.face_conn(orth[0], 1), oc1 = cc.face_conn(orth[1], 1);
    const bool bigger = (oc0 > 0 && direct == G.face_conn(oc0, face, 1)) ||
                        (oc1 > 0 && direct == G.face_conn(oc1, face, 1));
    if (bigger) {
      real_forced(direct, parentBuf, true);
      const int row = direct == G.face_conn(oc0, face, 1) ? 0 : 1;
      if (!cc.isParent) {
        from_sub(parentBuf, sci[row][0], sci[row][1], nbuf[0], force);
        nbuf[0].rfd = dir;
      } else {
        from_sub(parentBuf, sci[row][0], sci[row][1], auxBuf, true);
        auxBuf.isParent = true;
        auxBuf.cell = parentBuf.cell;
        from_sub(auxBuf, sci[row][0], sci[row][1], nbuf[0], force);
        nbuf[0].rfd = dir;
      }
    } else if (!cc.isParent) {
      real_forced(direct, nbuf[0], force);
      nbuf[0].rfd = dir;
    } else {
      int subRow = 0, subCol = 0;
      if (dir == 2) {
        subCol = cc.parentCol;
        subRow = face == 3 ? 1 : 2;
      } else if (dir == 1) {
        subRow = cc.parentRow;
        subCol = face == 1 ? 2 : 1;
      }
      real_forced(direct, parentBuf, true);
      from_sub(parentBuf, subRow, subCol, nbuf[0], force);
      nbuf[0].rfd = dir;
    }
  }

  // pr_FillNeighborCellsSubBuffer (:1300-1646); nb[slot][buffer], slots
  // 0-based, d the 0-based slot of the direct connection
  void sub_buffer(const SBuf &cc, int face, int first, SBuf (*nb)[2], int d, bool force) {
    SBuf auxBuf = SBuf::zero();
    SBuf *D = nb[d];
    if (face >= 5) {
      real_forced(cc.face_conn(face, 1), D[0], force);
      D[0].rfd = 3;
      from_horizontal(D[0], first, nb[d + 1], force);
      from_horizontal(D[0], first + 1, nb[d + 2], force);
      return;
    }
    if (cc.sfc[face - 1] > 1) {
      from_horizontal(cc, face, D, force);
      if (first >= 5) {
        for (int m = 0; m < 2; ++m) {
          real_forced(D[m].face_conn(first, 1), nb[d + 1][m], force);
          nb[d + 1][m].rfd = 3;
        }
        for (int m = 0; m < 2; ++m) {
          real_forced(D[m].face_conn(first + 1, 1), nb[d + 2][m], force);
          nb[d + 2][m].rfd = 3;
        }
        return;
      }
      const int i0 = first <= 2 ? 0 : 1, i1 = first <= 2 ? 1 : 0;
      from_horizontal(D[i0], first, nb[d + 1], false);
      from_horizontal(D[i1], first + 1, nb[d + 2], false);
      return;
    }
    if (cc.face_conn(face, 1) <= 0) return;
    from_horizontal(cc, face, D, force);
    if (D[0].fromSub) {
      if (first >= 5) {
        real_forced(G.face_conn(D[0].parentCell, first, 1), auxBuf, true);
        if (auxBuf.cell > 0) {
          from_sub(auxBuf, D[0].parentRow, D[0].parentCol, nb[d + 1][0], force);
          nb[d + 1][0].rfd = 3;
        }
        real_forced(G.face_conn(D[0].parentCell, first + 1, 1), auxBuf, true);
        if (auxBuf.cell > 0) {
          from_sub(auxBuf, D[0].parentRow, D[0].parentCol, nb[d + 2][0], force);
          nb[d + 2][0].rfd = 3;
        }
        return;
      }
      real_forced(D[0].parentCell, auxBuf, true);
      auxBuf.isParent = true;
      auxBuf.parentRow = D[0].parentRow;
      auxBuf.parentCol = D[0].parentCol;
      if (first < 3) {  // unset indexes in the reference (not reached)
        error = true;
        return;
      }
      const bool row1 = D[0].parentRow == 1;
      const int sameIdx = row1 ? d + 1 : d + 2, notSameIdx = row1 ? d + 2 : d + 1;
      const int notSameFace = row1 ? first + 1 : first;
      const int sameRow = row1 ? 2 : 1, sameCol = D[0].parentCol;
      from_horizontal(auxBuf, notSameFace, nb[notSameIdx], force);
      from_sub(auxBuf, sameRow, sameCol, nb[sameIdx][0], force);
      nb[sameIdx][0].rfd = 2;
      return;
    }
    if (first >= 5) {
      real_forced(D[0].face_conn(first, 1), nb[d + 1][0], force);
      nb[d + 1][0].rfd = 3;
      real_forced(D[0].face_conn(first + 1, 1), nb[d + 2][0], force);
      nb[d + 2][0].rfd = 3;
    } else {
      from_horizontal(D[0], first, nb[d + 1], force);
      from_horizontal(D[0], first + 1, nb[d + 2], force);
    }
  }

  // pr_FillNeighborCellDataUnstructured (:1257-1297), buffers reset first
  void fill(const SBuf &center, SBuf (*nb)[2]) {
    for (int m = 0; m < 18; ++m) nb[m][0] = nb[m][1] = SBuf::reset();
    const bool force = center.subcount() > 1;
    for (int n = 1; n <= 6; ++n) {
      const int first = n <= 2 ? 3 : (n <= 4 ? 5 : 1);
      sub_buffer(center, n, first, nb, (n - 1) * 3, force);
    }
  }
};

// pr_GetNeighborSubCellIndexes (ModpathCellData.f90:2875-3030): per slot
// (buffer 1..18, or 0 for a sub-cell of the cell itself; sub-row;
// sub-column) for the sub-cells (1,1), (1,2), (2,1), (2,2)
const int kNsi[4][18][3] = {
    {{1, 1, 2}, {1, 2, 2}, {3, 2, 2}, {0, 1, 2}, {0, 2, 2}, {10, 2, 2},
     {0, 2, 1}, {13, 2, 1}, {16, 2, 1}, {10, 2, 1}, {11, 2, 1}, {12, 2, 1},
     {13, 1, 1}, {14, 1, 2}, {13, 1, 2}, {16, 1, 1}, {17, 1, 2}, {16, 1, 2}},
    {{0, 1, 1}, {0, 2, 1}, {10, 2, 1}, {4, 1, 1}, {4, 2, 1}, {6, 2, 1},
     {0, 2, 2}, {13, 2, 2}, {16, 2, 2}, {10, 2, 2}, {11, 2, 2}, {12, 2, 2},
     {13, 1, 2}, {13, 1, 1}, {15, 1, 1}, {16, 1, 2}, {16, 1, 1}, {18, 1, 1}},
    {{1, 2, 2}, {2, 1, 2}, {1, 1, 2}, {0, 2, 2}, {7, 1, 2}, {0, 1, 2},
     {7, 1, 1}, {8, 1, 1}, {9, 1, 1}, {0, 1, 1}, {13, 1, 1}, {16, 1, 1},
     {13, 2, 1}, {14, 2, 2}, {13, 2, 2}, {16, 2, 1}, {17, 2, 2}, {16, 2, 2}},
    {{0, 2, 1}, {7, 1, 1}, {0, 1, 1}, {4, 2, 1}, {5, 1, 1}, {4, 1, 1},
     {7, 1, 2}, {8, 1, 2}, {9, 1, 2}, {0, 1, 2}, {13, 1, 2}, {16, 1, 2},
     {13, 2, 2}, {13, 2, 1}, {15, 2, 1}, {16, 2, 2}, {16, 2, 1}, {18, 2, 1}}};

// Slot descriptor bits (with the base cell in the other word)
enum : int32_t {
  SL_KIND = 0,     // 2 bits: 0 none, 1 a sub-cell of the unit's own cell, 2 buffer
  SL_ROW = 2,      // 2 bits: sub-row read
  SL_COL = 4,      // 2 bits: sub-column read
  SL_SKIP = 6,     // sub-cell indexing skipped (whole-cell flows and sizes)
  SL_FORCE0 = 7,   // the real cell forced into 2 x 2 sub-cells
  SL_NSUB = 8,     // 2 bits: sub-cell selections
  SL_R1 = 10, SL_C1 = 12, SL_F1 = 14,
  SL_R2 = 15, SL_C2 = 17, SL_F2 = 19,
};

int32_t pack_chain(const SBuf &b) {
  return (b.force0 ? 1 << SL_FORCE0 : 0) | (b.nsub << SL_NSUB) | (b.r[0] << SL_R1) |
         (b.c[0] << SL_C1) | (b.f[0] ? 1 << SL_F1 : 0) | (b.r[1] << SL_R2) |
         (b.c[1] << SL_C2) | (b.f[1] ? 1 << SL_F2 : 0);
}

}  // namespace

// Host side of rwpt_create_usg: checks the grid, assigns faces, lists the
// sub-face connections, numbers the units and forms every unit's stencil
// descriptors.  Returns an error text, or an empty string.
std::string usg_build_topology(const rwpt_usg_grid &g, UsgHost &H) {
  HostGrid G;
  G.ncell = g.ncell;
  G.ia = g.ia;
  G.ja = g.ja;
  G.layer = g.layer;
  G.cx = g.cell_x;
  G.cy = g.cell_y;
  G.dx = g.del_x;
  G.dy = g.del_y;
  const int n = g.ncell;
  G.jaface.assign((size_t)g.nja, 0);
  for (int c = 1; c <= n; ++c) {
    const int o = G.off(c), k = G.cnt(c);
    for (int i = 1; i < k; ++i) {
      const int code = face_assignment(G, c, g.ja[o + i]);
      if (code < 0) return "cell spacing not a smoothed rectangular quadtree";
      G.jaface[o + i] = code;
    }
  }
  G.pcount.assign(6 * (size_t)n, 0);
  G.sfc.assign(6 * (size_t)n, 1);
  G.conn.assign(24 * (size_t)n, 0);
  G.qidx.assign(24 * (size_t)n, -1);
  H.cell_unit.assign((size_t)n + 1, 0);
  for (int c = 1; c <= n; ++c) {
    bool split = false;
    for (int f = 1; f <= 6; ++f) {
      const int count = G.potential(c, f);
      if (count > (f <= 4 ? 2 : 4)) return "more sub-faces on a face than a quadtree has";
      G.pcount[6 * (c - 1) + f - 1] = count;
      if (count > 0) {
        G.sfc[6 * (c - 1) + f - 1] = count;
        for (int s = 1; s <= count; ++s) {
          const int m = G.face_conn(c, f, s);
          G.conn[24 * (c - 1) + 4 * (f - 1) + s - 1] = m;
          if (m > 0) G.qidx[24 * (c - 1) + 4 * (f - 1) + s - 1] = G.find_index(c, m);
        }
      }
      split |= count > 1;
    }
    H.cell_unit[c] = H.cell_unit[c - 1] + (split ? 4 : 1);
  }
  const int nunit = H.cell_unit[n];
  H.nunit = nunit;
  H.topo.assign((size_t)USG_TOPO * n, 0);
  for (int c = 0; c < n; ++c) {
    int32_t *t = &H.topo[(size_t)USG_TOPO * c];
    for (int f = 0; f < 6; ++f) {
      t[UT_PCOUNT + f] = G.pcount[6 * c + f];
      t[UT_SFC + f] = G.sfc[6 * c + f];
    }
    for (int k = 0; k < 24; ++k) {
      t[UT_CONN + k] = G.conn[24 * c + k];
      t[UT_QIDX + k] = G.qidx[24 * c + k];
    }
  }
  H.side.assign((size_t)USG_SIDE * nunit, 0);
  H.off.assign(4 * (size_t)nunit, 0.0);
  H.slots.assign((size_t)36 * nunit, 0);
  Filler F{G};
  SBuf nb[18][2];
  static const int opp[6] = {2, 1, 4, 3, 6, 5};
  for (int c = 1; c <= n; ++c) {
    SBuf center = SBuf::reset();
    F.real(c, center);
    F.fill(center, nb);
    if (F.error) return "neighbourhood the reference does not define (three refinement levels)";
    const int nsub = center.subcount();
    const int u0 = H.cell_unit[c - 1];
    for (int k = 0; k < nsub; ++k) {
      const int u = u0 + k, r = nsub == 1 ? 1 : 1 + k / 2, col = nsub == 1 ? 1 : 1 + k % 2;
      int32_t *sd = &H.side[(size_t)USG_SIDE * u];
      sd[US_CELL] = c;
      sd[US_ROW] = r;
      sd[US_COL] = col;
      sd[US_NSUB] = nsub;
      // offsets of the sub-cell in its cell (FillMassSubCellDataBuffer :2569-2584)
      const int scc = nsub == 1 ? 1 : 2, scr = scc;
      const double xinc = 1.0 / (double)scc, yinc = 1.0 / (double)scr;
      double *of = &H.off[4 * (size_t)u];
      of[0] = (col - 1) * xinc;
      of[1] = col < scc ? col * xinc : 1.0;
      of[2] = (scr - r) * yinc;
      of[3] = r > 1 ? (scr - r + 1) * yinc : 1.0;
      // where each face leads: a sibling unit, or the face of the neighbour
      // that meets this cell with ConvertFromNeighbor's full-face test
      for (int f = 1; f <= 6; ++f) {
        int code = 0;
        const bool internal = nsub == 4 && ((f == 1 && col == 2) || (f == 2 && col == 1) ||
                                            (f == 3 && r == 1) || (f == 4 && r == 2));
        if (internal) {
          const int r2 = f == 3 ? 2 : (f == 4 ? 1 : r);
          const int c2 = f == 1 ? 1 : (f == 2 ? 2 : col);
          code = u0 + (r2 - 1) * 2 + (c2 - 1);
        } else {
          // the sub-cell's connection (FillMassSubCellDataBuffer :2596-2666)
          const int cnt = G.sfc[6 * (c - 1) + f - 1];
          int m = 0;
          if (f <= 4) {
            const int sub = (f <= 2) ? r : col;
            m = cnt == 1 ? G.conn[24 * (c - 1) + 4 * (f - 1)]
                         : (cnt > 1 ? G.conn[24 * (c - 1) + 4 * (f - 1) + sub - 1] : 0);
          } else if (cnt == 1) {
            m = G.conn[24 * (c - 1) + 4 * (f - 1)];
          } else if (cnt == nsub) {
            m = G.conn[24 * (c - 1) + 4 * (f - 1) + (r - 1) * scc + col - 1];
          }
          if (m > 0) {
            const int tf = G.connection_face(m, c);
            const bool simple = tf > 0 && G.potential(m, tf) == 1 &&
                                G.potential(c, opp[tf - 1]) == 1;
            code = tf | (simple ? 8 : 0);
          }
        }
        sd[US_FACE + f - 1] = code;
      }
      // stencil descriptors (pr_FillNeighborSubCellVariables, TrackSubCell.F90:3856-4012)
      int32_t *sl = &H.slots[(size_t)36 * u];
      for (int s = 0; s < 18; ++s) {
        int32_t base = 0, bits = 0;
        if (nsub == 1) {
          const SBuf &b = nb[s][0];
          if (b.cell != 0) {
            base = b.base;
            bits = (2 << SL_KIND) | (1 << SL_ROW) | (1 << SL_COL) | (1 << SL_SKIP) |
                   pack_chain(b);
          }
        } else {
          const int *ix = kNsi[k][s];
          if (ix[0] == 0) {
            base = c;
            bits = (1 << SL_KIND) | (ix[1] << SL_ROW) | (ix[2] << SL_COL);
          } else {
            const SBuf *P = nb[ix[0] - 1];
            int which = -1, row = 1, colr = 1;
            bool skip = false;
            if (P[1].cell > 0) {
              if (P[0].subcount() > 1 || P[1].subcount() > 1) skip = true;
              which = P[1].rfd == 1 ? ix[1] - 1 : ix[2] - 1;
            } else if (P[0].cell > 0 || P[0].fromSub) {
              which = 0;
              if (P[0].subcount() > 1) {
                row = ix[1];
                colr = ix[2];
              }
            }
            if (which >= 0) {
              const SBuf &b = P[which];
              base = b.base;
              bits = (2 << SL_KIND) | (row << SL_ROW) | (colr << SL_COL) |
                     (skip ? 1 << SL_SKIP : 0) | pack_chain(b);
            }
          }
        }
        sl[2 * s] = base;
        sl[2 * s + 1] = bits;
      }
    }
  }
  return "";
}

// ---------------------------------------------------------------------------
// Device: cell assembly, stencil numbers, unit records
// ---------------------------------------------------------------------------

namespace {

struct UsgView {
  int32_t ncell, nunit;
  const int32_t *topo, *cell_unit, *side, *slots;
  const double *off, *cx, *cy, *dx, *dy, *top, *bot, *flowja, *storage, *bflow;
  const double *porosity, *retardation, *head, *source, *sink, *disp;
  const int32_t *ibound, *icbound, *layer_type, *zone;
  int64_t ndisp;
  int32_t default_icbound, uniform_porosity;
  double *casm;
};

constexpr int ASM_Q = 0, ASM_SRC = 24, ASM_SNK = 25, ASM_STO = 26, ASM_N = 28;

__device__ __forceinline__ int ib(const UsgView &V, int n) {
  return V.ibound ? V.ibound[n - 1] : 1;
}

// SetMassTransportDataUnstructured's face flows (FLOW-JA-FACE, MODFLOW-2005
// signs: faces 1, 3, 5 as given, 2, 4, 6 negated) and boundary flows
// (:674-1096), one thread per cell
__global__ void usg_assemble(const UsgView V) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= V.ncell) return;
  const int32_t *t = V.topo + (int64_t)USG_TOPO * c;
  double *a = V.casm + (int64_t)ASM_N * c;
  double Q[6][4];
  for (int f = 0; f < 6; ++f)
    for (int s = 0; s < 4; ++s) Q[f][s] = 0.0;
  double src = V.source ? V.source[c] : 0.0;
  double snk = V.sink ? V.sink[c] : 0.0;
  for (int f = 0; f < 6; ++f) {
    const int cnt = t[UT_SFC + f];
    for (int s = 0; s < cnt; ++s) {
      const int m = t[UT_CONN + 4 * f + s], q = t[UT_QIDX + 4 * f + s];
      if (m > 0 && q >= 0) Q[f][s] = (f % 2 == 0) ? V.flowja[q] : -V.flowja[q];
    }
  }
  if (V.bflow) {
    for (int f = 0; f < 6; ++f) {
      const double bf = V.bflow[6 * (int64_t)c + f];
      if (bf == 0.0) continue;
      const int cnt = t[UT_SFC + f];
      int nb = 1;  // SubFaceBoundaryCounts: 1 on a face without connections
      if (t[UT_PCOUNT + f] > 0) {
        nb = 0;
        for (int s = 0; s < cnt; ++s) {
          const int m = t[UT_CONN + 4 * f + s];
          if (m == 0 || ib(V, m) == 0) nb = nb + 1;
        }
      }
      if (nb > 0) {
        const double flow = bf / nb;
        for (int s = 0; s < cnt; ++s) {
          const int m = t[UT_CONN + 4 * f + s];
          if (m == 0 || ib(V, m) == 0) Q[f][s] = (f % 2 == 0) ? flow : -flow;
        }
      } else if (bf > 0.0) {
        src = src + bf;
      } else {
        snk = snk + bf;
      }
    }
  }
  for (int f = 0; f < 6; ++f)
    for (int s = 0; s < 4; ++s) a[ASM_Q + 4 * f + s] = Q[f][s];
  a[ASM_SRC] = src;
  a[ASM_SNK] = snk;
  a[ASM_STO] = V.storage ? V.storage[c] : 0.0;
  a[ASM_N - 1] = 0.0;
}

// A cell buffer's numbers (ModpathCellDataType's flow and size fields).
struct NBuf {
  double Q[6][4];
  int sfc[6];
  int scr, scc;
  double sf[4];  // SubCellFlows
  double DX, DY, top, bottom, head;
  int lt;
  double por, ret, src, snk, sto;
};

__device__ __forceinline__ double sub_boundary_flow(const NBuf &b, int face, int sub) {
  if (b.sfc[face - 1] == 1) return b.Q[face - 1][0] / (face <= 4 ? 2.0 : 4.0);
  return b.Q[face - 1][sub - 1];
}

// ComputeSubCellFlows (:2340-2393)
__device__ void subcell_flows(NBuf &b) {
  if (b.scr * b.scc == 1) return;
  const double qsrc = b.src / 4.0, qsnk = b.snk / 4.0, qsto = b.sto / 4.0;
  double q = 0.0, b1, b2, b3;
  q = q + sub_boundary_flow(b, 1, 1);
  q = q - sub_boundary_flow(b, 4, 1);
  q = q + sub_boundary_flow(b, 5, 1);
  q = q - sub_boundary_flow(b, 6, 1);
  b1 = q + qsrc + qsnk + qsto;
  q = 0.0;
  q = q - sub_boundary_flow(b, 2, 1);
  q = q - sub_boundary_flow(b, 4, 2);
  q = q + sub_boundary_flow(b, 5, 2);
  q = q - sub_boundary_flow(b, 6, 2);
  b2 = q + qsrc + qsnk + qsto;
  q = 0.0;
  q = q + sub_boundary_flow(b, 1, 2);
  q = q + sub_boundary_flow(b, 3, 1);
  q = q + sub_boundary_flow(b, 5, 3);
  q = q - sub_boundary_flow(b, 6, 3);
  b3 = q + qsrc + qsnk + qsto;
  b.sf[0] = 0.5 * (b1 + 0.5 * (b3 - b2));
  b.sf[2] = -b1 + b.sf[0];
  b.sf[1] = b3 - b.sf[2];
  b.sf[3] = -b2 - b.sf[0];
}

// FillSubCellFaceFlowsBuffer (:2817-2867)
__device__ void subcell_face_flows(const NBuf &b, int r, int c, double fl[6]) {
  if (b.scr * b.scc == 1) {
    for (int f = 0; f < 6; ++f) fl[f] = b.Q[f][0];
    return;
  }
  switch ((r - 1) * b.scc + c) {
    case 1:
      fl[0] = sub_boundary_flow(b, 1, 1); fl[1] = b.sf[0]; fl[2] = b.sf[2];
      fl[3] = sub_boundary_flow(b, 4, 1); fl[4] = sub_boundary_flow(b, 5, 1);
      fl[5] = sub_boundary_flow(b, 6, 1);
      break;
    case 2:
      fl[0] = b.sf[0]; fl[1] = sub_boundary_flow(b, 2, 1); fl[2] = b.sf[3];
      fl[3] = sub_boundary_flow(b, 4, 2); fl[4] = sub_boundary_flow(b, 5, 2);
      fl[5] = sub_boundary_flow(b, 6, 2);
      break;
    case 3:
      fl[0] = sub_boundary_flow(b, 1, 2); fl[1] = b.sf[1];
      fl[2] = sub_boundary_flow(b, 3, 1); fl[3] = b.sf[2];
      fl[4] = sub_boundary_flow(b, 5, 3); fl[5] = sub_boundary_flow(b, 6, 3);
      break;
    case 4:
      fl[0] = b.sf[1]; fl[1] = sub_boundary_flow(b, 2, 2);
      fl[2] = sub_boundary_flow(b, 3, 2); fl[3] = b.sf[3];
      fl[4] = sub_boundary_flow(b, 5, 4); fl[5] = sub_boundary_flow(b, 6, 4);
      break;
    default:  // the reference leaves the array as it was: zeros here
      for (int f = 0; f < 6; ++f) fl[f] = 0.0;
      break;
  }
}

// GetAveragedFaceFlow (:2444-2505): the sum over the face's whole array
__device__ __forceinline__ double averaged_face_flow(const NBuf &b, int f) {
  double s = 0.0;
  const int n = f < 4 ? 2 : 4;
  for (int k = 0; k < n; ++k) s = s + b.Q[f][k];
  return s;
}

// GetDZRW (:3208-3256): the thickness and the dry flag
__device__ double dzrw(const NBuf &b, bool &dry) {
  double dz = b.top - b.bottom;
  const double dzc = b.top - b.bottom;
  dry = false;
  if (b.lt == 1 && b.head < b.top) {
    dz = b.head - b.bottom;
    if ((dz <= 0.0) || ((b.head - b.bottom) / dzc < 0.01)) {
      dz = b.top - b.bottom;
      bool any = false;
      for (int f = 0; f < 6; ++f) any |= averaged_face_flow(b, f) != 0.0;
      dry = !any;
    }
  }
  return dz;
}

// The numbers FillCellBuffer gives cell n (1-based), with its own
// sub-cell flows when it is split
__device__ void real_cell(const UsgView &V, int n, NBuf &b) {
  const int32_t *t = V.topo + (int64_t)USG_TOPO * (n - 1);
  const double *a = V.casm + (int64_t)ASM_N * (n - 1);
  for (int f = 0; f < 6; ++f) {
    for (int s = 0; s < 4; ++s) b.Q[f][s] = a[ASM_Q + 4 * f + s];
    b.sfc[f] = t[UT_SFC + f];
  }
  b.src = a[ASM_SRC];
  b.snk = a[ASM_SNK];
  b.sto = a[ASM_STO];
  b.DX = V.dx[n - 1];
  b.DY = V.dy[n - 1];
  b.top = V.top[n - 1];
  b.bottom = V.bot[n - 1];
  b.head = V.head ? V.head[n - 1] : V.top[n - 1];
  b.lt = V.layer_type ? V.layer_type[n - 1] : 0;
  b.por = V.porosity[n - 1];
  b.ret = V.retardation ? V.retardation[n - 1] : 1.0;
  b.scr = b.scc = 1;
  for (int f = 0; f < 6; ++f)
    if (b.sfc[f] > 1) b.scr = b.scc = 2;
  for (int s = 0; s < 4; ++s) b.sf[s] = 0.0;
  subcell_flows(b);
}

// pr_FillCellBufferFromSubCell (ParticleTrackingEngine.f90:1907-1971) on a
// zeroed buffer
__device__ void from_sub(const NBuf &p, int r, int c, bool force, NBuf &o) {
  double fl[6];
  subcell_face_flows(p, r, c, fl);
  for (int f = 0; f < 6; ++f) {
    o.sfc[f] = 0;
    for (int s = 0; s < 4; ++s) o.Q[f][s] = 0.0;
    o.Q[f][0] = fl[f];
  }
  for (int s = 0; s < 4; ++s) o.sf[s] = 0.0;
  o.src = o.snk = o.sto = 0.0;
  o.scr = o.scc = 0;
  if (force) {
    o.src = p.src / 4.0;
    o.snk = p.snk / 4.0;
    o.sto = p.sto / 4.0;
    o.scr = o.scc = 2;
    subcell_flows(o);
  }
  o.DX = 0.5 * p.DX;
  o.DY = 0.5 * p.DY;
  o.bottom = p.bottom;
  o.top = p.top;
  o.head = 0.0;
  o.lt = 0;
  o.por = p.por;
  o.ret = p.ret;
}

// One stencil entry of a unit: face flows, face areas, volume, porosity
__device__ void stencil_entry(const UsgView &V, int32_t base, int32_t bits,
                              const NBuf &center, double fl[6], double ar[3],
                              double &vol, double &por) {
  const int kind = (bits >> SL_KIND) & 3;
  for (int f = 0; f < 6; ++f) fl[f] = 0.0;
  ar[0] = ar[1] = ar[2] = 0.0;
  vol = 0.0;
  por = 0.0;
  if (kind == 0) return;
  const int row = (bits >> SL_ROW) & 3, col = (bits >> SL_COL) & 3;
  const bool skip = (bits >> SL_SKIP) & 1;
  NBuf b;
  if (kind == 1) {
    b = center;
  } else {
    real_cell(V, base, b);
    if ((bits >> SL_FORCE0) & 1) {
      b.scr = b.scc = 2;
      subcell_flows(b);
    }
    const int nsub = (bits >> SL_NSUB) & 3;
    for (int k = 0; k < nsub; ++k) {
      const int r = (bits >> (k ? SL_R2 : SL_R1)) & 3;
      const int c = (bits >> (k ? SL_C2 : SL_C1)) & 3;
      const bool f = (bits >> (k ? SL_F2 : SL_F1)) & 1;
      NBuf o;
      from_sub(b, r, c, f, o);
      b = o;
    }
  }
  double dx, dy;
  if (kind == 1 || !skip) {
    if (kind == 1) subcell_face_flows(b, row, col, fl);
    else subcell_face_flows(b, row, col, fl);
    dx = b.DX / (double)b.scc;
    dy = b.DY / (double)b.scr;
  } else {
    for (int f = 0; f < 6; ++f) fl[f] = averaged_face_flow(b, f);
    dx = b.DX;
    dy = b.DY;
  }
  bool dry;
  const double dz = dzrw(b, dry);
  ar[0] = dy * dz;
  ar[1] = dx * dz;
  ar[2] = dx * dy;
  vol = dx * dy * dz;
  por = b.por;
}

// One unit's record (the sub-cell buffer of FillMassSubCellDataBuffer,
// ModpathCellData.f90:2522-2684, and its corner variables), one thread per
// unit; dbg (optional) receives the sub-cell tables in the checking layout
__global__ void __launch_bounds__(64) usg_units(const UsgView V, const TableParams tp,
                                                double *__restrict__ cells,
                                                double *__restrict__ dbg) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= V.nunit) return;
  const int32_t *sd = V.side + (int64_t)USG_SIDE * u;
  const int c = sd[US_CELL], r = sd[US_ROW], col = sd[US_COL];
  NBuf center;
  real_cell(V, c, center);
  double Qc[6];
  subcell_face_flows(center, r, col, Qc);
  const double DX = center.DX / (double)center.scc;
  const double DY = center.DY / (double)center.scr;
  bool dry = false;
  const double DZ = dzrw(center, dry);
  const int64_t d = V.ndisp == 1 ? 0 : (int64_t)(c - 1);
  const double aLH = V.disp[d], aTH = V.disp[V.ndisp + d];
  const double aLV = tp.disp_model == 1 ? V.disp[2 * V.ndisp + d] : 0.0;
  const double aTV = tp.disp_model == 1 ? V.disp[3 * V.ndisp + d] : 0.0;
  const double dm = V.disp[4 * V.ndisp + d];
  double nQ[18][6], nA[18][3], nV[18], nP[18];
  const int32_t *sl = V.slots + (int64_t)36 * u;
  for (int s = 0; s < 18; ++s)
    stencil_entry(V, sl[2 * s], sl[2 * s + 1], center, nQ[s], nA[s], nV[s], nP[s]);
  double *rec = cells + (int64_t)u * CELL_DOUBLES;
  RecordDbg rd;
  record_from_stencil(rec, tp, DX, DY, DZ, center.por, center.ret, Qc, aLH, aTH, aLV,
                      aTV, dm, V.uniform_porosity != 0, nQ, nA, nV, nP, rd);
  // connections and mass boundaries of the sub-cell's faces: internal faces
  // -1 and 0; the others the cell's sub-face of the sub-cell's row or column
  const int32_t *t = V.topo + (int64_t)USG_TOPO * (c - 1);
  const int nsub = sd[US_NSUB], scc = nsub == 1 ? 1 : 2, scr = scc;
  int conn[6], mb[6];
  for (int f = 0; f < 6; ++f) {
    conn[f] = -1;
    mb[f] = 0;
  }
  auto set_face = [&](int f, int sub) {
    const int cnt = t[UT_SFC + f];
    conn[f] = 0;
    if (cnt == 1) sub = 1;
    else if (cnt < 1) return;
    const int m = t[UT_CONN + 4 * f + sub - 1];
    conn[f] = m;
    if (t[UT_PCOUNT + f] == 0 || m == 0) mb[f] = V.default_icbound;
    else mb[f] = V.icbound ? V.icbound[m - 1] : 0;
  };
  if (r == 1) set_face(3, col);
  if (r == scr) set_face(2, col);
  if (col == 1) set_face(0, r);
  if (col == scc) set_face(1, r);
  for (int f = 4; f < 6; ++f) {
    const int cnt = t[UT_SFC + f];
    if (cnt == 1) set_face(f, 1);
    else if (cnt == nsub) set_face(f, (r - 1) * scc + col);
  }
  int32_t *ic = reinterpret_cast<int32_t *>(rec + F_CONN);
  int32_t *im = reinterpret_cast<int32_t *>(rec + F_MASSBND);
  for (int f = 0; f < 6; ++f) {
    ic[f] = conn[f];
    im[f] = mb[f];
  }
  // flags: IBoundTS of the cell, dry, cell has exit face (HasExitFace over
  // the sub-faces), VerifyDryCell's partially-dry test, weak source/sink
  const double sign = tp.backward ? -1.0 : 1.0;
  bool cell_exit = false;
  for (int f = 0; f < 6; ++f)
    for (int s = 0; s < center.sfc[f]; ++s) {
      const double q = sign * center.Q[f][s];
      cell_exit |= (f % 2 == 0) ? (q < 0.0) : (q > 0.0);
    }
  bool vpdry = false;
  if (center.lt == 1 && center.head < center.top) {
    vpdry = true;
    if ((center.head - center.bottom) <= 0.0) {
      bool any = false;
      for (int f = 0; f < 6; ++f) any |= averaged_face_flow(center, f) != 0.0;
      vpdry = any;
    }
  }
  int32_t *fl = reinterpret_cast<int32_t *>(rec + F_FLAGS);
  fl[0] = ib(V, c);
  fl[1] = (fl[1] & (CF_SUBCELL_EXIT | CF_DT_NONPOS)) | (dry ? CF_DRY : 0) |
          (cell_exit ? CF_CELL_EXIT : 0) | (vpdry ? CF_VERIFY_PDRY : 0) |
          (center.src != 0.0 ? CF_SOURCE : 0) | (center.snk != 0.0 ? CF_SINK : 0);
  int32_t *zn = reinterpret_cast<int32_t *>(rec + F_ZONE);
  zn[0] = V.zone ? V.zone[c - 1] : 0;
  zn[1] = 0;
  rec[F_SOURCE] = center.src;
  rec[F_SINK] = center.snk;
  if (dbg) {
    double *o = dbg + (int64_t)USG_DEBUG_FIELDS * u;
    const double *of = V.off + 4 * (int64_t)u;
    o[0] = c; o[1] = r; o[2] = col; o[3] = nsub;
    for (int f = 0; f < 6; ++f) {
      o[4 + f] = conn[f];
      o[10 + f] = mb[f];
    }
    o[16] = rd.sub_exit ? 1.0 : 0.0;
    o[17] = cell_exit ? 1.0 : 0.0;
    o[18] = dry ? 1.0 : 0.0;
    o[19] = (center.lt == 1 && center.head < center.top && !dry) ? 1.0 : 0.0;
    for (int f = 0; f < 6; ++f) o[20 + f] = rec[F_VX1 + f];
    o[26] = DX; o[27] = DY; o[28] = DZ;
    o[29] = of[0]; o[30] = of[1]; o[31] = of[2]; o[32] = of[3];
    o[33] = center.por; o[34] = center.ret; o[35] = rd.dtc;
    o[36] = center.src; o[37] = center.snk;
    for (int ci = 0; ci < 8; ++ci)
      for (int m = 0; m < 4; ++m) o[38 + 4 * ci + m] = rd.q[ci][m];
    for (int ci = 0; ci < 8; ++ci) o[70 + ci] = rd.pc[ci];
    for (int b6 = 0; b6 < 6; ++b6)
      for (int ci = 0; ci < 8; ++ci) o[78 + 8 * b6 + ci] = rd.Dc[b6][ci];
  }
}

// ConvertToModelXYZ (ModflowRectangularGrid.f90:524-549) with the cell
// bounds of GetLeft/Right/Front/Back (:275-309); top is SaturatedTop
__global__ void usg_global_xyz(int64_t n, const int32_t *__restrict__ cellv,
                               const double *__restrict__ lxv, const double *__restrict__ lyv,
                               const double *__restrict__ lzv, const double *__restrict__ cx,
                               const double *__restrict__ cy, const double *__restrict__ dx,
                               const double *__restrict__ dy, const double *__restrict__ top,
                               const double *__restrict__ bot, double *__restrict__ gx,
                               double *__restrict__ gy, double *__restrict__ gz) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int c0 = cellv[p] - 1;
  const double left = cx[c0] - (dx[c0] / 2.0), right = cx[c0] + (dx[c0] / 2.0);
  const double front = cy[c0] - (dy[c0] / 2.0), back = cy[c0] + (dy[c0] / 2.0);
  const double bt = bot[c0], tp = top[c0];
  const double x = lxv[p], y = lyv[p], z = lzv[p];
  gx[p] = fmin(fmax(((1.0 - x) * left) + (x * right), left), right);
  gy[p] = fmin(fmax(((1.0 - y) * front) + (y * back), front), back);
  gz[p] = fmin(fmax(((1.0 - z) * bt) + (z * tp), bt), tp);
}

UsgView make_view(rwpt_engine *e) {
  const DeviceProblem &P = e->prob;
  const UsgData &U = *e->usg;
  UsgView V{};
  V.ncell = U.ncell;
  V.nunit = U.nunit;
  V.topo = U.topo;
  V.cell_unit = U.cell_unit;
  V.side = U.side;
  V.slots = U.slots;
  V.off = U.off;
  V.cx = U.cx; V.cy = U.cy; V.dx = U.dx; V.dy = U.dy;
  V.top = P.top; V.bot = P.bot;
  V.flowja = U.flowja;
  V.storage = U.storage;
  V.bflow = P.bflow;
  V.porosity = P.porosity;
  V.retardation = P.retardation;
  V.head = P.head;
  V.source = P.source;
  V.sink = P.sink;
  V.disp = P.disp;
  V.ibound = P.ibound;
  V.icbound = P.icbound;
  V.layer_type = P.layer_type ? P.layer_type : U.cell_type;
  V.zone = P.zone;
  V.ndisp = P.ndisp;
  V.default_icbound = P.default_icbound;
  V.uniform_porosity = P.uniform_porosity;
  V.casm = U.casm;
  return V;
}

}  // namespace

hipError_t build_usg_tables(rwpt_engine *e, double *debug_out) {
  const UsgView V = make_view(e);
  const TableParams tp{e->prob.disp_model, e->opt.backward_tracking,
                       e->opt.time_step_kind, e->opt.time_step_params[0],
                       e->opt.time_step_params[1]};
  hipLaunchKernelGGL(usg_assemble, dim3((unsigned)((V.ncell + 255) / 256)), dim3(256), 0,
                     e->stream, V);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(usg_units, dim3((unsigned)((V.nunit + 63) / 64)), dim3(64), 0,
                     e->stream, V, tp, e->prob.cells, debug_out);
  err = hipGetLastError();
  e->prob.pow2_cells = false;
  return err;
}

hipError_t launch_usg_global_xyz(rwpt_engine *e, double *gx, double *gy, double *gz) {
  const int64_t n = e->part.n;
  const UsgData &U = *e->usg;
  if (n > 0)
    hipLaunchKernelGGL(usg_global_xyz, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       e->stream, n, e->part.cell, e->part.lx, e->part.ly, e->part.lz, U.cx,
                       U.cy, U.dx, U.dy, model_top(e->prob), e->prob.bot, gx, gy, gz);
  return hipGetLastError();
}

}  // namespace rwpt
