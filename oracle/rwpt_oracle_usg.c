/*
 * CPU oracle, unstructured part -- TEST INFRASTRUCTURE (included at the end
 * of rwpt_oracle.c, whose static sub-cell walk and corner formulas it
 * shares).
 *
 * The random walk on MODFLOW-USG / MODFLOW 6 DISV grids of rectangular
 * cells (a smoothed quadtree: a horizontal face meets one cell or two of
 * half its size), restated from:
 *  - ModflowRectangularGrid.f90: ComputeFaceAssignments (:98-253), the
 *    JaFace queries (:321-364, :479-513), ConvertFromNeighbor (:366-477),
 *    ConvertToModelXY / ConvertToLocalXY / SnapToValue (:551-630);
 *  - ModpathCellData.f90: SetMassTransportDataUnstructured (:674-1096),
 *    ComputeSubCellFlows (:2340-2441), FillMassSubCellDataBuffer
 *    (:2522-2684), FillSubCellFaceFlowsBuffer (:2817-2867),
 *    GetNeighborSubCellIndexes (:2875-3030), the sub-cell areas, flows and
 *    volumes (:3035-3130), GetDZRW (:3208-3256);
 *  - ParticleTrackingEngine.f90: pr_FillNeighborCellDataUnstructured and
 *    its helpers (:1257-2005), pr_RWPTrackPath (:1051-1254);
 *  - TrackSubCell.F90: pr_FillNeighborSubCellVariables (:3856-4012);
 *  - TrackCell.f90: the sub-cell loop (:375-669), FindSubCell (:47-90).
 *
 * Buffer states (what the reference's compiler, amdflang, gives): a
 * neighbour buffer is Reset before every neighbour fill (counts 1, flags
 * off); a buffer filled from a sub-cell (an intent(out) dummy) and every
 * local buffer start all zero, sub-face and sub-cell counts included,
 * because the compiler initialises such objects from the type's default
 * initialisation, in which the components without a default are zero.
 */

/* ------------------------------------------------------------------------ */
/* Grid: face assignments and connection queries                            */
/* ------------------------------------------------------------------------ */

typedef struct {
  const orc_usg_problem *U;
  int *jaface; /* nja, ComputeFaceAssignments */
} ugrid;

static inline int u_off(const ugrid *G, int n) { return G->U->ia[n - 1] - 1; }
static inline int u_cnt(const ugrid *G, int n) {
  return G->U->ia[n] - G->U->ia[n - 1];
}
static inline double u_left(const ugrid *G, int n) {
  return G->U->cell_x[n - 1] - (G->U->del_x[n - 1] / 2.0);
}
static inline double u_right(const ugrid *G, int n) {
  return G->U->cell_x[n - 1] + (G->U->del_x[n - 1] / 2.0);
}
static inline double u_front(const ugrid *G, int n) {
  return G->U->cell_y[n - 1] - (G->U->del_y[n - 1] / 2.0);
}
static inline double u_back(const ugrid *G, int n) {
  return G->U->cell_y[n - 1] + (G->U->del_y[n - 1] / 2.0);
}

static double snap_ratio(double r, double tol) {
  if (fabs(r - 0.5) <= tol) return 0.5;
  if (fabs(r - 1.0) <= tol) return 1.0;
  if (fabs(r - 2.0) <= tol) return 2.0;
  return r;
}

/* ComputeFaceAssignments, ModflowRectangularGrid.f90:98-253: the JaFace
 * code face*10000 + subface*100 + subface count of every connection; -1
 * for a spacing a smoothed quadtree cannot have (the reference stops). */
static int face_code(const ugrid *G, int n, int conn) {
  const orc_usg_problem *U = G->U;
  const double tol = (double)0.0001f;
  const double *cx = U->cell_x, *cy = U->cell_y;
  double left = u_left(G, n), right = u_right(G, n);
  double front = u_front(G, n), back = u_back(G, n);
  double cl = u_left(G, conn), cr = u_right(G, conn);
  double cf = u_front(G, conn), cb = u_back(G, conn);
  int layer = U->layer[n - 1], cl_ = U->layer[conn - 1];
  double rx = snap_ratio(U->del_x[conn - 1] / U->del_x[n - 1], tol);
  double ry = snap_ratio(U->del_y[conn - 1] / U->del_y[n - 1], tol);
  int m = 0;
  if (cl_ != layer) {
    m = cl_ > layer ? 50000 : 60000;
    if (rx == 0.5 && ry == 0.5) {
      if (cy[conn - 1] > cy[n - 1] && cx[conn - 1] < cx[n - 1]) m += 104;
      else if (cy[conn - 1] > cy[n - 1] && cx[conn - 1] > cx[n - 1]) m += 204;
      else if (cy[conn - 1] < cy[n - 1] && cx[conn - 1] < cx[n - 1]) m += 304;
      else if (cy[conn - 1] < cy[n - 1] && cx[conn - 1] > cx[n - 1]) m += 404;
    } else if ((rx == 1.0 || rx == 2.0) && rx == ry) {
      m += 101;
    } else {
      return -1;
    }
  } else if (fabs(left - cr) <= tol || fabs(right - cl) <= tol) {
    m = fabs(left - cr) <= tol ? 10000 : 20000;
    if (ry == 0.5) {
      if (cy[conn - 1] > cy[n - 1]) m += 102;
      else if (cy[conn - 1] < cy[n - 1]) m += 202;
    } else if (ry == 1.0 || ry == 2.0) {
      m += 101;
    } else {
      return -1;
    }
  } else if (fabs(front - cb) <= tol || fabs(back - cf) <= tol) {
    m = fabs(front - cb) <= tol ? 30000 : 40000;
    if (rx <= 0.5) {
      if (cx[conn - 1] < cx[n - 1]) m += 102;
      else if (cx[conn - 1] > cx[n - 1]) m += 202;
    } else if (rx == 1.0 || rx == 2.0) {
      m += 101;
    } else {
      return -1;
    }
  }
  return m;
}

static int ugrid_init(ugrid *G, const orc_usg_problem *U) {
  G->U = U;
  G->jaface = (int *)malloc(sizeof(int) * (size_t)(U->nja > 0 ? U->nja : 1));
  if (!G->jaface) return -1;
  for (int n = 1; n <= U->ncell; ++n) {
    int off = U->ia[n - 1] - 1, cnt = U->ia[n] - U->ia[n - 1];
    if (cnt == 0) continue;
    G->jaface[off] = 0;
    for (int i = 1; i < cnt; ++i) {
      int m = face_code(G, n, U->ja[off + i]);
      if (m < 0) {
        free(G->jaface);
        G->jaface = NULL;
        return -3;
      }
      G->jaface[off + i] = m;
    }
  }
  return 0;
}

/* ExtractConnectionCount, :47-62 */
static int jf_count(int v) { return v - 100 * (int)((double)v / 100.0); }

/* GetPotentialFaceConnectionCount, :321-343 */
static int u_potential_count(const ugrid *G, int n, int face) {
  int off = u_off(G, n), cnt = u_cnt(G, n);
  int f1 = 10000 * face, f2 = 10000 * (face + 1);
  for (int i = 1; i < cnt; ++i) {
    int v = G->jaface[off + i];
    if (v > f1 && v < f2) return jf_count(v);
  }
  return 0;
}

/* GetFaceConnection, :345-364 */
static int u_face_conn(const ugrid *G, int n, int face, int sub) {
  if (n <= 0) return 0;
  int off = u_off(G, n), cnt = u_cnt(G, n);
  int flag1 = 10000 * face + 100 * sub, flag2 = 10000 * face + 100 * (sub + 1);
  for (int i = 0; i < cnt; ++i) {
    int v = G->jaface[off + i];
    if (v >= flag1 && v < flag2) return G->U->ja[off + i];
  }
  return 0;
}

/* FindConnectionIndex, :498-513 (0-based position in JA, or -1) */
static int u_find_index(const ugrid *G, int n, int conn) {
  int off = u_off(G, n), cnt = u_cnt(G, n);
  for (int i = 1; i < cnt; ++i)
    if (G->U->ja[off + i] == conn) return off + i;
  return -1;
}

/* FindConnectionPosition, :479-496: the face of n that conn meets */
static int u_connection_face(const ugrid *G, int n, int conn) {
  if (n < 1 || n > G->U->ncell || conn == n) return 0;
  int idx = u_find_index(G, n, conn);
  if (idx < 0) return 0;
  int v = G->jaface[idx];
  int nn = 100 * (int)((double)v / 100.0);
  if (nn > 10000 && nn < 20000) return 1;
  if (nn > 20000 && nn < 30000) return 2;
  if (nn > 30000 && nn < 40000) return 3;
  if (nn > 40000 && nn < 50000) return 4;
  if (nn > 50000 && nn < 60000) return 5;
  if (nn > 60000 && nn < 70000) return 6;
  return -1;
}

/* ------------------------------------------------------------------------ */
/* Cell buffers                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
  int cell, layer;
  double DX, DY, top, bottom, head;
  int layer_type, ibound, zone;
  double porosity, retardation, source, sink, storage;
  int sfc[6], sfbc[6];
  int conn[6][4], mb[6][4];
  double Q[6][4]; /* faces 1-4 use two entries, faces 5-6 four */
  double subflows[4];
  int scr, scc;
  int fromSub, isParent, parentCell, parentRow, parentCol, rfd;
  double aLH, aTH, aLV, aTV, dmeff;
  int dry, pdry;
} ucbuf;

static const int q_size[6] = {2, 2, 2, 2, 4, 4};

/* pr_ResetArrays, ModpathCellData.f90:433-461 */
static void u_reset_arrays(ucbuf *b) {
  for (int f = 0; f < 6; ++f) {
    b->sfc[f] = 1;
    b->sfbc[f] = 1;
    for (int s = 0; s < 4; ++s) {
      b->Q[f][s] = 0.0;
      b->conn[f][s] = 0;
      b->mb[f][s] = 0;
    }
  }
  for (int s = 0; s < 4; ++s) b->subflows[s] = 0.0;
}

/* pr_Reset, :391-429 (the layer type and head keep their values; they are
 * zero in every buffer the tests reach) */
static void u_reset(ucbuf *b) {
  int lt = b->layer_type;
  double hd = b->head;
  memset(b, 0, sizeof *b);
  b->layer_type = lt;
  b->head = hd;
  b->scr = b->scc = 1;
  u_reset_arrays(b);
}

static inline int u_subcount(const ucbuf *b) { return b->scc * b->scr; }

/* GetSubCellBoundaryFlow, :2396-2441 */
static double u_sub_boundary_flow(const ucbuf *b, int face, int sub) {
  if (face <= 4) {
    if (b->sfc[face - 1] == 1) return b->Q[face - 1][0] / 2.0;
    return b->Q[face - 1][sub - 1];
  }
  if (b->sfc[face - 1] == 1) return b->Q[face - 1][0] / 4.0;
  return b->Q[face - 1][sub - 1];
}

/* pr_ComputeSubCellFlows, :2340-2393 */
static void u_compute_subcell_flows(ucbuf *b) {
  double bb[4] = {0, 0, 0, 0}, qf;
  if (u_subcount(b) == 1) return;
  double qsrc = b->source / 4.0, qsink = b->sink / 4.0, qsto = b->storage / 4.0;
  qf = 0.0;
  qf = qf + u_sub_boundary_flow(b, 1, 1);
  qf = qf - u_sub_boundary_flow(b, 4, 1);
  qf = qf + u_sub_boundary_flow(b, 5, 1);
  qf = qf - u_sub_boundary_flow(b, 6, 1);
  bb[0] = qf + qsrc + qsink + qsto;
  qf = 0.0;
  qf = qf - u_sub_boundary_flow(b, 2, 1);
  qf = qf - u_sub_boundary_flow(b, 4, 2);
  qf = qf + u_sub_boundary_flow(b, 5, 2);
  qf = qf - u_sub_boundary_flow(b, 6, 2);
  bb[1] = qf + qsrc + qsink + qsto;
  qf = 0.0;
  qf = qf + u_sub_boundary_flow(b, 1, 2);
  qf = qf + u_sub_boundary_flow(b, 3, 1);
  qf = qf + u_sub_boundary_flow(b, 5, 3);
  qf = qf - u_sub_boundary_flow(b, 6, 3);
  bb[2] = qf + qsrc + qsink + qsto;
  b->subflows[0] = 0.5 * (bb[0] + 0.5 * (bb[2] - bb[1]));
  b->subflows[2] = -bb[0] + b->subflows[0];
  b->subflows[1] = bb[2] - b->subflows[2];
  b->subflows[3] = -bb[1] - b->subflows[0];
}

/* pr_FillSubCellFaceFlowsBuffer, :2817-2867 */
static void u_subcell_face_flows(const ucbuf *b, int r, int c, double fl[6]) {
  if (u_subcount(b) == 1) {
    for (int f = 0; f < 6; ++f) fl[f] = b->Q[f][0];
    return;
  }
  int k = (r - 1) * b->scc + c;
  switch (k) {
    case 1:
      fl[0] = u_sub_boundary_flow(b, 1, 1); fl[1] = b->subflows[0];
      fl[2] = b->subflows[2]; fl[3] = u_sub_boundary_flow(b, 4, 1);
      fl[4] = u_sub_boundary_flow(b, 5, 1); fl[5] = u_sub_boundary_flow(b, 6, 1);
      break;
    case 2:
      fl[0] = b->subflows[0]; fl[1] = u_sub_boundary_flow(b, 2, 1);
      fl[2] = b->subflows[3]; fl[3] = u_sub_boundary_flow(b, 4, 2);
      fl[4] = u_sub_boundary_flow(b, 5, 2); fl[5] = u_sub_boundary_flow(b, 6, 2);
      break;
    case 3:
      fl[0] = u_sub_boundary_flow(b, 1, 2); fl[1] = b->subflows[1];
      fl[2] = u_sub_boundary_flow(b, 3, 1); fl[3] = b->subflows[2];
      fl[4] = u_sub_boundary_flow(b, 5, 3); fl[5] = u_sub_boundary_flow(b, 6, 3);
      break;
    case 4:
      fl[0] = b->subflows[1]; fl[1] = u_sub_boundary_flow(b, 2, 2);
      fl[2] = u_sub_boundary_flow(b, 3, 2); fl[3] = b->subflows[3];
      fl[4] = u_sub_boundary_flow(b, 5, 4); fl[5] = u_sub_boundary_flow(b, 6, 4);
      break;
    default: /* the reference leaves the array as it was */
      break;
  }
}

/* pr_GetAveragedFaceFlow, :2444-2505: the sum over the whole array */
static double u_averaged_face_flow(const ucbuf *b, int face) {
  double s = 0.0;
  for (int k = 0; k < q_size[face - 1]; ++k) s = s + b->Q[face - 1][k];
  return s;
}

/* pr_GetDZRW, :3208-3256 (sets the dry flags) */
static double u_dzrw(ucbuf *b) {
  double dz = b->top - b->bottom, dzc = b->top - b->bottom;
  b->dry = 0;
  b->pdry = 0;
  if (b->layer_type == 1 && b->head < b->top) {
    dz = b->head - b->bottom;
    b->pdry = 1;
    if ((dz <= 0.0) || ((b->head - b->bottom) / dzc < 0.01)) {
      dz = b->top - b->bottom;
      b->dry = 1;
      b->pdry = 0;
      int any = 0;
      for (int f = 1; f <= 6; ++f) any |= (u_averaged_face_flow(b, f) != 0.0);
      if (any) {
        b->dry = 0;
        b->pdry = 1;
      }
    }
  }
  return dz;
}

/* pr_VerifyDryCell, :3161-3205 */
static void u_verify_dry(ucbuf *b) {
  b->dry = 0;
  b->pdry = 0;
  if (b->layer_type == 1 && b->head < b->top) {
    double dz = b->head - b->bottom;
    b->pdry = 1;
    if (dz <= 0.0) {
      b->dry = 1;
      b->pdry = 0;
      int any = 0;
      for (int f = 1; f <= 6; ++f) any |= (u_averaged_face_flow(b, f) != 0.0);
      if (any) {
        b->dry = 0;
        b->pdry = 1;
      }
    }
  }
}

/* pr_HasExitFace, :320-388 */
static int u_cell_has_exit(const ucbuf *b, int backward) {
  double sign = backward ? -1.0 : 1.0;
  for (int f = 0; f < 6; ++f)
    for (int k = 0; k < b->sfc[f]; ++k) {
      double q = sign * b->Q[f][k];
      if ((f % 2 == 0) ? (q < 0.0) : (q > 0.0)) return 1;
    }
  return 0;
}

static inline int u_ibound(const orc_usg_problem *U, int n) {
  return U->base.ibound ? U->base.ibound[n - 1] : 1;
}

/* pr_FillTransportCellBufferUnstructured (ParticleTrackingEngine.f90:711-748)
 * = pr_SetMassTransportDataUnstructured (ModpathCellData.f90:674-1096) and
 * the dispersion parameters of the cell (:535-570) */
static void u_fill_cell(const ugrid *G, int n, ucbuf *b) {
  const orc_usg_problem *U = G->U;
  const orc_problem *P = &U->base;
  u_reset_arrays(b);
  b->cell = n;
  b->layer = U->layer[n - 1];
  b->DX = U->del_x[n - 1];
  b->DY = U->del_y[n - 1];
  b->bottom = P->bot[n - 1];
  b->top = P->top[n - 1];
  b->zone = P->zone ? P->zone[n - 1] : 0;
  b->layer_type = P->layer_type ? P->layer_type[n - 1] : 0;
  b->head = P->head ? P->head[n - 1] : P->top[n - 1];
  b->ibound = u_ibound(U, n);
  b->porosity = P->porosity[n - 1];
  b->retardation = P->retardation ? P->retardation[n - 1] : 1.0;
  b->source = P->source_flow ? P->source_flow[n - 1] : 0.0;
  b->sink = P->sink_flow ? P->sink_flow[n - 1] : 0.0;
  b->storage = U->storage ? U->storage[n - 1] : 0.0;
  for (int f = 0; f < 6; ++f) {
    int count = u_potential_count(G, n, f + 1);
    if (count > 0) {
      b->sfc[f] = count;
      int i = 0;
      for (int s = 0; s < count; ++s) {
        int conn = u_face_conn(G, n, f + 1, s + 1);
        b->conn[f][s] = conn;
        if (conn == 0) {
          i = i + 1;
          b->mb[f][s] = P->default_icbound;
        } else {
          if (u_ibound(U, conn) == 0) i = i + 1;
          b->mb[f][s] = P->icbound ? P->icbound[conn - 1] : 0;
        }
      }
      b->sfbc[f] = i;
    } else {
      for (int s = 0; s < q_size[f]; ++s) b->mb[f][s] = P->default_icbound;
    }
  }
  for (int f = 0; f < 6; ++f)
    for (int s = 0; s < b->sfc[f]; ++s) {
      int conn = b->conn[f][s];
      if (conn > 0) {
        int idx = u_find_index(G, n, conn);
        double q = idx >= 0 ? U->flowja[idx] : 0.0;
        b->Q[f][s] = (f % 2 == 0) ? q : -q;
      }
    }
  if (P->bflow) {
    for (int f = 0; f < 6; ++f) {
      double bf = P->bflow[6 * (n - 1) + f];
      if (bf == 0.0) continue;
      if (b->sfbc[f] > 0) {
        double flow = bf / b->sfbc[f];
        for (int s = 0; s < b->sfc[f]; ++s) {
          int conn = b->conn[f][s];
          if (conn == 0 || u_ibound(U, conn) == 0)
            b->Q[f][s] = (f % 2 == 0) ? flow : -flow;
        }
      } else if (bf > 0.0) {
        b->source = b->source + bf;
      } else {
        b->sink = b->sink + bf;
      }
    }
  }
  b->scr = 1;
  b->scc = 1;
  for (int f = 0; f < 6; ++f)
    if (b->sfc[f] > 1) {
      b->scr = 2;
      b->scc = 2;
      u_compute_subcell_flows(b);
      break;
    }
  int d = P->disp_uniform ? 0 : n - 1;
  b->aLH = P->aL[d];
  b->aTH = P->aT[d];
  b->dmeff = P->dmeff[d];
  b->aLV = P->disp_model == 1 ? P->aLV[d] : 0.0;
  b->aTV = P->disp_model == 1 ? P->aTV[d] : 0.0;
}

/* GetFaceConnection of a buffer (:2257-2283) */
static int u_buf_conn(const ucbuf *b, int face, int sub) {
  if (sub < 1 || sub > b->sfc[face - 1]) return 0;
  return b->conn[face - 1][sub - 1];
}

/* ------------------------------------------------------------------------ */
/* Neighbour buffers: ParticleTrackingEngine.f90:1257-2005                  */
/* ------------------------------------------------------------------------ */

typedef ucbuf upair[2]; /* neighborCellData(1:2, m) */

/* pr_FillCellFromRealData, :1974-2002 */
static void u_fill_real(const ugrid *G, int n, ucbuf *b, int force) {
  if (n <= 0) return;
  u_fill_cell(G, n, b);
  if (force) {
    b->scr = 2;
    b->scc = 2;
    u_compute_subcell_flows(b);
  }
}

/* pr_FillCellBufferFromSubCell, :1907-1971 (the output starts all zero) */
static void u_fill_from_sub(const ucbuf *parent, int r, int c, ucbuf *out,
                            int force) {
  double fl[6];
  memset(out, 0, sizeof *out);
  u_subcell_face_flows(parent, r, c, fl);
  for (int f = 0; f < 6; ++f) out->Q[f][0] = fl[f];
  if (force) {
    out->source = parent->source / 4.0;
    out->sink = parent->sink / 4.0;
    out->storage = parent->storage / 4.0;
    out->scr = 2;
    out->scc = 2;
    u_compute_subcell_flows(out);
  }
  out->fromSub = 1;
  out->parentCell = parent->cell;
  out->parentRow = r;
  out->parentCol = c;
  out->DX = 0.5 * parent->DX;
  out->DY = 0.5 * parent->DY;
  out->bottom = parent->bottom;
  out->top = parent->top;
  out->porosity = parent->porosity;
  out->retardation = parent->retardation;
}

/* pr_FillNeighborCellsConnectionFromHorizontalFace, :1649-1903 */
static int u_from_horizontal(const ugrid *G, const ucbuf *cc, int face,
                             ucbuf nbuf[2], int force) {
  ucbuf parentBuf, auxBuf;
  memset(&parentBuf, 0, sizeof parentBuf);
  memset(&auxBuf, 0, sizeof auxBuf);
  if (cc->cell <= 0) return 0;
  int direct = u_buf_conn(cc, face, 1);
  int dir = face <= 2 ? 1 : (face <= 4 ? 2 : 3);
  if (cc->sfc[face - 1] > 1) {
    if (!cc->isParent) {
      /* (a sub-face without a neighbour would be read as cell 0 there) */
      for (int m = 0; m < cc->sfc[face - 1] && m < 2; ++m) {
        int conn = u_buf_conn(cc, face, m + 1);
        if (conn > 0) u_fill_cell(G, conn, &nbuf[m]);
        nbuf[m].rfd = dir;
      }
    } else {
      int m;
      if (dir == 2) m = cc->parentCol;
      else return -4; /* the reference stops ("inconsistency") */
      int conn = u_buf_conn(cc, face, m);
      if (conn > 0) u_fill_cell(G, conn, &nbuf[0]);
      nbuf[0].rfd = dir;
      if (force) {
        nbuf[0].scr = 2;
        nbuf[0].scc = 2;
        u_compute_subcell_flows(&nbuf[0]);
      }
    }
    return 0;
  }
  if (direct <= 0) return 0;
  int orth[2], sci[2][2];
  switch (face) {
    case 1: orth[0] = 3; orth[1] = 4; sci[0][0] = 1; sci[0][1] = 2; sci[1][0] = 2; sci[1][1] = 2; break;
    case 2: orth[0] = 3; orth[1] = 4; sci[0][0] = 1; sci[0][1] = 1; sci[1][0] = 2; sci[1][1] = 1; break;
    case 3: orth[0] = 1; orth[1] = 2; sci[0][0] = 1; sci[0][1] = 2; sci[1][0] = 1; sci[1][1] = 1; break;
    default: orth[0] = 1; orth[1] = 2; sci[0][0] = 2; sci[0][1] = 2; sci[1][0] = 2; sci[1][1] = 1; break;
  }
  int oc[2] = {u_buf_conn(cc, orth[0], 1), u_buf_conn(cc, orth[1], 1)};
  int bigger = 0;
  for (int k = 0; k < 2; ++k)
    if (oc[k] > 0 && direct == u_face_conn(G, oc[k], face, 1)) {
      bigger = 1;
      break;
    }
  if (bigger) {
    u_fill_real(G, direct, &parentBuf, 1);
    int row = (direct == u_face_conn(G, oc[0], face, 1)) ? 0 : 1;
    if (!cc->isParent) {
      u_fill_from_sub(&parentBuf, sci[row][0], sci[row][1], &nbuf[0], force);
      nbuf[0].rfd = dir;
    } else {
      u_fill_from_sub(&parentBuf, sci[row][0], sci[row][1], &auxBuf, 1);
      auxBuf.isParent = 1;
      auxBuf.cell = parentBuf.cell;
      u_fill_from_sub(&auxBuf, sci[row][0], sci[row][1], &nbuf[0], force);
      nbuf[0].rfd = dir;
    }
  } else if (!cc->isParent) {
    u_fill_real(G, direct, &nbuf[0], force);
    nbuf[0].rfd = dir;
  } else {
    int subRow = 0, subCol = 0;
    if (dir == 2) {
      subCol = cc->parentCol;
      subRow = face == 3 ? 1 : 2;
    } else if (dir == 1) {
      subRow = cc->parentRow;
      subCol = face == 1 ? 2 : 1;
    }
    u_fill_real(G, direct, &parentBuf, 1);
    u_fill_from_sub(&parentBuf, subRow, subCol, &nbuf[0], force);
    nbuf[0].rfd = dir;
  }
  return 0;
}

/* pr_FillNeighborCellsSubBuffer, :1300-1646.  nb is neighborCellData as
 * nb[slot - 1][buffer - 1]. */
static int u_sub_buffer(const ugrid *G, const ucbuf *cc, int face, int first,
                        upair nb[18], int d, int force) {
  ucbuf auxBuf;
  memset(&auxBuf, 0, sizeof auxBuf);
  int e = 0;
  ucbuf *D = nb[d - 1];
  if (face >= 5) {
    u_fill_real(G, u_buf_conn(cc, face, 1), &D[0], force);
    D[0].rfd = 3;
    e |= u_from_horizontal(G, &D[0], first, nb[d], force);
    e |= u_from_horizontal(G, &D[0], first + 1, nb[d + 1], force);
    return e;
  }
  if (cc->sfc[face - 1] > 1) {
    e |= u_from_horizontal(G, cc, face, D, force);
    if (first >= 5) {
      for (int m = 0; m < 2; ++m) {
        u_fill_real(G, u_buf_conn(&D[m], first, 1), &nb[d][m], force);
        nb[d][m].rfd = 3;
      }
      for (int m = 0; m < 2; ++m) {
        u_fill_real(G, u_buf_conn(&D[m], first + 1, 1), &nb[d + 1][m], force);
        nb[d + 1][m].rfd = 3;
      }
      return e;
    }
    int i0 = first <= 2 ? 0 : 1, i1 = first <= 2 ? 1 : 0;
    e |= u_from_horizontal(G, &D[i0], first, nb[d], 0);
    e |= u_from_horizontal(G, &D[i1], first + 1, nb[d + 1], 0);
    return e;
  }
  if (u_buf_conn(cc, face, 1) <= 0) return 0;
  e |= u_from_horizontal(G, cc, face, D, force);
  if (D[0].fromSub) {
    if (first >= 5) {
      u_fill_real(G, u_face_conn(G, D[0].parentCell, first, 1), &auxBuf, 1);
      if (auxBuf.cell > 0) {
        u_fill_from_sub(&auxBuf, D[0].parentRow, D[0].parentCol, &nb[d][0], force);
        nb[d][0].rfd = 3;
      }
      u_fill_real(G, u_face_conn(G, D[0].parentCell, first + 1, 1), &auxBuf, 1);
      if (auxBuf.cell > 0) {
        u_fill_from_sub(&auxBuf, D[0].parentRow, D[0].parentCol, &nb[d + 1][0], force);
        nb[d + 1][0].rfd = 3;
      }
      return e;
    }
    u_fill_real(G, D[0].parentCell, &auxBuf, 1);
    auxBuf.isParent = 1;
    auxBuf.parentRow = D[0].parentRow;
    auxBuf.parentCol = D[0].parentCol;
    int dirId = 0, sameIdx = 0, notSameIdx = 0, notSameFace = 0, sameRow = 0, sameCol = 0;
    if (first >= 3) {
      dirId = 2;
      if (D[0].parentRow == 1) {
        sameIdx = d + 1;
        notSameIdx = d + 2;
        notSameFace = first + 1;
        sameRow = 2;
      } else {
        sameIdx = d + 2;
        notSameIdx = d + 1;
        notSameFace = first;
        sameRow = 1;
      }
      sameCol = D[0].parentCol;
    } else {
      /* the reference leaves the indexes unset here (it notes that the
       * case does not occur); the tests never reach it */
      return -5;
    }
    e |= u_from_horizontal(G, &auxBuf, notSameFace, nb[notSameIdx - 1], force);
    u_fill_from_sub(&auxBuf, sameRow, sameCol, &nb[sameIdx - 1][0], force);
    nb[sameIdx - 1][0].rfd = dirId;
    return e;
  }
  if (first >= 5) {
    u_fill_real(G, u_buf_conn(&D[0], first, 1), &nb[d][0], force);
    nb[d][0].rfd = 3;
    u_fill_real(G, u_buf_conn(&D[0], first + 1, 1), &nb[d + 1][0], force);
    nb[d + 1][0].rfd = 3;
  } else {
    e |= u_from_horizontal(G, &D[0], first, nb[d], force);
    e |= u_from_horizontal(G, &D[0], first + 1, nb[d + 1], force);
  }
  return e;
}

/* pr_FillNeighborCellDataUnstructured, :1257-1297, with every buffer reset
 * first */
static int u_fill_neighbors(const ugrid *G, const ucbuf *center, upair nb[18]) {
  for (int m = 0; m < 18; ++m) {
    u_reset(&nb[m][0]);
    u_reset(&nb[m][1]);
  }
  int force = u_subcount(center) > 1;
  int e = 0;
  for (int n = 1; n <= 6; ++n) {
    int first = n <= 2 ? 3 : (n <= 4 ? 5 : 1);
    e |= u_sub_buffer(G, center, n, first, nb, (n - 1) * 3 + 1, force);
  }
  return e;
}

/* ------------------------------------------------------------------------ */
/* Sub-cell buffer and corner variables                                     */
/* ------------------------------------------------------------------------ */

/* pr_GetNeighborSubCellIndexes, ModpathCellData.f90:2875-3030: per slot
 * (buffer index, sub-row, sub-column); buffer index 0 is a sub-cell of the
 * cell itself */
static const int nsi_11[18][3] = {
    {1, 1, 2}, {1, 2, 2}, {3, 2, 2}, {0, 1, 2}, {0, 2, 2}, {10, 2, 2},
    {0, 2, 1}, {13, 2, 1}, {16, 2, 1}, {10, 2, 1}, {11, 2, 1}, {12, 2, 1},
    {13, 1, 1}, {14, 1, 2}, {13, 1, 2}, {16, 1, 1}, {17, 1, 2}, {16, 1, 2}};
static const int nsi_12[18][3] = {
    {0, 1, 1}, {0, 2, 1}, {10, 2, 1}, {4, 1, 1}, {4, 2, 1}, {6, 2, 1},
    {0, 2, 2}, {13, 2, 2}, {16, 2, 2}, {10, 2, 2}, {11, 2, 2}, {12, 2, 2},
    {13, 1, 2}, {13, 1, 1}, {15, 1, 1}, {16, 1, 2}, {16, 1, 1}, {18, 1, 1}};
static const int nsi_21[18][3] = {
    {1, 2, 2}, {2, 1, 2}, {1, 1, 2}, {0, 2, 2}, {7, 1, 2}, {0, 1, 2},
    {7, 1, 1}, {8, 1, 1}, {9, 1, 1}, {0, 1, 1}, {13, 1, 1}, {16, 1, 1},
    {13, 2, 1}, {14, 2, 2}, {13, 2, 2}, {16, 2, 1}, {17, 2, 2}, {16, 2, 2}};
static const int nsi_22[18][3] = {
    {0, 2, 1}, {7, 1, 1}, {0, 1, 1}, {4, 2, 1}, {5, 1, 1}, {4, 1, 1},
    {7, 1, 2}, {8, 1, 2}, {9, 1, 2}, {0, 1, 2}, {13, 1, 2}, {16, 1, 2},
    {13, 2, 2}, {13, 2, 1}, {15, 2, 1}, {16, 2, 2}, {16, 2, 1}, {18, 2, 1}};

static void u_neighbor_indexes(const ucbuf *b, int r, int c, int out[18][3]) {
  if (u_subcount(b) == 1) {
    for (int n = 0; n < 18; ++n) {
      out[n][0] = n + 1;
      out[n][1] = 1;
      out[n][2] = 1;
    }
    return;
  }
  const int(*t)[3] = r == 1 ? (c == 1 ? nsi_11 : nsi_12) : (c == 1 ? nsi_21 : nsi_22);
  memcpy(out, t, sizeof(int) * 54);
}

/* FillSubCellFaceAreas / FillSubCellFaceFlows / GetVolume, :3035-3130 */
static void u_face_areas(ucbuf *b, int skip, double a[3]) {
  double dx, dy;
  if (!skip) {
    dx = b->DX / (double)b->scc;
    dy = b->DY / (double)b->scr;
  } else {
    dx = b->DX;
    dy = b->DY;
  }
  double dz = u_dzrw(b);
  a[0] = dy * dz;
  a[1] = dx * dz;
  a[2] = dx * dy;
}
static void u_face_flows(const ucbuf *b, int r, int c, int skip, double fl[6]) {
  if (!skip) {
    u_subcell_face_flows(b, r, c, fl);
  } else {
    for (int f = 0; f < 6; ++f) fl[f] = u_averaged_face_flow(b, f + 1);
  }
}
static double u_volume(ucbuf *b, int skip) {
  double dx, dy;
  if (!skip) {
    dx = b->DX / (double)b->scc;
    dy = b->DY / (double)b->scr;
  } else {
    dx = b->DX;
    dy = b->DY;
  }
  double dz = u_dzrw(b);
  return dx * dy * dz;
}

/* pr_FillNeighborSubCellVariables, TrackSubCell.F90:3856-4012 */
static void u_stencil(ucbuf *center, upair nb[18], const int nsi[18][3],
                      double nflow[18][6], double narea[18][3], double nvol[18],
                      double nphi[18]) {
  memset(nflow, 0, sizeof(double) * 108);
  memset(narea, 0, sizeof(double) * 54);
  memset(nvol, 0, sizeof(double) * 18);
  memset(nphi, 0, sizeof(double) * 18);
  if (u_subcount(center) == 1) {
    for (int n = 0; n < 18; ++n) {
      ucbuf *b = &nb[nsi[n][0] - 1][0];
      if (b->cell == 0) continue;
      u_face_flows(b, 1, 1, 1, nflow[n]);
      u_face_areas(b, 1, narea[n]);
      nvol[n] = u_volume(b, 1);
      nphi[n] = b->porosity;
    }
    return;
  }
  for (int n = 0; n < 18; ++n) {
    int skip = 0;
    if (nsi[n][0] == 0) {
      u_subcell_face_flows(center, nsi[n][1], nsi[n][2], nflow[n]);
      u_face_areas(center, 0, narea[n]);
      nvol[n] = u_volume(center, 0);
      nphi[n] = center->porosity;
      continue;
    }
    upair *P = &nb[nsi[n][0] - 1];
    int sci, row, col;
    if ((*P)[1].cell > 0) {
      if (u_subcount(&(*P)[0]) > 1 || u_subcount(&(*P)[1]) > 1) skip = 1;
      if ((*P)[1].rfd == 1) sci = nsi[n][1];
      else sci = nsi[n][2];
      row = 1;
      col = 1;
    } else if ((*P)[0].cell > 0 || (*P)[0].fromSub) {
      sci = 1;
      if (u_subcount(&(*P)[0]) > 1) {
        row = nsi[n][1];
        col = nsi[n][2];
      } else {
        row = 1;
        col = 1;
      }
    } else {
      continue;
    }
    ucbuf *b = &(*P)[sci - 1];
    u_face_flows(b, row, col, skip, nflow[n]);
    u_face_areas(b, skip, narea[n]);
    nvol[n] = u_volume(b, skip);
    nphi[n] = b->porosity;
  }
}

/* pr_FillMassSubCellDataBuffer, ModpathCellData.f90:2522-2684, with the
 * sub-cell offsets in the parent cell (offx, offy) */
static void u_fill_subcell(ucbuf *b, int r, int c, int backward, scbuf *s,
                           double offx[2], double offy[2]) {
  double fl[6];
  memset(s, 0, sizeof *s);
  u_subcell_face_flows(b, r, c, fl);
  s->DX = b->DX / (double)b->scc;
  s->DY = b->DY / (double)b->scr;
  s->DZ = u_dzrw(b);
  s->dry = b->dry;
  s->pdry = b->pdry;
  double sign = backward ? -1.0 : 1.0;
  s->VX1 = sign * fl[0] / s->DY / s->DZ / b->porosity / b->retardation;
  s->VX2 = sign * fl[1] / s->DY / s->DZ / b->porosity / b->retardation;
  s->VY1 = sign * fl[2] / s->DX / s->DZ / b->porosity / b->retardation;
  s->VY2 = sign * fl[3] / s->DX / s->DZ / b->porosity / b->retardation;
  s->VZ1 = sign * fl[4] / s->DX / s->DY / b->porosity / b->retardation;
  s->VZ2 = sign * fl[5] / s->DX / s->DY / b->porosity / b->retardation;
  double xinc = 1.0 / (double)b->scc;
  offx[0] = (c - 1) * xinc;
  offx[1] = 1.0;
  if (c < b->scc) offx[1] = c * xinc;
  double yinc = 1.0 / (double)b->scr;
  offy[0] = (b->scr - r) * yinc;
  offy[1] = 1.0;
  if (r > 1) offy[1] = (b->scr - r + 1) * yinc;
  for (int f = 0; f < 6; ++f) {
    s->conn[f] = -1;
    s->mass_bnd[f] = 0;
  }
  if (r == 1) {
    int count = b->sfc[3];
    s->conn[3] = 0;
    if (count == 1) { s->conn[3] = b->conn[3][0]; s->mass_bnd[3] = b->mb[3][0]; }
    else if (count > 1) { s->conn[3] = b->conn[3][c - 1]; s->mass_bnd[3] = b->mb[3][c - 1]; }
  }
  if (r == b->scr) {
    int count = b->sfc[2];
    s->conn[2] = 0;
    if (count == 1) { s->conn[2] = b->conn[2][0]; s->mass_bnd[2] = b->mb[2][0]; }
    else if (count > 1) { s->conn[2] = b->conn[2][c - 1]; s->mass_bnd[2] = b->mb[2][c - 1]; }
  }
  if (c == 1) {
    int count = b->sfc[0];
    s->conn[0] = 0;
    if (count == 1) { s->conn[0] = b->conn[0][0]; s->mass_bnd[0] = b->mb[0][0]; }
    else if (count > 1) { s->conn[0] = b->conn[0][r - 1]; s->mass_bnd[0] = b->mb[0][r - 1]; }
  }
  if (c == b->scc) {
    int count = b->sfc[1];
    s->conn[1] = 0;
    if (count == 1) { s->conn[1] = b->conn[1][0]; s->mass_bnd[1] = b->mb[1][0]; }
    else if (count > 1) { s->conn[1] = b->conn[1][r - 1]; s->mass_bnd[1] = b->mb[1][r - 1]; }
  }
  int rc = b->scr * b->scc;
  for (int f = 4; f < 6; ++f) {
    if (b->sfc[f] == 1) {
      s->conn[f] = b->conn[f][0];
      s->mass_bnd[f] = b->mb[f][0];
    } else if (b->sfc[f] == rc) {
      int k = (r - 1) * b->scc + c;
      s->conn[f] = b->conn[f][k - 1];
      s->mass_bnd[f] = b->mb[f][k - 1];
    }
  }
  s->porosity = b->porosity;
  s->retardation = b->retardation;
  s->aLH = b->aLH;
  s->aLV = b->aLV;
  s->aTH = b->aTH;
  s->aTV = b->aTV;
  s->dmeff = b->dmeff;
}

/* the sub-cell's corner variables (pr_ComputeCornerVariables,
 * TrackSubCell.F90:3667-3850) */
static void u_corners(const orc_problem *P, ucbuf *center, upair nb[18], int r,
                      int c, const scbuf *s, corners *C) {
  int nsi[18][3];
  double nflow[18][6], narea[18][3], nvol[18], nphi[18], cq[6];
  u_neighbor_indexes(center, r, c, nsi);
  u_stencil(center, nb, nsi, nflow, narea, nvol, nphi);
  u_subcell_face_flows(center, r, c, cq);
  corners_from_stencil(P, s, cq, nflow, narea, nvol, nphi, C);
}

/* ------------------------------------------------------------------------ */
/* Tables                                                                   */
/* ------------------------------------------------------------------------ */

/* One record per sub-cell of every cell, cells in order and sub-cells
 * row-major, ORC_USG_FIELDS doubles each:
 *  [0] cell  [1] row  [2] column  [3] sub-cells of the cell
 *  [4..9] connection of faces 1-6  [10..15] mass boundary  [16] sub-cell
 *  has exit  [17] cell has exit  [18] dry  [19] partially dry
 *  [20..25] VX1..VZ2  [26..28] DX DY DZ  [29..30] offset x  [31..32]
 *  offset y  [33] porosity  [34] retardation  [35] dtcell  [36] source
 *  [37] sink  [38..69] corner q (8 corners x 4)  [70..77] corner porosity
 *  [78..125] Dxxx Dyyy Dzzz Dxyx Dxzx Dyzy (8 each)
 * Returns the number of records, or a negative error code; out may be NULL
 * to count them. */
long orc_usg_tables(const orc_usg_problem *U, double *out) {
  ugrid G;
  int e = ugrid_init(&G, U);
  if (e) return e;
  const orc_problem *P = &U->base;
  int nd = P->dim_mask[0] + P->dim_mask[1] + P->dim_mask[2];
  upair *nb = (upair *)calloc(18, sizeof(upair));
  ucbuf center;
  memset(&center, 0, sizeof center);
  tsc T;
  memset(&T, 0, sizeof T);
  T.P = P;
  T.ndim = nd;
  long k = 0;
  for (int n = 1; n <= U->ncell && e == 0; ++n) {
    u_fill_cell(&G, n, &center);
    e = u_fill_neighbors(&G, &center, nb);
    int chx = u_cell_has_exit(&center, P->backward);
    for (int r = 1; r <= center.scr && e == 0; ++r)
      for (int c = 1; c <= center.scc; ++c, ++k) {
        if (!out) continue;
        double offx[2], offy[2];
        u_fill_subcell(&center, r, c, P->backward, &T.s, offx, offy);
        u_corners(P, &center, nb, r, c, &T.s, &T.C);
        double *o = out + k * ORC_USG_FIELDS;
        o[0] = n; o[1] = r; o[2] = c; o[3] = u_subcount(&center);
        for (int f = 0; f < 6; ++f) {
          o[4 + f] = T.s.conn[f];
          o[10 + f] = T.s.mass_bnd[f];
        }
        o[16] = subcell_has_exit(&T.s);
        o[17] = chx;
        o[18] = T.s.dry;
        o[19] = T.s.pdry;
        o[20] = T.s.VX1; o[21] = T.s.VX2; o[22] = T.s.VY1;
        o[23] = T.s.VY2; o[24] = T.s.VZ1; o[25] = T.s.VZ2;
        o[26] = T.s.DX; o[27] = T.s.DY; o[28] = T.s.DZ;
        o[29] = offx[0]; o[30] = offx[1]; o[31] = offy[0]; o[32] = offy[1];
        o[33] = T.s.porosity; o[34] = T.s.retardation;
        o[35] = rw_time_step(&T);
        o[36] = center.source; o[37] = center.sink;
        for (int ci = 0; ci < 8; ++ci)
          for (int m = 0; m < 4; ++m) o[38 + 4 * ci + m] = T.C.q[ci][m];
        for (int ci = 0; ci < 8; ++ci) o[70 + ci] = T.C.phi[ci];
        const double *D[6] = {T.C.Dxxx, T.C.Dyyy, T.C.Dzzz, T.C.Dxyx, T.C.Dxzx, T.C.Dyzy};
        for (int b6 = 0; b6 < 6; ++b6)
          for (int ci = 0; ci < 8; ++ci) o[78 + 8 * b6 + ci] = D[b6][ci];
      }
  }
  free(nb);
  free(G.jaface);
  return e ? e : k;
}

/* ------------------------------------------------------------------------ */
/* Walk                                                                     */
/* ------------------------------------------------------------------------ */

/* TrackCell%ExecuteRandomWalkParticleTracking, TrackCell.f90:375-669, on a
 * cell that may be split in sub-cells.  (x, y, z) is the location in the
 * parent cell; the result location too. */
static void u_track_cell(tsc *T, ucbuf *center, upair nb[18], double lx,
                         double ly, double lz, double lt, double maximumTime,
                         cellresult *R, long long *nsteps) {
  const orc_problem *P = T->P;
  int stopIfNoSubCellExit = 1;
  R->status = TC_UNDEFINED;
  R->exitFace = 0;
  R->next = 0;
  R->zero_dt = 0;
  R->x = lx; R->y = ly; R->z = lz; R->t = lt;
  if ((lx < 0.0) || (lx > 1.0) || (ly < 0.0) || (ly > 1.0) || (lz < 0.0) ||
      (lz > 1.0)) {
    R->status = TC_INVALID;
    return;
  }
  if (center->ibound == 0) {
    R->status = TC_INACTIVE;
    return;
  }
  if (P->stop_zone > 0 && center->zone == P->stop_zone) {
    R->status = TC_STOPZONE;
    return;
  }
  if (P->backward) {
    if (P->stop_weak_sources) {
      if (center->source != 0.0) {
        R->status = TC_WEAKSOURCE;
        return;
      } else if (!P->steady_state) stopIfNoSubCellExit = 0;
    }
  } else if (P->stop_weak_sinks) {
    if (center->sink != 0.0) {
      R->status = TC_WEAKSINK;
      return;
    } else if (!P->steady_state) stopIfNoSubCellExit = 0;
  }
  if (!u_cell_has_exit(center, P->backward)) {
    if (P->steady_state || (P->backward && center->source != 0.0) ||
        (!P->backward && center->sink != 0.0)) {
      R->status = TC_NOEXIT;
      return;
    }
  }
  /* FindSubCell, :47-90 */
  int row = 0, col = 0;
  if (u_subcount(center) == 1) {
    row = col = 1;
  } else if (u_subcount(center) == 4) {
    const double off[3] = {0.0, 1.0 / 2.0, 1.0};
    for (int i = 1; i <= 2 && !row; ++i)
      for (int j = 1; j <= 2; ++j) {
        double xmin = off[j - 1], xmax = off[j], ymin = off[2 - i], ymax = off[3 - i];
        if (lx >= xmin && lx <= xmax && ly >= ymin && ly <= ymax) {
          row = i;
          col = j;
          break;
        }
      }
  }
  if (row == 0 || col == 0) {
    R->status = TC_UNDEFINED;
    return;
  }
  double offx[2], offy[2];
  u_fill_subcell(center, row, col, P->backward, &T->s, offx, offy);
  /* ConvertFromLocalParentCoordinate (ModpathSubCellData.f90:129-149) */
  double sx = 0.0, sy = 0.0, sz = 0.0, st = 0.0;
  if (!((lx < offx[0]) || (lx > offx[1]) || (ly < offy[0]) || (ly > offy[1]) ||
        (lz < 0.0) || (lz > 1.0))) {
    sx = (lx - offx[0]) / (offx[1] - offx[0]);
    sy = (ly - offy[0]) / (offy[1] - offy[0]);
    sz = (lz - 0.0) / (1.0 - 0.0);
    st = lt;
  }
  u_corners(P, center, nb, row, col, &T->s, &T->C);
  for (int count = 1; count <= 100; ++count) {
    subresult S;
    track_subcell(T, stopIfNoSubCellExit, sx, sy, sz, st, maximumTime, &S, nsteps);
    /* ConvertToLocalParentCoordinate (ModpathSubCellData.f90:114-126) */
    double px = ((1.0 - S.x) * offx[0]) + (S.x * offx[1]);
    double py = ((1.0 - S.y) * offy[0]) + (S.y * offy[1]);
    double pz = ((1.0 - S.z) * 0.0) + (S.z * 1.0);
    switch (S.status) {
      case SC_UNDEFINED:
        R->status = TC_UNDEFINED;
        return;
      case SC_EXIT_INTERNAL: {
        int ok = 1;
        switch (S.exitFace) {
          case 1: ok = col == 2; sx = 1.0; sy = S.y; sz = S.z; col = 1; break;
          case 2: ok = col == 1; sx = 0.0; sy = S.y; sz = S.z; col = 2; break;
          case 3: ok = row == 1; sx = S.x; sy = 1.0; sz = S.z; row = 2; break;
          case 4: ok = row == 2; sx = S.x; sy = 0.0; sz = S.z; row = 1; break;
          default: ok = 0;
        }
        if (!ok) {
          R->status = TC_UNDEFINED;
          return;
        }
        st = S.t;
        u_fill_subcell(center, row, col, P->backward, &T->s, offx, offy);
        u_corners(P, center, nb, row, col, &T->s, &T->C);
        break;
      }
      case SC_EXIT_CELLFACE:
        R->status = TC_EXIT_CELLFACE;
        R->exitFace = S.exitFace;
        R->next = S.conn;
        R->zero_dt = S.zero_dt;
        R->x = px; R->y = py; R->z = pz; R->t = S.t;
        return;
      case SC_MAXTIME:
        R->status = TC_STOPTIME;
        R->exitFace = S.exitFace;
        R->x = px; R->y = py; R->z = pz; R->t = S.t;
        return;
      case SC_INACTIVE:
        R->status = TC_INACTIVE;
        return;
      case SC_NOEXIT:
        R->status = TC_NOEXIT;
        return;
      default:
        R->status = TC_UNDEFINED;
        return;
    }
  }
  R->status = TC_UNDEFINED;
}

/* ConvertFromNeighbor, ModflowRectangularGrid.f90:366-477: the location in
 * `to` of the point (fx, fy, fz) of `from` on their shared face; 0 when
 * the reference finds no defined location (it returns cell 0). */
static int u_convert_from_neighbor(const ugrid *G, int to, int from, double fx,
                                   double fy, double fz, double *x, double *y,
                                   double *z) {
  int face = u_connection_face(G, to, from);
  int bad = 0;
  switch (face) {
    case 1: bad = fx != 1.0; break;
    case 2: bad = fx != 0.0; break;
    case 3: bad = fy != 1.0; break;
    case 4: bad = fy != 0.0; break;
    case 5: bad = fz != 1.0; break;
    case 6: bad = fz != 0.0; break;
    default: bad = 1;
  }
  if (bad) return 0;
  int count2 = u_potential_count(G, to, face);
  static const int opp[6] = {2, 1, 4, 3, 6, 5};
  int count1 = u_potential_count(G, from, opp[face - 1]);
  if (count1 == 1 && count2 == 1) {
    *x = fx; *y = fy; *z = fz;
    switch (face) {
      case 1: *x = 0.0; break;
      case 2: *x = 1.0; break;
      case 3: *y = 0.0; break;
      case 4: *y = 1.0; break;
      case 5: *z = 0.0; break;
      default: *z = 1.0; break;
    }
    return 1;
  }
  /* ConvertToModelXY of `from`, ConvertToLocalXY of `to` */
  double left = u_left(G, from), right = u_right(G, from);
  double front = u_front(G, from), back = u_back(G, from);
  double gx = ((1.0 - fx) * left) + (fx * right);
  if (gx < left) gx = left;
  if (gx > right) gx = right;
  double gy = ((1.0 - fy) * front) + (fy * back);
  if (gy < front) gy = front;
  if (gy > back) gy = back;
  left = u_left(G, to); right = u_right(G, to);
  front = u_front(G, to); back = u_back(G, to);
  double lx = (gx - left) / (right - left);
  if (lx > 1.0) lx = 1.0;
  if (lx < 0.0) lx = 0.0;
  double ly = (gy - front) / (back - front);
  if (ly > 1.0) ly = 1.0;
  if (ly < 0.0) ly = 0.0;
  const double tol = 1.0e-3;
  double snapv = 0.0, *sv = NULL;
  switch (face) {
    case 1: sv = &lx; snapv = 0.0; break;
    case 2: sv = &lx; snapv = 1.0; break;
    case 3: sv = &ly; snapv = 0.0; break;
    case 4: sv = &ly; snapv = 1.0; break;
    default: break;
  }
  if (sv) {
    double mn = snapv - (tol / 2.0), mx = snapv + (tol / 2.0);
    if (*sv >= mn && *sv <= mx) *sv = snapv;
  }
  *x = lx;
  *y = ly;
  double lz = fz;
  if (face == 5) lz = 0.0;
  if (face == 6) lz = 1.0;
  *z = lz;
  return 1;
}

/* pr_RWPTrackPath (ParticleTrackingEngine.f90:1051-1254) on an
 * unstructured grid, mapped to MPathRW's particle status */
static int u_track_path(const ugrid *G, tsc *T, ucbuf *center, upair nb[18],
                        int *cell, double *lx, double *ly, double *lz,
                        double *t, double t_stop, long long *nsteps) {
  const orc_usg_problem *U = G->U;
  int loc_cell = *cell, stalls = 0;
  double x = *lx, y = *ly, z = *lz, tt = *t;
  u_fill_cell(G, loc_cell, center);
  if (u_fill_neighbors(G, center, nb)) return 9;
  for (;;) {
    if (loc_cell != center->cell) {
      u_fill_cell(G, loc_cell, center);
      if (u_fill_neighbors(G, center, nb)) return 9;
    }
    cellresult R;
    u_track_cell(T, center, nb, x, y, z, tt, t_stop, &R, nsteps);
    switch (R.status) {
      case TC_EXIT_CELLFACE: {
        *cell = center->cell;
        *lx = R.x; *ly = R.y; *lz = R.z; *t = R.t;
        stalls = R.zero_dt ? stalls + 1 : 0;
        if (stalls > ORC_MAX_STALLS) return 9;
        int next = R.next;
        if (next > 0 && u_ibound(U, next) != 0) {
          double nx, ny, nz;
          if (!u_convert_from_neighbor(G, next, center->cell, R.x, R.y, R.z, &nx, &ny,
                                       &nz))
            return 9;
          loc_cell = next;
          x = nx; y = ny; z = nz;
          tt = R.t;
        } else {
          return 2;
        }
        break;
      }
      case TC_STOPTIME:
        *cell = center->cell;
        *lx = R.x; *ly = R.y; *lz = R.z; *t = R.t;
        return 1;
      case TC_NOEXIT:
      case TC_STOPZONE:
      case TC_WEAKSINK:
      case TC_WEAKSOURCE:
      case TC_INACTIVE:
        *cell = loc_cell;
        *lx = x; *ly = y; *lz = z; *t = tt;
        return R.status == TC_NOEXIT       ? 5
               : R.status == TC_STOPZONE   ? 6
               : R.status == TC_WEAKSINK   ? 3
               : R.status == TC_WEAKSOURCE ? 4
                                           : 7;
      default:
        return 9;
    }
  }
}

int orc_usg_advance(const orc_usg_problem *U, long n, int *cell, double *lx,
                    double *ly, double *lz, double *t, int *status,
                    uint64_t *rng, double t_engine, double t_stop,
                    long long *substeps, int nthreads) {
  if (!U || n < 0) return -1;
  const orc_problem *P = &U->base;
  int nd = P->dim_mask[0] + P->dim_mask[1] + P->dim_mask[2];
  if (nd <= 0) return -2;
  int dims[3] = {0, 0, 0}, c = 0;
  for (int d = 0; d < 3; ++d)
    if (P->dim_mask[d]) dims[c++] = d + 1;
  ugrid G;
  int e = ugrid_init(&G, U);
  if (e) return e;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    ucbuf center;
    upair *nb = (upair *)calloc(18, sizeof(upair));
    tsc T;
    memset(&center, 0, sizeof center);
    memset(&T, 0, sizeof T);
    T.P = P;
    T.ndim = nd;
    for (int d = 0; d < 3; ++d) T.dims[d] = dims[d];
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (long p = 0; p < n; ++p) {
      if (status[p] == 0) {
        if (t[p] < t_engine) status[p] = 8;
        else if (t[p] <= t_stop) {
          status[p] = 1;
          if (u_ibound(U, cell[p]) == 0) status[p] = 7;
        }
      }
      if (status[p] == 7) {
        ucbuf b;
        memset(&b, 0, sizeof b);
        u_fill_cell(&G, cell[p], &b);
        u_verify_dry(&b);
        if (b.pdry) status[p] = 1;
      }
      if (status[p] != 1) continue;
      T.rng = rng + 2 * p;
      long long ns = 0;
      status[p] = u_track_path(&G, &T, &center, nb, &cell[p], &lx[p], &ly[p], &lz[p],
                               &t[p], t_stop, &ns);
      if (substeps) substeps[p] += ns;
    }
    free(nb);
  }
  free(G.jaface);
  return 0;
}

/* ConvertToModelXYZ (ModflowRectangularGrid.f90:524-549) on the cell
 * bounds, SaturatedTop = Top */
void orc_usg_global_xyz(const orc_usg_problem *U, long n, const int *cell,
                        const double *lx, const double *ly, const double *lz,
                        double *gx, double *gy, double *gz) {
  ugrid G;
  G.U = U;
  G.jaface = NULL;
  for (long p = 0; p < n; ++p) {
    int c = cell[p];
    double left = u_left(&G, c), right = u_right(&G, c);
    double front = u_front(&G, c), back = u_back(&G, c);
    double bottom = U->base.bot[c - 1], top = U->base.top[c - 1];
    double x = ((1.0 - lx[p]) * left) + (lx[p] * right);
    if (x < left) x = left;
    if (x > right) x = right;
    double y = ((1.0 - ly[p]) * front) + (ly[p] * back);
    if (y < front) y = front;
    if (y > back) y = back;
    double z = ((1.0 - lz[p]) * bottom) + (lz[p] * top);
    if (z < bottom) z = bottom;
    if (z > top) z = top;
    gx[p] = x;
    gy[p] = y;
    gz[p] = z;
  }
}
