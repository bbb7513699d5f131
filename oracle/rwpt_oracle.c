/*
 * CPU oracle for the MODPATH-RW random-walk hot path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference algorithm, written function by
 * function against the Fortran sources (citations are file:line relative to
 * the reference's src/).  It re-assembles the cell buffer and the 18-cell
 * neighbour stencil on every cell entry exactly as the reference does, so it
 * is independent of the precomputed per-cell tables the HIP product uses.
 * Floating-point expressions keep the reference's operand order; build with
 * -ffp-contract=off so that no fused multiply-add changes a rounding.
 *
 * Only tests/, the smoke check and bench.py's cpu_baseline leg use this file.
 */
#include "rwpt_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* RNG: lib/rngzig/rng_par_zig.f90                                          */
/* ------------------------------------------------------------------------ */

static inline uint64_t rotl64(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}

/* rng_xoroshiro128plus, rng_par_zig.f90:253-281 */
uint64_t orc_xoroshiro_next(uint64_t s[2]) {
  uint64_t s1 = s[0], s2 = s[1];
  uint64_t ival = s1 + s2;
  s2 = s2 ^ s1;
  s1 = (rotl64(s1, 24) ^ s2) ^ (s2 << 16);
  s2 = rotl64(s2, 37);
  s[0] = s1;
  s[1] = s2;
  return ival;
}

/* rng_uni -> rng_int64_to_real64, rng_par_zig.f90:67-74, 321-335 */
double orc_uniform(uint64_t s[2]) {
  uint64_t x = orc_xoroshiro_next(s);
  uint64_t b = 0x3FF0000000000000ull | (x >> 12);
  double d;
  memcpy(&d, &b, sizeof b);  /* the value's 8 bytes (a dual build's value) */
  return d - 1.0;
}

/* pr_GenerateStandardNormalRandomZig, TrackSubCell.F90:3631-3663 */
double orc_normal12(uint64_t s[2]) {
  double r = 0.0;
  for (int m = 0; m < 12; ++m) r = r + orc_uniform(s);
  return r - 6.0;
}

/* rng_jump, rng_par_zig.f90:284-307 */
void orc_rng_jump(uint64_t s[2]) {
  static const uint64_t jmp[2] = {~0x206ffd6b270aab5aull,
                                  0x170865df4b3201fcull};
  uint64_t s1 = 0, s2 = 0;
  for (int i = 0; i < 2; ++i)
    for (int b = 0; b < 64; ++b) {
      if (jmp[i] & (1ull << b)) {
        s1 ^= s[0];
        s2 ^= s[1];
      }
      (void)orc_xoroshiro_next(s);
    }
  s[0] = s1;
  s[1] = s2;
}

/* Stream k of rng_zigset (rng_par_zig.f90:183-191) is jump^k(base). */
void orc_rng_streams(const uint64_t base[2], long first, long n,
                     uint64_t *out) {
  uint64_t s[2] = {base[0], base[1]};
  for (long k = 0; k < first; ++k) orc_rng_jump(s);
  for (long p = 0; p < n; ++p) {
    out[2 * p] = s[0];
    out[2 * p + 1] = s[1];
    orc_rng_jump(s);
  }
}

/* ------------------------------------------------------------------------ */
/* Grid: RectangularGridDis.f90, ModflowRectangularGrid.f90                 */
/* ------------------------------------------------------------------------ */

static inline int ncells(const orc_problem *P) {
  return P->ncol * P->nrow * P->nlay;
}

static inline void cell_krc(const orc_problem *P, int c, int *k, int *i,
                            int *j) {
  int lc = P->nrow * P->ncol, c0 = c - 1;
  *k = c0 / lc + 1;
  int r = c0 % lc;
  *i = r / P->ncol + 1;
  *j = r % P->ncol + 1;
}

/* ComputeFaceConnection, RectangularGridDis.f90 (faces 1..6) */
static int face_conn(const orc_problem *P, int c, int f) {
  int k, i, j;
  cell_krc(P, c, &k, &i, &j);
  switch (f) {
    case 1: return j > 1 ? c - 1 : 0;
    case 2: return j < P->ncol ? c + 1 : 0;
    case 3: return i < P->nrow ? c + P->ncol : 0;
    case 4: return i > 1 ? c - P->ncol : 0;
    case 5: return k < P->nlay ? c + P->nrow * P->ncol : 0;
    case 6: return k > 1 ? c - P->nrow * P->ncol : 0;
  }
  return 0;
}

static inline int ibound_of(const orc_problem *P, int c) {
  return P->ibound ? P->ibound[c - 1] : 1;
}

/* ------------------------------------------------------------------------ */
/* Cell buffer: ModpathCellData.f90                                         */
/* ------------------------------------------------------------------------ */

typedef struct {
  int cell; /* 0 => empty (reset) buffer */
  double DX, DY, top, bottom, head;
  int layer_type, ibound, zone;
  double porosity, retardation, source, sink;
  double Q[6];
  int conn[6], mass_bnd[6], bnd_count[6];
  double aLH, aTH, aLV, aTV, dmeff;
  int dry, pdry;
} cbuf;

/* pr_SetMassTransportDataStructured, ModpathCellData.f90:1099-1487, plus
 * pr_FillDispersionParameters*, ParticleTrackingEngine.f90:535-612 */
static void fill_cell(const orc_problem *P, int c, cbuf *b) {
  memset(b, 0, sizeof *b);
  int k, i, j;
  cell_krc(P, c, &k, &i, &j);
  for (int f = 0; f < 6; ++f) b->bnd_count[f] = 1; /* pr_ResetArrays */
  b->cell = c;
  b->DX = P->delr[j - 1];
  b->DY = P->delc[i - 1];
  b->top = P->top[c - 1];
  b->bottom = P->bot[c - 1];
  b->head = P->head ? P->head[c - 1] : P->top[c - 1];
  b->layer_type = P->layer_type ? P->layer_type[c - 1] : 0;
  b->zone = P->zone ? P->zone[c - 1] : 0;
  b->ibound = ibound_of(P, c);
  b->porosity = P->porosity[c - 1];
  b->retardation = P->retardation ? P->retardation[c - 1] : 1.0;
  b->source = P->source_flow ? P->source_flow[c - 1] : 0.0;
  b->sink = P->sink_flow ? P->sink_flow[c - 1] : 0.0;
  for (int f = 0; f < 6; ++f) {
    int conn = face_conn(P, c, f + 1);
    if (conn > 0) {
      b->conn[f] = conn;
      b->bnd_count[f] = ibound_of(P, conn) == 0 ? 1 : 0;
      b->mass_bnd[f] = P->icbound ? P->icbound[conn - 1] : 0;
    } else {
      b->mass_bnd[f] = P->default_icbound;
    }
  }
  /* face flows, :1315-1351 */
  if (b->conn[0] > 0) b->Q[0] = P->frf[b->conn[0] - 1];
  b->Q[1] = P->frf[c - 1];
  b->Q[2] = -P->fff[c - 1];
  if (b->conn[3] > 0) b->Q[3] = -P->fff[b->conn[3] - 1];
  b->Q[4] = -P->flf[c - 1];
  if (b->conn[5] > 0) b->Q[5] = -P->flf[b->conn[5] - 1];
  /* boundary flows, :1354-1472 */
  if (P->bflow) {
    for (int f = 0; f < 6; ++f) {
      double bf = P->bflow[6 * (c - 1) + f];
      if (bf != 0.0) {
        if (b->bnd_count[f] > 0) {
          double flow = bf / b->bnd_count[f];
          int conn = b->conn[f];
          if (conn == 0 || ibound_of(P, conn) == 0)
            b->Q[f] = (f % 2 == 0) ? flow : -flow;
        } else {
          if (bf > 0.0) b->source = b->source + bf;
          else b->sink = b->sink + bf;
        }
      }
    }
  }
  int d = P->disp_uniform ? 0 : c - 1;
  b->aLH = P->aL[d];
  b->aTH = P->aT[d];
  b->dmeff = P->dmeff[d];
  if (P->disp_model == 1) {
    b->aLV = P->aLV[d];
    b->aTV = P->aTV[d];
  }
}

/* pr_GetDZRW, ModpathCellData.f90:3208-3256 (mutates the dry flags) */
static double get_dzrw(cbuf *b) {
  double dz = b->top - b->bottom;
  double dzc = b->top - b->bottom;
  b->dry = 0;
  b->pdry = 0;
  if (b->layer_type == 1) {
    if (b->head < b->top) {
      dz = b->head - b->bottom;
      b->dry = 0;
      b->pdry = 1;
      if ((dz <= 0.0) || ((b->head - b->bottom) / dzc < 0.01)) {
        dz = b->top - b->bottom;
        b->dry = 1;
        b->pdry = 0;
        int any = 0;
        for (int f = 0; f < 6; ++f) any |= (b->Q[f] != 0.0);
        if (any) {
          b->dry = 0;
          b->pdry = 1;
        }
      }
    }
  }
  return dz;
}

/* pr_VerifyDryCell, ModpathCellData.f90:3180-3205 */
static void verify_dry(cbuf *b) {
  double dz = b->top - b->bottom;
  b->dry = 0;
  b->pdry = 0;
  if (b->layer_type == 1) {
    if (b->head < b->top) {
      dz = b->head - b->bottom;
      b->dry = 0;
      b->pdry = 1;
      if (dz <= 0.0) {
        b->dry = 1;
        b->pdry = 0;
        int any = 0;
        for (int f = 0; f < 6; ++f) any |= (b->Q[f] != 0.0);
        if (any) {
          b->dry = 0;
          b->pdry = 1;
        }
      }
    }
  }
}

/* pr_HasExitFace (cell), ModpathCellData.f90:320-388 */
static int cell_has_exit(const cbuf *b, int backward) {
  double sign = backward ? -1.0 : 1.0;
  if (sign * b->Q[0] < 0.0) return 1;
  if (sign * b->Q[1] > 0.0) return 1;
  if (sign * b->Q[2] < 0.0) return 1;
  if (sign * b->Q[3] > 0.0) return 1;
  if (sign * b->Q[4] < 0.0) return 1;
  if (sign * b->Q[5] > 0.0) return 1;
  return 0;
}

/* pr_FillNeighborCellDataStructured, ParticleTrackingEngine.f90:2009-2126.
 * Slot order: W, W-front, W-back, E, E-front, E-back, front, front-bottom,
 * front-top, back, back-bottom, back-top, bottom, bottom-W, bottom-E, top,
 * top-W, top-E. */
static void fill_neighbors(const orc_problem *P, const cbuf *center,
                           cbuf nb[18]) {
  static const int sub[6][2] = {{3, 4}, {3, 4}, {5, 6},
                                {5, 6}, {1, 2}, {1, 2}};
  for (int n = 0; n < 6; ++n) {
    int slot = 3 * n;
    int conn = center->conn[n];
    if (conn > 0) {
      fill_cell(P, conn, &nb[slot]);
      for (int m = 0; m < 2; ++m) {
        int c2 = nb[slot].conn[sub[n][m] - 1];
        if (c2 > 0) fill_cell(P, c2, &nb[slot + 1 + m]);
        else memset(&nb[slot + 1 + m], 0, sizeof(cbuf));
      }
    } else {
      memset(&nb[slot], 0, 3 * sizeof(cbuf));
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Sub-cell buffer: ModpathSubCellData.f90, FillMassSubCellDataBuffer       */
/* ------------------------------------------------------------------------ */

typedef struct {
  double VX1, VX2, VY1, VY2, VZ1, VZ2, DX, DY, DZ;
  int conn[6], mass_bnd[6];
  double porosity, retardation, aLH, aTH, aLV, aTV, dmeff;
  int dry, pdry;
} scbuf;

/* pr_FillMassSubCellDataBuffer, ModpathCellData.f90:2522-2684 */
static void fill_subcell(cbuf *b, scbuf *s, int backward) {
  memset(s, 0, sizeof *s);
  s->DX = b->DX / 1.0;
  s->DY = b->DY / 1.0;
  s->DZ = get_dzrw(b);
  s->dry = b->dry;
  s->pdry = b->pdry;
  double sign = 1.0;
  if (backward) sign = -sign;
  s->VX1 = sign * b->Q[0] / s->DY / s->DZ / b->porosity / b->retardation;
  s->VX2 = sign * b->Q[1] / s->DY / s->DZ / b->porosity / b->retardation;
  s->VY1 = sign * b->Q[2] / s->DX / s->DZ / b->porosity / b->retardation;
  s->VY2 = sign * b->Q[3] / s->DX / s->DZ / b->porosity / b->retardation;
  s->VZ1 = sign * b->Q[4] / s->DX / s->DY / b->porosity / b->retardation;
  s->VZ2 = sign * b->Q[5] / s->DX / s->DY / b->porosity / b->retardation;
  for (int f = 0; f < 6; ++f) {
    s->conn[f] = b->conn[f];
    s->mass_bnd[f] = b->mass_bnd[f];
  }
  s->porosity = b->porosity;
  s->retardation = b->retardation;
  s->aLH = b->aLH;
  s->aLV = b->aLV;
  s->aTH = b->aTH;
  s->aTV = b->aTV;
  s->dmeff = b->dmeff;
}

/* pr_HasExitFace (sub-cell), ModpathSubCellData.f90 */
static int subcell_has_exit(const scbuf *s) {
  return (s->VX1 < 0.0) || (s->VX2 > 0.0) || (s->VY1 < 0.0) ||
         (s->VY2 > 0.0) || (s->VZ1 < 0.0) || (s->VZ2 > 0.0);
}

/* ------------------------------------------------------------------------ */
/* Corner variables: TrackSubCell.F90:3667-4363, 2557-2691, 3117-3303       */
/* ------------------------------------------------------------------------ */

typedef struct {
  double q[8][4];
  double phi[8];
  double Dxxx[8], Dyyy[8], Dzzz[8], Dxyx[8], Dxzx[8], Dxyy[8], Dyzy[8],
      Dxzz[8], Dyzz[8];
} corners;

/* pr_SetCornerComponentsIndexes, TrackSubCell.F90:4142-4224 */
static const int cx_idx[8][4] = {{7, 8, 13, 1},   {7, 8, 13, 2},
                                 {10, 11, 13, 1}, {10, 11, 13, 2},
                                 {7, 9, 16, 1},   {7, 9, 16, 2},
                                 {10, 12, 16, 1}, {10, 12, 16, 2}};
static const int cy_idx[8][4] = {{1, 13, 14, 3}, {4, 13, 15, 3},
                                 {1, 13, 14, 4}, {4, 13, 15, 4},
                                 {1, 16, 17, 3}, {4, 16, 18, 3},
                                 {1, 16, 17, 4}, {4, 16, 18, 4}};
static const int cz_idx[8][4] = {{1, 2, 7, 5},  {4, 5, 7, 5}, {1, 3, 10, 5},
                                 {4, 6, 10, 5}, {1, 2, 7, 6}, {4, 5, 7, 6},
                                 {1, 3, 10, 6}, {4, 6, 10, 6}};
/* pr_SetCornerPorosityIndexes, TrackSubCell.F90:4329-4363 (kept as written:
 * corners 110 and 111 use the listed slots) */
static const int cp_idx[8][6] = {{1, 2, 7, 8, 13, 14},   {4, 5, 7, 8, 13, 15},
                                 {1, 3, 10, 11, 13, 14}, {4, 6, 10, 12, 16, 18},
                                 {1, 2, 7, 9, 16, 17},   {4, 5, 7, 9, 16, 18},
                                 {1, 3, 10, 12, 16, 17}, {4, 6, 10, 11, 13, 15}};

static void corners_from_stencil(const orc_problem *P, const scbuf *s,
                                 const double cq[6], double nflow[18][6],
                                 double narea[18][3], const double nvol[18],
                                 const double nphi[18], corners *C);

static void compute_corners(const orc_problem *P, cbuf *center, cbuf nb[18],
                            const scbuf *s, corners *C) {
  double nflow[18][6], narea[18][3], nvol[18], nphi[18];
  /* pr_FillNeighborSubCellVariables, :3856-4012 (unrefined branch) */
  memset(nflow, 0, sizeof nflow);
  memset(narea, 0, sizeof narea);
  memset(nvol, 0, sizeof nvol);
  memset(nphi, 0, sizeof nphi);
  for (int n = 0; n < 18; ++n) {
    if (nb[n].cell == 0) continue;
    for (int f = 0; f < 6; ++f) nflow[n][f] = nb[n].Q[f];
    double dx = nb[n].DX, dy = nb[n].DY, dz = get_dzrw(&nb[n]);
    narea[n][0] = dy * dz;
    narea[n][1] = dx * dz;
    narea[n][2] = dx * dy;
    double dz2 = get_dzrw(&nb[n]);
    nvol[n] = dx * dy * dz2;
    nphi[n] = nb[n].porosity;
  }
  corners_from_stencil(P, s, center->Q, nflow, narea, nvol, nphi, C);
}

/* Corner variables of a sub-cell from its face flows cq and the face
 * flows, face areas, volumes and porosities of its 18 stencil cells
 * (pr_ComputeCornerVariables after pr_FillNeighborSubCellVariables). */
static void corners_from_stencil(const orc_problem *P, const scbuf *s,
                                 const double cq[6], double nflow[18][6],
                                 double narea[18][3], const double nvol[18],
                                 const double nphi[18], corners *C) {
  /* pr_ComputeCornerDischarge, :4016-4070 + GetInterpolatedCornerDischarge */
  for (int dir = 0; dir < 3; ++dir) {
    for (int ci = 0; ci < 8; ++ci) {
      const int *ix = dir == 0 ? cx_idx[ci] : dir == 1 ? cy_idx[ci] : cz_idx[ci];
      double sumArea = dir == 0   ? s->DY * s->DZ
                       : dir == 1 ? s->DX * s->DZ
                                  : s->DX * s->DY;
      double sumFlow = cq[ix[3] - 1];
      for (int m = 0; m < 3; ++m) {
        sumArea = sumArea + narea[ix[m] - 1][dir];
        sumFlow = sumFlow + nflow[ix[m] - 1][ix[3] - 1];
      }
      C->q[ci][dir] = sumFlow / sumArea;
    }
  }
  for (int ci = 0; ci < 8; ++ci)
    C->q[ci][3] = sqrt(C->q[ci][0] * C->q[ci][0] + C->q[ci][1] * C->q[ci][1] +
                       C->q[ci][2] * C->q[ci][2]);
  /* corner porosity, :4227-4326 */
  if (P->uniform_porosity) {
    for (int ci = 0; ci < 8; ++ci) C->phi[ci] = s->porosity;
  } else {
    for (int ci = 0; ci < 8; ++ci) {
      double sumV = s->DX * s->DY * s->DZ;
      double sumW = s->porosity * sumV;
      for (int m = 0; m < 6; ++m) {
        int n = cp_idx[ci][m] - 1;
        sumV = sumV + nvol[n];
        sumW = sumW + nvol[n] * nphi[n];
      }
      C->phi[ci] = sumW / sumV;
    }
  }
  /* discharge products, :3720-3823 */
  double qp[8][8];
  for (int ci = 0; ci < 8; ++ci) {
    const double *q = C->q[ci];
    for (int m = 0; m < 8; ++m) qp[ci][m] = 0.0;
    if (q[3] > 0.0) {
      qp[ci][0] = q[0] * q[0];
      qp[ci][1] = q[0] * q[1];
      qp[ci][2] = q[0] * q[2];
      qp[ci][3] = q[1] * q[1];
      qp[ci][4] = q[1] * q[2];
      qp[ci][5] = q[2] * q[2];
      qp[ci][6] = qp[ci][0] + qp[ci][3];
      if (qp[ci][6] > 0.0) qp[ci][7] = qp[ci][5] / qp[ci][6];
      for (int m = 0; m < 7; ++m) qp[ci][m] = qp[ci][m] / q[3];
    }
  }
  double dm = s->dmeff;
  if (P->disp_model == 0) {
    /* pr_CornerDispersionLinearIsotropic, :2557-2691.  Corners 010..111
     * use qCorner100(4) as their |q| (:2622-2677). */
    double aL = s->aLH, aT = s->aTH, aLmaT = aL - aT;
    for (int ci = 0; ci < 8; ++ci) {
      double qn = (ci == 0) ? C->q[0][3] : C->q[1][3];
      double aTQ = aT * qn + C->phi[ci] * dm;
      C->Dxxx[ci] = aTQ + aLmaT * qp[ci][0];
      C->Dyyy[ci] = aTQ + aLmaT * qp[ci][3];
      C->Dzzz[ci] = aTQ + aLmaT * qp[ci][5];
      C->Dxyx[ci] = aLmaT * qp[ci][1];
      C->Dxzx[ci] = aLmaT * qp[ci][2];
      C->Dxyy[ci] = C->Dxyx[ci];
      C->Dyzy[ci] = aLmaT * qp[ci][4];
      C->Dxzz[ci] = C->Dxzx[ci];
      C->Dyzz[ci] = C->Dyzy[ci];
    }
  } else {
    /* pr_CornerDispersionAxisymmetric, :3117-3303 */
    double aLH = s->aLH, aLV = s->aLV, aTH = s->aTH, aTV = s->aTV;
    for (int ci = 0; ci < 8; ++ci) {
      double pD = C->phi[ci] * dm;
      double cos2 = 0.0;
      if (C->q[ci][3] > 0.0) cos2 = C->q[ci][2] * C->q[ci][2] / C->q[ci][3];
      double aL = aLH + cos2 * (aLV - aLH);
      double aT = aTV + cos2 * (aTH - aTV);
      double aLmaT = aL - aT;
      const double *p = qp[ci];
      C->Dxxx[ci] = pD + (aL + aT * p[7]) * p[0] + aTH * p[3] * (1.0 + p[7]);
      C->Dyyy[ci] = pD + aTH * p[0] * (1.0 + p[7]) + p[3] * (aL + aT * p[7]);
      C->Dxyx[ci] = (aL - aTH * (1.0 + p[7]) + aT * p[7]) * p[0];
      C->Dxyy[ci] = C->Dxyx[ci];
      C->Dzzz[ci] = pD + aT * p[6] + aL * p[5];
      C->Dxzx[ci] = aLmaT * p[2];
      C->Dyzy[ci] = aLmaT * p[4];
      C->Dxzz[ci] = C->Dxzx[ci];
      C->Dyzz[ci] = C->Dyzy[ci];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Interpolation and displacements: TrackSubCell.F90                        */
/* ------------------------------------------------------------------------ */

/* pr_Trilinear, :4546-4580 */
static double trilinear(double x, double y, double z, const double v[8]) {
  double v00 = (1.0 - x) * v[0] + x * v[1];
  double v01 = (1.0 - x) * v[4] + x * v[5];
  double v10 = (1.0 - x) * v[2] + x * v[3];
  double v11 = (1.0 - x) * v[6] + x * v[7];
  double v0 = (1.0 - y) * v00 + y * v10;
  double v1 = (1.0 - y) * v01 + y * v11;
  return (1.0 - z) * v0 + z * v1;
}

/* pr_TrilinearDerivativeX/Y/Z, :4432-4543 */
static double tri_dx(double x, double y, double z, const double v[8],
                     double DX) {
  (void)x;
  double v00 = (v[1] - v[0]) / DX, v01 = (v[5] - v[4]) / DX;
  double v10 = (v[3] - v[2]) / DX, v11 = (v[7] - v[6]) / DX;
  double v0 = (1.0 - y) * v00 + y * v10;
  double v1 = (1.0 - y) * v01 + y * v11;
  return (1.0 - z) * v0 + z * v1;
}
static double tri_dy(double x, double y, double z, const double v[8],
                     double DY) {
  (void)y;
  double v00 = (v[2] - v[0]) / DY, v01 = (v[6] - v[4]) / DY;
  double v10 = (v[3] - v[1]) / DY, v11 = (v[7] - v[5]) / DY;
  double v0 = (1.0 - x) * v00 + x * v10;
  double v1 = (1.0 - x) * v01 + x * v11;
  return (1.0 - z) * v0 + z * v1;
}
static double tri_dz(double x, double y, double z, const double v[8],
                     double DZ) {
  (void)z;
  double v00 = (v[4] - v[0]) / DZ, v01 = (v[6] - v[2]) / DZ;
  double v10 = (v[5] - v[1]) / DZ, v11 = (v[7] - v[3]) / DZ;
  double v0 = (1.0 - x) * v00 + x * v10;
  double v1 = (1.0 - x) * v01 + x * v11;
  return (1.0 - y) * v0 + y * v1;
}

typedef struct {
  const orc_problem *P;
  scbuf s;
  corners C;
  int ndim, dims[3];
  uint64_t *rng;
} tsc; /* TrackSubCellType state for one particle */

/* pr_DispersionDivergence, :2695-2818 */
static void divergence(const tsc *T, double x, double y, double z,
                       double *dx, double *dy, double *dz) {
  const corners *C = &T->C;
  const scbuf *s = &T->s;
  double dDxxdx = tri_dx(x, y, z, C->Dxxx, s->DX);
  double dDyydy = tri_dy(x, y, z, C->Dyyy, s->DY);
  double dDzzdz = tri_dz(x, y, z, C->Dzzz, s->DZ);
  double dDxydx = tri_dx(x, y, z, C->Dxyx, s->DX);
  double dDxzdx = tri_dx(x, y, z, C->Dxzx, s->DX);
  double dDxydy = tri_dy(x, y, z, C->Dxyy, s->DY);
  double dDyzdy = tri_dy(x, y, z, C->Dyzy, s->DY);
  double dDxzdz = tri_dz(x, y, z, C->Dxzz, s->DZ);
  double dDyzdz = tri_dz(x, y, z, C->Dyzz, s->DZ);
  *dx = (dDxxdx + dDxydy + dDxzdz) / s->porosity / s->retardation;
  *dy = (dDxydx + dDyydy + dDyzdz) / s->porosity / s->retardation;
  *dz = (dDxzdx + dDyzdy + dDzzdz) / s->porosity / s->retardation;
}

static void corner_velocity(const tsc *T, int comp, double v[8]) {
  for (int ci = 0; ci < 8; ++ci) v[ci] = T->C.q[ci][comp] / T->C.phi[ci];
}

/* pr_DisplacementRandomDischarge (3-D), :2821-2928, and the axisymmetric
 * variant, :3306-3420 (alphaTH replaces alphaT in B13/B23) */
static void random_disp_3d(tsc *T, double x, double y, double z,
                           double alphaL, double alphaT, double alphaT13,
                           double dMEff, double *dBx, double *dBy,
                           double *dBz) {
  double v[8], vBx, vBy, vBz;
  corner_velocity(T, 0, v);
  vBx = trilinear(x, y, z, v);
  corner_velocity(T, 1, v);
  vBy = trilinear(x, y, z, v);
  corner_velocity(T, 2, v);
  vBz = trilinear(x, y, z, v);
  double R = T->s.retardation;
  double vBnorm = sqrt(vBx * vBx + vBy * vBy + vBz * vBz);
  double vBnormxy = sqrt(vBx * vBx + vBy * vBy);
  double B11 = 0, B12 = 0, B13 = 0, B21 = 0, B22 = 0, B23 = 0, B31 = 0,
         B32 = 0;
  if (vBnorm > 0.0) {
    B11 = vBx * sqrt(2 * (alphaL * vBnorm + dMEff) / R) / vBnorm;
    B21 = vBy * sqrt(2 * (alphaL * vBnorm + dMEff) / R) / vBnorm;
    B31 = vBz * sqrt(2 * (alphaL * vBnorm + dMEff) / R) / vBnorm;
    B32 = vBnormxy * sqrt(2 * (alphaT * vBnorm + dMEff) / R) / vBnorm;
    if (vBnormxy > 0.0) {
      B12 = -(vBx * vBz * sqrt(2 * (alphaT * vBnorm + dMEff) / R) / vBnorm /
              vBnormxy);
      B13 = -(vBy * sqrt(2 * (alphaT13 * vBnorm + dMEff) / R) / vBnormxy);
      B22 = -(vBy * vBz * sqrt(2 * (alphaT * vBnorm + dMEff) / R) / vBnorm /
              vBnormxy);
      B23 = vBx * sqrt(2 * (alphaT13 * vBnorm + dMEff) / R) / vBnormxy;
    }
  }
  double r1 = orc_normal12(T->rng);
  double r2 = orc_normal12(T->rng);
  double r3 = orc_normal12(T->rng);
  *dBx = B11 * r1 + B12 * r2 + B13 * r3;
  *dBy = B21 * r1 + B22 * r2 + B23 * r3;
  *dBz = B31 * r1 + B32 * r2;
}

/* pr_DisplacementRandomDischarge2D, :2931-3028 (and the axisymmetric 2-D
 * form, :3423-3500, which passes alphaTH as the transverse dispersivity) */
static void random_disp_2d(tsc *T, double x, double y, double z,
                           double alphaL, double alphaT, double dMEff,
                           double *dBx, double *dBy, double *dBz) {
  int d1 = T->dims[0] - 1, d2 = T->dims[1] - 1;
  double v[8], dB[3] = {0.0, 0.0, 0.0};
  corner_velocity(T, d1, v);
  double vB1 = trilinear(x, y, z, v);
  corner_velocity(T, d2, v);
  double vB2 = trilinear(x, y, z, v);
  double R = T->s.retardation;
  double vBnorm = sqrt(vB1 * vB1 + vB2 * vB2);
  double B11 = 0, B12 = 0, B21 = 0, B22 = 0;
  if (vBnorm > 0.0) {
    B11 = vB1 * sqrt(2 * (alphaL * vBnorm + dMEff) / R) / vBnorm;
    B21 = vB2 * sqrt(2 * (alphaL * vBnorm + dMEff) / R) / vBnorm;
    B12 = -(vB2 * sqrt(2 * (alphaT * vBnorm + dMEff) / R) / vBnorm);
    B22 = vB1 * sqrt(2 * (alphaT * vBnorm + dMEff) / R) / vBnorm;
  }
  double r1 = orc_normal12(T->rng);
  double r2 = orc_normal12(T->rng);
  dB[d1] = B11 * r1 + B12 * r2;
  dB[d2] = B21 * r1 + B22 * r2;
  *dBx = dB[0];
  *dBy = dB[1];
  *dBz = dB[2];
}

/* pr_DisplacementRandomDischarge1D, :3031-3113 */
static void random_disp_1d(tsc *T, double x, double y, double z,
                           double alphaL, double dMEff, double *dBx,
                           double *dBy, double *dBz) {
  int d1 = T->dims[0] - 1;
  double v[8], dB[3] = {0.0, 0.0, 0.0};
  corner_velocity(T, d1, v);
  double vB1 = trilinear(x, y, z, v);
  double R = T->s.retardation;
  double vBnorm = sqrt(vB1 * vB1);
  double B11 = 0.0;
  if (vBnorm > 0.0) B11 = sqrt(2 * (alphaL * vBnorm + dMEff) / R);
  double r1 = orc_normal12(T->rng);
  dB[d1] = B11 * r1;
  *dBx = dB[0];
  *dBy = dB[1];
  *dBz = dB[2];
}

/* pr_AdvectionDisplacementExponential / Eulerian, :1487-1573 */
static void advection(const tsc *T, double dt, double vx, double vy,
                      double vz, double *ax, double *ay, double *az) {
  const scbuf *s = &T->s;
  if (T->P->advection_kind == 1) {
    const double dvtol = 1.0e-10;
    double dvxdx = (s->VX2 - s->VX1) / s->DX;
    if (fabs(dvxdx) > dvtol) *ax = vx * (exp(dvxdx * dt) - 1.0) / dvxdx;
    else *ax = vx * dt;
    double dvydy = (s->VY2 - s->VY1) / s->DY;
    if (fabs(dvydy) > dvtol) *ay = vy * (exp(dvydy * dt) - 1.0) / dvydy;
    else *ay = vy * dt;
    double dvzdz = (s->VZ2 - s->VZ1) / s->DZ;
    if (fabs(dvzdz) > dvtol) *az = vz * (exp(dvzdz * dt) - 1.0) / dvzdz;
    else *az = vz * dt;
  } else {
    *ax = vx * dt;
    *ay = vy * dt;
    *az = vz * dt;
  }
}

/* pr_RWPTDisplacementsLinear :673-706 / ...Axisymmetric :764-807 */
static void rwpt_displacements(tsc *T, double x, double y, double z,
                               double dt, double *vx, double *vy, double *vz,
                               double *ax, double *ay, double *az,
                               double *dBx, double *dBy, double *dBz,
                               double *ddx, double *ddy, double *ddz) {
  const scbuf *s = &T->s;
  /* pr_LinearInterpolationVelocities, :1461-1484 */
  *vx = (1.0 - x) * s->VX1 + x * s->VX2;
  *vy = (1.0 - y) * s->VY1 + y * s->VY2;
  *vz = (1.0 - z) * s->VZ1 + z * s->VZ2;
  double aL, aT, aTH = s->aTH;
  if (T->P->disp_model == 1) {
    double vnorm = sqrt(*vx * *vx + *vy * *vy + *vz * *vz);
    double cos2 = 0.0;
    if (vnorm > 0.0) cos2 = *vz * *vz / vnorm;
    aL = s->aLH + cos2 * (s->aLV - s->aLH);
    aT = s->aTV + cos2 * (s->aTH - s->aTV);
  } else {
    aL = s->aLH;
    aT = s->aTH;
    aTH = aT;
  }
  divergence(T, x, y, z, ddx, ddy, ddz);
  if (T->ndim == 3)
    random_disp_3d(T, x, y, z, aL, aT, aTH, s->dmeff, dBx, dBy, dBz);
  else if (T->ndim == 2) /* the axisymmetric 2-D form uses alphaTH */
    random_disp_2d(T, x, y, z, aL, aTH, s->dmeff, dBx, dBy, dBz);
  else
    random_disp_1d(T, x, y, z, aL, s->dmeff, dBx, dBy, dBz);
  advection(T, dt, *vx, *vy, *vz, ax, ay, az);
}

/* pr_ComputeRandomWalkTimeStep, :1376-1458 */
static double rw_time_step(const tsc *T) {
  const scbuf *s = &T->s;
  const orc_problem *P = T->P;
  double mx = fmax(fabs(s->VX1), fabs(s->VX2));
  double my = fmax(fabs(s->VY1), fabs(s->VY2));
  double mz = fmax(fabs(s->VZ1), fabs(s->VZ2));
  double aL = s->aLH, aT = s->aTH;
  double dts[2] = {0.0, 0.0};
  switch (P->time_step_kind) {
    case 1:
      return P->ts_params[0] / (mx / s->DX + my / s->DY + mz / s->DZ);
    case 2:
      return P->ts_params[1] / (aL * mx / (s->DX * s->DX) +
                                aT * my / (s->DY * s->DY) +
                                aT * mz / (s->DZ * s->DZ));
    case 3: {
      dts[0] = P->ts_params[0] / (mx / s->DX + my / s->DY + mz / s->DZ);
      dts[1] = P->ts_params[1] / (aL * mx / (s->DX * s->DX) +
                                  aT * my / (s->DY * s->DY) +
                                  aT * mz / (s->DZ * s->DZ));
      double m = DBL_MAX; /* minval over an empty mask is huge() */
      for (int i = 0; i < 2; ++i)
        if (dts[i] > 0.0 && dts[i] < m) m = dts[i];
      return m;
    }
    default:
      return P->ts_params[0];
  }
}

/* pr_DetectExitFaceAndUpdateTimeStepQuadratic, :1576-1892 (one axis) */
static int quad_axis(double n, double p, double d, double v, double divD,
                     double dB, int f_lo, int f_hi, double dtcur,
                     double *dtaxis, int *exitFace, double *dt) {
  /* returns 1 if the caller must return immediately (interface case) */
  double dInterface, AFace, BFace, z1, z2, zsqrt, zsqrtarg;
  int exitF;
  if (n > 1.0) {
    dInterface = d * (1.0 - p);
    exitF = f_hi;
  } else {
    dInterface = -d * p;
    exitF = f_lo;
  }
  if (dInterface == 0.0) {
    *exitFace = exitF;
    *dt = 0.0;
    return 1;
  }
  AFace = v + divD;
  BFace = dB;
  zsqrtarg = BFace * BFace + 4 * dInterface * AFace;
  if ((zsqrtarg >= 0.0) && (AFace != 0.0)) {
    zsqrt = sqrt(zsqrtarg);
    z1 = (-BFace + zsqrt) / (2 * AFace);
    z2 = (-BFace - zsqrt) / (2 * AFace);
    if (z1 > 0.0 || z2 > 0.0) {
      double m = (z1 > 0.0 && z2 > 0.0) ? (z1 < z2 ? z1 : z2)
                                        : (z1 > 0.0 ? z1 : z2);
      *dtaxis = m * m;
      if (*dtaxis == 0.0) {
        *exitFace = exitF;
        *dt = 0.0;
        return 1;
      }
    } else if (z1 == 0.0 || z2 == 0.0) {
      *exitFace = exitF;
      *dt = 0.0;
      return 1;
    }
  } else if (AFace == 0.0) {
    if (BFace != 0.0) {
      double r = dInterface / BFace;
      *dtaxis = r * r;
      if (*dtaxis == 0.0) {
        *exitFace = exitF;
        *dt = 0.0;
        return 1;
      }
    }
  }
  if (*dtaxis > dtcur) *dtaxis = 0.0;
  return 0;
}

/* pr_NewtonRaphsonVariablesExponential, :2082-2125 */
static void nr_vars(double dt, double v, double v1, double v2, double dx,
                    double dI, double divD, double dB, double *f0, double *f1,
                    double *f2, double *f3) {
  if (dt == 0.0) return;
  const double dvtol = 1.0e-10;
  double dvdx = (v2 - v1) / dx, dAdv;
  if (fabs(dvdx) > dvtol) dAdv = v * (exp(dvdx * dt) - 1.0) / dvdx;
  else dAdv = v * dt;
  *f0 = dAdv + divD * dt + dB * sqrt(dt) - dI;
  *f1 = v * exp(dvdx * dt) + divD + 0.5 * dB / sqrt(dt);
  *f2 = dvdx * v * exp(dvdx * dt) - 0.25 * dB * pow(dt, -1.5);
  *f3 = dvdx * dvdx * v * exp(dvdx * dt) + 0.375 * dB * pow(dt, -2.5);
}

/* pr_GetConvergenceFunction(+Derivative), :2128-2159 */
static double conv_g(double f0, double f1, double f2) {
  return fabs(f2 * f0 / (f1 * f1));
}
static double conv_gd(double f0, double f1, double f2, double f3) {
  double gna = f2 * f0 / (f1 * f1);
  double gpp = -2.0 * (f0 * (f2 * f2)) / (f1 * f1 * f1) +
               (f1 * f2 + f0 * f3) / (f1 * f1);
  return gna / fabs(gna) * gpp;
}

/* Literals the reference writes as default REAL (single precision). */
#define R4(x) ((double)(float)(x))

/* 10d0**n with an integer n: repeated products (exact for |n| <= 22) and a
 * reciprocal for n < 0, as an integer power is evaluated. */
static double pow10i(int n) {
  double p = 1.0;
  for (int i = 0; i < (n < 0 ? -n : n); ++i) p = p * 10.0;
  return n < 0 ? 1.0 / p : p;
}

/* pr_SetInitialGuess, :2162-2448 */
static double nr_initial_guess(double dt, double v, double v1, double v2,
                               double dx, double dI, double divD, double dB) {
  double dtnew = 0.0, dtold, dt0, f0 = 0, f1 = 0, f2 = 0, f3 = 0;
  double gprimdernew, gprimnew, gprimder, gprim, gprimmin, gamman;
  double dtfraction = 0.0, dtoldfraction, mindtfrac = 0.0;
  const double gdtol = R4(1e-6);
  const int maxter = 50;
  int counter, ccount, bcount, zcount, rcount, gcount;
  if (dt == 0.0) return dtnew;
  gprimmin = R4(1e10);
  gprim = R4(1e10);
  gcount = 0;
  while ((gprim > 1) && (gcount < 7)) {
    gcount = gcount + 1;
    dtfraction = pow10i(-8 + gcount);
    dt0 = dt * dtfraction;
    nr_vars(dt0, v, v1, v2, dx, dI, divD, dB, &f0, &f1, &f2, &f3);
    gprim = conv_g(f0, f1, f2);
    if (gprim < gprimmin) {
      gprimmin = gprim;
      mindtfrac = dtfraction;
    }
  }
  if ((gprim < 1) && (gcount == 1)) {
    dtoldfraction = 1e-8;
  } else if (gprim < 1) {
    dtoldfraction = pow10i(-8 + gcount - 1);
  } else {
    dtfraction = mindtfrac;
    dtoldfraction = 0.5 * dtfraction / 10.0;
  }
  dtnew = dtfraction * dt;
  dtold = dtoldfraction * dt;
  nr_vars(dtnew, v, v1, v2, dx, dI, divD, dB, &f0, &f1, &f2, &f3);
  gprim = conv_g(f0, f1, f2);
  if (gprim < 0.5) return dtnew;

  gprimnew = 10.0;
  gprim = 10.0;
  ccount = bcount = zcount = rcount = counter = 0;
  while (counter < maxter) {
    counter = counter + 1;
    nr_vars(dtold, v, v1, v2, dx, dI, divD, dB, &f0, &f1, &f2, &f3);
    gprimder = conv_gd(f0, f1, f2, f3);
    nr_vars(dtnew, v, v1, v2, dx, dI, divD, dB, &f0, &f1, &f2, &f3);
    gprimdernew = conv_gd(f0, f1, f2, f3);
    gprim = conv_g(f0, f1, f2);
    if ((gprimdernew - gprimder) > 0.0) {
      double dg = gprimdernew - gprimder;
      gamman = fabs((dtnew - dtold) * dg) / (dg * dg);
    } else {
      gamman = 0.0;
    }
    dtold = dtnew;
    dtnew = dtnew - gamman * gprimdernew;
    if (isnan(dtnew)) {
      dtnew = 0.0;
      break;
    }
    if ((gprim < 1) && (dtnew / dt > 1)) {
      dtnew = dtold;
      return dtnew;
    } else if (dtnew / dt > 1) {
      dtnew = R4(0.99) * dt;
    }
    if ((gprim < 1) && (dtnew < 0) && (dtold > 0)) {
      dtnew = dtold;
      return dtnew;
    } else if (dtnew < 0) {
      dtnew = R4(1e-8) * dt;
      zcount = zcount + 1;
      if (zcount > 1) {
        dtnew = 0.0;
        break;
      }
    }
    nr_vars(dtnew, v, v1, v2, dx, dI, divD, dB, &f0, &f1, &f2, &f3);
    gprimnew = conv_g(f0, f1, f2);
    if ((gprimnew < 1) && (dtnew > 0) && (dtnew / dt < 1)) {
      rcount = rcount + 1;
      if (rcount > 1) return dtnew;
    } else {
      rcount = 0;
    }
    if ((gprim < 1) && (gprimnew > gprim)) {
      dtnew = dtold;
      return dtnew;
    }
    if (counter > maxter) break;
    if ((fabs((dtold - dtnew) / dtold) < gdtol) && (gprim < 1)) {
      ccount = ccount + 1;
      if (ccount > 1) return dtnew;
    } else {
      ccount = 0;
    }
    if ((fabs((dtold - dtnew) / dtold) < gdtol) && (gprim > 1)) {
      bcount = bcount + 1;
      if (bcount > 1) break;
    } else {
      bcount = 0;
    }
  }
  if (gprim < 1) return dtnew;

  gprim = R4(1e10);
  gprimmin = R4(1e10);
  mindtfrac = 0.0;
  gcount = 1;
  while ((gprim > 1) && (gcount < 9)) {
    gcount = gcount + 1;
    dtfraction = (double)(0.1f * (float)gcount);
    dt0 = dt * dtfraction;
    nr_vars(dt0, v, v1, v2, dx, dI, divD, dB, &f0, &f1, &f2, &f3);
    gprim = conv_g(f0, f1, f2);
    if (gprim < gprimmin) {
      gprimmin = gprim;
      mindtfrac = dtfraction;
    }
  }
  if (gprimmin < 1) dtnew = dt * mindtfrac;
  if ((dtnew <= 0.0) && (fabs(v) > 0.0)) dtnew = fabs(dI / v);
  return dtnew;
}

/* NewtonRaphsonTimeStepExponentialAdvection, :2451-2554 */
static double nr_time_step(double dt, double v, double v1, double v2,
                           double dx, double dI, double divD, double dB) {
  const double dvtol = 1.0e-10, nrtol = R4(1e-6);
  const int maxIter = 50;
  double nrf0, nrfprim, nrerror = 1e6, dAdv;
  double dvdx = (v2 - v1) / dx;
  double dt0 = nr_initial_guess(dt, v, v1, v2, dx, dI, divD, dB);
  if (dt0 == 0.0) return 0.0;
  int countIter = 0;
  while ((fabs(nrerror / dt0) > nrtol) && (countIter < maxIter)) {
    countIter = countIter + 1;
    if (fabs(dvdx) > dvtol) dAdv = v * (exp(dvdx * dt0) - 1.0) / dvdx;
    else dAdv = v * dt0;
    nrf0 = dAdv + divD * dt0 + dB * sqrt(dt0) - dI;
    nrfprim = v * exp(dvdx * dt0) + divD + 0.5 * dB / sqrt(dt0);
    nrerror = -nrf0 / nrfprim;
    dt0 = dt0 + nrerror;
    if (dt0 < 0.0) dt0 = R4(-0.9) * dt0;
  }
  double dtnr = dt0;
  if (dt0 > dt) return dt;
  if ((countIter == maxIter) && (fabs(nrerror / dt0) > nrtol)) return 0.0;
  return dtnr;
}

/* ExitFaceAndUpdateTimeStep: Quadratic (:1576-1892) or Newton (:1895-2079) */
static void exit_face_and_dt(const tsc *T, double x, double y, double z,
                             double nx, double ny, double nz, double vx,
                             double vy, double vz, double divDx, double divDy,
                             double divDz, double dBx, double dBy, double dBz,
                             double *t, double *dt, double dtxyz[3],
                             int *exitFace) {
  const scbuf *s = &T->s;
  int exX = 0, exY = 0, exZ = 0;
  *t = *t - *dt;
  const double pn[3] = {nx, ny, nz}, pp[3] = {x, y, z};
  const double d[3] = {s->DX, s->DY, s->DZ}, v[3] = {vx, vy, vz};
  const double dD[3] = {divDx, divDy, divDz}, dB[3] = {dBx, dBy, dBz};
  const double v1[3] = {s->VX1, s->VY1, s->VZ1};
  const double v2[3] = {s->VX2, s->VY2, s->VZ2};
  int *ex[3] = {&exX, &exY, &exZ};
  for (int a = 0; a < 3; ++a) {
    if ((pn[a] > 1.0) || (pn[a] < 0.0)) {
      if (T->P->advection_kind == 2) {
        if (quad_axis(pn[a], pp[a], d[a], v[a], dD[a], dB[a], 2 * a + 1,
                      2 * a + 2, *dt, &dtxyz[a], exitFace, dt))
          return;
        *ex[a] = (pn[a] > 1.0) ? 2 * a + 2 : 2 * a + 1;
      } else {
        double dI;
        if (pn[a] > 1.0) {
          dI = d[a] * (1.0 - pp[a]);
          *ex[a] = 2 * a + 2;
        } else {
          dI = -d[a] * pp[a];
          *ex[a] = 2 * a + 1;
        }
        if (dI == 0.0) {
          *exitFace = *ex[a];
          *dt = 0.0;
          return;
        }
        dtxyz[a] = nr_time_step(*dt, v[a], v1[a], v2[a], d[a], dI, dD[a],
                                dB[a]);
        if (dtxyz[a] > *dt) dtxyz[a] = 0.0;
      }
    }
  }
  if (dtxyz[0] == 0.0 && dtxyz[1] == 0.0 && dtxyz[2] == 0.0) {
    *t = *t + *dt;
    return;
  }
  int imin = -1;
  for (int a = 0; a < 3; ++a)
    if (dtxyz[a] > 0.0 && (imin < 0 || dtxyz[a] < dtxyz[imin])) imin = a;
  *dt = dtxyz[imin];
  *exitFace = *ex[imin];
  *t = *t + *dt;
}

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */

/* TrackSubCellResult.f90:48-105 */
enum { SC_UNDEFINED = 0, SC_EXIT_INTERNAL = 1, SC_EXIT_CELLFACE = 2,
       SC_MAXTIME = 3, SC_NOEXIT = 4, SC_INACTIVE = 7 };
/* TrackCellResult.f90:53-125 */
enum { TC_UNDEFINED = 0, TC_EXIT_CELLFACE = 1, TC_STOPTIME = 2,
       TC_NOEXIT = 3, TC_STOPZONE = 4, TC_INACTIVE = 5, TC_INVALID = 6,
       TC_WEAKSINK = 7, TC_WEAKSOURCE = 8 };

typedef struct {
  int status, exitFace, conn, zero_dt;
  double x, y, z, t;
} subresult;

/* Guards against walks the reference never ends (not in the reference):
 * sub-steps per particle and advance call, and consecutive cell exits made
 * with a zero time step (a particle sitting on a face whose drift points
 * into the face from both sides, with no random component across it,
 * ping-pongs between the two cells forever in pr_RWPTrackPath).  Either
 * reports status 9. */
#define ORC_MAX_STEPS 20000000LL
#define ORC_MAX_STALLS 1000

/* pr_ExecuteRandomWalkParticleTracking, TrackSubCell.F90:812-1373 */
static void track_subcell(tsc *T, int stopIfNoExit, double ix, double iy,
                          double iz, double initialTime, double maximumTime,
                          subresult *R, long long *nsteps) {
  const double maxRelativeJump = 0.5;
  const double dtReductionFactor = R4(0.1);
  const int maxInterfaceLoopCounter = 5, maxRestartPositionCounter = 10;
  scbuf *s = &T->s;
  double vx = 0, vy = 0, vz = 0, dt, t = 0.0, x, y, z, nx, ny, nz;
  double divDx = 0, divDy = 0, divDz = 0, dAdvx = 0, dAdvy = 0, dAdvz = 0;
  double dBx = 0, dBy = 0, dBz = 0, xi, yi, zi, dxrw, dyrw, dzrw;
  double tinit, dtcell, dtxyz[3];
  int exitFace = 0, posRestartCounter, reboundCounter, intLoopCounter;
  int reachedMaximumTime;

  R->status = SC_UNDEFINED;
  R->exitFace = 0;
  R->conn = 0;
  R->zero_dt = 0;
  R->x = ix;
  R->y = iy;
  R->z = iz;
  R->t = initialTime;
  if (stopIfNoExit && !subcell_has_exit(s)) {
    R->status = SC_NOEXIT;
    return;
  }
  double dx = s->DX, dy = s->DY, dz = s->DZ;
  x = ix; y = iy; z = iz;
  nx = ix; ny = iy; nz = iz;
  if (s->dry) {
    R->status = SC_INACTIVE; /* final time is not defined by the reference */
    return;
  }
  dtcell = rw_time_step(T);
  t = initialTime;
  tinit = initialTime;
  if (dtcell <= 0.0) {
    R->status = SC_UNDEFINED;
    return;
  }
  posRestartCounter = 0;
  exitFace = 0;
  reachedMaximumTime = 0;
  for (;;) {
    /* guard against a walk that never ends (not in the reference) */
    if (nsteps && *nsteps >= ORC_MAX_STEPS) {
      R->status = SC_UNDEFINED;
      return;
    }
    if (nsteps) *nsteps += 1;
    dt = dtcell;
    t = t + dt;
    if (maximumTime < t) {
      t = t - dt;
      dt = fmax(maximumTime - t, 0.0);
      t = maximumTime;
      reachedMaximumTime = 1;
    }
    rwpt_displacements(T, x, y, z, dt, &vx, &vy, &vz, &dAdvx, &dAdvy, &dAdvz,
                       &dBx, &dBy, &dBz, &divDx, &divDy, &divDz);
    dxrw = dAdvx + divDx * dt + dBx * sqrt(dt);
    dyrw = dAdvy + divDy * dt + dBy * sqrt(dt);
    dzrw = dAdvz + divDz * dt + dBz * sqrt(dt);
    while ((fabs(dxrw / dx) > maxRelativeJump) ||
           (fabs(dyrw / dy) > maxRelativeJump) ||
           (fabs(dzrw / dz) > maxRelativeJump)) {
      t = t - dt;
      dt = dtReductionFactor * dt;
      advection(T, dt, vx, vy, vz, &dAdvx, &dAdvy, &dAdvz);
      dxrw = dAdvx + divDx * dt + dBx * sqrt(dt);
      dyrw = dAdvy + divDy * dt + dBy * sqrt(dt);
      dzrw = dAdvz + divDz * dt + dBz * sqrt(dt);
      t = t + dt;
      if (reachedMaximumTime) reachedMaximumTime = 0;
    }
    nx = x + dxrw / dx;
    ny = y + dyrw / dy;
    nz = z + dzrw / dz;

    intLoopCounter = 0;
    while ((nx > 1.0) || (nx < 0.0) || (ny > 1.0) || (ny < 0.0) ||
           (nz > 1.0) || (nz < 0.0)) {
      intLoopCounter = intLoopCounter + 1;
      dtxyz[0] = dtxyz[1] = dtxyz[2] = 0.0;
      exit_face_and_dt(T, x, y, z, nx, ny, nz, vx, vy, vz, divDx, divDy, divDz,
                       dBx, dBy, dBz, &t, &dt, dtxyz, &exitFace);
      if (intLoopCounter > maxInterfaceLoopCounter) {
        nx = ix; ny = iy; nz = iz;
        t = tinit;
        dt = dtcell;
        posRestartCounter = posRestartCounter + 1;
        if (posRestartCounter > maxRestartPositionCounter) {
          R->exitFace = 0;
          R->status = SC_UNDEFINED;
          R->x = x; R->y = y; R->z = z; R->t = t;
          return;
        }
        break;
      }
      advection(T, dt, vx, vy, vz, &dAdvx, &dAdvy, &dAdvz);
      if (reachedMaximumTime) reachedMaximumTime = 0;
      switch (exitFace) {
        case 1:
        case 2:
          nx = 1.0;
          if (exitFace == 1) nx = 0.0;
          dyrw = dAdvy + divDy * dt + dBy * sqrt(dt);
          ny = y + dyrw / dy;
          dzrw = dAdvz + divDz * dt + dBz * sqrt(dt);
          nz = z + dzrw / dz;
          break;
        case 3:
        case 4:
          dxrw = dAdvx + divDx * dt + dBx * sqrt(dt);
          nx = x + dxrw / dx;
          ny = 1.0;
          if (exitFace == 3) ny = 0.0;
          dzrw = dAdvz + divDz * dt + dBz * sqrt(dt);
          nz = z + dzrw / dz;
          break;
        case 5:
        case 6:
          dxrw = dAdvx + divDx * dt + dBx * sqrt(dt);
          nx = x + dxrw / dx;
          dyrw = dAdvy + divDy * dt + dBy * sqrt(dt);
          ny = y + dyrw / dy;
          nz = 1.0;
          if (exitFace == 5) nz = 0.0;
          break;
        default: /* case(0): restart */
          nx = ix; ny = iy; nz = iz;
          t = tinit;
          dt = dtcell;
          posRestartCounter = posRestartCounter + 1;
          if (posRestartCounter > maxRestartPositionCounter) {
            R->exitFace = 0;
            R->status = SC_UNDEFINED;
            R->x = x; R->y = y; R->z = z; R->t = t;
            return;
          }
          goto interface_done;
      }
      if ((nx > 1.0) || (nx < 0.0) || (ny > 1.0) || (ny < 0.0) ||
          (nz > 1.0) || (nz < 0.0)) {
        continue;
      }
      /* elasticRebound, :1157-1312 */
      reboundCounter = 0;
      while (exitFace > 0) {
        if (s->mass_bnd[exitFace - 1] != 1) break;
        reboundCounter = reboundCounter + 1;
        if (reboundCounter > maxRestartPositionCounter) {
          R->exitFace = 0;
          R->status = SC_UNDEFINED;
          R->x = x; R->y = y; R->z = z; R->t = t;
          return;
        }
        xi = nx; yi = ny; zi = nz;
        if (dt == 0.0) {
          nx = ix; ny = iy; nz = iz;
          t = t - dt;
          dt = dtcell;
          exitFace = 0;
          posRestartCounter = posRestartCounter + 1;
          break;
        }
        t = t + dt;
        if (maximumTime < t) {
          t = t - dt;
          dt = fmax(maximumTime - t, 0.0);
          t = maximumTime;
          reachedMaximumTime = 1;
          advection(T, dt, vx, vy, vz, &dAdvx, &dAdvy, &dAdvz);
          dxrw = dAdvx + divDx * dt + dBx * sqrt(dt);
          dyrw = dAdvy + divDy * dt + dBy * sqrt(dt);
          dzrw = dAdvz + divDz * dt + dBz * sqrt(dt);
        }
        switch (exitFace) {
          case 1:
          case 2:
            nx = nx - dxrw / dx;
            ny = ny + dyrw / dy;
            nz = nz + dzrw / dz;
            break;
          case 3:
          case 4:
            nx = nx + dxrw / dx;
            ny = ny - dyrw / dy;
            nz = nz + dzrw / dz;
            break;
          case 5:
          case 6:
            nx = nx + dxrw / dx;
            ny = ny + dyrw / dy;
            nz = nz - dzrw / dz;
            break;
        }
        if ((nx > 1.0) || (nx < 0.0) || (ny > 1.0) || (ny < 0.0) ||
            (nz > 1.0) || (nz < 0.0)) {
          x = xi; y = yi; z = zi;
          switch (exitFace) {
            case 1: case 2: vx = -vx; divDx = -divDx; dBx = -dBx; break;
            case 3: case 4: vy = -vy; divDy = -divDy; dBy = -dBy; break;
            case 5: case 6: vz = -vz; divDz = -divDz; dBz = -dBz; break;
          }
          exitFace = 0;
        } else {
          dt = dtcell;
          exitFace = 0;
          if ((nx == 0.0) || (nx == 1.0) || (ny == 0.0) || (ny == 1.0) ||
              (nz == 0.0) || (nz == 1.0)) {
            if (nx == 0.0) exitFace = 1;
            if (nx == 1.0) exitFace = 2;
            if (ny == 0.0) exitFace = 3;
            if (ny == 1.0) exitFace = 4;
            if (nz == 0.0) exitFace = 5;
            if (nz == 1.0) exitFace = 6;
          }
        }
      }
    }
  interface_done:
    if (reachedMaximumTime) {
      R->status = SC_MAXTIME;
      R->exitFace = exitFace;
      R->x = nx; R->y = ny; R->z = nz; R->t = t;
      return;
    }
    if (exitFace > 0) {
      R->conn = s->conn[exitFace - 1];
      R->status = R->conn < 0 ? SC_EXIT_INTERNAL : SC_EXIT_CELLFACE;
      R->exitFace = exitFace;
      R->zero_dt = dt == 0.0;
      R->x = nx; R->y = ny; R->z = nz; R->t = t;
      return;
    }
    x = nx;
    y = ny;
    z = nz;
  }
}

/* ------------------------------------------------------------------------ */
/* Cell and path level                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  int status, exitFace, next, zero_dt;
  double x, y, z, t;
} cellresult;

/* pr_ExecuteRandomWalkParticleTracking, TrackCell.f90:375-669 */
static void track_cell(tsc *T, cbuf *center, cbuf nb[18], double lx,
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
  } else {
    if (P->stop_weak_sinks) {
      if (center->sink != 0.0) {
        R->status = TC_WEAKSINK;
        return;
      } else if (!P->steady_state) stopIfNoSubCellExit = 0;
    }
  }
  if (!cell_has_exit(center, P->backward)) {
    if (P->steady_state) {
      R->status = TC_NOEXIT;
      return;
    } else if (P->backward && center->source != 0.0) {
      R->status = TC_NOEXIT;
      return;
    } else if (!P->backward && center->sink != 0.0) {
      R->status = TC_NOEXIT;
      return;
    }
  }
  fill_subcell(center, &T->s, P->backward);
  compute_corners(P, center, nb, &T->s, &T->C);
  subresult S;
  track_subcell(T, stopIfNoSubCellExit, lx, ly, lz, lt, maximumTime, &S,
                nsteps);
  switch (S.status) {
    case SC_UNDEFINED:
      R->status = TC_UNDEFINED;
      return;
    case SC_EXIT_CELLFACE:
      R->status = TC_EXIT_CELLFACE;
      R->exitFace = S.exitFace;
      R->next = S.conn;
      R->zero_dt = S.zero_dt;
      R->x = S.x; R->y = S.y; R->z = S.z; R->t = S.t;
      return;
    case SC_MAXTIME:
      R->status = TC_STOPTIME;
      R->exitFace = S.exitFace;
      R->x = S.x; R->y = S.y; R->z = S.z; R->t = S.t;
      return;
    case SC_INACTIVE:
      R->status = TC_INACTIVE;
      return;
    case SC_NOEXIT:
      R->status = TC_NOEXIT;
      return;
    default: /* internal faces do not occur on structured grids */
      R->status = TC_UNDEFINED;
      return;
  }
}

/* pr_RWPTrackPath (ParticleTrackingEngine.f90:1051-1254) mapped to the
 * particle status of MPathRW.F90:1337-1403.  Returns the new status; the
 * location written back is the last pathline point the reference records
 * (its pCoordLast): the exit point for a boundary exit, the final point at
 * the stopping time, the cell-entry location for statuses 3-7, and for an
 * undefined result the last point recorded before it. */
static int track_path(const orc_problem *P, tsc *T, cbuf *center,
                      cbuf nb[18], int *cell, double *lx, double *ly,
                      double *lz, double *t, double t_stop,
                      long long *nsteps) {
  int loc_cell = *cell, stalls = 0;
  double x = *lx, y = *ly, z = *lz, tt = *t;
  fill_cell(P, loc_cell, center);
  fill_neighbors(P, center, nb);
  for (;;) {
    if (loc_cell != center->cell) {
      fill_cell(P, loc_cell, center);
      fill_neighbors(P, center, nb);
    }
    cellresult R;
    track_cell(T, center, nb, x, y, z, tt, t_stop, &R, nsteps);
    switch (R.status) {
      case TC_EXIT_CELLFACE: {
        *cell = center->cell;
        *lx = R.x; *ly = R.y; *lz = R.z; *t = R.t;
        stalls = R.zero_dt ? stalls + 1 : 0;
        if (stalls > ORC_MAX_STALLS) return 9;
        int next = R.next;
        if (next > 0 && ibound_of(P, next) != 0) {
          /* ConvertFromNeighbor, ModflowRectangularGrid.f90:366-477: the
           * face coordinate must sit on the shared face and flips */
          double fx = R.x, fy = R.y, fz = R.z;
          int bad = 0;
          switch (R.exitFace) {
            case 1: bad = fx != 0.0; fx = 1.0; break;
            case 2: bad = fx != 1.0; fx = 0.0; break;
            case 3: bad = fy != 0.0; fy = 1.0; break;
            case 4: bad = fy != 1.0; fy = 0.0; break;
            case 5: bad = fz != 0.0; fz = 1.0; break;
            case 6: bad = fz != 1.0; fz = 0.0; break;
            default: bad = 1;
          }
          if (bad) return 9; /* the reference has no defined continuation */
          loc_cell = next;
          x = fx; y = fy; z = fz;
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
        return R.status == TC_NOEXIT     ? 5
               : R.status == TC_STOPZONE ? 6
               : R.status == TC_WEAKSINK ? 3
               : R.status == TC_WEAKSOURCE ? 4
                                         : 7;
      default: /* Undefined and InvalidLocation */
        return 9;
    }
  }
}

int orc_advance(const orc_problem *P, long n, int *cell, double *lx,
                double *ly, double *lz, double *t, int *status,
                uint64_t *rng, double t_engine, double t_stop,
                long long *substeps, int nthreads) {
  if (!P || n < 0) return -1;
  int nd = P->dim_mask[0] + P->dim_mask[1] + P->dim_mask[2];
  if (nd <= 0) return -2;
  int dims[3] = {0, 0, 0}, c = 0;
  for (int d = 0; d < 3; ++d)
    if (P->dim_mask[d]) dims[c++] = d + 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    cbuf center, *nb = (cbuf *)malloc(18 * sizeof(cbuf));
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
      /* release, MPathRW.F90:1256-1302 (no drape) */
      if (status[p] == 0) {
        if (t[p] < t_engine) status[p] = 8;
        else if (t[p] <= t_stop) {
          status[p] = 1;
          if (ibound_of(P, cell[p]) == 0) status[p] = 7;
        }
      }
      if (status[p] == 7) {
        cbuf b;
        fill_cell(P, cell[p], &b);
        verify_dry(&b);
        if (b.pdry) status[p] = 1;
      }
      if (status[p] != 1) continue;
      T.rng = rng + 2 * p;
      long long ns = 0;
      status[p] = track_path(P, &T, &center, nb, &cell[p], &lx[p], &ly[p],
                             &lz[p], &t[p], t_stop, &ns);
      if (substeps) substeps[p] += ns;
    }
    free(nb);
  }
  return 0;
}

/* Per-cell quantities the device tables hold, for table parity checks.
 * Layout per cell (ORC_CELL_FIELDS doubles):
 *  [0..5]  VX1 VX2 VY1 VY2 VZ1 VZ2   [6..8] DX DY DZ   [9] porosity
 *  [10] retardation  [11] dtcell  [12..43] corner q (8 corners x 4)
 *  [44..51] corner porosity  [52..99] Dxxx Dyyy Dzzz Dxyx Dxzx Dyzy
 *  (6 blocks of 8; Dxyy=Dxyx, Dxzz=Dxzx, Dyzz=Dyzy)  [100] dry
 *  [101] sub-cell has exit face  [102] source flow  [103] sink flow (after
 *  boundary flows on faces with an active neighbour were added) */
int orc_cell_tables(const orc_problem *P, double *out) {
  int nc = ncells(P);
  int nd = P->dim_mask[0] + P->dim_mask[1] + P->dim_mask[2];
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    cbuf center, *nb = (cbuf *)malloc(18 * sizeof(cbuf));
    tsc T;
    memset(&T, 0, sizeof T);
    T.P = P;
    T.ndim = nd;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int c = 1; c <= nc; ++c) {
      double *o = out + (size_t)(c - 1) * ORC_CELL_FIELDS;
      fill_cell(P, c, &center);
      fill_neighbors(P, &center, nb);
      fill_subcell(&center, &T.s, P->backward);
      compute_corners(P, &center, nb, &T.s, &T.C);
      const scbuf *s = &T.s;
      o[0] = s->VX1; o[1] = s->VX2; o[2] = s->VY1; o[3] = s->VY2;
      o[4] = s->VZ1; o[5] = s->VZ2; o[6] = s->DX; o[7] = s->DY; o[8] = s->DZ;
      o[9] = s->porosity; o[10] = s->retardation; o[11] = rw_time_step(&T);
      for (int ci = 0; ci < 8; ++ci)
        for (int m = 0; m < 4; ++m) o[12 + 4 * ci + m] = T.C.q[ci][m];
      for (int ci = 0; ci < 8; ++ci) o[44 + ci] = T.C.phi[ci];
      for (int ci = 0; ci < 8; ++ci) {
        o[52 + ci] = T.C.Dxxx[ci];
        o[60 + ci] = T.C.Dyyy[ci];
        o[68 + ci] = T.C.Dzzz[ci];
        o[76 + ci] = T.C.Dxyx[ci];
        o[84 + ci] = T.C.Dxzx[ci];
        o[92 + ci] = T.C.Dyzy[ci];
      }
      o[100] = s->dry;
      o[101] = subcell_has_exit(s);
      o[102] = center.source;
      o[103] = center.sink;
    }
    free(nb);
  }
  return 0;
}

/* ConvertToModelXYZ, ModflowRectangularGrid.f90:524-549 with the DIS
 * coordinates of RectangularGridDis.f90:295-361 */
void orc_global_xyz(const orc_problem *P, long n, const int *cell,
                    const double *lx, const double *ly, const double *lz,
                    double *gx, double *gy, double *gz) {
  double *X = (double *)malloc((P->ncol + 1) * sizeof(double));
  double *Y = (double *)malloc((P->nrow + 1) * sizeof(double));
  X[0] = 0.0;
  for (int j = 0; j < P->ncol; ++j) X[j + 1] = X[j] + P->delr[j];
  Y[0] = 0.0;
  for (int i = 0; i < P->nrow; ++i) Y[0] = Y[0] + P->delc[i];
  for (int i = 0; i < P->nrow; ++i) Y[i + 1] = Y[i] - P->delc[i];
  for (long p = 0; p < n; ++p) {
    int k, i, j, c = cell[p];
    cell_krc(P, c, &k, &i, &j);
    double cx = (X[j - 1] + X[j]) / 2.0, cy = (Y[i - 1] + Y[i]) / 2.0;
    double left = cx - (P->delr[j - 1] / 2.0);
    double right = cx + (P->delr[j - 1] / 2.0);
    double front = cy - (P->delc[i - 1] / 2.0);
    double back = cy + (P->delc[i - 1] / 2.0);
    double bottom = P->bot[c - 1], top = P->top[c - 1];
    double g = ((1.0 - lx[p]) * left) + (lx[p] * right);
    if (g < left) g = left;
    if (g > right) g = right;
    gx[p] = g;
    g = ((1.0 - ly[p]) * front) + (ly[p] * back);
    if (g < front) g = front;
    if (g > back) g = back;
    gy[p] = g;
    g = ((1.0 - lz[p]) * bottom) + (lz[p] * top);
    if (g < bottom) g = bottom;
    if (g > top) g = top;
    gz[p] = g;
  }
  free(X);
  free(Y);
}

/* prComputeCounts / prComputeCountsWeighted, Histogram.F90:213-350
 * (the weighted form here accumulates raw weights; the effective-mass
 * normalisation is applied by the caller) */
void orc_histogram(long n, const double *x, const double *y, const double *z,
                   const double *w, const double origin[3],
                   const double binsize[3], const int nbins[3],
                   double *counts) {
  const double *pts[3] = {x, y, z};
  size_t nb = (size_t)nbins[0] * nbins[1] * nbins[2];
  memset(counts, 0, nb * sizeof(double));
  for (long p = 0; p < n; ++p) {
    int gi[3] = {1, 1, 1}, inside = 1;
    for (int d = 0; d < 3; ++d) {
      /* the histogram checks every dimension with a bin size, a single
         bin included (Histogram.F90:121, 306-317) */
      if (!(binsize[d] > 0.0)) continue;
      gi[d] = (int)floor((pts[d][p] - origin[d]) / binsize[d]) + 1;
      if (gi[d] > nbins[d] || gi[d] <= 0) {
        inside = 0;
        break;
      }
    }
    if (!inside) continue;
    size_t idx = (size_t)(gi[0] - 1) +
                 (size_t)nbins[0] * ((gi[1] - 1) + (size_t)nbins[1] * (gi[2] - 1));
    counts[idx] = counts[idx] + (w ? w[p] : 1.0);
  }
}

/* Direct entry points to the exit-time solvers, for root checks. */
double orc_newton_axis(double dt, double v, double v1, double v2, double dx,
                       double dI, double divD, double dB) {
  return nr_time_step(dt, v, v1, v2, dx, dI, divD, dB);
}

double orc_quadratic_axis(double dt, double n, double p, double d, double v,
                          double divD, double dB, int *exit_face) {
  double dtaxis = 0.0, dtout = dt;
  int ef = 0;
  if (quad_axis(n, p, d, v, divD, dB, 1, 2, dt, &dtaxis, &ef, &dtout)) {
    *exit_face = ef;
    return 0.0;
  }
  *exit_face = 0;
  return dtaxis;
}


/* ------------------------------------------------------------------------ */
/* gpkde: one density pass with the Silverman bandwidth                      */
/* (GridProjectedKDE.F90 prComputeDensity :3598-4392 with                    */
/*  nOptimizationLoops = 0 and the brute-force kernel prSetKernelBrute       */
/*  :6224-6253; KernelMultiGaussian.f90 SetupMatrix :549-590, W :720-790,    */
/*  spans and reflection :209-390)                                           */
/* ------------------------------------------------------------------------ */

typedef struct {
  int m[3];      /* matrixPositiveShape */
  int s[3];      /* full shape 2m+1 */
  double *mat;   /* column-major (x fastest) */
} kmat;

static void kmat_setup(kmat *K, const double h[3], const double bin[3],
                       int ndim) {
  for (int d = 0; d < 3; ++d) {
    K->m[d] = bin[d] != 0.0 ? (int)ceil(3 * h[d] / bin[d]) : 0;
    K->s[d] = 2 * K->m[d] + 1;
  }
  size_t n = (size_t)K->s[0] * K->s[1] * K->s[2];
  K->mat = (double *)malloc(n * sizeof(double));
  if (h[0] == 0.0 && h[1] == 0.0 && h[2] == 0.0) {
    memset(K->mat, 0, n * sizeof(double));
    return;
  }
  double hd[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d)
    if (bin[d] != 0.0) hd[d] = h[d] / bin[d];
  const double sqrt2 = sqrt(2.0);
  /* zero-positive octant: 0.5^nDim times the erf differences per active
   * dimension, in dimension order (the (i +- 0.5) are exact) */
  int mx = K->m[0], my = K->m[1], mz = K->m[2];
  double *zp = (double *)malloc((size_t)(mx + 1) * (my + 1) * (mz + 1) *
                                sizeof(double));
  for (int k = 0; k <= mz; ++k)
    for (int j = 0; j <= my; ++j)
      for (int i = 0; i <= mx; ++i) {
        double v = pow(0.5, (double)ndim);
        int idx[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
          if (hd[d] <= 0.0) continue;
          double a = (double)((float)idx[d] + 0.5f);
          double b = (double)((float)idx[d] - 0.5f);
          v = v * (erf(a / (hd[d] * sqrt2)) - erf(b / (hd[d] * sqrt2)));
        }
        zp[(size_t)i + (mx + 1) * ((size_t)j + (my + 1) * k)] = v;
      }
  /* unfold to the full matrix (prUnfoldZeroPositiveMatrix) */
  for (int k = 0; k < K->s[2]; ++k)
    for (int j = 0; j < K->s[1]; ++j)
      for (int i = 0; i < K->s[0]; ++i) {
        int a = abs(i - mx), b = abs(j - my), c = abs(k - mz);
        K->mat[(size_t)i + K->s[0] * ((size_t)j + K->s[1] * k)] =
            zp[(size_t)a + (mx + 1) * ((size_t)b + (my + 1) * c)];
      }
  free(zp);
  double sum = 0.0;
  for (size_t q = 0; q < n; ++q) sum = sum + K->mat[q];
  if (sum != 0.0)
    for (size_t q = 0; q < n; ++q) K->mat[q] = K->mat[q] / sum;
}

/* prKernelReflection along one axis of the kernel matrix: kernel indices
 * [lo_out] fold back into the grid (1-based boundLoc semantics). */
static void kmat_fold(kmat *K, int axis, int lower, int boundLoc) {
  int s0 = K->s[0], s1 = K->s[1], s2 = K->s[2];
  int len = K->s[axis];
#define AT(i, j, k) K->mat[(size_t)(i) + s0 * ((size_t)(j) + (size_t)s1 * (k))]
  (void)s2;
  /* as 1-D lines along `axis` */
  int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
  int n1 = K->s[o1], n2 = K->s[o2];
  double *line = (double *)malloc(len * sizeof(double));
  for (int b = 0; b < n2; ++b)
    for (int a = 0; a < n1; ++a) {
      int idx[3];
      for (int q = 0; q < len; ++q) {
        idx[axis] = q; idx[o1] = a; idx[o2] = b;
        line[q] = AT(idx[0], idx[1], idx[2]);
      }
      if (lower) {
        int L = boundLoc; /* kernel 1..L lie outside; fold onto L+1..2L */
        for (int t = 0; t < L; ++t) line[L + t] = line[L + t] + line[L - 1 - t];
        for (int t = 0; t < L; ++t) line[t] = 0;
      } else {
        int lenb = len - boundLoc + 1; /* kernel boundLoc..len outside */
        for (int t = 0; t < lenb; ++t)
          line[boundLoc - 1 - lenb + t] =
              line[boundLoc - 1 - lenb + t] + line[len - 1 - t];
        for (int t = boundLoc - 1; t < len; ++t) line[t] = 0;
      }
      for (int q = 0; q < len; ++q) {
        idx[axis] = q; idx[o1] = a; idx[o2] = b;
        AT(idx[0], idx[1], idx[2]) = line[q];
      }
    }
  free(line);
#undef AT
}

int orc_kde_density(long n, const double *x, const double *y,
                    const double *z, const double *w, const double origin[3],
                    const double binsize[3], const int nbins[3],
                    double *density, double *counts, double hout[3]) {
  return orc_kde_density_sel(n, x, y, z, w, origin, binsize, nbins, 0, 1.0,
                             density, counts, hout);
}

int orc_kde_density_sel(long n, const double *x, const double *y,
                        const double *z, const double *w,
                        const double origin[3], const double binsize[3],
                        const int nbins[3], int selection, double factor,
                        double *density, double *counts, double hout[3]) {
  const double *pts[3] = {x, y, z};
  int dmask[3], ndim = 0;
  for (int d = 0; d < 3; ++d) {
    dmask[d] = nbins[d] > 1;
    ndim += dmask[d];
  }
  size_t nb = (size_t)nbins[0] * nbins[1] * nbins[2];
  /* histogram (prComputeCounts / prComputeCountsWeighted, format 0) */
  orc_histogram(n, x, y, z, w, origin, binsize, nbins, counts);
  double binVolume = 1.0;
  int nbd = 0;
  for (int d = 0; d < 3; ++d)
    if (binsize[d] > 0.0) {
      binVolume = binVolume * binsize[d];
      ++nbd;
    }
  /* binDistance over the dimensions with a bin size (Histogram.F90:155) */
  const double binDistance = pow(binVolume, 1.0 / (double)(nbd > 0 ? nbd : 1));
  double effMass = 1.0;
  long nPoints;
  if (w) {
    double tm = 0.0, tm2 = 0.0;
    for (long p = 0; p < n; ++p) tm = tm + w[p];
    for (long p = 0; p < n; ++p) tm2 = tm2 + w[p] * w[p];
    double neff = tm * tm / tm2;
    effMass = tm / neff;
    for (size_t q = 0; q < nb; ++q) counts[q] = counts[q] / effMass;
    nPoints = n;
  } else {
    double tm = 0.0;
    for (size_t q = 0; q < nb; ++q) tm = tm + counts[q];
    nPoints = (long)tm;
  }
  /* Silverman bandwidth (:4034-4054) over all points passed */
  double mean[3], sd[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    double s = 0.0;
    for (long p = 0; p < n; ++p) s = s + pts[d][p];
    mean[d] = s / (double)n;
  }
  for (int d = 0; d < 3; ++d) {
    if (!dmask[d]) continue;
    double s = 0.0;
    for (long p = 0; p < n; ++p) {
      double e = pts[d][p] - mean[d];
      s = s + e * e;
    }
    sd[d] = sqrt(s / (double)n);
  }
  double sscale = 1.0;
  for (int d = 0; d < 3; ++d)
    if (dmask[d]) sscale = sscale * sd[d];
  sscale = pow(sscale, 1.0 / (double)ndim);
  double hs = sscale * pow(4.0 / ((ndim + 2.0) * (double)nPoints),
                           1.0 / (ndim + 4.0));
  if (sscale == 0.0)
    hs = 2.0 * binDistance;
  else if (factor > 0.0 && factor != 1.0)
    hs = hs * factor;
  if (selection == 1) hs = factor * binDistance;
  double h[3];
  for (int d = 0; d < 3; ++d) h[d] = dmask[d] ? hs : 0.0;
  if (hout)
    for (int d = 0; d < 3; ++d) hout[d] = h[d];
  memset(density, 0, nb * sizeof(double));
  /* scatter of every active bin's reflected kernel (:4623-4661) */
  for (int k = 0; k < nbins[2]; ++k)
    for (int j = 0; j < nbins[1]; ++j)
      for (int i = 0; i < nbins[0]; ++i) {
        size_t b = (size_t)i + nbins[0] * ((size_t)j + (size_t)nbins[1] * k);
        double c = counts[b];
        if (c == 0.0) continue;
        kmat K;
        kmat_setup(&K, h, binsize, ndim);
        int id[3] = {i + 1, j + 1, k + 1}, gs[3], ks[3];
        for (int d = 0; d < 3; ++d) {
          gs[d] = id[d] - K.m[d] < 1 ? 1 : id[d] - K.m[d];
          ks[d] = gs[d] + K.m[d] - id[d] + 1;
        }
        for (int d = 0; d < 3; ++d) {
          if (!dmask[d]) continue;
          if (id[d] - K.m[d] < 1) kmat_fold(&K, d, 1, 1 - id[d] + K.m[d]);
          if (id[d] + K.m[d] > nbins[d])
            kmat_fold(&K, d, 0,
                      2 * K.m[d] + 1 - (id[d] + K.m[d] - nbins[d]) + 1);
        }
        int ge[3];
        for (int d = 0; d < 3; ++d)
          ge[d] = id[d] + K.m[d] > nbins[d] ? nbins[d] : id[d] + K.m[d];
        for (int gk = gs[2]; gk <= ge[2]; ++gk)
          for (int gj = gs[1]; gj <= ge[1]; ++gj)
            for (int gi = gs[0]; gi <= ge[0]; ++gi) {
              int mi = ks[0] + gi - gs[0] - 1, mj = ks[1] + gj - gs[1] - 1,
                  mk = ks[2] + gk - gs[2] - 1;
              size_t g = (size_t)(gi - 1) +
                         nbins[0] * ((size_t)(gj - 1) + (size_t)nbins[1] * (gk - 1));
              density[g] = density[g] +
                           c * K.mat[(size_t)mi + K.s[0] * ((size_t)mj + (size_t)K.s[1] * mk)];
            }
        free(K.mat);
      }
  for (size_t q = 0; q < nb; ++q) density[q] = density[q] / binVolume;
  if (w) {
    for (size_t q = 0; q < nb; ++q) {
      counts[q] = counts[q] * effMass;
      density[q] = density[q] * effMass;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Unstructured grids (shares the static sub-cell walk and corner formulas) */
/* ------------------------------------------------------------------------ */
#ifndef ORC_DUAL_BUILD
#include "rwpt_oracle_usg.c"
#endif
