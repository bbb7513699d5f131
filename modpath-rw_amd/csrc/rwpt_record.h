// The per-cell (or per-sub-cell) random-walk record from the cell's own
// face flows and the face flows, face areas, volumes and porosities of its
// 18-cell stencil: corner discharge, corner porosity, discharge products,
// corner dispersion and their derivative terms, corner velocities, the RW
// time step, face pore velocities and flags (TrackSubCell.F90:3667-4363,
// 2557-2691, 3117-3303, 4432-4543, 1376-1458; ModpathCellData.f90:2557-2564).
// Shared by the structured table build (rwpt_cells.hip) and the
// unstructured one (rwpt_usg.hip); both translation units are built with
// -ffp-contract=off and keep the reference's operand order.
#pragma once

#include "rwpt_internal.h"

namespace rwpt {
namespace {

// Corner component and porosity index tables (1-based slots, 4th entry of
// the component tables is the face), TrackSubCell.F90:4142-4224, 4329-4363.
__constant__ int kCx[8][4] = {{7, 8, 13, 1},   {7, 8, 13, 2},   {10, 11, 13, 1},
                              {10, 11, 13, 2}, {7, 9, 16, 1},   {7, 9, 16, 2},
                              {10, 12, 16, 1}, {10, 12, 16, 2}};
__constant__ int kCy[8][4] = {{1, 13, 14, 3}, {4, 13, 15, 3}, {1, 13, 14, 4},
                              {4, 13, 15, 4}, {1, 16, 17, 3}, {4, 16, 18, 3},
                              {1, 16, 17, 4}, {4, 16, 18, 4}};
__constant__ int kCz[8][4] = {{1, 2, 7, 5},  {4, 5, 7, 5}, {1, 3, 10, 5},
                              {4, 6, 10, 5}, {1, 2, 7, 6}, {4, 5, 7, 6},
                              {1, 3, 10, 6}, {4, 6, 10, 6}};
__constant__ int kCp[8][6] = {{1, 2, 7, 8, 13, 14},   {4, 5, 7, 8, 13, 15},
                              {1, 3, 10, 11, 13, 14}, {4, 6, 10, 12, 16, 18},
                              {1, 2, 7, 9, 16, 17},   {4, 5, 7, 9, 16, 18},
                              {1, 3, 10, 12, 16, 17}, {4, 6, 10, 11, 13, 15}};

struct TableParams {
  int32_t disp_model, backward, time_step_kind;
  double ts0, ts1;
};

// 1/d when d is a power of two (then exact), else 0.
__device__ __forceinline__ double pow2_recip(double d) {
  int e = 0;
  const double m = frexp(d, &e);
  if (m != 0.5 || 1 - e < -1022 || 1 - e > 1023) return 0.0;
  return ldexp(1.0, 1 - e);
}

// What a record build computes beside the record, for the checking copies
// of the tables.
struct RecordDbg {
  double q[8][4], pc[8], Dc[6][8], dtc;
  bool sub_exit;
};

// Fills rec's float fields and the sub-cell-exit and time-step flags (the
// other integer fields are the caller's).  Qc: the cell's own face flows;
// nQ/nA/nV/nP: its stencil (pr_FillNeighborSubCellVariables).
__device__ __forceinline__ void record_from_stencil(
    double *rec, const TableParams &tp, double DX, double DY, double DZ, double phi,
    double R, const double Qc[6], double aLH, double aTH, double aLV, double aTV,
    double dm, bool uniform_porosity, const double (*nQ)[6], const double (*nA)[3],
    const double *nV, const double *nP, RecordDbg &dbg) {
  // face pore velocities (FillMassSubCellDataBuffer, ModpathCellData.f90:2557-2564)
  const double sign = tp.backward ? -1.0 : 1.0;
  const double VX1 = sign * Qc[0] / DY / DZ / phi / R;
  const double VX2 = sign * Qc[1] / DY / DZ / phi / R;
  const double VY1 = sign * Qc[2] / DX / DZ / phi / R;
  const double VY2 = sign * Qc[3] / DX / DZ / phi / R;
  const double VZ1 = sign * Qc[4] / DX / DY / phi / R;
  const double VZ2 = sign * Qc[5] / DX / DY / phi / R;
  // corner discharge (pr_ComputeCornerDischarge, GetInterpolatedCornerDischarge)
  double q[8][4];
  for (int dir = 0; dir < 3; ++dir) {
    for (int ci = 0; ci < 8; ++ci) {
      const int *ix = dir == 0 ? kCx[ci] : (dir == 1 ? kCy[ci] : kCz[ci]);
      double sumA = dir == 0 ? DY * DZ : (dir == 1 ? DX * DZ : DX * DY);
      double sumQ = Qc[ix[3] - 1];
      for (int m = 0; m < 3; ++m) {
        sumA = sumA + nA[ix[m] - 1][dir];
        sumQ = sumQ + nQ[ix[m] - 1][ix[3] - 1];
      }
      q[ci][dir] = sumQ / sumA;
    }
  }
  for (int ci = 0; ci < 8; ++ci)
    q[ci][3] = sqrt(q[ci][0] * q[ci][0] + q[ci][1] * q[ci][1] +
                    q[ci][2] * q[ci][2]);
  // corner porosity (Uniform / Interpolated)
  double pc[8];
  for (int ci = 0; ci < 8; ++ci) {
    if (uniform_porosity) {
      pc[ci] = phi;
    } else {
      double sV = DX * DY * DZ;
      double sW = phi * sV;
      for (int m = 0; m < 6; ++m) {
        const int n = kCp[ci][m] - 1;
        sV = sV + nV[n];
        sW = sW + nV[n] * nP[n];
      }
      pc[ci] = sW / sV;
    }
  }
  // discharge products and corner dispersion
  double Dc[6][8];  // Dxxx Dyyy Dzzz Dxyx Dxzx Dyzy
  for (int ci = 0; ci < 8; ++ci) {
    double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (q[ci][3] > 0.0) {
      p[0] = q[ci][0] * q[ci][0];
      p[1] = q[ci][0] * q[ci][1];
      p[2] = q[ci][0] * q[ci][2];
      p[3] = q[ci][1] * q[ci][1];
      p[4] = q[ci][1] * q[ci][2];
      p[5] = q[ci][2] * q[ci][2];
      p[6] = p[0] + p[3];
      if (p[6] > 0.0) p[7] = p[5] / p[6];
      for (int m = 0; m < 7; ++m) p[m] = p[m] / q[ci][3];
    }
    if (tp.disp_model == 0) {
      // corners 010..111 take |q| of corner 100 (TrackSubCell.F90:2622-2677)
      const double qn = ci == 0 ? q[0][3] : q[1][3];
      const double aLmaT = aLH - aTH;
      const double aTQ = aTH * qn + pc[ci] * dm;
      Dc[0][ci] = aTQ + aLmaT * p[0];
      Dc[1][ci] = aTQ + aLmaT * p[3];
      Dc[2][ci] = aTQ + aLmaT * p[5];
      Dc[3][ci] = aLmaT * p[1];
      Dc[4][ci] = aLmaT * p[2];
      Dc[5][ci] = aLmaT * p[4];
    } else {
      const double pD = pc[ci] * dm;
      double cos2 = 0.0;
      if (q[ci][3] > 0.0) cos2 = q[ci][2] * q[ci][2] / q[ci][3];
      const double aL = aLH + cos2 * (aLV - aLH);
      const double aT = aTV + cos2 * (aTH - aTV);
      const double aLmaT = aL - aT;
      Dc[0][ci] = pD + (aL + aT * p[7]) * p[0] + aTH * p[3] * (1.0 + p[7]);
      Dc[1][ci] = pD + aTH * p[0] * (1.0 + p[7]) + p[3] * (aL + aT * p[7]);
      Dc[2][ci] = pD + aT * p[6] + aL * p[5];
      Dc[3][ci] = (aL - aTH * (1.0 + p[7]) + aT * p[7]) * p[0];
      Dc[4][ci] = aLmaT * p[2];
      Dc[5][ci] = aLmaT * p[4];
    }
  }
  // derivative terms: X(a) uses corner pairs along x, etc.
  // term order: Xxx Yyy Zzz Xxy Xxz Yxy Yyz Zxz Zyz
  const int src[9] = {0, 1, 2, 3, 4, 3, 5, 4, 5};
  const int axis[9] = {0, 1, 2, 0, 0, 1, 1, 2, 2};
  for (int tix = 0; tix < 9; ++tix) {
    const double *v = Dc[src[tix]];
    double a, b, c, dd;
    if (axis[tix] == 0) {
      a = (v[1] - v[0]) / DX;   // 100-000
      b = (v[3] - v[2]) / DX;   // 110-010
      c = (v[5] - v[4]) / DX;   // 101-001
      dd = (v[7] - v[6]) / DX;  // 111-011
    } else if (axis[tix] == 1) {
      a = (v[2] - v[0]) / DY;   // 010-000
      b = (v[3] - v[1]) / DY;   // 110-100
      c = (v[6] - v[4]) / DY;   // 011-001
      dd = (v[7] - v[5]) / DY;  // 111-101
    } else {
      a = (v[4] - v[0]) / DZ;   // 001-000
      b = (v[5] - v[1]) / DZ;   // 101-100
      c = (v[6] - v[2]) / DZ;   // 011-010
      dd = (v[7] - v[3]) / DZ;  // 111-110
    }
    rec[F_DD + 4 * tix + 0] = a;
    rec[F_DD + 4 * tix + 1] = b;
    rec[F_DD + 4 * tix + 2] = c;
    rec[F_DD + 4 * tix + 3] = dd;
  }
  for (int comp = 0; comp < 3; ++comp)
    for (int ci = 0; ci < 8; ++ci)
      rec[F_VC + 8 * comp + ci] = q[ci][comp] / pc[ci];
  // cell time step (pr_ComputeRandomWalkTimeStep, TrackSubCell.F90:1376-1458)
  const double mx = fmax(fabs(VX1), fabs(VX2));
  const double my = fmax(fabs(VY1), fabs(VY2));
  const double mz = fmax(fabs(VZ1), fabs(VZ2));
  double dtc;
  const double adv = tp.ts0 / (mx / DX + my / DY + mz / DZ);
  const double dis = tp.ts1 / (aLH * mx / (DX * DX) + aTH * my / (DY * DY) +
                               aTH * mz / (DZ * DZ));
  switch (tp.time_step_kind) {
    case 1: dtc = adv; break;
    case 2: dtc = dis; break;
    case 3: {
      double m = 1.7976931348623157e308;
      if (adv > 0.0 && adv < m) m = adv;
      if (dis > 0.0 && dis < m) m = dis;
      dtc = m;
      break;
    }
    default: dtc = tp.ts0;
  }
  rec[F_VX1] = VX1; rec[F_VX2] = VX2; rec[F_VY1] = VY1;
  rec[F_VY2] = VY2; rec[F_VZ1] = VZ1; rec[F_VZ2] = VZ2;
  rec[F_DX] = DX; rec[F_DY] = DY; rec[F_DZ] = DZ;
  rec[F_PHI] = phi; rec[F_R] = R; rec[F_DTCELL] = dtc;
  rec[F_DVDX] = (VX2 - VX1) / DX;
  rec[F_DVDY] = (VY2 - VY1) / DY;
  rec[F_DVDZ] = (VZ2 - VZ1) / DZ;
  rec[F_ALH] = aLH; rec[F_ATH] = aTH; rec[F_ALV] = aLV; rec[F_ATV] = aTV;
  rec[F_DMEFF] = dm;
  const bool sub_exit = (VX1 < 0.0) || (VX2 > 0.0) || (VY1 < 0.0) ||
                        (VY2 > 0.0) || (VZ1 < 0.0) || (VZ2 > 0.0);
  int32_t *fl = reinterpret_cast<int32_t *>(rec + F_FLAGS);
  fl[1] = (fl[1] & ~(CF_SUBCELL_EXIT | CF_DT_NONPOS)) | (sub_exit ? CF_SUBCELL_EXIT : 0) |
          (dtc <= 0.0 ? CF_DT_NONPOS : 0);
  rec[F_RDX] = pow2_recip(DX);
  rec[F_RDY] = pow2_recip(DY);
  rec[F_RDZ] = pow2_recip(DZ);
  rec[F_RR] = pow2_recip(R);
  for (int pad = F_RR + 1; pad < CELL_DOUBLES; ++pad) rec[pad] = 0.0;
  for (int ci = 0; ci < 8; ++ci) {
    for (int m = 0; m < 4; ++m) dbg.q[ci][m] = q[ci][m];
    dbg.pc[ci] = pc[ci];
    for (int b6 = 0; b6 < 6; ++b6) dbg.Dc[b6][ci] = Dc[b6][ci];
  }
  dbg.dtc = dtc;
  dbg.sub_exit = sub_exit;
}

// The structured tables' checking copy (DEBUG_FIELDS doubles).
__device__ __forceinline__ void record_debug(const double *rec, const RecordDbg &dbg,
                                             double *o) {
  for (int f = 0; f < 6; ++f) o[f] = rec[F_VX1 + f];
  o[6] = rec[F_DX]; o[7] = rec[F_DY]; o[8] = rec[F_DZ]; o[9] = rec[F_PHI];
  o[10] = rec[F_R]; o[11] = dbg.dtc;
  for (int ci = 0; ci < 8; ++ci)
    for (int m = 0; m < 4; ++m) o[12 + 4 * ci + m] = dbg.q[ci][m];
  for (int ci = 0; ci < 8; ++ci) o[44 + ci] = dbg.pc[ci];
  for (int b6 = 0; b6 < 6; ++b6)
    for (int ci = 0; ci < 8; ++ci) o[52 + 8 * b6 + ci] = dbg.Dc[b6][ci];
  const int32_t *fl = reinterpret_cast<const int32_t *>(rec + F_FLAGS);
  o[100] = (fl[1] & CF_DRY) ? 1.0 : 0.0;
  o[101] = dbg.sub_exit ? 1.0 : 0.0;
  o[102] = rec[F_SOURCE];
  o[103] = rec[F_SINK];
}

}  // namespace
}  // namespace rwpt
