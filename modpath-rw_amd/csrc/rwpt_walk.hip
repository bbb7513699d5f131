// Random-walk advance kernel (plain FP64 form) and particle utilities.
//
// One thread carries one particle from its current location to t_stop,
// through as many cells as it crosses, entirely in registers; the state is
// read once and written once per launch (structure of arrays in HBM).  Per
// sub-step the thread reads its cell's CellRec (built once per flow step by
// rwpt_cells.hip) instead of re-assembling the cell and its 18 neighbours on
// every cell entry as pr_RWPTrackPath does.
//
// The sub-step logic is the reference's (TrackSubCell.F90:812-1373): the
// displacement, the jump limiter that reuses the same draw, the interface
// loop with the quadratic (EULERIAN) or Newton (EXPONENTIAL) exit-time
// solve, restarts and elastic rebound.  Floating-point expressions keep the
// reference operand order and the file is built with -ffp-contract=off, so
// with the EULERIAN solver every result is the double the reference forms.
#include "rwpt_internal.h"

namespace rwpt {
namespace {

// ---------------------------------------------------------------------------
// xoroshiro128+ as rng_par_zig.f90 defines it (shifts 24/16/37, jump
// polynomial 0xdf900294d8f554a5/0x170865df4b3201fc, mantissa fill to [0,1)).
// ---------------------------------------------------------------------------
struct Stream {
  uint64_t a, b;
};

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}

__device__ __forceinline__ double draw_uniform(Stream &s) {
  const uint64_t s1 = s.a, s2 = s.b ^ s.a;
  const uint64_t out = s.a + s.b;
  s.a = (rotl64(s1, 24) ^ s2) ^ (s2 << 16);
  s.b = rotl64(s2, 37);
  return __longlong_as_double(
             (long long)(0x3FF0000000000000ull | (out >> 12))) - 1.0;
}

// Irwin-Hall normal of pr_GenerateStandardNormalRandomZig: 12 uniforms
// summed in order, minus 6.
__device__ __forceinline__ double draw_normal(Stream &s) {
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < 12; ++m) acc = acc + draw_uniform(s);
  return acc - 6.0;
}

__device__ __forceinline__ double f32lit(float v) { return (double)v; }

// Guards against walks the reference never ends (not in the reference):
// sub-steps per particle and launch, and consecutive cell exits made with a
// zero time step (a particle on a face whose drift points into the face from
// both sides, with no random component across it, ping-pongs between the two
// cells forever in pr_RWPTrackPath).  Either reports status 9.
constexpr int64_t kMaxSteps = 20000000;
constexpr int kMaxStalls = 1000;

// 10d0**n for an integer n, evaluated as an integer power (exact products,
// reciprocal for n < 0).
__device__ __forceinline__ double pow10i(int n) {
  double p = 1.0;
  for (int i = 0; i < (n < 0 ? -n : n); ++i) p = p * 10.0;
  return n < 0 ? 1.0 / p : p;
}

// ---------------------------------------------------------------------------
// Interpolation on the cell record
// ---------------------------------------------------------------------------

// trilinear interpolation of 8 corner values (pr_Trilinear)
__device__ __forceinline__ double tri8(const double *__restrict__ v, double x,
                                       double y, double z) {
  const double v00 = (1.0 - x) * v[0] + x * v[1];
  const double v01 = (1.0 - x) * v[4] + x * v[5];
  const double v10 = (1.0 - x) * v[2] + x * v[3];
  const double v11 = (1.0 - x) * v[6] + x * v[7];
  const double v0 = (1.0 - y) * v00 + y * v10;
  const double v1 = (1.0 - y) * v01 + y * v11;
  return (1.0 - z) * v0 + z * v1;
}

// derivative term stored as the four edge differences (pr_TrilinearDerivative)
__device__ __forceinline__ double deriv4(const double *__restrict__ d, double u,
                                         double w) {
  const double v0 = (1.0 - u) * d[0] + u * d[1];
  const double v1 = (1.0 - u) * d[2] + u * d[3];
  return (1.0 - w) * v0 + w * v1;
}

struct StepTerms {
  double vx, vy, vz;        // interpolated face velocities
  double dAx, dAy, dAz;     // advection displacement
  double dBx, dBy, dBz;     // random displacement (per sqrt(dt))
  double ddx, ddy, ddz;     // dispersion divergence (drift correction)
};

__device__ __forceinline__ void advect(const double *__restrict__ rec,
                                       int kind, double dt, StepTerms &s) {
  if (kind == 1) {  // pr_AdvectionDisplacementExponential
    const double dvtol = 1.0e-10;
    const double gx = rec[F_DVDX], gy = rec[F_DVDY], gz = rec[F_DVDZ];
    s.dAx = fabs(gx) > dvtol ? s.vx * (exp(gx * dt) - 1.0) / gx : s.vx * dt;
    s.dAy = fabs(gy) > dvtol ? s.vy * (exp(gy * dt) - 1.0) / gy : s.vy * dt;
    s.dAz = fabs(gz) > dvtol ? s.vz * (exp(gz * dt) - 1.0) / gz : s.vz * dt;
  } else {          // pr_AdvectionDisplacementEulerian
    s.dAx = s.vx * dt;
    s.dAy = s.vy * dt;
    s.dAz = s.vz * dt;
  }
}

// pr_RWPTDisplacementsLinear / ...Axisymmetric for one sub-step.
__device__ __forceinline__ void displacements(const double *__restrict__ rec,
                                              const WalkParams &W, double x,
                                              double y, double z, double dt,
                                              Stream &rng, StepTerms &s) {
  s.vx = (1.0 - x) * rec[F_VX1] + x * rec[F_VX2];
  s.vy = (1.0 - y) * rec[F_VY1] + y * rec[F_VY2];
  s.vz = (1.0 - z) * rec[F_VZ1] + z * rec[F_VZ2];
  double aL, aT, aTH;
  if (W.disp_model == 1) {
    const double vn = sqrt(s.vx * s.vx + s.vy * s.vy + s.vz * s.vz);
    double c2 = 0.0;
    if (vn > 0.0) c2 = s.vz * s.vz / vn;
    aL = rec[F_ALH] + c2 * (rec[F_ALV] - rec[F_ALH]);
    aT = rec[F_ATV] + c2 * (rec[F_ATH] - rec[F_ATV]);
    aTH = rec[F_ATH];
  } else {
    aL = rec[F_ALH];
    aT = rec[F_ATH];
    aTH = aT;
  }
  // drift correction, pr_DispersionDivergence
  const double *dd = rec + F_DD;
  const double t0 = deriv4(dd + 0, y, z), t1 = deriv4(dd + 4, x, z);
  const double t2 = deriv4(dd + 8, x, y), t3 = deriv4(dd + 12, y, z);
  const double t4 = deriv4(dd + 16, y, z), t5 = deriv4(dd + 20, x, z);
  const double t6 = deriv4(dd + 24, x, z), t7 = deriv4(dd + 28, x, y);
  const double t8 = deriv4(dd + 32, x, y);
  const double phi = rec[F_PHI], R = rec[F_R];
  s.ddx = (t0 + t5 + t7) / phi / R;
  s.ddy = (t3 + t1 + t8) / phi / R;
  s.ddz = (t4 + t6 + t2) / phi / R;
  // random displacement, pr_DisplacementRandomDischarge{,2D,1D}
  const double dm = rec[F_DMEFF];
  const double *vc = rec + F_VC;
  if (W.ndim == 3) {
    const double bx = tri8(vc, x, y, z), by = tri8(vc + 8, x, y, z);
    const double bz = tri8(vc + 16, x, y, z);
    const double vn = sqrt(bx * bx + by * by + bz * bz);
    const double vxy = sqrt(bx * bx + by * by);
    double B11 = 0, B12 = 0, B13 = 0, B21 = 0, B22 = 0, B23 = 0, B31 = 0,
           B32 = 0;
    if (vn > 0.0) {
      const double sl = sqrt(2 * (aL * vn + dm) / R);
      const double st = sqrt(2 * (aT * vn + dm) / R);
      B11 = bx * sl / vn;
      B21 = by * sl / vn;
      B31 = bz * sl / vn;
      B32 = vxy * st / vn;
      if (vxy > 0.0) {
        const double sh = sqrt(2 * (aTH * vn + dm) / R);
        B12 = -(bx * bz * st / vn / vxy);
        B13 = -(by * sh / vxy);
        B22 = -(by * bz * st / vn / vxy);
        B23 = bx * sh / vxy;
      }
    }
    const double r1 = draw_normal(rng), r2 = draw_normal(rng);
    const double r3 = draw_normal(rng);
    s.dBx = B11 * r1 + B12 * r2 + B13 * r3;
    s.dBy = B21 * r1 + B22 * r2 + B23 * r3;
    s.dBz = B31 * r1 + B32 * r2;
  } else if (W.ndim == 2) {
    const int d1 = W.dims[0] - 1, d2 = W.dims[1] - 1;
    const double b1 = tri8(vc + 8 * d1, x, y, z);
    const double b2 = tri8(vc + 8 * d2, x, y, z);
    const double vn = sqrt(b1 * b1 + b2 * b2);
    double B11 = 0, B12 = 0, B21 = 0, B22 = 0;
    if (vn > 0.0) {
      const double sl = sqrt(2 * (aL * vn + dm) / R);
      const double st = sqrt(2 * (aTH * vn + dm) / R);
      B11 = b1 * sl / vn;
      B21 = b2 * sl / vn;
      B12 = -(b2 * st / vn);
      B22 = b1 * st / vn;
    }
    const double r1 = draw_normal(rng), r2 = draw_normal(rng);
    double dB[3] = {0.0, 0.0, 0.0};
    dB[d1] = B11 * r1 + B12 * r2;
    dB[d2] = B21 * r1 + B22 * r2;
    s.dBx = dB[0];
    s.dBy = dB[1];
    s.dBz = dB[2];
  } else {
    const int d1 = W.dims[0] - 1;
    const double b1 = tri8(vc + 8 * d1, x, y, z);
    const double vn = sqrt(b1 * b1);
    double B11 = 0.0;
    if (vn > 0.0) B11 = sqrt(2 * (aL * vn + dm) / R);
    const double r1 = draw_normal(rng);
    double dB[3] = {0.0, 0.0, 0.0};
    dB[d1] = B11 * r1;
    s.dBx = dB[0];
    s.dBy = dB[1];
    s.dBz = dB[2];
  }
  advect(rec, W.advection_kind, dt, s);
}

// ---------------------------------------------------------------------------
// Exit-time solvers
// ---------------------------------------------------------------------------

// One axis of pr_DetectExitFaceAndUpdateTimeStepQuadratic.  Returns true when
// the particle sits on the interface (the caller returns dt = 0 at once).
__device__ __forceinline__ bool quadratic_axis(double dI, double A, double B,
                                               double dtcur, double &dtaxis) {
  const double arg = B * B + 4 * dI * A;
  if ((arg >= 0.0) && (A != 0.0)) {
    const double sq = sqrt(arg);
    const double z1 = (-B + sq) / (2 * A);
    const double z2 = (-B - sq) / (2 * A);
    if (z1 > 0.0 || z2 > 0.0) {
      const double m = (z1 > 0.0 && z2 > 0.0) ? fmin(z1, z2)
                                              : (z1 > 0.0 ? z1 : z2);
      dtaxis = m * m;
      if (dtaxis == 0.0) return true;
    } else if (z1 == 0.0 || z2 == 0.0) {
      return true;
    }
  } else if (A == 0.0) {
    if (B != 0.0) {
      const double r = dI / B;
      dtaxis = r * r;
      if (dtaxis == 0.0) return true;
    }
  }
  if (dtaxis > dtcur) dtaxis = 0.0;
  return false;
}

struct NRTerms {
  double f0, f1, f2, f3;
};

// pr_NewtonRaphsonVariablesExponential (leaves the terms untouched at dt=0)
__device__ __forceinline__ void nr_terms(double dt, double v, double g,
                                         double dI, double divD, double dB,
                                         NRTerms &T) {
  if (dt == 0.0) return;
  const double e = exp(g * dt);
  const double dA = fabs(g) > 1.0e-10 ? v * (exp(g * dt) - 1.0) / g : v * dt;
  T.f0 = dA + divD * dt + dB * sqrt(dt) - dI;
  T.f1 = v * e + divD + 0.5 * dB / sqrt(dt);
  T.f2 = g * v * e - 0.25 * dB * pow(dt, -1.5);
  T.f3 = g * g * v * e + 0.375 * dB * pow(dt, -2.5);
}

__device__ __forceinline__ double conv(const NRTerms &T) {
  return fabs(T.f2 * T.f0 / (T.f1 * T.f1));
}
__device__ __forceinline__ double conv_der(const NRTerms &T) {
  const double g = T.f2 * T.f0 / (T.f1 * T.f1);
  const double gp = -2.0 * (T.f0 * (T.f2 * T.f2)) / (T.f1 * T.f1 * T.f1) +
                    (T.f1 * T.f2 + T.f0 * T.f3) / (T.f1 * T.f1);
  return g / fabs(g) * gp;
}

// pr_SetInitialGuess (TrackSubCell.F90:2162-2448): decade scan, gradient
// descent on the convergence function, then a 0.2..0.9 fraction scan.
__device__ double nr_guess(double dt, double v, double g, double dI,
                           double divD, double dB) {
  double dtnew = 0.0;
  if (dt == 0.0) return dtnew;
  NRTerms T{0.0, 0.0, 0.0, 0.0};
  const double big = f32lit(1e10f), gdtol = f32lit(1e-6f);
  double gmin = big, gp = big, mfrac = 0.0, frac = 0.0;
  int gc = 0;
  while ((gp > 1) && (gc < 7)) {
    gc = gc + 1;
    frac = pow10i(-8 + gc);
    nr_terms(dt * frac, v, g, dI, divD, dB, T);
    gp = conv(T);
    if (gp < gmin) {
      gmin = gp;
      mfrac = frac;
    }
  }
  double ofrac;
  if ((gp < 1) && (gc == 1)) ofrac = 1e-8;
  else if (gp < 1) ofrac = pow10i(-8 + gc - 1);
  else {
    frac = mfrac;
    ofrac = 0.5 * frac / 10.0;
  }
  dtnew = frac * dt;
  double dtold = ofrac * dt;
  nr_terms(dtnew, v, g, dI, divD, dB, T);
  gp = conv(T);
  if (gp < 0.5) return dtnew;
  double gnew = 10.0;
  gp = 10.0;
  int cc = 0, bc = 0, zc = 0, rc = 0, it = 0;
  while (it < 50) {
    it = it + 1;
    nr_terms(dtold, v, g, dI, divD, dB, T);
    const double gd_old = conv_der(T);
    nr_terms(dtnew, v, g, dI, divD, dB, T);
    const double gd_new = conv_der(T);
    gp = conv(T);
    double gam = 0.0;
    if ((gd_new - gd_old) > 0.0) {
      const double dg = gd_new - gd_old;
      gam = fabs((dtnew - dtold) * dg) / (dg * dg);
    }
    dtold = dtnew;
    dtnew = dtnew - gam * gd_new;
    if (isnan(dtnew)) {
      dtnew = 0.0;
      break;
    }
    if ((gp < 1) && (dtnew / dt > 1)) return dtold;
    else if (dtnew / dt > 1) dtnew = f32lit(0.99f) * dt;
    if ((gp < 1) && (dtnew < 0) && (dtold > 0)) return dtold;
    else if (dtnew < 0) {
      dtnew = f32lit(1e-8f) * dt;
      zc = zc + 1;
      if (zc > 1) {
        dtnew = 0.0;
        break;
      }
    }
    nr_terms(dtnew, v, g, dI, divD, dB, T);
    gnew = conv(T);
    if ((gnew < 1) && (dtnew > 0) && (dtnew / dt < 1)) {
      rc = rc + 1;
      if (rc > 1) return dtnew;
    } else {
      rc = 0;
    }
    if ((gp < 1) && (gnew > gp)) return dtold;
    if (it > 50) break;
    const double rel = fabs((dtold - dtnew) / dtold);
    if ((rel < gdtol) && (gp < 1)) {
      cc = cc + 1;
      if (cc > 1) return dtnew;
    } else {
      cc = 0;
    }
    if ((rel < gdtol) && (gp > 1)) {
      bc = bc + 1;
      if (bc > 1) break;
    } else {
      bc = 0;
    }
  }
  if (gp < 1) return dtnew;
  gp = big;
  gmin = big;
  mfrac = 0.0;
  gc = 1;
  while ((gp > 1) && (gc < 9)) {
    gc = gc + 1;
    frac = (double)(0.1f * (float)gc);
    nr_terms(dt * frac, v, g, dI, divD, dB, T);
    gp = conv(T);
    if (gp < gmin) {
      gmin = gp;
      mfrac = frac;
    }
  }
  if (gmin < 1) dtnew = dt * mfrac;
  if ((dtnew <= 0.0) && (fabs(v) > 0.0)) dtnew = fabs(dI / v);
  return dtnew;
}

// NewtonRaphsonTimeStepExponentialAdvection (TrackSubCell.F90:2451-2554)
__device__ double newton_axis(double dt, double v, double g, double dI,
                              double divD, double dB) {
  const double nrtol = f32lit(1e-6f);
  double d0 = nr_guess(dt, v, g, dI, divD, dB);
  if (d0 == 0.0) return 0.0;
  double err = 1e6;
  int it = 0;
  while ((fabs(err / d0) > nrtol) && (it < 50)) {
    it = it + 1;
    const double dA =
        fabs(g) > 1.0e-10 ? v * (exp(g * d0) - 1.0) / g : v * d0;
    const double f0 = dA + divD * d0 + dB * sqrt(d0) - dI;
    const double f1 = v * exp(g * d0) + divD + 0.5 * dB / sqrt(d0);
    err = -f0 / f1;
    d0 = d0 + err;
    if (d0 < 0.0) d0 = f32lit(-0.9f) * d0;
  }
  if (d0 > dt) return dt;
  if ((it == 50) && (fabs(err / d0) > nrtol)) return 0.0;
  return d0;
}

// ExitFaceAndUpdateTimeStep: finds the axis and time at which the
// displacement reaches the first crossed face; updates t, dt, exitFace
// exactly as the reference (including leaving exitFace untouched when no
// axis yields a time).
__device__ void exit_time(const double *__restrict__ rec, int kind, double x,
                          double y, double z, double nx, double ny, double nz,
                          const StepTerms &s, double &t, double &dt,
                          int &exitFace) {
  t = t - dt;
  const double pn[3] = {nx, ny, nz}, pp[3] = {x, y, z};
  const double dd[3] = {rec[F_DX], rec[F_DY], rec[F_DZ]};
  const double v[3] = {s.vx, s.vy, s.vz};
  const double dv[3] = {s.ddx, s.ddy, s.ddz};
  const double db[3] = {s.dBx, s.dBy, s.dBz};
  const double gr[3] = {rec[F_DVDX], rec[F_DVDY], rec[F_DVDZ]};
  double dtx[3] = {0.0, 0.0, 0.0};
  int fx[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if ((pn[a] > 1.0) || (pn[a] < 0.0)) {
      double dI;
      if (pn[a] > 1.0) {
        dI = dd[a] * (1.0 - pp[a]);
        fx[a] = 2 * a + 2;
      } else {
        dI = -dd[a] * pp[a];
        fx[a] = 2 * a + 1;
      }
      if (dI == 0.0) {
        exitFace = fx[a];
        dt = 0.0;
        return;
      }
      if (kind == 2) {
        if (quadratic_axis(dI, v[a] + dv[a], db[a], dt, dtx[a])) {
          exitFace = fx[a];
          dt = 0.0;
          return;
        }
      } else {
        dtx[a] = newton_axis(dt, v[a], gr[a], dI, dv[a], db[a]);
        if (dtx[a] > dt) dtx[a] = 0.0;
      }
    }
  }
  if (dtx[0] == 0.0 && dtx[1] == 0.0 && dtx[2] == 0.0) {
    t = t + dt;
    return;
  }
  int im = -1;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (dtx[a] > 0.0 && (im < 0 || dtx[a] < dtx[im])) im = a;
  dt = dtx[im];
  exitFace = fx[im];
  t = t + dt;
}

__device__ __forceinline__ bool outside(double x, double y, double z) {
  return (x > 1.0) || (x < 0.0) || (y > 1.0) || (y < 0.0) || (z > 1.0) ||
         (z < 0.0);
}

// outcomes of one cell visit (TrackSubCellResult status values)
enum : int { OUT_UNDEFINED = 0, OUT_EXIT = 2, OUT_MAXTIME = 3,
             OUT_NOEXIT = 4, OUT_INACTIVE = 7 };

// pr_ExecuteRandomWalkParticleTracking (TrackSubCell.F90:812-1373) for one
// cell visit starting at (ix, iy, iz, t0).  On OUT_EXIT/OUT_MAXTIME the
// final location is returned in (fx, fy, fz, ft) and the face in face.
__device__ int visit_cell(const double *__restrict__ rec, const WalkParams &W,
                          bool stopIfNoExit, double ix, double iy, double iz,
                          double t0, Stream &rng, int64_t &steps, double &fx,
                          double &fy, double &fz, double &ft, int &face,
                          bool &zero_dt) {
  const int32_t flags = reinterpret_cast<const int32_t *>(rec + F_FLAGS)[1];
  if (stopIfNoExit && !(flags & CF_SUBCELL_EXIT)) return OUT_NOEXIT;
  if (flags & CF_DRY) return OUT_INACTIVE;
  const double dtcell = rec[F_DTCELL];
  if (dtcell <= 0.0) return OUT_UNDEFINED;
  const double maxRelJump = 0.5, dtRed = f32lit(0.1f);
  const double DX = rec[F_DX], DY = rec[F_DY], DZ = rec[F_DZ];
  const int32_t *mb = reinterpret_cast<const int32_t *>(rec + F_MASSBND);
  double x = ix, y = iy, z = iz, nx = ix, ny = iy, nz = iz;
  double t = t0, dt;
  int exitFace = 0, restarts = 0;
  bool atMax = false;
  StepTerms s;
  for (;;) {
    // guard against a walk that never ends (not in the reference): the
    // particle is reported undefined instead of holding the GPU
    if (steps >= kMaxSteps) return OUT_UNDEFINED;
    steps += 1;
    dt = dtcell;
    t = t + dt;
    if (W.t_stop < t) {
      t = t - dt;
      dt = fmax(W.t_stop - t, 0.0);
      t = W.t_stop;
      atMax = true;
    }
    displacements(rec, W, x, y, z, dt, rng, s);
    double sq = sqrt(dt);
    double dxr = s.dAx + s.ddx * dt + s.dBx * sq;
    double dyr = s.dAy + s.ddy * dt + s.dBy * sq;
    double dzr = s.dAz + s.ddz * dt + s.dBz * sq;
    // jump limiter: same draw, dt cut by 0.1 until each jump <= half a cell
    while ((fabs(dxr / DX) > maxRelJump) || (fabs(dyr / DY) > maxRelJump) ||
           (fabs(dzr / DZ) > maxRelJump)) {
      t = t - dt;
      dt = dtRed * dt;
      advect(rec, W.advection_kind, dt, s);
      sq = sqrt(dt);
      dxr = s.dAx + s.ddx * dt + s.dBx * sq;
      dyr = s.dAy + s.ddy * dt + s.dBy * sq;
      dzr = s.dAz + s.ddz * dt + s.dBz * sq;
      t = t + dt;
      atMax = false;
    }
    nx = x + dxr / DX;
    ny = y + dyr / DY;
    nz = z + dzr / DZ;
    int loops = 0;
    while (outside(nx, ny, nz)) {
      loops = loops + 1;
      exit_time(rec, W.advection_kind, x, y, z, nx, ny, nz, s, t, dt,
                exitFace);
      if (loops > 5) {  // maxInterfaceLoopCounter: restart the visit
        nx = ix; ny = iy; nz = iz;
        t = t0;
        dt = dtcell;
        restarts = restarts + 1;
        if (restarts > 10) return OUT_UNDEFINED;
        break;
      }
      advect(rec, W.advection_kind, dt, s);
      atMax = false;
      if (exitFace == 0) {  // no face found: restart with new draws
        nx = ix; ny = iy; nz = iz;
        t = t0;
        dt = dtcell;
        restarts = restarts + 1;
        if (restarts > 10) return OUT_UNDEFINED;
        break;
      }
      sq = sqrt(dt);
      if (exitFace <= 2) {
        nx = exitFace == 1 ? 0.0 : 1.0;
        dyr = s.dAy + s.ddy * dt + s.dBy * sq;
        ny = y + dyr / DY;
        dzr = s.dAz + s.ddz * dt + s.dBz * sq;
        nz = z + dzr / DZ;
      } else if (exitFace <= 4) {
        dxr = s.dAx + s.ddx * dt + s.dBx * sq;
        nx = x + dxr / DX;
        ny = exitFace == 3 ? 0.0 : 1.0;
        dzr = s.dAz + s.ddz * dt + s.dBz * sq;
        nz = z + dzr / DZ;
      } else {
        dxr = s.dAx + s.ddx * dt + s.dBx * sq;
        nx = x + dxr / DX;
        dyr = s.dAy + s.ddy * dt + s.dBy * sq;
        ny = y + dyr / DY;
        nz = exitFace == 5 ? 0.0 : 1.0;
      }
      if (outside(nx, ny, nz)) continue;
      // elastic rebound on faces whose mass boundary is 1
      int rebounds = 0;
      while (exitFace > 0) {
        if (mb[exitFace - 1] != 1) break;
        rebounds = rebounds + 1;
        if (rebounds > 10) return OUT_UNDEFINED;
        const double xi = nx, yi = ny, zi = nz;
        if (dt == 0.0) {
          nx = ix; ny = iy; nz = iz;
          t = t - dt;
          dt = dtcell;
          exitFace = 0;
          restarts = restarts + 1;
          break;
        }
        t = t + dt;
        if (W.t_stop < t) {
          t = t - dt;
          dt = fmax(W.t_stop - t, 0.0);
          t = W.t_stop;
          atMax = true;
          advect(rec, W.advection_kind, dt, s);
          sq = sqrt(dt);
          dxr = s.dAx + s.ddx * dt + s.dBx * sq;
          dyr = s.dAy + s.ddy * dt + s.dBy * sq;
          dzr = s.dAz + s.ddz * dt + s.dBz * sq;
        }
        if (exitFace <= 2) {
          nx = nx - dxr / DX;
          ny = ny + dyr / DY;
          nz = nz + dzr / DZ;
        } else if (exitFace <= 4) {
          nx = nx + dxr / DX;
          ny = ny - dyr / DY;
          nz = nz + dzr / DZ;
        } else {
          nx = nx + dxr / DX;
          ny = ny + dyr / DY;
          nz = nz - dzr / DZ;
        }
        if (outside(nx, ny, nz)) {
          x = xi; y = yi; z = zi;
          if (exitFace <= 2) {
            s.vx = -s.vx; s.ddx = -s.ddx; s.dBx = -s.dBx;
          } else if (exitFace <= 4) {
            s.vy = -s.vy; s.ddy = -s.ddy; s.dBy = -s.dBy;
          } else {
            s.vz = -s.vz; s.ddz = -s.ddz; s.dBz = -s.dBz;
          }
          exitFace = 0;
        } else {
          dt = dtcell;
          exitFace = 0;
          if (nx == 0.0) exitFace = 1;
          if (nx == 1.0) exitFace = 2;
          if (ny == 0.0) exitFace = 3;
          if (ny == 1.0) exitFace = 4;
          if (nz == 0.0) exitFace = 5;
          if (nz == 1.0) exitFace = 6;
        }
      }
    }
    if (atMax) {
      fx = nx; fy = ny; fz = nz; ft = t;
      face = exitFace;
      return OUT_MAXTIME;
    }
    if (exitFace > 0) {
      fx = nx; fy = ny; fz = nz; ft = t;
      face = exitFace;
      zero_dt = dt == 0.0;
      return OUT_EXIT;
    }
    x = nx;
    y = ny;
    z = nz;
  }
}

// pr_RWPTrackPath + TrackCell (ParticleTrackingEngine.f90:1051-1254,
// TrackCell.f90:375-669) mapped to MPathRW particle status codes.  The
// location written back is the last pathline point the reference records.
__device__ int track_path(const double *__restrict__ cells,
                          const WalkParams &W, int &cell, double &lx,
                          double &ly, double &lz, double &lt, Stream &rng,
                          int64_t &steps) {
  int c = cell, stalls = 0;
  double x = lx, y = ly, z = lz, t = lt;
  for (;;) {
    const double *rec = cells + (int64_t)(c - 1) * CELL_DOUBLES;
    if (outside(x, y, z)) return 9;  // InvalidLocation -> Undefined
    const int32_t *fl = reinterpret_cast<const int32_t *>(rec + F_FLAGS);
    // checks before the first tracking point report the entry location
    if (fl[0] == 0) goto stop_here_7;
    {
      const int32_t zone = reinterpret_cast<const int32_t *>(rec + F_ZONE)[0];
      if (W.stop_zone > 0 && zone == W.stop_zone) {
        cell = c; lx = x; ly = y; lz = z; lt = t;
        return 6;
      }
      bool stopIfNoExit = true;
      const double src = rec[F_SOURCE], snk = rec[F_SINK];
      if (W.backward) {
        if (W.stop_weak_sources) {
          if (src != 0.0) {
            cell = c; lx = x; ly = y; lz = z; lt = t;
            return 4;
          } else if (!W.steady_state) stopIfNoExit = false;
        }
      } else if (W.stop_weak_sinks) {
        if (snk != 0.0) {
          cell = c; lx = x; ly = y; lz = z; lt = t;
          return 3;
        } else if (!W.steady_state) stopIfNoExit = false;
      }
      if (!(fl[1] & CF_CELL_EXIT)) {
        if (W.steady_state || (W.backward && src != 0.0) ||
            (!W.backward && snk != 0.0)) {
          cell = c; lx = x; ly = y; lz = z; lt = t;
          return 5;
        }
      }
      double fx, fy, fz, ft;
      int face = 0;
      bool zero_dt = false;
      const int out = visit_cell(rec, W, stopIfNoExit, x, y, z, t, rng, steps,
                                 fx, fy, fz, ft, face, zero_dt);
      if (out == OUT_MAXTIME) {
        cell = c; lx = fx; ly = fy; lz = fz; lt = ft;
        return 1;
      }
      if (out == OUT_EXIT) {
        // exit point in this cell is the last recorded point
        cell = c; lx = fx; ly = fy; lz = fz; lt = ft;
        stalls = zero_dt ? stalls + 1 : 0;
        if (stalls > kMaxStalls) return 9;
        const int next = reinterpret_cast<const int32_t *>(rec + F_CONN)[face - 1];
        if (next <= 0) return 2;
        const double *nrec = cells + (int64_t)(next - 1) * CELL_DOUBLES;
        if (reinterpret_cast<const int32_t *>(nrec + F_FLAGS)[0] == 0) return 2;
        // ConvertFromNeighbor: full-face connection, face coordinate flips
        bool bad;
        switch (face) {
          case 1: bad = fx != 0.0; fx = 1.0; break;
          case 2: bad = fx != 1.0; fx = 0.0; break;
          case 3: bad = fy != 0.0; fy = 1.0; break;
          case 4: bad = fy != 1.0; fy = 0.0; break;
          case 5: bad = fz != 0.0; fz = 1.0; break;
          default: bad = fz != 1.0; fz = 0.0; break;
        }
        if (bad) return 9;
        c = next;
        x = fx; y = fy; z = fz; t = ft;
        continue;
      }
      if (out == OUT_UNDEFINED) return 9;
      cell = c; lx = x; ly = y; lz = z; lt = t;
      return out == OUT_NOEXIT ? 5 : 7;
    }
  stop_here_7:
    cell = c; lx = x; ly = y; lz = z; lt = t;
    return 7;
  }
}

__global__ void __launch_bounds__(256)
    advance_f64(const double *__restrict__ cells, int64_t n,
                int32_t *__restrict__ cellv, int32_t *__restrict__ statusv,
                double *__restrict__ lxv, double *__restrict__ lyv,
                double *__restrict__ lzv, double *__restrict__ tv,
                uint64_t *__restrict__ rngv, int64_t *__restrict__ stepv,
                const WalkParams W, unsigned long long *__restrict__ counters) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int st = statusv[p];
  int c = cellv[p];
  double t = tv[p];
  int64_t steps = 0;
  if (st == 0) {  // release, MPathRW.F90:1256-1302
    if (t < W.t_engine) st = 8;
    else if (t <= W.t_stop) {
      st = 1;
      const double *rec = cells + (int64_t)(c - 1) * CELL_DOUBLES;
      if (reinterpret_cast<const int32_t *>(rec + F_FLAGS)[0] == 0) st = 7;
    }
  }
  if (st == 7) {
    const double *rec = cells + (int64_t)(c - 1) * CELL_DOUBLES;
    if (reinterpret_cast<const int32_t *>(rec + F_FLAGS)[1] & CF_VERIFY_PDRY)
      st = 1;
  }
  if (st == 1) {
    double x = lxv[p], y = lyv[p], z = lzv[p];
    Stream rng{rngv[2 * p], rngv[2 * p + 1]};
    st = track_path(cells, W, c, x, y, z, t, rng, steps);
    cellv[p] = c;
    lxv[p] = x;
    lyv[p] = y;
    lzv[p] = z;
    tv[p] = t;
    rngv[2 * p] = rng.a;
    rngv[2 * p + 1] = rng.b;
    stepv[p] += steps;
  }
  statusv[p] = st;
  // per-wave totals of active particles and sub-steps
  unsigned long long act = st == 1 ? 1ull : 0ull;
  unsigned long long ns = (unsigned long long)steps;
  for (int off = 32; off > 0; off >>= 1) {
    act += __shfl_down(act, off, 64);
    ns += __shfl_down(ns, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&counters[0], act);
    atomicAdd(&counters[1], ns);
  }
}

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

hipError_t launch_advance(rwpt_engine *e, double t_stop) {
  Particles &P = e->part;
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
  hipError_t err = hipMemsetAsync(e->counters, 0, 4 * sizeof(unsigned long long),
                                  e->stream);
  if (err != hipSuccess) return err;
  if (P.n == 0) return hipSuccess;
  const int block = 256;
  KernelTimer &T = e->timers[0];
  (void)hipEventRecord(T.start, e->stream);
  hipLaunchKernelGGL(advance_f64, dim3((unsigned)((P.n + block - 1) / block)),
                     dim3(block), 0, e->stream, e->prob.cells, P.n, P.cell,
                     P.status, P.lx, P.ly, P.lz, P.t, P.rng, P.substeps, W,
                     e->counters);
  err = hipGetLastError();
  (void)hipEventRecord(T.stop, e->stream);
  return err;
}

hipError_t launch_global_xyz(rwpt_engine *e, double *gx, double *gy,
                             double *gz) {
  // face coordinates X(1..ncol+1), Y(1..nrow+1) as RectangularGridDis builds
  DeviceProblem &P = e->prob;
  std::vector<double> hdelr(P.ncol), hdelc(P.nrow);
  (void)hipMemcpy(hdelr.data(), P.delr, P.ncol * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipMemcpy(hdelc.data(), P.delc, P.nrow * sizeof(double), hipMemcpyDeviceToHost);
  std::vector<double> X(P.ncol + 1), Y(P.nrow + 1);
  X[0] = 0.0;
  for (int j = 0; j < P.ncol; ++j) X[j + 1] = X[j] + hdelr[j];
  Y[0] = 0.0;
  for (int i = 0; i < P.nrow; ++i) Y[0] = Y[0] + hdelc[i];
  for (int i = 0; i < P.nrow; ++i) Y[i + 1] = Y[i] - hdelc[i];
  double *dX = nullptr, *dY = nullptr;
  hipError_t err = hipMalloc(&dX, X.size() * sizeof(double));
  if (err != hipSuccess) return err;
  err = hipMalloc(&dY, Y.size() * sizeof(double));
  if (err != hipSuccess) {
    (void)hipFree(dX);
    return err;
  }
  (void)hipMemcpyAsync(dX, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice,
                 e->stream);
  (void)hipMemcpyAsync(dY, Y.data(), Y.size() * sizeof(double), hipMemcpyHostToDevice,
                 e->stream);
  const int block = 256;
  const int64_t n = e->part.n;
  if (n > 0)
    hipLaunchKernelGGL(global_xyz_kernel, dim3((unsigned)((n + block - 1) / block)),
                       dim3(block), 0, e->stream, n, e->part.cell, e->part.lx,
                       e->part.ly, e->part.lz, P.ncol, P.nrow, dX, dY, P.delr,
                       P.delc, P.top, P.bot, gx, gy, gz);
  err = hipGetLastError();
  (void)hipStreamSynchronize(e->stream);
  (void)hipFree(dX);
  (void)hipFree(dY);
  return err;
}

}  // namespace rwpt
