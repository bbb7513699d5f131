// Device code of the random-walk advance, shared by the two translation
// units that instantiate it: rwpt_walk.hip (plain FP64, built without
// floating-point contraction so that every result is the double the
// reference forms) and rwpt_walk32.hip (the FP32-local walk, built with
// contraction into fused multiply-add).
//
// One thread carries one particle from its current location to t_stop,
// through as many cells as it crosses, entirely in registers; the state is
// read once and written once per launch (structure of arrays in HBM).  Per
// sub-step the thread reads its cell's record (built once per flow step by
// rwpt_cells.hip) instead of re-assembling the cell and its 18 neighbours on
// every cell entry as pr_RWPTrackPath does.
//
// The sub-step logic is the reference's (TrackSubCell.F90:812-1373): the
// displacement, the jump limiter that reuses the same draw, the interface
// loop with the quadratic (EULERIAN) or Newton (EXPONENTIAL) exit-time
// solve, restarts and elastic rebound.
#pragma once

#include <algorithm>

#include "rwpt_internal.h"
#include "rwpt_walk_common.h"

namespace rwpt {
namespace {


// Arithmetic by type.  FP64: IEEE division and square root in the
// reference's operand order; a / d uses the record's reciprocal of d only
// when d is a power of two (then the product rounds the same exact real as
// the quotient, so the result is bit-identical).  FP32: reciprocal and square
// root instructions (1 ulp), every division by a cell size, R or phi a
// product with the record's reciprocal.
template <typename R> struct M;
template <> struct M<double> {
  static __device__ __forceinline__ double sq(double x) { return sqrt(x); }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  static __device__ __forceinline__ double divr(double a, double d, double r) {
    return r != 0.0 ? a * r : a / d;
  }
  static __device__ __forceinline__ double div_phi(double a, const double *rec) {
    return a / rec[F_PHI];
  }
  static __device__ __forceinline__ double expm1_over(double v, double g,
                                                      double dt) {
    return v * (exp(g * dt) - 1.0) / g;
  }
};
template <> struct M<float> {
  static __device__ __forceinline__ float sq(float x) {
    return __builtin_amdgcn_sqrtf(x);
  }
  static __device__ __forceinline__ float div(float a, float b) {
    return a * __builtin_amdgcn_rcpf(b);
  }
  static __device__ __forceinline__ float divr(float a, float, float r) {
    return a * r;
  }
  static __device__ __forceinline__ float div_phi(float a, const float *rec) {
    return a * rec[G_RPHI];
  }
  static __device__ __forceinline__ float expm1_over(float v, float g, float dt) {
    return v * expm1f(g * dt) * __builtin_amdgcn_rcpf(g);
  }
};

// P2 = 1: every cell's DX, DY, DZ and R are powers of two (a flag of the
// table build), so every a / d of the sub-step is the exact product a * (1/d)
// with no per-lane test left to the compiler.
template <typename R, int P2> struct Ops : M<R> {};
template <> struct Ops<double, 1> : M<double> {
  static __device__ __forceinline__ double divr(double a, double, double r) {
    return a * r;
  }
};


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
template <typename R>
__device__ __forceinline__ R tri8(const R *__restrict__ v, R x, R y, R z) {
  const R v00 = (R(1) - x) * v[0] + x * v[1];
  const R v01 = (R(1) - x) * v[4] + x * v[5];
  const R v10 = (R(1) - x) * v[2] + x * v[3];
  const R v11 = (R(1) - x) * v[6] + x * v[7];
  const R v0 = (R(1) - y) * v00 + y * v10;
  const R v1 = (R(1) - y) * v01 + y * v11;
  return (R(1) - z) * v0 + z * v1;
}

// derivative term stored as the four edge differences (pr_TrilinearDerivative)
template <typename R>
__device__ __forceinline__ R deriv4(const R *__restrict__ d, R u, R w) {
  const R v0 = (R(1) - u) * d[0] + u * d[1];
  const R v1 = (R(1) - u) * d[2] + u * d[3];
  return (R(1) - w) * v0 + w * v1;
}

// FP32 record forms (fused multiply-adds): the trilinear interpolant of
// the corner velocities from its 8 polynomial coefficients, and a drift
// term's bilinear interpolant from its 4 coefficients.
__device__ __forceinline__ float tri_poly(const float *__restrict__ c, float x, float y,
                                          float z) {
  const float a = c[4] + c[7] * z;
  const float b = c[1] + c[5] * z + y * a;
  const float d = c[2] + c[6] * z;
  return c[0] + c[3] * z + y * d + x * b;
}
__device__ __forceinline__ float bil_poly(const float *__restrict__ e, float u, float w) {
  return e[0] + e[2] * w + u * (e[1] + e[3] * w);
}

// exponential-advection velocity gradient (v2 - v1) / D along axis a:
// stored in the FP64 record, slope times reciprocal in the FP32 one
template <typename R>
__device__ __forceinline__ R vel_gradient(const R *__restrict__ rec, int a) {
  if constexpr (sizeof(R) == 4)
    return rec[G_DV + a] * rec[Lay<R>::RDX + a];
  else
    return rec[F_DVDX + a];
}

template <typename R> struct StepTerms {
  R vx, vy, vz;        // interpolated face velocities
  R dAx, dAy, dAz;     // advection displacement
  R dBx, dBy, dBz;     // random displacement (per sqrt(dt))
  R ddx, ddy, ddz;     // dispersion divergence (drift correction)
};

template <typename R, int ADV>
__device__ __forceinline__ void advect(const R *__restrict__ rec, R dt,
                                       StepTerms<R> &s) {
  if constexpr (ADV == 1) {  // pr_AdvectionDisplacementExponential
    const R dvtol = R(1.0e-10);
    const R gx = vel_gradient(rec, 0), gy = vel_gradient(rec, 1),
            gz = vel_gradient(rec, 2);
    s.dAx = fabs(gx) > dvtol ? M<R>::expm1_over(s.vx, gx, dt) : s.vx * dt;
    s.dAy = fabs(gy) > dvtol ? M<R>::expm1_over(s.vy, gy, dt) : s.vy * dt;
    s.dAz = fabs(gz) > dvtol ? M<R>::expm1_over(s.vz, gz, dt) : s.vz * dt;
  } else {          // pr_AdvectionDisplacementEulerian
    s.dAx = s.vx * dt;
    s.dAy = s.vy * dt;
    s.dAz = s.vz * dt;
  }
}

// pr_RWPTDisplacementsLinear / ...Axisymmetric for one sub-step on the
// FP32 record: the same quantities from the record's polynomial forms, each
// product with a reciprocal where the FP64 form divides.
template <int NRM, int NDIM, int MODEL, int ADV>
__device__ __forceinline__ void displacements32(const float *__restrict__ rec,
                                                const WalkParams &W, float x, float y,
                                                float z, float dt, Stream &rng,
                                                StepTerms<float> &s) {
  using Mf = M<float>;
  s.vx = rec[F_VX1] + x * rec[G_DV];
  s.vy = rec[F_VY1] + y * rec[G_DV + 1];
  s.vz = rec[F_VZ1] + z * rec[G_DV + 2];
  float aL, aT, aTH;  // 2 alpha / R
  if constexpr (MODEL == 1) {
    const float vn = Mf::sq(s.vx * s.vx + s.vy * s.vy + s.vz * s.vz);
    float c2 = 0.0f;
    if (vn > 0.0f) c2 = Mf::div(s.vz * s.vz, vn);
    aL = rec[G_S2ALH] + c2 * (rec[G_S2ALV] - rec[G_S2ALH]);
    aT = rec[G_S2ATV] + c2 * (rec[G_S2ATH] - rec[G_S2ATV]);
    aTH = rec[G_S2ATH];
  } else {
    aL = rec[G_S2ALH];
    aT = rec[G_S2ATH];
    aTH = aT;
  }
  // drift correction, pr_DispersionDivergence (terms already over phi R)
  const float *dd = rec + G_DD;
  s.ddx = bil_poly(dd + 0, y, z) + bil_poly(dd + 20, x, z) + bil_poly(dd + 28, x, y);
  s.ddy = bil_poly(dd + 12, y, z) + bil_poly(dd + 4, x, z) + bil_poly(dd + 32, x, y);
  s.ddz = bil_poly(dd + 16, y, z) + bil_poly(dd + 24, x, z) + bil_poly(dd + 8, x, y);
  const float dm = rec[G_S2DM];
  const float *vc = rec + G_VC;
  if constexpr (NDIM == 3) {
    const float bx = tri_poly(vc, x, y, z), by = tri_poly(vc + 8, x, y, z);
    const float bz = tri_poly(vc + 16, x, y, z);
    const float vxy2 = bx * bx + by * by;
    const float vn = Mf::sq(vxy2 + bz * bz);
    const float vxy = Mf::sq(vxy2);
    float B11 = 0, B12 = 0, B13 = 0, B21 = 0, B22 = 0, B23 = 0, B31 = 0, B32 = 0;
    if (vn > 0.0f) {
      const float rvn = __builtin_amdgcn_rcpf(vn);
      const float sl = Mf::sq(aL * vn + dm) * rvn;
      const float st = Mf::sq(aT * vn + dm);
      B11 = bx * sl;
      B21 = by * sl;
      B31 = bz * sl;
      B32 = vxy * st * rvn;
      if (vxy > 0.0f) {
        const float rvxy = __builtin_amdgcn_rcpf(vxy);
        const float sh = Mf::sq(aTH * vn + dm) * rvxy;
        const float stc = st * rvn * rvxy;
        B12 = -(bx * bz) * stc;
        B13 = -by * sh;
        B22 = -(by * bz) * stc;
        B23 = bx * sh;
      }
    }
    const float r1 = Normal<float, NRM>::draw(rng), r2 = Normal<float, NRM>::draw(rng);
    const float r3 = Normal<float, NRM>::draw(rng);
    s.dBx = B11 * r1 + B12 * r2 + B13 * r3;
    s.dBy = B21 * r1 + B22 * r2 + B23 * r3;
    s.dBz = B31 * r1 + B32 * r2;
  } else if constexpr (NDIM == 2) {
    const int d1 = W.dims[0] - 1, d2 = W.dims[1] - 1;
    const float b1 = tri_poly(vc + 8 * d1, x, y, z);
    const float b2 = tri_poly(vc + 8 * d2, x, y, z);
    const float vn = Mf::sq(b1 * b1 + b2 * b2);
    float B11 = 0, B12 = 0, B21 = 0, B22 = 0;
    if (vn > 0.0f) {
      const float rvn = __builtin_amdgcn_rcpf(vn);
      const float sl = Mf::sq(aL * vn + dm) * rvn;
      const float st = Mf::sq(aTH * vn + dm) * rvn;
      B11 = b1 * sl;
      B21 = b2 * sl;
      B12 = -b2 * st;
      B22 = b1 * st;
    }
    const float r1 = Normal<float, NRM>::draw(rng), r2 = Normal<float, NRM>::draw(rng);
    float dB[3] = {0.0f, 0.0f, 0.0f};
    dB[d1] = B11 * r1 + B12 * r2;
    dB[d2] = B21 * r1 + B22 * r2;
    s.dBx = dB[0];
    s.dBy = dB[1];
    s.dBz = dB[2];
  } else {
    const int d1 = W.dims[0] - 1;
    const float b1 = tri_poly(vc + 8 * d1, x, y, z);
    const float vn = Mf::sq(b1 * b1);
    float B11 = 0.0f;
    if (vn > 0.0f) B11 = Mf::sq(aL * vn + dm);
    const float r1 = Normal<float, NRM>::draw(rng);
    float dB[3] = {0.0f, 0.0f, 0.0f};
    dB[d1] = B11 * r1;
    s.dBx = dB[0];
    s.dBy = dB[1];
    s.dBz = dB[2];
  }
  advect<float, ADV>(rec, dt, s);
}

// pr_RWPTDisplacementsLinear / ...Axisymmetric for one sub-step.
template <typename R, int NRM, int P2, int NDIM, int MODEL, int ADV>
__device__ __forceinline__ void displacements(const R *__restrict__ rec,
                                              const WalkParams &W, R x, R y,
                                              R z, R dt, Stream &rng,
                                              StepTerms<R> &s) {
  if constexpr (sizeof(R) == 4) {
    displacements32<NRM, NDIM, MODEL, ADV>(rec, W, x, y, z, dt, rng, s);
    return;
  } else {
  using Mx = Ops<R, P2>;
  s.vx = (R(1) - x) * rec[F_VX1] + x * rec[F_VX2];
  s.vy = (R(1) - y) * rec[F_VY1] + y * rec[F_VY2];
  s.vz = (R(1) - z) * rec[F_VZ1] + z * rec[F_VZ2];
  R aL, aT, aTH;
  if constexpr (MODEL == 1) {
    const R vn = Mx::sq(s.vx * s.vx + s.vy * s.vy + s.vz * s.vz);
    R c2 = R(0);
    if (vn > R(0)) c2 = Mx::div(s.vz * s.vz, vn);
    aL = rec[F_ALH] + c2 * (rec[F_ALV] - rec[F_ALH]);
    aT = rec[F_ATV] + c2 * (rec[F_ATH] - rec[F_ATV]);
    aTH = rec[F_ATH];
  } else {
    aL = rec[F_ALH];
    aT = rec[F_ATH];
    aTH = aT;
  }
  // drift correction, pr_DispersionDivergence
  const R *dd = rec + F_DD;
  const R t0 = deriv4(dd + 0, y, z), t1 = deriv4(dd + 4, x, z);
  const R t2 = deriv4(dd + 8, x, y), t3 = deriv4(dd + 12, y, z);
  const R t4 = deriv4(dd + 16, y, z), t5 = deriv4(dd + 20, x, z);
  const R t6 = deriv4(dd + 24, x, z), t7 = deriv4(dd + 28, x, y);
  const R t8 = deriv4(dd + 32, x, y);
  const R Rt = rec[F_R], rR = rec[Lay<R>::RR];
  s.ddx = Mx::divr(Mx::div_phi(t0 + t5 + t7, rec), Rt, rR);
  s.ddy = Mx::divr(Mx::div_phi(t3 + t1 + t8, rec), Rt, rR);
  s.ddz = Mx::divr(Mx::div_phi(t4 + t6 + t2, rec), Rt, rR);
  // random displacement, pr_DisplacementRandomDischarge{,2D,1D}
  const R dm = rec[F_DMEFF];
  const R *vc = rec + F_VC;
  if constexpr (NDIM == 3) {
    const R bx = tri8(vc, x, y, z), by = tri8(vc + 8, x, y, z);
    const R bz = tri8(vc + 16, x, y, z);
    const R vn = Mx::sq(bx * bx + by * by + bz * bz);
    const R vxy = Mx::sq(bx * bx + by * by);
    R B11 = 0, B12 = 0, B13 = 0, B21 = 0, B22 = 0, B23 = 0, B31 = 0, B32 = 0;
    if (vn > R(0)) {
      const R sl = Mx::sq(Mx::divr(2 * (aL * vn + dm), Rt, rR));
      const R st = Mx::sq(Mx::divr(2 * (aT * vn + dm), Rt, rR));
      B11 = Mx::div(bx * sl, vn);
      B21 = Mx::div(by * sl, vn);
      B31 = Mx::div(bz * sl, vn);
      B32 = Mx::div(vxy * st, vn);
      if (vxy > R(0)) {
        const R sh = Mx::sq(Mx::divr(2 * (aTH * vn + dm), Rt, rR));
        B12 = -Mx::div(Mx::div(bx * bz * st, vn), vxy);
        B13 = -Mx::div(by * sh, vxy);
        B22 = -Mx::div(Mx::div(by * bz * st, vn), vxy);
        B23 = Mx::div(bx * sh, vxy);
      }
    }
    const R r1 = Normal<R, NRM>::draw(rng), r2 = Normal<R, NRM>::draw(rng);
    const R r3 = Normal<R, NRM>::draw(rng);
    s.dBx = B11 * r1 + B12 * r2 + B13 * r3;
    s.dBy = B21 * r1 + B22 * r2 + B23 * r3;
    s.dBz = B31 * r1 + B32 * r2;
  } else if constexpr (NDIM == 2) {
    const int d1 = W.dims[0] - 1, d2 = W.dims[1] - 1;
    const R b1 = tri8(vc + 8 * d1, x, y, z);
    const R b2 = tri8(vc + 8 * d2, x, y, z);
    const R vn = Mx::sq(b1 * b1 + b2 * b2);
    R B11 = 0, B12 = 0, B21 = 0, B22 = 0;
    if (vn > R(0)) {
      const R sl = Mx::sq(Mx::divr(2 * (aL * vn + dm), Rt, rR));
      const R st = Mx::sq(Mx::divr(2 * (aTH * vn + dm), Rt, rR));
      B11 = Mx::div(b1 * sl, vn);
      B21 = Mx::div(b2 * sl, vn);
      B12 = -Mx::div(b2 * st, vn);
      B22 = Mx::div(b1 * st, vn);
    }
    const R r1 = Normal<R, NRM>::draw(rng), r2 = Normal<R, NRM>::draw(rng);
    R dB[3] = {R(0), R(0), R(0)};
    dB[d1] = B11 * r1 + B12 * r2;
    dB[d2] = B21 * r1 + B22 * r2;
    s.dBx = dB[0];
    s.dBy = dB[1];
    s.dBz = dB[2];
  } else {
    const int d1 = W.dims[0] - 1;
    const R b1 = tri8(vc + 8 * d1, x, y, z);
    const R vn = Mx::sq(b1 * b1);
    R B11 = R(0);
    if (vn > R(0)) B11 = Mx::sq(Mx::divr(2 * (aL * vn + dm), Rt, rR));
    const R r1 = Normal<R, NRM>::draw(rng);
    R dB[3] = {R(0), R(0), R(0)};
    dB[d1] = B11 * r1;
    s.dBx = dB[0];
    s.dBy = dB[1];
    s.dBz = dB[2];
  }
  advect<R, ADV>(rec, dt, s);
  }
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

// FP32 form: the same roots from the cancellation-free pair q/A, -dI/q with
// q = -(B + sign(B) sqrt(B^2 + 4 dI A))/2.
__device__ __forceinline__ bool quadratic_axis(float dI, float A, float B,
                                               float dtcur, float &dtaxis) {
  const float arg = B * B + 4 * dI * A;
  if ((arg >= 0.0f) && (A != 0.0f)) {
    const float sq = __builtin_amdgcn_sqrtf(arg);
    const float q = -0.5f * (B + copysignf(sq, B));
    const float z1 = q * __builtin_amdgcn_rcpf(A);
    const float z2 = q != 0.0f ? -dI * __builtin_amdgcn_rcpf(q) : 0.0f;
    if (z1 > 0.0f || z2 > 0.0f) {
      const float m = (z1 > 0.0f && z2 > 0.0f) ? fminf(z1, z2)
                                               : (z1 > 0.0f ? z1 : z2);
      dtaxis = m * m;
      if (dtaxis == 0.0f) return true;
    } else if (z1 == 0.0f || z2 == 0.0f) {
      return true;
    }
  } else if (A == 0.0f) {
    if (B != 0.0f) {
      const float r = dI * __builtin_amdgcn_rcpf(B);
      dtaxis = r * r;
      if (dtaxis == 0.0f) return true;
    }
  }
  if (dtaxis > dtcur) dtaxis = 0.0f;
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
__device__ __noinline__ double nr_guess(double dt, double v, double g, double dI,
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

// NewtonRaphsonTimeStepExponentialAdvection (TrackSubCell.F90:2451-2554);
// kept out of line: it runs only on exit attempts of EXPONENTIAL runs
__device__ __noinline__ double newton_axis(double dt, double v, double g, double dI,
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
// axis yields a time).  The Newton solve of the exponential form runs in
// FP64 for either arithmetic type.
// FP32 quadratic exit time without branches: the time for one axis (0:
// none) and, in hit, whether the particle sits on the interface (the cases
// of the branched form above, evaluated with selects; the values a case
// does not use may be infinite or NaN and are never selected).
__device__ __forceinline__ float quadratic_axis_flat(float dI, float A, float B,
                                                     float dtcur, bool &hit) {
  const float arg = B * B + 4 * dI * A;
  const bool caseA = (arg >= 0.0f) && (A != 0.0f);
  const bool caseB = (A == 0.0f) && (B != 0.0f);
  const float sq = __builtin_amdgcn_sqrtf(fmaxf(arg, 0.0f));
  const float q = -0.5f * (B + copysignf(sq, B));
  const float z1 = q * __builtin_amdgcn_rcpf(A);
  const float z2 = q != 0.0f ? -dI * __builtin_amdgcn_rcpf(q) : 0.0f;
  const bool p1 = z1 > 0.0f, p2 = z2 > 0.0f, pos = p1 || p2;
  const float m = (p1 && p2) ? fminf(z1, z2) : (p1 ? z1 : z2);
  const float dA = m * m;
  const float r = dI * __builtin_amdgcn_rcpf(B);
  const float dB = r * r;
  const bool h = caseA ? (pos ? dA == 0.0f : (z1 == 0.0f || z2 == 0.0f))
                       : (caseB && dB == 0.0f);
  float dta = caseA ? (pos ? dA : 0.0f) : (caseB ? dB : 0.0f);
  if (!h && dta > dtcur) dta = 0.0f;
  hit = h;
  return dta;
}

template <typename R, int ADV>
__device__ void exit_time(const R *__restrict__ rec, R x, R y, R z, R nx, R ny,
                          R nz, const StepTerms<R> &s, double &t, R &dt,
                          int &exitFace) {
  if constexpr (sizeof(R) == 4 && ADV == 2) {
    // every axis evaluated, the reference's first early return (the
    // particle on an interface) chosen afterwards in axis order
    t = t - dt;
    const float pn[3] = {nx, ny, nz}, pp[3] = {x, y, z};
    const float dd[3] = {rec[F_DX], rec[F_DY], rec[F_DZ]};
    const float A[3] = {s.vx + s.ddx, s.vy + s.ddy, s.vz + s.ddz};
    const float db[3] = {s.dBx, s.dBy, s.dBz};
    float dtx[3];
    int fx[3], ret = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const bool up = pn[a] > 1.0f, out = up || (pn[a] < 0.0f);
      const float dI = up ? dd[a] * (1.0f - pp[a]) : -dd[a] * pp[a];
      fx[a] = up ? 2 * a + 2 : 2 * a + 1;
      bool hit;
      const float d = quadratic_axis_flat(dI, A[a], db[a], dt, hit);
      const bool early = out && (dI == 0.0f || hit);
      if (ret == 0 && early) ret = fx[a];
      dtx[a] = (out && !early) ? d : 0.0f;
    }
    if (ret != 0) {
      exitFace = ret;
      dt = 0.0f;
      return;
    }
    if (dtx[0] == 0.0f && dtx[1] == 0.0f && dtx[2] == 0.0f) {
      t = t + dt;
      return;
    }
    int im = -1;
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (dtx[a] > 0.0f && (im < 0 || dtx[a] < dtx[im])) im = a;
    dt = im == 0 ? dtx[0] : (im == 1 ? dtx[1] : dtx[2]);
    exitFace = im == 0 ? fx[0] : (im == 1 ? fx[1] : fx[2]);
    t = t + dt;
    return;
  } else {
  t = t - dt;
  const R pn[3] = {nx, ny, nz}, pp[3] = {x, y, z};
  const R dd[3] = {rec[F_DX], rec[F_DY], rec[F_DZ]};
  const R v[3] = {s.vx, s.vy, s.vz};
  const R dv[3] = {s.ddx, s.ddy, s.ddz};
  const R db[3] = {s.dBx, s.dBy, s.dBz};
  const R gr[3] = {vel_gradient(rec, 0), vel_gradient(rec, 1), vel_gradient(rec, 2)};
  R dtx[3] = {R(0), R(0), R(0)};
  int fx[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if ((pn[a] > R(1)) || (pn[a] < R(0))) {
      R dI;
      if (pn[a] > R(1)) {
        dI = dd[a] * (R(1) - pp[a]);
        fx[a] = 2 * a + 2;
      } else {
        dI = -dd[a] * pp[a];
        fx[a] = 2 * a + 1;
      }
      if (dI == R(0)) {
        exitFace = fx[a];
        dt = R(0);
        return;
      }
      if constexpr (ADV == 2) {
        if (quadratic_axis(dI, v[a] + dv[a], db[a], dt, dtx[a])) {
          exitFace = fx[a];
          dt = R(0);
          return;
        }
      } else {
        dtx[a] = (R)newton_axis((double)dt, (double)v[a], (double)gr[a],
                                (double)dI, (double)dv[a], (double)db[a]);
        if (dtx[a] > dt) dtx[a] = R(0);
      }
    }
  }
  if (dtx[0] == R(0) && dtx[1] == R(0) && dtx[2] == R(0)) {
    t = t + dt;
    return;
  }
  int im = -1;
#pragma unroll
  for (int a = 0; a < 3; ++a)
    if (dtx[a] > R(0) && (im < 0 || dtx[a] < dtx[im])) im = a;
  dt = dtx[im];
  exitFace = fx[im];
  t = t + dt;
  }
}


// Persistent loop state of a cell visit: everything
// pr_ExecuteRandomWalkParticleTracking (TrackSubCell.F90:812-1373) carries
// from one iteration of its time loop to the next, so a visit can be paused
// between iterations and resumed in a later launch with identical results.
// Times are FP64 for either arithmetic type (global time accumulation).
template <typename R> struct Visit {
  R x, y, z;       // current location
  double t;        // and time
  R ix, iy, iz;    // visit start (initialLocation, tinit)
  double t0;
  int restarts;    // posRestartCounter
  int exitFace;    // exitFace (kept across iterations)
};

// One iteration of the time loop (TrackSubCell.F90:956-1368).  On IT_EXIT
// and IT_MAXTIME the final point is left in V.x..V.t; *zero_dt reports an
// exit taken with a zero time step.
template <typename R, int NRM, int P2, int NDIM, int MODEL, int ADV>
__device__ int iterate(const R *__restrict__ rec, const WalkParams &W,
                       Visit<R> &V, Stream &rng, bool &zero_dt) {
  using Mx = Ops<R, P2>;
  const R dtcell = rec[F_DTCELL];
  const R maxRelJump = R(0.5), dtRed = R(f32lit(0.1f));
  const R DX = rec[F_DX], DY = rec[F_DY], DZ = rec[F_DZ];
  const R rX = rec[Lay<R>::RDX], rY = rec[Lay<R>::RDY], rZ = rec[Lay<R>::RDZ];
  const int32_t *mb = ints(rec) + Lay<R>::MASSBND;
  R x = V.x, y = V.y, z = V.z;
  double t = V.t;
  R nx, ny, nz, dt;
  int exitFace = V.exitFace;
  bool atMax = false;
  StepTerms<R> s;
  dt = dtcell;
  t = t + dt;
  if (W.t_stop < t) {
    t = t - dt;
    dt = R(fmax(W.t_stop - t, 0.0));
    t = W.t_stop;
    atMax = true;
  }
  displacements<R, NRM, P2, NDIM, MODEL, ADV>(rec, W, x, y, z, dt, rng, s);
  R sq = Mx::sq(dt);
  R dxr = s.dAx + s.ddx * dt + s.dBx * sq;
  R dyr = s.dAy + s.ddy * dt + s.dBy * sq;
  R dzr = s.dAz + s.ddz * dt + s.dBz * sq;
  // jump limiter: same draw, dt cut by 0.1 until each jump <= half a cell
  while ((fabs(Mx::divr(dxr, DX, rX)) > maxRelJump) ||
         (fabs(Mx::divr(dyr, DY, rY)) > maxRelJump) ||
         (fabs(Mx::divr(dzr, DZ, rZ)) > maxRelJump)) {
    t = t - dt;
    dt = dtRed * dt;
    advect<R, ADV>(rec, dt, s);
    sq = Mx::sq(dt);
    dxr = s.dAx + s.ddx * dt + s.dBx * sq;
    dyr = s.dAy + s.ddy * dt + s.dBy * sq;
    dzr = s.dAz + s.ddz * dt + s.dBz * sq;
    t = t + dt;
    atMax = false;
  }
  nx = x + Mx::divr(dxr, DX, rX);
  ny = y + Mx::divr(dyr, DY, rY);
  nz = z + Mx::divr(dzr, DZ, rZ);
  int loops = 0;
  while (outside(nx, ny, nz)) {
    loops = loops + 1;
    exit_time<R, ADV>(rec, x, y, z, nx, ny, nz, s, t, dt, exitFace);
    if (loops > 5) {  // maxInterfaceLoopCounter: restart the visit
      nx = V.ix; ny = V.iy; nz = V.iz;
      t = V.t0;
      dt = dtcell;
      V.restarts = V.restarts + 1;
      if (V.restarts > 10) return IT_UNDEFINED;
      break;
    }
    advect<R, ADV>(rec, dt, s);
    atMax = false;
    if (exitFace == 0) {  // no face found: restart with new draws
      nx = V.ix; ny = V.iy; nz = V.iz;
      t = V.t0;
      dt = dtcell;
      V.restarts = V.restarts + 1;
      if (V.restarts > 10) return IT_UNDEFINED;
      break;
    }
    sq = Mx::sq(dt);
    // the point at the exit time, the exit axis on its face
    const R ex = s.dAx + s.ddx * dt + s.dBx * sq;
    const R ey = s.dAy + s.ddy * dt + s.dBy * sq;
    const R ez = s.dAz + s.ddz * dt + s.dBz * sq;
    const int ax = (exitFace - 1) >> 1;
    const R side = (exitFace & 1) ? R(0) : R(1);
    if (ax != 0) dxr = ex;
    if (ax != 1) dyr = ey;
    if (ax != 2) dzr = ez;
    nx = ax == 0 ? side : x + Mx::divr(ex, DX, rX);
    ny = ax == 1 ? side : y + Mx::divr(ey, DY, rY);
    nz = ax == 2 ? side : z + Mx::divr(ez, DZ, rZ);
    if (outside(nx, ny, nz)) continue;
    // elastic rebound on faces whose mass boundary is 1
    int rebounds = 0;
    while (exitFace > 0) {
      if (mb[exitFace - 1] != 1) break;
      rebounds = rebounds + 1;
      if (rebounds > 10) return IT_UNDEFINED;
      const R xi = nx, yi = ny, zi = nz;
      if (dt == R(0)) {
        nx = V.ix; ny = V.iy; nz = V.iz;
        t = t - dt;
        dt = dtcell;
        exitFace = 0;
        V.restarts = V.restarts + 1;
        break;
      }
      t = t + dt;
      if (W.t_stop < t) {
        t = t - dt;
        dt = R(fmax(W.t_stop - t, 0.0));
        t = W.t_stop;
        atMax = true;
        advect<R, ADV>(rec, dt, s);
        sq = Mx::sq(dt);
        dxr = s.dAx + s.ddx * dt + s.dBx * sq;
        dyr = s.dAy + s.ddy * dt + s.dBy * sq;
        dzr = s.dAz + s.ddz * dt + s.dBz * sq;
      }
      if (exitFace <= 2) {
        nx = nx - Mx::divr(dxr, DX, rX);
        ny = ny + Mx::divr(dyr, DY, rY);
        nz = nz + Mx::divr(dzr, DZ, rZ);
      } else if (exitFace <= 4) {
        nx = nx + Mx::divr(dxr, DX, rX);
        ny = ny - Mx::divr(dyr, DY, rY);
        nz = nz + Mx::divr(dzr, DZ, rZ);
      } else {
        nx = nx + Mx::divr(dxr, DX, rX);
        ny = ny + Mx::divr(dyr, DY, rY);
        nz = nz - Mx::divr(dzr, DZ, rZ);
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
        if (nx == R(0)) exitFace = 1;
        if (nx == R(1)) exitFace = 2;
        if (ny == R(0)) exitFace = 3;
        if (ny == R(1)) exitFace = 4;
        if (nz == R(0)) exitFace = 5;
        if (nz == R(1)) exitFace = 6;
      }
    }
  }
  V.x = nx; V.y = ny; V.z = nz; V.t = t;
  V.exitFace = exitFace;
  if (atMax) return IT_MAXTIME;
  if (exitFace > 0) {
    zero_dt = dt == R(0);
    return IT_EXIT;
  }
  return IT_CONTINUE;
}


// Wave-aggregated append of (key, value) to a work list.
__device__ __forceinline__ void append_work(bool want, int32_t key,
                                            int32_t val, int32_t *keys,
                                            int32_t *vals,
                                            unsigned long long *count) {
  const unsigned long long mask = __ballot(want);
  if (mask == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (want) {
    const unsigned long long below = mask & ((1ull << lane) - 1ull);
    const unsigned long long idx = base + __popcll(below);
    keys[idx] = key;
    vals[idx] = val;
  }
}

// Wave sums of two counters into global memory.
__device__ __forceinline__ void wave_count(unsigned long long a,
                                           unsigned long long b,
                                           unsigned long long *c0,
                                           unsigned long long *c1) {
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  if (__lane_id() == 0) {
    if (a) atomicAdd(c0, a);
    if (b) atomicAdd(c1, b);
  }
}

// XCD-aware block order: consecutive chunks of the (cell-sorted) work list
// go to the same XCD, whose L2 then holds neighbouring cell records
// (workgroups are dealt to the 8 XCDs round robin).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nblocks) {
  const int64_t q = nblocks / 8, r = nblocks % 8;
  const int64_t x = b % 8, j = b / 8;
  return x * q + (x < r ? x : r) + j;
}

// Particle state as the walk of arithmetic type R sees it: positions in R,
// times in FP64.
template <typename R> struct ParticlePtrs {
  int32_t *cell, *status, *aux;
  R *lx, *ly, *lz, *ix, *iy, *iz;
  double *t, *t0;
  uint64_t *rng;
  int64_t *substeps;
};

// Start of an advance: release pending particles (MPathRW.F90:1256-1302),
// reset the per-call loop state and list the active particles.
template <typename R>
__global__ void prepare_advance(const R *__restrict__ cells,
                                const int32_t *__restrict__ cell_unit, int64_t n,
                                ParticlePtrs<R> S, const WalkParams W,
                                int32_t *__restrict__ keys,
                                int32_t *__restrict__ vals,
                                unsigned long long *__restrict__ counters) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool track = false;
  int c = 0;
  if (p < n) {
    int st = S.status[p];
    c = S.cell[p];
    // the cell's record (its first unit's on an unstructured grid)
    const R *rec = cells + (int64_t)(cell_unit ? cell_unit[c - 1] : c - 1) * Lay<R>::STRIDE;
    if (st == 0) {
      const double t = S.t[p];
      if (t < W.t_engine) st = 8;
      else if (t <= W.t_stop) {
        st = 1;
        if (ints(rec)[Lay<R>::FLAGS] == 0) st = 7;
      }
    }
    if (st == 7) {
      if (ints(rec)[Lay<R>::FLAGS + 1] & CF_VERIFY_PDRY) st = 1;
    }
    S.status[p] = st;
    if (st == 1) {
      S.aux[p] = 0;
      track = true;
    }
  }
  append_work(track, c, (int32_t)p, keys, vals, &counters[2]);
}

// One launch of the walk: every listed particle runs up to `budget`
// time-loop iterations; particles not done are appended to the next list.
// A launch captured in a graph reads the list length from nwork_dev (its
// grid covers every particle; the blocks beyond the list exit at once).
template <typename R, int NRM, int P2, int NDIM, int MODEL, int ADV, int MINW>
__global__ void __launch_bounds__(256, MINW)
    walk_kernel(const R *__restrict__ cells,
                const int32_t *__restrict__ work, int64_t nwork,
                ParticlePtrs<R> S, const WalkParams W, int budget,
                int32_t *__restrict__ nkeys, int32_t *__restrict__ nvals,
                unsigned long long *__restrict__ counters,
                const unsigned long long *__restrict__ nwork_dev) {
  if (nwork_dev) nwork = (int64_t)*nwork_dev;
  const int64_t nblk = (nwork + blockDim.x - 1) / blockDim.x;
  if ((int64_t)blockIdx.x >= nblk) return;
  const int64_t blk = xcd_block(blockIdx.x, nblk);
  const int64_t i = blk * blockDim.x + threadIdx.x;
  bool more = false;
  int c = 0, p = 0;
  unsigned long long steps = 0, ended_active = 0;
  if (i < nwork) {
    p = work[i];
    c = S.cell[p];
    Visit<R> V;
    V.x = S.lx[p]; V.y = S.ly[p]; V.z = S.lz[p]; V.t = S.t[p];
    Aux A = unpack_aux(S.aux[p]);
    if (A.phase) {
      V.ix = S.ix[p]; V.iy = S.iy[p]; V.iz = S.iz[p]; V.t0 = S.t0[p];
    } else {
      V.ix = V.x; V.iy = V.y; V.iz = V.z; V.t0 = V.t;
    }
    V.restarts = A.restarts;
    V.exitFace = A.exitFace;
    Stream rng{S.rng[2 * p], S.rng[2 * p + 1]};
    int status = 1;
    bool done = false;
    const R *rec = cells + (int64_t)(c - 1) * Lay<R>::STRIDE;
    // entering a cell (a released particle): TrackCell checks, new visit;
    // after a transfer the same checks run at once, in the exit branch
    auto enter = [&]() -> bool {
      const int st = entry_checks(rec, W, V.x, V.y, V.z);
      if (st != 0) {
        status = st;
        if (st == 9)  // InvalidLocation: last recorded point
          last_point(cells, A.entryFace, c, V.x, V.y, V.z);
        return false;
      }
      A.phase = 1;
      V.ix = V.x; V.iy = V.y; V.iz = V.z; V.t0 = V.t;
      V.restarts = 0;
      V.exitFace = 0;
      return true;
    };
    if (!A.phase) done = !enter();
    while (!done) {
      if ((int)steps >= budget) {
        more = true;
        break;
      }
      steps += 1;
      bool zero_dt = false;
      const int out = iterate<R, NRM, P2, NDIM, MODEL, ADV>(rec, W, V, rng, zero_dt);
      if (out == IT_CONTINUE) continue;
      if (out == IT_MAXTIME) {
        status = 1;
        done = true;
      } else if (out == IT_UNDEFINED) {
        status = 9;
        V.x = V.ix; V.y = V.iy; V.z = V.iz; V.t = V.t0;
        last_point(cells, A.entryFace, c, V.x, V.y, V.z);
        done = true;
      } else {  // IT_EXIT: the exit point is the last recorded point
        const int face = V.exitFace;
        A.stalls = zero_dt ? A.stalls + 1 : 0;
        if (A.stalls > kMaxStalls) {
          status = 9;
          done = true;
          break;
        }
        const int next = ints(rec)[Lay<R>::CONN + face - 1];
        // no neighbour, or one with IBOUND 0 (a flag of this cell's record)
        if (next <= 0 ||
            (ints(rec)[Lay<R>::FLAGS + 1] & (CF_NB_INACTIVE << (face - 1)))) {
          status = 2;
          done = true;
          break;
        }
        // ConvertFromNeighbor: full-face connection, the face coordinate
        // (0 on odd faces, 1 on even ones) flips
        const int ax = (face - 1) >> 1;
        const R side = (face & 1) ? R(0) : R(1);
        const bool bad = (ax == 0 ? V.x : (ax == 1 ? V.y : V.z)) != side;
        const R fx = ax == 0 ? R(1) - side : V.x;
        const R fy = ax == 1 ? R(1) - side : V.y;
        const R fz = ax == 2 ? R(1) - side : V.z;
        if (bad) {  // no defined continuation in the reference
          status = 9;
          done = true;
          break;
        }
        c = next;
        rec = cells + (int64_t)(c - 1) * Lay<R>::STRIDE;
        V.x = fx; V.y = fy; V.z = fz;
        A.phase = 0;
        A.entryFace = face <= 2 ? 3 - face : (face <= 4 ? 7 - face : 11 - face);
        done = !enter();
      }
    }
    // statuses 3-7 come from the entry checks and report the entry point,
    // which is the current point
    S.cell[p] = c;
    S.lx[p] = V.x; S.ly[p] = V.y; S.lz[p] = V.z; S.t[p] = V.t;
    S.rng[2 * p] = rng.a;
    S.rng[2 * p + 1] = rng.b;
    S.substeps[p] += (int64_t)steps;
    if (more) {
      A.restarts = V.restarts;
      A.exitFace = V.exitFace;
      S.aux[p] = pack_aux(A);
      S.ix[p] = V.ix; S.iy[p] = V.iy; S.iz[p] = V.iz; S.t0[p] = V.t0;
    } else {
      S.status[p] = status;
      ended_active = status == 1 ? 1ull : 0ull;
    }
  }
  append_work(more, c, p, nkeys, nvals, &counters[2]);
  wave_count(ended_active, steps, &counters[0], &counters[1]);
}

// Particles still listed after the step guard: status 9 at the last point.
template <typename R>
__global__ void terminate_walk(const R *__restrict__ cells,
                               const int32_t *__restrict__ work, int64_t nwork,
                               ParticlePtrs<R> S) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwork) return;
  const int p = work[i];
  Aux A = unpack_aux(S.aux[p]);
  int c = S.cell[p];
  R x = S.ix[p], y = S.iy[p], z = S.iz[p];
  double t = S.t0[p];
  if (!A.phase) {
    x = S.lx[p]; y = S.ly[p]; z = S.lz[p]; t = S.t[p];
  }
  last_point(cells, A.entryFace, c, x, y, z);
  S.cell[p] = c;
  S.lx[p] = x; S.ly[p] = y; S.lz[p] = z; S.t[p] = t;
  S.status[p] = 9;
}

template <typename R> ParticlePtrs<R> particle_ptrs(Particles &P);
template <> ParticlePtrs<double> particle_ptrs<double>(Particles &P) {
  return ParticlePtrs<double>{P.cell, P.status, P.aux, P.lx, P.ly, P.lz, P.ix,
                              P.iy,   P.iz,     P.t,   P.t0, P.rng, P.substeps};
}
template <> ParticlePtrs<float> particle_ptrs<float>(Particles &P) {
  return ParticlePtrs<float>{P.cell, P.status, P.aux, P.flx, P.fly, P.flz, P.fix,
                             P.fiy,  P.fiz,    P.t,   P.t0,  P.rng, P.substeps};
}

template <typename R> const R *cell_table(const rwpt_engine *e);
template <> const double *cell_table<double>(const rwpt_engine *e) { return e->prob.cells; }
template <> const float *cell_table<float>(const rwpt_engine *e) { return e->prob.cells32; }

// Launch of the walk variant for (precision, normal generator, power-of-two
// cells, dimensions, model, advection, occupancy hint); the occupancy
// variants exist for the 3-D linear EULERIAN walk only.
template <typename R, int NRM, int P2, int ND, int MO, int AD>
void launch_walk_variant(rwpt_engine *e, const WalkParams &W, dim3 grid, int block,
                         const int32_t *work, int64_t cnt, int budget, int32_t *nk,
                         int32_t *nv) {
  const R *cells = cell_table<R>(e);
  const ParticlePtrs<R> S = particle_ptrs<R>(e->part);
  const int mw = e->tune.min_waves;
  if constexpr (ND == 3 && MO == 0 && AD == 2) {
    if (mw == 3) {
      hipLaunchKernelGGL((walk_kernel<R, NRM, P2, ND, MO, AD, 3>), grid, dim3(block), 0,
                         e->stream, cells, work, cnt, S, W, budget, nk, nv,
                         e->counters, e->nwork_dev);
      return;
    }
    if (mw == 4) {
      hipLaunchKernelGGL((walk_kernel<R, NRM, P2, ND, MO, AD, 4>), grid, dim3(block), 0,
                         e->stream, cells, work, cnt, S, W, budget, nk, nv,
                         e->counters, e->nwork_dev);
      return;
    }
  }
  hipLaunchKernelGGL((walk_kernel<R, NRM, P2, ND, MO, AD, 1>), grid, dim3(block), 0,
                     e->stream, cells, work, cnt, S, W, budget, nk, nv, e->counters,
                     e->nwork_dev);
}

// The walk variant for the engine's dimensions, dispersion model and
// advection kind.
template <typename R, int NRM, int P2>
hipError_t launch_walk_typed(rwpt_engine *e, const WalkParams &W, dim3 grid, int block,
                             const int32_t *work, int64_t cnt, int budget, int32_t *nk,
                             int32_t *nv) {
  const int mode = (W.ndim - 1) * 4 + W.disp_model * 2 + (W.advection_kind - 1);
#define RWPT_WALK(ND, MO, AD)                                                       \
  case (ND - 1) * 4 + MO * 2 + (AD - 1):                                            \
    launch_walk_variant<R, NRM, P2, ND, MO, AD>(e, W, grid, block, work, cnt, budget, \
                                                nk, nv);                            \
    break;
  switch (mode) {
    RWPT_WALK(1, 0, 1) RWPT_WALK(1, 0, 2) RWPT_WALK(1, 1, 1) RWPT_WALK(1, 1, 2)
    RWPT_WALK(2, 0, 1) RWPT_WALK(2, 0, 2) RWPT_WALK(2, 1, 1) RWPT_WALK(2, 1, 2)
    RWPT_WALK(3, 0, 1) RWPT_WALK(3, 0, 2) RWPT_WALK(3, 1, 1) RWPT_WALK(3, 1, 2)
    default: return hipErrorInvalidValue;
  }
#undef RWPT_WALK
  return hipGetLastError();
}

template <typename R>
hipError_t launch_prepare_typed(rwpt_engine *e, const WalkParams &W) {
  Particles &P = e->part;
  const int pb = 256;
  hipLaunchKernelGGL(prepare_advance<R>, dim3((unsigned)((P.n + pb - 1) / pb)), dim3(pb),
                     0, e->stream, cell_table<R>(e),
                     (const int32_t *)(e->usg ? e->usg->cell_unit : nullptr), P.n,
                     particle_ptrs<R>(P), W,
                     P.keys[0], P.vals[0], e->counters);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_terminate_typed(rwpt_engine *e, const int32_t *work, int64_t cnt) {
  hipLaunchKernelGGL(terminate_walk<R>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0,
                     e->stream, cell_table<R>(e), work, cnt, particle_ptrs<R>(e->part));
  return hipGetLastError();
}

}  // namespace
}  // namespace rwpt
