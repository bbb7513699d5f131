// Pathwise adjoint of an observation misfit through the random walk.
//
// With common random numbers a particle's path is a deterministic, piecewise
// smooth function of the flow and transport parameters: for fixed draws the
// branches of a sub-step (jump limiter, interface loop, face choice) are
// fixed near almost every parameter value, and inside a branch the position
// after the sub-step is a smooth function of the position before it, of the
// time, and of the current cell's record.  The misfit
//   J = sum_o (c_o - target_o)^2,  c_o = sum_p m_p K_h(X_p(T) - x_o)
// smooths the particles with a Gaussian kernel, so dJ/dX_p is defined for
// every active particle, also across cell crossings.  The gradient of J with
// respect to every cell's dispersivities and diffusion and to a common
// porosity scale is then the sum over particles and sub-steps of the local
// sensitivities, taken in reverse order (adjoint) from dJ/dX_p(T).
//
// Forward: the advance of the engine itself (rwpt_walk.hip) through the
// caller's output times, with the particle state copied at every window
// start (checkpoints).  Reverse, window by window from the last: the
// window's walk is replayed from its checkpoint by a walk kernel that
// records the input of every time-loop iteration (a tape) and must end
// bit-identical to the forward state after the window (checked); the tape
// also holds each iteration's outcome (final step, jump-limiter factor,
// exit axis and root, the normal draws).  Then one thread per active
// particle walks its tape backwards and applies the transpose of each
// iteration's Jacobian (state x, y, z, t and the cell's aLH, aTH, Dm and
// porosity scale) to the adjoint of the iteration's output: by a
// hand-written reverse step for the common iterations, and, for the rare
// ones (restarts to the visit start, elastic rebounds), in a second pass
// that re-runs the iteration in forward-mode dual numbers of four
// directions twice.  Restarts carry their adjoint to the visit's first
// iteration.  The record's derivatives with respect to the cell parameters
// come from central differences of the table build (the records are linear
// in aLH, aTH, Dm and rational in the porosity).
//
// Scope: FP64 engines, EULERIAN advection, the linear dispersion model,
// 2-D and 3-D walks.
#include <stdio.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "rwpt_internal.h"
#include "rwpt_walk_common.h"

namespace rwpt {
namespace {

constexpr int kTan = F_CONN;  // record fields carrying parameter tangents
constexpr int kBlk = 64;

// ---------------------------------------------------------------------------
// Forward-mode dual numbers: the value is formed by the same double
// operation as in the plain walk, so a dual iteration takes the plain
// iteration's branches bit for bit.
// ---------------------------------------------------------------------------
template <int N> struct Dual {
  double v;
  double d[N];
};

template <int N> __device__ __forceinline__ Dual<N> dconst(double v) {
  Dual<N> r;
  r.v = v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = 0.0;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N> &a, double b) {
  Dual<N> r = a;
  r.v = a.v + b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(double a, const Dual<N> &b) {
  Dual<N> r = b;
  r.v = a + b.v;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N> &a, double b) {
  Dual<N> r = a;
  r.v = a.v - b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(double a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int N> __device__ __forceinline__ Dual<N> operator-(const Dual<N> &a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N> &a, double b) {
  Dual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a.v / b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N> &a, double b) {
  Dual<N> r;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, const Dual<N> &b) {
  Dual<N> r;
  r.v = a / b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -(r.v * b.d[k]) / b.v;
  return r;
}
// At 0 (a zero time step after an exit from a face, a zero velocity) the
// derivative is taken as 0: such points have measure zero on a path.
template <int N> __device__ __forceinline__ Dual<N> dsqrt(const Dual<N> &a) {
  Dual<N> r;
  r.v = sqrt(a.v);
  const double q = r.v > 0.0 ? 1.0 / (2.0 * r.v) : 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * q;
  return r;
}

__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ double val(double a) { return a; }
template <int N> __device__ __forceinline__ double val(const Dual<N> &a) { return a.v; }
template <typename T> struct Cst;
template <> struct Cst<double> {
  static __device__ __forceinline__ double of(double v) { return v; }
};
template <int N> struct Cst<Dual<N>> {
  static __device__ __forceinline__ Dual<N> of(double v) { return dconst<N>(v); }
};

// Cell record as the arithmetic type sees it: plain doubles, or duals whose
// parameter directions carry the record's tangents.
template <typename T> struct RecT;
template <> struct RecT<double> {
  const double *rec;
  __device__ __forceinline__ double operator()(int f) const { return rec[f]; }
};
// Directions P0..N-1 of a dual record carry the tangents of the cell
// parameters (aLH, aTH, Dm, porosity scale); P0 = N: no parameter
// directions.
template <int N, int P0> struct RecD {
  const double *rec;
  const double *tan;  // this cell's tangents: tan[f * 4 + k]
  __device__ __forceinline__ Dual<N> operator()(int f) const {
    Dual<N> r = dconst<N>(rec[f]);
    if (P0 < N && f < kTan) {
#pragma unroll
      for (int k = 0; k < N - P0; ++k) r.d[P0 + k] = tan[f * 4 + k];
    }
    return r;
  }
};

// ---------------------------------------------------------------------------
// One time-loop iteration of a cell visit for the EULERIAN linear-model walk,
// written once for doubles and duals.  Every double operation, its operand
// order and every branch test are those of iterate() in rwpt_walk.hip
// (TrackSubCell.F90:956-1368); the replay checks the equality bit for bit.
// ---------------------------------------------------------------------------
// The visit start is a constant of every iteration (its adjoint is carried
// separately), so it stays a plain double.
template <typename T> struct VisitG {
  T x, y, z, t;            // location and time
  double ix, iy, iz, t0;   // visit start
  int restarts, exitFace;
};
// The advective displacement v * dt is formed where it is used: every use
// follows an update of v or dt exactly as the plain walk's advect() does.
template <typename T> struct TermsG {
  T vx, vy, vz, dBx, dBy, dBz, ddx, ddy, ddz;
};

template <typename T, typename Rec>
__device__ __forceinline__ T tri8g(const Rec &rec, int b, const T &x, const T &y,
                                   const T &z) {
  const T v00 = (1.0 - x) * rec(b + 0) + x * rec(b + 1);
  const T v01 = (1.0 - x) * rec(b + 4) + x * rec(b + 5);
  const T v10 = (1.0 - x) * rec(b + 2) + x * rec(b + 3);
  const T v11 = (1.0 - x) * rec(b + 6) + x * rec(b + 7);
  const T v0 = (1.0 - y) * v00 + y * v10;
  const T v1 = (1.0 - y) * v01 + y * v11;
  return (1.0 - z) * v0 + z * v1;
}

template <typename T, typename Rec>
__device__ __forceinline__ T deriv4g(const Rec &rec, int b, const T &u, const T &w) {
  const T v0 = (1.0 - u) * rec(b + 0) + u * rec(b + 1);
  const T v1 = (1.0 - u) * rec(b + 2) + u * rec(b + 3);
  return (1.0 - w) * v0 + w * v1;
}


template <typename T, typename Rec, int NDIM, typename Obs>
__device__ void displacements_g(const Rec &rec, const WalkParams &W, const T &x,
                                const T &y, const T &z, const T &dt, Stream &rng,
                                TermsG<T> &s, Obs &obs) {
  s.vx = (1.0 - x) * rec(F_VX1) + x * rec(F_VX2);
  s.vy = (1.0 - y) * rec(F_VY1) + y * rec(F_VY2);
  s.vz = (1.0 - z) * rec(F_VZ1) + z * rec(F_VZ2);
  const T aL = rec(F_ALH), aT = rec(F_ATH), aTH = aT;
  const T t0 = deriv4g<T>(rec, F_DD + 0, y, z), t1 = deriv4g<T>(rec, F_DD + 4, x, z);
  const T t2 = deriv4g<T>(rec, F_DD + 8, x, y), t3 = deriv4g<T>(rec, F_DD + 12, y, z);
  const T t4 = deriv4g<T>(rec, F_DD + 16, y, z), t5 = deriv4g<T>(rec, F_DD + 20, x, z);
  const T t6 = deriv4g<T>(rec, F_DD + 24, x, z), t7 = deriv4g<T>(rec, F_DD + 28, x, y);
  const T t8 = deriv4g<T>(rec, F_DD + 32, x, y);
  const T Rt = rec(F_R), phi = rec(F_PHI);
  s.ddx = ((t0 + t5 + t7) / phi) / Rt;
  s.ddy = ((t3 + t1 + t8) / phi) / Rt;
  s.ddz = ((t4 + t6 + t2) / phi) / Rt;
  const T dm = rec(F_DMEFF);
  const T zero = Cst<T>::of(0.0);
  if constexpr (NDIM == 3) {
    const T bx = tri8g<T>(rec, F_VC, x, y, z), by = tri8g<T>(rec, F_VC + 8, x, y, z);
    const T bz = tri8g<T>(rec, F_VC + 16, x, y, z);
    const T vn = dsqrt(bx * bx + by * by + bz * bz);
    const T vxy = dsqrt(bx * bx + by * by);
    T B11 = zero, B12 = zero, B13 = zero, B21 = zero, B22 = zero, B23 = zero, B31 = zero,
      B32 = zero;
    if (val(vn) > 0.0) {
      const T sl = dsqrt((2.0 * (aL * vn + dm)) / Rt);
      const T st = dsqrt((2.0 * (aT * vn + dm)) / Rt);
      B11 = (bx * sl) / vn;
      B21 = (by * sl) / vn;
      B31 = (bz * sl) / vn;
      B32 = (vxy * st) / vn;
      if (val(vxy) > 0.0) {
        const T sh = dsqrt((2.0 * (aTH * vn + dm)) / Rt);
        B12 = -((((bx * bz) * st) / vn) / vxy);
        B13 = -((by * sh) / vxy);
        B22 = -((((by * bz) * st) / vn) / vxy);
        B23 = (bx * sh) / vxy;
      }
    }
    const double r1 = Normal<double, 0>::draw(rng), r2 = Normal<double, 0>::draw(rng);
    const double r3 = Normal<double, 0>::draw(rng);
    obs.draws(r1, r2, r3);
    s.dBx = B11 * r1 + B12 * r2 + B13 * r3;
    s.dBy = B21 * r1 + B22 * r2 + B23 * r3;
    s.dBz = B31 * r1 + B32 * r2;
  } else {
    const int d1 = W.dims[0] - 1, d2 = W.dims[1] - 1;
    const T b1 = tri8g<T>(rec, F_VC + 8 * d1, x, y, z);
    const T b2 = tri8g<T>(rec, F_VC + 8 * d2, x, y, z);
    const T vn = dsqrt(b1 * b1 + b2 * b2);
    T B11 = zero, B12 = zero, B21 = zero, B22 = zero;
    if (val(vn) > 0.0) {
      const T sl = dsqrt((2.0 * (aL * vn + dm)) / Rt);
      const T st = dsqrt((2.0 * (aTH * vn + dm)) / Rt);
      B11 = (b1 * sl) / vn;
      B21 = (b2 * sl) / vn;
      B12 = -((b2 * st) / vn);
      B22 = (b1 * st) / vn;
    }
    const double r1 = Normal<double, 0>::draw(rng), r2 = Normal<double, 0>::draw(rng);
    obs.draws(r1, r2, 0.0);
    const T u1 = B11 * r1 + B12 * r2, u2 = B21 * r1 + B22 * r2;
    s.dBx = d1 == 0 ? u1 : (d2 == 0 ? u2 : zero);
    s.dBy = d1 == 1 ? u1 : (d2 == 1 ? u2 : zero);
    s.dBz = d1 == 2 ? u1 : (d2 == 2 ? u2 : zero);
  }
}

// One axis of pr_DetectExitFaceAndUpdateTimeStepQuadratic.
template <typename T>
__device__ bool quadratic_axis_g(const T &dI, const T &A, const T &B, const T &dtcur,
                                 T &dtaxis, T &root) {
  const T arg = B * B + 4.0 * dI * A;
  if ((val(arg) >= 0.0) && (val(A) != 0.0)) {
    const T sq = dsqrt(arg);
    const T z1 = (-B + sq) / (2.0 * A);
    const T z2 = (-B - sq) / (2.0 * A);
    if (val(z1) > 0.0 || val(z2) > 0.0) {
      const T m = (val(z1) > 0.0 && val(z2) > 0.0) ? (val(z2) < val(z1) ? z2 : z1)
                                                   : (val(z1) > 0.0 ? z1 : z2);
      dtaxis = m * m;
      root = m;
      if (val(dtaxis) == 0.0) return true;
    } else if (val(z1) == 0.0 || val(z2) == 0.0) {
      return true;
    }
  } else if (val(A) == 0.0) {
    if (val(B) != 0.0) {
      const T r = dI / B;
      dtaxis = r * r;
      root = r;
      if (val(dtaxis) == 0.0) return true;
    }
  }
  if (val(dtaxis) > val(dtcur)) dtaxis = Cst<T>::of(0.0);
  return false;
}

template <typename T, typename Rec, typename Obs>
__device__ void exit_time_g(const Rec &rec, const T &x, const T &y, const T &z,
                            const T &nx, const T &ny, const T &nz, const TermsG<T> &s,
                            T &t, T &dt, int &exitFace, Obs &obs) {
  t = t - dt;
  const T pn[3] = {nx, ny, nz}, pp[3] = {x, y, z};
  const T dd[3] = {rec(F_DX), rec(F_DY), rec(F_DZ)};
  const T v[3] = {s.vx, s.vy, s.vz};
  const T dv[3] = {s.ddx, s.ddy, s.ddz};
  const T db[3] = {s.dBx, s.dBy, s.dBz};
  T dtx[3] = {Cst<T>::of(0.0), Cst<T>::of(0.0), Cst<T>::of(0.0)};
  T mr[3] = {Cst<T>::of(0.0), Cst<T>::of(0.0), Cst<T>::of(0.0)};
  double dIv[3] = {0.0, 0.0, 0.0};
  int fx[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if ((val(pn[a]) > 1.0) || (val(pn[a]) < 0.0)) {
      T dI;
      if (val(pn[a]) > 1.0) {
        dI = dd[a] * (1.0 - pp[a]);
        fx[a] = 2 * a + 2;
      } else {
        dI = (-dd[a]) * pp[a];
        fx[a] = 2 * a + 1;
      }
      if (val(dI) == 0.0) {
        exitFace = fx[a];
        dt = Cst<T>::of(0.0);
        obs.zero_exit(a);
        return;
      }
      dIv[a] = val(dI);
      if (quadratic_axis_g(dI, v[a] + dv[a], db[a], dt, dtx[a], mr[a])) {
        exitFace = fx[a];
        dt = Cst<T>::of(0.0);
        obs.zero_exit(a);
        return;
      }
    }
  }
  if (val(dtx[0]) == 0.0 && val(dtx[1]) == 0.0 && val(dtx[2]) == 0.0) {
    t = t + dt;
    obs.rare();
    return;
  }
  int im = -1;
  double best = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double va = val(dtx[a]);
    if (va > 0.0 && (im < 0 || va < best)) {
      im = a;
      best = va;
    }
  }
  dt = im == 0 ? dtx[0] : (im == 1 ? dtx[1] : dtx[2]);
  exitFace = im == 0 ? fx[0] : (im == 1 ? fx[1] : fx[2]);
  t = t + dt;
  if constexpr (Obs::kRecords) {
    const T &m = im == 0 ? mr[0] : (im == 1 ? mr[1] : mr[2]);
    obs.exit_axis(im, val(m), val(v[im] + dv[im]), val(db[im]), dIv[im]);
  }
}

// Observers of an iteration's branches: the plain and dual iterations record
// nothing; the hand-written reverse step records the branch taken.
struct NoObs {
  static constexpr bool kRecords = false;
  __device__ void clip() {}
  __device__ void limit(double) {}
  __device__ void zero_exit(int) {}
  __device__ void exit_axis(int, double, double, double, double) {}
  __device__ void rare() {}
  __device__ void final_dt(double) {}
  __device__ void draws(double, double, double) {}
};

struct StepObs {
  static constexpr bool kRecords = true;
  bool clipped = false, exited = false, zero = false, odd = false;
  double c = 1.0;                       // product of the jump-limiter cuts
  int im = -1;                          // exit axis of the last exit round
  double m = 0.0, A = 0.0, B = 0.0, dI = 0.0, dt = 0.0;
  __device__ void clip() { clipped = true; }
  __device__ void limit(double f) { c = c * f; }
  __device__ void zero_exit(int a) {
    exited = true;
    zero = true;
    im = a;
  }
  __device__ void exit_axis(int a, double mm, double AA, double BB, double dII) {
    exited = true;
    zero = false;
    im = a;
    m = mm;
    A = AA;
    B = BB;
    dI = dII;
  }
  double r[3] = {0.0, 0.0, 0.0};         // the sub-step's normal draws
  __device__ void rare() { odd = true; }
  __device__ void final_dt(double d) { dt = d; }
  __device__ void draws(double a, double b, double c3) {
    r[0] = a;
    r[1] = b;
    r[2] = c3;
  }
};

template <typename T>
__device__ __forceinline__ bool outside_g(const T &x, const T &y, const T &z) {
  return outside(val(x), val(y), val(z));
}

// *restart: 1 when the iteration returned to the visit start (location and
// time), 2 when an elastic rebound with a zero time step returned it to the
// visit start location only.
template <typename T, typename Rec, int NDIM, typename Obs = NoObs>
__device__ int iterate_g(const Rec &rec, const WalkParams &W, VisitG<T> &V,
                         Stream &rng, bool &zero_dt, int &restart, Obs &&obs = Obs{}) {
  const T dtcell = rec(F_DTCELL);
  const double maxRelJump = 0.5, dtRed = f32lit(0.1f);
  const T DX = rec(F_DX), DY = rec(F_DY), DZ = rec(F_DZ);
  const int32_t *mb = ints(rec.rec) + Lay<double>::MASSBND;
  T x = V.x, y = V.y, z = V.z;
  T t = V.t;
  T nx, ny, nz, dt;
  int exitFace = V.exitFace;
  bool atMax = false;
  TermsG<T> s;
  dt = dtcell;
  t = t + dt;
  if (W.t_stop < val(t)) {
    t = t - dt;
    const T rem = W.t_stop - t;
    dt = val(rem) > 0.0 ? rem : Cst<T>::of(fmax(val(rem), 0.0));
    t = Cst<T>::of(W.t_stop);
    atMax = true;
    obs.clip();
  }
  displacements_g<T, Rec, NDIM>(rec, W, x, y, z, dt, rng, s, obs);
  T sq = dsqrt(dt);
  T dxr = s.vx * dt + s.ddx * dt + s.dBx * sq;
  T dyr = s.vy * dt + s.ddy * dt + s.dBy * sq;
  T dzr = s.vz * dt + s.ddz * dt + s.dBz * sq;
  while ((fabs(val(dxr) / val(DX)) > maxRelJump) || (fabs(val(dyr) / val(DY)) > maxRelJump) ||
         (fabs(val(dzr) / val(DZ)) > maxRelJump)) {
    t = t - dt;
    dt = dtRed * dt;
    obs.limit(dtRed);
    sq = dsqrt(dt);
    dxr = s.vx * dt + s.ddx * dt + s.dBx * sq;
    dyr = s.vy * dt + s.ddy * dt + s.dBy * sq;
    dzr = s.vz * dt + s.ddz * dt + s.dBz * sq;
    t = t + dt;
    atMax = false;
  }
  nx = x + dxr / DX;
  ny = y + dyr / DY;
  nz = z + dzr / DZ;
  int loops = 0;
  while (outside_g(nx, ny, nz)) {
    loops = loops + 1;
    exit_time_g<T>(rec, x, y, z, nx, ny, nz, s, t, dt, exitFace, obs);
    if (loops > 5) {
      nx = Cst<T>::of(V.ix); ny = Cst<T>::of(V.iy); nz = Cst<T>::of(V.iz);
      t = Cst<T>::of(V.t0);
      dt = dtcell;
      V.restarts = V.restarts + 1;
      restart = 1;
      obs.rare();
      if (V.restarts > 10) return IT_UNDEFINED;
      break;
    }
    atMax = false;
    if (exitFace == 0) {
      nx = Cst<T>::of(V.ix); ny = Cst<T>::of(V.iy); nz = Cst<T>::of(V.iz);
      t = Cst<T>::of(V.t0);
      dt = dtcell;
      V.restarts = V.restarts + 1;
      restart = 1;
      obs.rare();
      if (V.restarts > 10) return IT_UNDEFINED;
      break;
    }
    sq = dsqrt(dt);
    if (exitFace <= 2) {
      nx = Cst<T>::of(exitFace == 1 ? 0.0 : 1.0);
      dyr = s.vy * dt + s.ddy * dt + s.dBy * sq;
      ny = y + dyr / DY;
      dzr = s.vz * dt + s.ddz * dt + s.dBz * sq;
      nz = z + dzr / DZ;
    } else if (exitFace <= 4) {
      dxr = s.vx * dt + s.ddx * dt + s.dBx * sq;
      nx = x + dxr / DX;
      ny = Cst<T>::of(exitFace == 3 ? 0.0 : 1.0);
      dzr = s.vz * dt + s.ddz * dt + s.dBz * sq;
      nz = z + dzr / DZ;
    } else {
      dxr = s.vx * dt + s.ddx * dt + s.dBx * sq;
      nx = x + dxr / DX;
      dyr = s.vy * dt + s.ddy * dt + s.dBy * sq;
      ny = y + dyr / DY;
      nz = Cst<T>::of(exitFace == 5 ? 0.0 : 1.0);
    }
    if (outside_g(nx, ny, nz)) continue;
    int rebounds = 0;
    while (exitFace > 0) {
      if (mb[exitFace - 1] != 1) break;
      rebounds = rebounds + 1;
      obs.rare();
      if (rebounds > 10) return IT_UNDEFINED;
      const T xi = nx, yi = ny, zi = nz;
      if (val(dt) == 0.0) {
        nx = Cst<T>::of(V.ix); ny = Cst<T>::of(V.iy); nz = Cst<T>::of(V.iz);
        t = t - dt;
        dt = dtcell;
        exitFace = 0;
        V.restarts = V.restarts + 1;
        restart = 2;
        break;
      }
      t = t + dt;
      if (W.t_stop < val(t)) {
        t = t - dt;
        const T rem = W.t_stop - t;
        dt = val(rem) > 0.0 ? rem : Cst<T>::of(fmax(val(rem), 0.0));
        t = Cst<T>::of(W.t_stop);
        atMax = true;
            sq = dsqrt(dt);
        dxr = s.vx * dt + s.ddx * dt + s.dBx * sq;
        dyr = s.vy * dt + s.ddy * dt + s.dBy * sq;
        dzr = s.vz * dt + s.ddz * dt + s.dBz * sq;
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
      if (outside_g(nx, ny, nz)) {
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
        if (val(nx) == 0.0) exitFace = 1;
        if (val(nx) == 1.0) exitFace = 2;
        if (val(ny) == 0.0) exitFace = 3;
        if (val(ny) == 1.0) exitFace = 4;
        if (val(nz) == 0.0) exitFace = 5;
        if (val(nz) == 1.0) exitFace = 6;
      }
    }
  }
  V.x = nx; V.y = ny; V.z = nz; V.t = t;
  V.exitFace = exitFace;
  obs.final_dt(val(dt));
  if (atMax) return IT_MAXTIME;
  if (exitFace > 0) {
    zero_dt = val(dt) == 0.0;
    return IT_EXIT;
  }
  return IT_CONTINUE;
}

// ---------------------------------------------------------------------------
// Tape of one window: the input of every time-loop iteration of the
// particles that are active at the final time.
// ---------------------------------------------------------------------------
struct TapeEntry {
  double x, y, z, t, ix, iy, iz, t0;  // input: location, time, visit start
  uint64_t ra, rb;                    // input: rng state
  double dt, c, m, r[3];              // outcome: final step, limiter factor, exit root, draws
  int32_t cell;
  int32_t flags;  // input: bit 0 first iteration of a visit, 1-4 restarts,
                  // 5-7 exitFace; outcome: 8 exited, 9 zero-step exit,
                  // 10 cut at the window end, 11 rare branch, 12-13 exit
                  // axis, 14-16 result code, 17-19 exitFace
};

// The tape of one window is laid out by wavefront: the wavefront of
// threads 64w .. 64w+63 owns a block of len_w entries per lane (len_w its
// lanes' largest iteration count), entry k of the block stored field by
// field, each field as 64 consecutive words, one per lane.  A lane's column
// is block start + lane; its entry k begins k * kTapeWords * 64 words
// further on.  Every tape store and load of a wavefront then covers
// consecutive words.
constexpr int kTapeWords = 17;
static_assert(sizeof(TapeEntry) == kTapeWords * 8, "tape entry is 17 words");

__device__ __forceinline__ uint64_t dbits(double v) { return (uint64_t)__double_as_longlong(v); }
__device__ __forceinline__ double bitsd(uint64_t v) { return __longlong_as_double((long long)v); }

__device__ __forceinline__ void tape_put(uint64_t *__restrict__ tape, int64_t col, int64_t k,
                                         const TapeEntry &E) {
  uint64_t *q = tape + col + k * (int64_t)(kTapeWords * kBlk);
  q[0 * kBlk] = dbits(E.x);
  q[1 * kBlk] = dbits(E.y);
  q[2 * kBlk] = dbits(E.z);
  q[3 * kBlk] = dbits(E.t);
  q[4 * kBlk] = dbits(E.ix);
  q[5 * kBlk] = dbits(E.iy);
  q[6 * kBlk] = dbits(E.iz);
  q[7 * kBlk] = dbits(E.t0);
  q[8 * kBlk] = E.ra;
  q[9 * kBlk] = E.rb;
  q[10 * kBlk] = dbits(E.dt);
  q[11 * kBlk] = dbits(E.c);
  q[12 * kBlk] = dbits(E.m);
  q[13 * kBlk] = dbits(E.r[0]);
  q[14 * kBlk] = dbits(E.r[1]);
  q[15 * kBlk] = dbits(E.r[2]);
  q[16 * kBlk] = (uint64_t)(uint32_t)E.cell | ((uint64_t)(uint32_t)E.flags << 32);
}

__device__ __forceinline__ TapeEntry tape_get(const uint64_t *__restrict__ tape, int64_t col,
                                              int64_t k) {
  const uint64_t *q = tape + col + k * (int64_t)(kTapeWords * kBlk);
  TapeEntry E;
  E.x = bitsd(q[0 * kBlk]);
  E.y = bitsd(q[1 * kBlk]);
  E.z = bitsd(q[2 * kBlk]);
  E.t = bitsd(q[3 * kBlk]);
  E.ix = bitsd(q[4 * kBlk]);
  E.iy = bitsd(q[5 * kBlk]);
  E.iz = bitsd(q[6 * kBlk]);
  E.t0 = bitsd(q[7 * kBlk]);
  E.ra = q[8 * kBlk];
  E.rb = q[9 * kBlk];
  E.dt = bitsd(q[10 * kBlk]);
  E.c = bitsd(q[11 * kBlk]);
  E.m = bitsd(q[12 * kBlk]);
  E.r[0] = bitsd(q[13 * kBlk]);
  E.r[1] = bitsd(q[14 * kBlk]);
  E.r[2] = bitsd(q[15 * kBlk]);
  const uint64_t w = q[16 * kBlk];
  E.cell = (int32_t)(uint32_t)w;
  E.flags = (int32_t)(uint32_t)(w >> 32);
  return E;
}

// Column of thread j of a window's launch, from the block starts (entries).
__device__ __forceinline__ int64_t tape_col(const int64_t *__restrict__ wstart, int64_t j) {
  return wstart[j / kBlk] * (int64_t)(kTapeWords * kBlk) + j % kBlk;
}

// Particle arrays of a checkpoint (or of the engine after the window).
struct CkPtrs {
  int32_t *cell, *status;
  double *lx, *ly, *lz, *t;
  uint64_t *rng;
  int64_t *substeps;
};

__device__ __forceinline__ bool same(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b);
}

// ---------------------------------------------------------------------------
// Hand-written reverse of the common iteration (no restart, no rebound):
// the displacement terms at the iteration's start point are recomputed with
// their intermediates, and the output adjoint is carried back through
//   n_a = x_a + (v_a dt + dd_a dt + dB_a sqrt(dt)) / DX_a
// (dt the cell step times the jump-limiter cuts, or t_stop - t when the
// step is cut at the end of the window), or, after an exit on axis a*,
// through dt = m^2 with A m^2 + B m - dI = 0 (A = v + dd, B = dB along a*,
// dI = DX (1 - x) or -DX x), whose implicit derivative is
//   (2 A m + B) dm = d(dI) - m^2 dA - m dB,
// then through the velocity, drift and dispersion-matrix formulas to the
// start point and to the record fields, whose parameter tangents give the
// cell parameters' gradient.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ivn_safe(double v) { return v > 0.0 ? 1.0 / v : 0.0; }

// element i of a 3-array and an update of it without a dynamic index (which
// would place the array in scratch memory)
__device__ __forceinline__ double pick(const double a[3], int i) {
  return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]);
}
__device__ __forceinline__ void bump(double a[3], int i, double v) {
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (q == i) a[q] += v;
}

struct DispFull {
  double v[3], V1[3], V2[3], dd[3], dB[3], r[3];
  double b[3], vn, vxy, sl, st, sh;
  double B11, B12, B13, B21, B22, B23, B31, B32;
  double phi, R, aL, aT, dm;
};

__device__ __forceinline__ double tri8d(const double *c, double x, double y, double z) {
  const double v00 = (1.0 - x) * c[0] + x * c[1];
  const double v01 = (1.0 - x) * c[4] + x * c[5];
  const double v10 = (1.0 - x) * c[2] + x * c[3];
  const double v11 = (1.0 - x) * c[6] + x * c[7];
  const double v0 = (1.0 - y) * v00 + y * v10;
  const double v1 = (1.0 - y) * v01 + y * v11;
  return (1.0 - z) * v0 + z * v1;
}

__device__ __forceinline__ double der4d(const double *d, double u, double w) {
  const double v0 = (1.0 - u) * d[0] + u * d[1];
  const double v1 = (1.0 - u) * d[2] + u * d[3];
  return (1.0 - w) * v0 + w * v1;
}

// Record tangents of the linear dispersion model in closed form (the
// gradient they give is checked against central differences of the misfit
// by the tests): with
// porosity scaled by s, the face and corner velocities and the cell
// porosity scale as 1/s, s and s (d/ds at s = 1: -v, -v, phi); aLH, aTH and
// Dm are record fields of their own (tangent 1); the drift terms are linear
// in aLH, aTH and Dm, their Dm part carrying the corner porosity, so
// d/ds = Dm d/dDm.  Only the drift terms' three tangents are stored
// (dd3[cell][term][aLH, aTH, Dm]), rounded to FP32: they are the largest
// load of a reverse iteration (108 values), and in FP32 the sweep needs
// half the registers and load round trips for them (11 % of the gradient's
// time at 1e6 particles) while the gradient changes by less than 1e-7
// relative.

// Reverse of the trilinear interpolation of corner values at base f (value
// bval at the point): lam * db/d(x, y, z) to lx, and to the porosity scale
// the corner values' tangent -b.
__device__ __forceinline__ void tri8_rev(const double *rec, int f, double x, double y,
                                         double z, double bval, double lam, double lx[3],
                                         double g[4]) {
  const double *c = rec + f;
  const double v00 = (1.0 - x) * c[0] + x * c[1], v10 = (1.0 - x) * c[2] + x * c[3];
  const double v01 = (1.0 - x) * c[4] + x * c[5], v11 = (1.0 - x) * c[6] + x * c[7];
  lx[0] += lam * ((1.0 - z) * ((1.0 - y) * (c[1] - c[0]) + y * (c[3] - c[2])) +
                  z * ((1.0 - y) * (c[5] - c[4]) + y * (c[7] - c[6])));
  lx[1] += lam * ((1.0 - z) * (v10 - v00) + z * (v11 - v01));
  lx[2] += lam * (((1.0 - y) * v01 + y * v11) - ((1.0 - y) * v00 + y * v10));
  g[3] += -lam * bval;
}

// Reverse of der4d for drift term j (record base F_DD + 4j) over the
// coordinates (iu, iw); dd3 holds the cell's drift-term tangents.
__device__ __forceinline__ void der4_rev(const double *rec, const float *dd3, int j,
                                         const double p[3], int iu, int iw, double dm,
                                         double lam, double lx[3], double g[4]) {
  const double *d = rec + F_DD + 4 * j;
  const double u = p[iu], w = p[iw];
  lx[iu] += lam * ((1.0 - w) * (d[1] - d[0]) + w * (d[3] - d[2]));
  lx[iw] += lam * (((1.0 - u) * d[2] + u * d[3]) - ((1.0 - u) * d[0] + u * d[1]));
  const double wt[4] = {(1.0 - w) * (1.0 - u), (1.0 - w) * u, w * (1.0 - u), w * u};
  const float *t = dd3 + 12 * j;  // [corner pair q][param k] for q < 4, k < 3
  double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] += wt[q] * (double)t[3 * q + k];
  g[0] += lam * acc[0];
  g[1] += lam * acc[1];
  g[2] += lam * acc[2];
  g[3] += lam * dm * acc[2];
}

template <int NDIM>
__device__ void disp_full(const double *rec, const WalkParams &W, const double p[3],
                          const double r[3], DispFull &D) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    D.V1[a] = rec[F_VX1 + 2 * a];
    D.V2[a] = rec[F_VX2 + 2 * a];
    D.v[a] = (1.0 - p[a]) * D.V1[a] + p[a] * D.V2[a];
  }
  D.aL = rec[F_ALH];
  D.aT = rec[F_ATH];
  D.dm = rec[F_DMEFF];
  D.phi = rec[F_PHI];
  D.R = rec[F_R];
  const double *dd = rec + F_DD;
  const double x = p[0], y = p[1], z = p[2];
  const double t0 = der4d(dd + 0, y, z), t1 = der4d(dd + 4, x, z), t2 = der4d(dd + 8, x, y);
  const double t3 = der4d(dd + 12, y, z), t4 = der4d(dd + 16, y, z);
  const double t5 = der4d(dd + 20, x, z), t6 = der4d(dd + 24, x, z);
  const double t7 = der4d(dd + 28, x, y), t8 = der4d(dd + 32, x, y);
  D.dd[0] = ((t0 + t5 + t7) / D.phi) / D.R;
  D.dd[1] = ((t3 + t1 + t8) / D.phi) / D.R;
  D.dd[2] = ((t4 + t6 + t2) / D.phi) / D.R;
  D.B11 = D.B12 = D.B13 = D.B21 = D.B22 = D.B23 = D.B31 = D.B32 = 0.0;
  D.sl = D.st = D.sh = 0.0;
  if constexpr (NDIM == 3) {
#pragma unroll
    for (int a = 0; a < 3; ++a) D.b[a] = tri8d(rec + F_VC + 8 * a, x, y, z);
    D.vn = sqrt(D.b[0] * D.b[0] + D.b[1] * D.b[1] + D.b[2] * D.b[2]);
    D.vxy = sqrt(D.b[0] * D.b[0] + D.b[1] * D.b[1]);
    if (D.vn > 0.0) {
      D.sl = sqrt((2.0 * (D.aL * D.vn + D.dm)) / D.R);
      D.st = sqrt((2.0 * (D.aT * D.vn + D.dm)) / D.R);
      D.B11 = (D.b[0] * D.sl) / D.vn;
      D.B21 = (D.b[1] * D.sl) / D.vn;
      D.B31 = (D.b[2] * D.sl) / D.vn;
      D.B32 = (D.vxy * D.st) / D.vn;
      if (D.vxy > 0.0) {
        D.sh = sqrt((2.0 * (D.aT * D.vn + D.dm)) / D.R);
        D.B12 = -((((D.b[0] * D.b[2]) * D.st) / D.vn) / D.vxy);
        D.B13 = -((D.b[1] * D.sh) / D.vxy);
        D.B22 = -((((D.b[1] * D.b[2]) * D.st) / D.vn) / D.vxy);
        D.B23 = (D.b[0] * D.sh) / D.vxy;
      }
    }
    D.r[0] = r[0];
    D.r[1] = r[1];
    D.r[2] = r[2];
    D.dB[0] = D.B11 * D.r[0] + D.B12 * D.r[1] + D.B13 * D.r[2];
    D.dB[1] = D.B21 * D.r[0] + D.B22 * D.r[1] + D.B23 * D.r[2];
    D.dB[2] = D.B31 * D.r[0] + D.B32 * D.r[1];
  } else {
    const int d1 = W.dims[0] - 1, d2 = W.dims[1] - 1;
    D.b[0] = tri8d(rec + F_VC + 8 * d1, x, y, z);
    D.b[1] = tri8d(rec + F_VC + 8 * d2, x, y, z);
    D.b[2] = 0.0;
    D.vn = sqrt(D.b[0] * D.b[0] + D.b[1] * D.b[1]);
    D.vxy = 0.0;
    if (D.vn > 0.0) {
      D.sl = sqrt((2.0 * (D.aL * D.vn + D.dm)) / D.R);
      D.st = sqrt((2.0 * (D.aT * D.vn + D.dm)) / D.R);
      D.B11 = (D.b[0] * D.sl) / D.vn;
      D.B21 = (D.b[1] * D.sl) / D.vn;
      D.B12 = -((D.b[1] * D.st) / D.vn);
      D.B22 = (D.b[0] * D.st) / D.vn;
    }
    D.r[0] = r[0];
    D.r[1] = r[1];
    D.r[2] = 0.0;
    const double u1 = D.B11 * D.r[0] + D.B12 * D.r[1], u2 = D.B21 * D.r[0] + D.B22 * D.r[1];
#pragma unroll
    for (int a = 0; a < 3; ++a) D.dB[a] = a == d1 ? u1 : (a == d2 ? u2 : 0.0);
  }
}

// Options of the reverse sweep (rwpt_adjoint_options).
struct SweepOpts {
  double graze_cut;  // exits with kappa below it lose their exit-time derivative
  int frozen_noise;  // 1: the random displacement's derivative with respect
                     // to the sub-step duration is dropped
};

// Returns false (nothing changed) for an iteration on a rare branch; the
// branch taken and the draws come from the tape.  At a cell exit, kappa
// receives |2 A m + B| / (|2 A m| + |B|) of the exit root (1 otherwise), and
// a root with kappa < O.graze_cut contributes no exit-time derivative.
template <int NDIM>
__device__ bool reverse_step(const double *rec, const double *tan, const float *dd3,
                             const WalkParams &W, const TapeEntry &E, const double l[4],
                             double ln[4], double g[4], const SweepOpts &O, double &kappa) {
  const int32_t fl = E.flags;
  const int out = (fl >> 14) & 7;
  if ((fl & (1 << 11)) || out == IT_UNDEFINED) return false;
  const bool exited = (fl >> 8) & 1, zero = (fl >> 9) & 1, clipped = (fl >> 10) & 1;
  const int im = (fl >> 12) & 3, exitFace = (fl >> 17) & 7;
  const double p[3] = {E.x, E.y, E.z};
  const double DXv[3] = {rec[F_DX], rec[F_DY], rec[F_DZ]};
  DispFull D;
  disp_full<NDIM>(rec, W, p, E.r, D);
  // output adjoint; after an exit the face coordinate is the constant face
  const int ax = (out == IT_EXIT && exitFace > 0) ? (exitFace - 1) / 2 : -1;
  const double lo[3] = {ax == 0 ? 0.0 : l[0], ax == 1 ? 0.0 : l[1], ax == 2 ? 0.0 : l[2]};
  const double lt_out = l[3];
  double lx[3] = {0.0, 0.0, 0.0}, lv[3] = {0.0, 0.0, 0.0}, ldd[3] = {0.0, 0.0, 0.0};
  double ldB[3] = {0.0, 0.0, 0.0}, lt = 0.0, ldt = 0.0;
  const double dt = E.dt, sq = sqrt(dt);
  const double hs = sq > 0.0 ? 0.5 / sq : 0.0;
  const int skip = exited ? im : -1;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a == skip) continue;
    const double w = lo[a] / DXv[a];
    lx[a] += lo[a];
    lv[a] += w * dt;
    ldd[a] += w * dt;
    ldB[a] += w * sq;
    ldt += w * (D.v[a] + D.dd[a] + (O.frozen_noise ? 0.0 : D.dB[a] * hs));
  }
  if (exited) {
    // t_out = t + dt_exit
    lt += lt_out;
    ldt += lt_out;
    if (!zero) {
      // the root of A m^2 + B m - dI = 0 on the exit axis, A = v + dd, B = dB
      const double A = pick(D.v, im) + pick(D.dd, im), Bq = pick(D.dB, im), m = E.m;
      const double den = 2.0 * A * m + Bq;
      const double scale = fabs(2.0 * A * m) + fabs(Bq);
      kappa = scale > 0.0 ? fabs(den) / scale : 0.0;
      if (den != 0.0 && !(kappa < O.graze_cut)) {
        const double q = 2.0 * m * ldt / den;
        bump(lx, im, -pick(DXv, im) * q);   // d(dI)/dx = -DX on either face
        bump(lv, im, -m * m * q);
        bump(ldd, im, -m * m * q);
        bump(ldB, im, -m * q);
      }
    }
  } else if (!clipped) {
    // dt = c dt_cell, t_out = t + dt
    lt += lt_out;
    {
      // the cell step's tangents (zero for FIXED steps) from the full table
      const double lc = E.c * (ldt + lt_out);
#pragma unroll
      for (int k = 0; k < 4; ++k) g[k] += lc * tan[F_DTCELL * 4 + k];
    }
  } else if (E.c == 1.0) {
    // dt = t_stop - t, t_out = t_stop
    lt += -ldt;
  } else {
    // dt = c (t_stop - t), t_out = t + dt
    lt += lt_out * (1.0 - E.c) - E.c * ldt;
  }
  // velocities
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lx[a] += lv[a] * (D.V2[a] - D.V1[a]);
    g[3] += -lv[a] * D.v[a];  // dv/ds = -v
  }
  // drift: dd_a = S_a / phi / R
  {
    const double sphi = ldd[0] / D.phi / D.R, sphj = ldd[1] / D.phi / D.R,
                 sphk = ldd[2] / D.phi / D.R;
    g[3] += -(ldd[0] * D.dd[0] + ldd[1] * D.dd[1] + ldd[2] * D.dd[2]);  // d(1/phi)/ds
    // term k at F_DD + 4k over its coordinates: t0 yz, t1 xz, t2 xy, t3 yz,
    // t4 yz, t5 xz, t6 xz, t7 xy, t8 xy; dd_x = t0+t5+t7, dd_y = t3+t1+t8,
    // dd_z = t4+t6+t2
    der4_rev(rec, dd3, 0, p, 1, 2, D.dm, sphi, lx, g);
    der4_rev(rec, dd3, 5, p, 0, 2, D.dm, sphi, lx, g);
    der4_rev(rec, dd3, 7, p, 0, 1, D.dm, sphi, lx, g);
    der4_rev(rec, dd3, 3, p, 1, 2, D.dm, sphj, lx, g);
    der4_rev(rec, dd3, 1, p, 0, 2, D.dm, sphj, lx, g);
    der4_rev(rec, dd3, 8, p, 0, 1, D.dm, sphj, lx, g);
    der4_rev(rec, dd3, 4, p, 1, 2, D.dm, sphk, lx, g);
    der4_rev(rec, dd3, 6, p, 0, 2, D.dm, sphk, lx, g);
    der4_rev(rec, dd3, 2, p, 0, 1, D.dm, sphk, lx, g);
  }
  // dispersion matrix
  double lb[3] = {0.0, 0.0, 0.0}, lvn = 0.0, lvxy = 0.0, lsl = 0.0, lst = 0.0, lsh = 0.0;
  if (D.vn > 0.0) {
    if constexpr (NDIM == 3) {
      const double l11 = ldB[0] * D.r[0], l12 = ldB[0] * D.r[1], l13 = ldB[0] * D.r[2];
      const double l21 = ldB[1] * D.r[0], l22 = ldB[1] * D.r[1], l23 = ldB[1] * D.r[2];
      const double l31 = ldB[2] * D.r[0], l32 = ldB[2] * D.r[1];
      const double ivn = 1.0 / D.vn;
      lb[0] += l11 * D.sl * ivn;
      lb[1] += l21 * D.sl * ivn;
      lb[2] += l31 * D.sl * ivn;
      lsl += (l11 * D.b[0] + l21 * D.b[1] + l31 * D.b[2]) * ivn;
      lvn -= (l11 * D.B11 + l21 * D.B21 + l31 * D.B31 + l32 * D.B32) * ivn;
      lvxy += l32 * D.st * ivn;
      lst += l32 * D.vxy * ivn;
      if (D.vxy > 0.0) {
        const double ivxy = 1.0 / D.vxy, k = D.st * ivn * ivxy;
        lb[0] += -l12 * D.b[2] * k + l23 * D.sh * ivxy;
        lb[1] += -l22 * D.b[2] * k - l13 * D.sh * ivxy;
        lb[2] += -(l12 * D.b[0] + l22 * D.b[1]) * k;
        lst += -(l12 * D.b[0] * D.b[2] + l22 * D.b[1] * D.b[2]) * ivn * ivxy;
        lsh += (-l13 * D.b[1] + l23 * D.b[0]) * ivxy;
        lvn -= (l12 * D.B12 + l22 * D.B22) * ivn;
        lvxy -= (l12 * D.B12 + l22 * D.B22 + l13 * D.B13 + l23 * D.B23) * ivxy;
      }
    } else {
      const double b1 = pick(ldB, W.dims[0] - 1), b2 = pick(ldB, W.dims[1] - 1);
      const double l11 = b1 * D.r[0], l12 = b1 * D.r[1];
      const double l21 = b2 * D.r[0], l22 = b2 * D.r[1];
      const double ivn = 1.0 / D.vn;
      lb[0] += (l11 * D.sl + l22 * D.st) * ivn;
      lb[1] += (l21 * D.sl - l12 * D.st) * ivn;
      lsl += (l11 * D.b[0] + l21 * D.b[1]) * ivn;
      lst += (-l12 * D.b[1] + l22 * D.b[0]) * ivn;
      lvn -= (l11 * D.B11 + l21 * D.B21 + l12 * D.B12 + l22 * D.B22) * ivn;
    }
    // s = sqrt(2 (a vn + dm) / R): ds = (vn da + a dvn + ddm) / (s R)
    double laL = 0.0, laT = 0.0, ldm = 0.0;
    if (D.sl > 0.0) {
      const double f = lsl / (D.sl * D.R);
      laL += f * D.vn;
      lvn += f * D.aL;
      ldm += f;
    }
    if (D.st > 0.0) {
      const double f = lst / (D.st * D.R);
      laT += f * D.vn;
      lvn += f * D.aT;
      ldm += f;
    }
    if (D.sh > 0.0) {
      const double f = lsh / (D.sh * D.R);
      laT += f * D.vn;
      lvn += f * D.aT;
      ldm += f;
    }
    g[0] += laL;
    g[1] += laT;
    g[2] += ldm;
    // norms
    lb[0] += lvn * D.b[0] * ivn_safe(D.vn);
    lb[1] += lvn * D.b[1] * ivn_safe(D.vn);
    lb[2] += lvn * D.b[2] * ivn_safe(D.vn);
    if (D.vxy > 0.0) {
      lb[0] += lvxy * D.b[0] / D.vxy;
      lb[1] += lvxy * D.b[1] / D.vxy;
    }
    // corner velocities
    if constexpr (NDIM == 3) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
        tri8_rev(rec, F_VC + 8 * a, p[0], p[1], p[2], D.b[a], lb[a], lx, g);
    } else {
      tri8_rev(rec, F_VC + 8 * (W.dims[0] - 1), p[0], p[1], p[2], D.b[0], lb[0], lx, g);
      tri8_rev(rec, F_VC + 8 * (W.dims[1] - 1), p[0], p[1], p[2], D.b[1], lb[1], lx, g);
    }
  }
  ln[0] = lx[0];
  ln[1] = lx[1];
  ln[2] = lx[2];
  ln[3] = lt;
  return true;
}

// Replay of one window for the listed particles from checkpoint `in`,
// recording the tape; the end state must equal `want` bit for bit (the
// listed particle count of mismatches goes to *bad).  The start of an
// advance (prepare_advance) and the visit loop (walk_kernel) without launch
// budget: one thread walks its particle to the end of the window, thread j
// taking listed particle ord[j].
template <int NDIM>
__global__ void __launch_bounds__(kBlk) adj_replay(
    const double *__restrict__ cells, const int32_t *__restrict__ sel, int64_t nsel,
    const int32_t *__restrict__ ord, CkPtrs in, CkPtrs want, const WalkParams W,
    const int64_t *__restrict__ cnt, const int64_t *__restrict__ wstart,
    uint64_t *__restrict__ tape, unsigned long long *__restrict__ bad) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsel) return;
  const int64_t i = ord[j];
  const int p = sel[i];
  int st = in.status[p];
  int c = in.cell[p];
  VisitG<double> V;
  V.x = in.lx[p]; V.y = in.ly[p]; V.z = in.lz[p]; V.t = in.t[p];
  V.ix = V.x; V.iy = V.y; V.iz = V.z; V.t0 = V.t;
  V.restarts = 0;
  V.exitFace = 0;
  Stream rng{in.rng[2 * p], in.rng[2 * p + 1]};
  if (st == 0) {
    if (V.t < W.t_engine) st = 8;
    else if (V.t <= W.t_stop) {
      st = 1;
      if (ints(cells + (int64_t)(c - 1) * CELL_DOUBLES)[Lay<double>::FLAGS] == 0) st = 7;
    }
  }
  if (st == 7 && (ints(cells + (int64_t)(c - 1) * CELL_DOUBLES)[Lay<double>::FLAGS + 1] &
                  CF_VERIFY_PDRY))
    st = 1;
  const int64_t col = tape_col(wstart, j), cap = cnt[i];
  int64_t k = 0;
  int status = st;
  bool overflow = false;
  if (st == 1) {
    Aux A{0, 0, 0, 0, 0};
    bool first = false;
    while (true) {
      const double *rec = cells + (int64_t)(c - 1) * CELL_DOUBLES;
      if (!A.phase) {
        const int ec = entry_checks(rec, W, V.x, V.y, V.z);
        if (ec != 0) {
          status = ec;
          if (ec == 9) last_point(cells, A.entryFace, c, V.x, V.y, V.z);
          break;
        }
        A.phase = 1;
        V.ix = V.x; V.iy = V.y; V.iz = V.z; V.t0 = V.t;
        V.restarts = 0;
        V.exitFace = 0;
        first = true;
      }
      if (k >= cap) {
        overflow = true;
        break;
      }
      TapeEntry E;
      E.x = V.x; E.y = V.y; E.z = V.z; E.t = V.t;
      E.ix = V.ix; E.iy = V.iy; E.iz = V.iz; E.t0 = V.t0;
      E.ra = rng.a; E.rb = rng.b;
      E.cell = c;
      int32_t fl = (first ? 1 : 0) | (V.restarts << 1) | (V.exitFace << 5);
      first = false;
      bool zero_dt = false;
      int restart = 0;
      StepObs ob;
      const int out = iterate_g<double, RecT<double>, NDIM, StepObs &>(
          RecT<double>{rec}, W, V, rng, zero_dt, restart, ob);
      E.dt = ob.dt;
      E.c = ob.c;
      E.m = ob.m;
      E.r[0] = ob.r[0];
      E.r[1] = ob.r[1];
      E.r[2] = ob.r[2];
      fl |= (ob.exited ? 1 << 8 : 0) | (ob.zero ? 1 << 9 : 0) | (ob.clipped ? 1 << 10 : 0) |
            (ob.odd || restart != 0 ? 1 << 11 : 0) | ((ob.im < 0 ? 0 : ob.im) << 12) |
            (out << 14) | (V.exitFace << 17);
      E.flags = fl;
      tape_put(tape, col, k, E);
      k += 1;
      if (out == IT_CONTINUE) continue;
      if (out == IT_MAXTIME) {
        status = 1;
        break;
      }
      if (out == IT_UNDEFINED) {
        status = 9;
        V.x = V.ix; V.y = V.iy; V.z = V.iz; V.t = V.t0;
        last_point(cells, A.entryFace, c, V.x, V.y, V.z);
        break;
      }
      const int face = V.exitFace;
      A.stalls = zero_dt ? A.stalls + 1 : 0;
      if (A.stalls > kMaxStalls) {
        status = 9;
        break;
      }
      const int next = ints(rec)[Lay<double>::CONN + face - 1];
      bool inactive_next = next <= 0;
      if (!inactive_next)
        inactive_next =
            ints(cells + (int64_t)(next - 1) * CELL_DOUBLES)[Lay<double>::FLAGS] == 0;
      if (inactive_next) {
        status = 2;
        break;
      }
      bool badface;
      double fx = V.x, fy = V.y, fz = V.z;
      switch (face) {
        case 1: badface = fx != 0.0; fx = 1.0; break;
        case 2: badface = fx != 1.0; fx = 0.0; break;
        case 3: badface = fy != 0.0; fy = 1.0; break;
        case 4: badface = fy != 1.0; fy = 0.0; break;
        case 5: badface = fz != 0.0; fz = 1.0; break;
        default: badface = fz != 1.0; fz = 0.0; break;
      }
      if (badface) {
        status = 9;
        break;
      }
      c = next;
      V.x = fx; V.y = fy; V.z = fz;
      A.phase = 0;
      A.entryFace = face <= 2 ? 3 - face : (face <= 4 ? 7 - face : 11 - face);
    }
  }
  const bool ok = !overflow && k == cap && c == want.cell[p] && status == want.status[p] &&
                  same(V.x, want.lx[p]) && same(V.y, want.ly[p]) && same(V.z, want.lz[p]) &&
                  same(V.t, want.t[p]) && rng.a == want.rng[2 * p] &&
                  rng.b == want.rng[2 * p + 1];
  if (!ok) atomicAdd(bad, 1ull);
}

// Sort keys of the listed particles for one window: the particle's
// iteration count in the window (clamped to what fits above the cell
// bits), then the cell it starts the window in.  A thread replays or
// reverses its particle's whole window, so a wavefront lasts as long as
// its longest lane: equal counts keep the lanes busy together, and equal
// cells let them share cell records.
__global__ void adj_order_keys(const int32_t *__restrict__ sel, int64_t nsel,
                               const int32_t *__restrict__ cell,
                               const int64_t *__restrict__ cnt, int cell_bits,
                               int32_t *__restrict__ key, int32_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsel) return;
  const int64_t cmax = (int64_t(1) << (31 - cell_bits)) - 1;
  const int64_t c = cnt[i] < cmax ? cnt[i] : cmax;
  key[i] = (int32_t)((c << cell_bits) | (int64_t)cell[sel[i]]);
  idx[i] = (int32_t)i;
}

// The particles with status 1 (active at the final time) in sel, order not
// kept: one atomic per wavefront reserves its lanes' slots.
__global__ void adj_select(int64_t n, const int32_t *__restrict__ status,
                           int32_t *__restrict__ sel, unsigned long long *__restrict__ count) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = p < n && status[p] == 1;
  const unsigned long long mask = __ballot(on);
  if (mask == 0) return;
  const int lane = __lane_id();
  const int first = __ffsll((long long)mask) - 1;
  unsigned long long base = 0;
  if (lane == first) base = atomicAdd(count, (unsigned long long)__popcll(mask));
  base = __shfl(base, first, 64);
  if (on) sel[base + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)p;
}

// The listed particles whose adjoint is not zero, with their adjoints
// (order not kept).  An adjoint that is exactly zero at the final time (a
// particle beyond the reach of every observation kernel, where the
// Gaussian underflows) stays zero through every window, so such particles
// need neither replay nor reverse sweep.
__global__ void adj_keep(const int32_t *__restrict__ sel, const double *__restrict__ lam,
                         int64_t nsel, int32_t *__restrict__ sel2, double *__restrict__ lam2,
                         unsigned long long *__restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsel) return;
  const double l0 = lam[4 * i], l1 = lam[4 * i + 1], l2 = lam[4 * i + 2], l3 = lam[4 * i + 3];
  if (l0 == 0.0 && l1 == 0.0 && l2 == 0.0 && l3 == 0.0) return;
  const unsigned long long q = atomicAdd(count, 1ull);
  sel2[q] = sel[i];
  lam2[4 * q] = l0;
  lam2[4 * q + 1] = l1;
  lam2[4 * q + 2] = l2;
  lam2[4 * q + 3] = l3;
}

// Tape block length of each wavefront of a window's launch: the largest
// iteration count among its lanes (threads 64w .. 64w+63, particle ord[j]);
// entry nwaves is 0 so that an exclusive scan also gives the total.
__global__ void adj_wave_len(const int32_t *__restrict__ ord, int64_t nsel,
                             const int64_t *__restrict__ cnt, int64_t *__restrict__ wlen) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nw = (nsel + kBlk - 1) / kBlk;
  if (w > nw) return;
  int64_t m = 0;
  if (w < nw) {
    const int64_t e = (w + 1) * kBlk < nsel ? (w + 1) * kBlk : nsel;
    for (int64_t j = w * kBlk; j < e; ++j) m = cnt[ord[j]] > m ? cnt[ord[j]] : m;
  }
  wlen[w] = m;
}

// Iterations of each listed particle in the window, from the sub-step
// counters before and after it.
__global__ void adj_counts(const int32_t *__restrict__ sel, int64_t nsel,
                           const int64_t *__restrict__ before,
                           const int64_t *__restrict__ after, int64_t *__restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsel) cnt[i] = after[sel[i]] - before[sel[i]];
  if (i == nsel) cnt[i] = 0;
}

// Gradient sums of one particle's visit to one cell, added to the cell.
__device__ __forceinline__ void flush(int c, const double g[4], double *g_al, double *g_at,
                                      double *g_dm, double &g_s) {
  if (c <= 0) return;
  if (g_al) unsafeAtomicAdd(&g_al[c - 1], g[0]);
  if (g_at) unsafeAtomicAdd(&g_at[c - 1], g[1]);
  if (g_dm) unsafeAtomicAdd(&g_dm[c - 1], g[2]);
  g_s += g[3];
}

// Tallies of the reverse sweep for rwpt_adjoint_stats, per thread; the
// layout of the reduced array is the one rwpt_adjoint_stats documents.
struct SweepStats {
  int n[5] = {0, 0, 0, 0, 0};            // iterations per kind
  float lg[5] = {0.f, 0.f, 0.f, 0.f, 0.f};  // sum of log10(growth) > 0
  float mx[5] = {0.f, 0.f, 0.f, 0.f, 0.f};  // largest log10(growth)
  int big[5] = {0, 0, 0, 0, 0};          // growth > 10
  int kap[7] = {0, 0, 0, 0, 0, 0, 0};    // exits by kappa decade
  int cut = 0;                           // exit-time derivatives dropped

  __device__ void add(int kind, const double lo[4], const double ln[4]) {
    const double a = fabs(lo[0]) + fabs(lo[1]) + fabs(lo[2]) + fabs(lo[3]);
    const double b = fabs(ln[0]) + fabs(ln[1]) + fabs(ln[2]) + fabs(ln[3]);
    n[kind] += 1;
    if (a > 0.0 && b > a) {
      const float g = (float)log10(b / a);
      lg[kind] += g;
      mx[kind] = fmaxf(mx[kind], g);
      big[kind] += g > 1.f ? 1 : 0;
    }
  }
  __device__ void exit_kappa(double k, double cut_at) {
    const int b = k < 1e-10 ? 0 : k < 1e-8 ? 1 : k < 1e-6 ? 2 : k < 1e-4 ? 3
                : k < 1e-2 ? 4 : k < 1e-1 ? 5 : 6;
    kap[b] += 1;
    cut += k < cut_at ? 1 : 0;
  }
};

// Wave sums (maxima for the log growth maxima) of the tallies, added to
// out[0 .. 27] by lane 0.  Every lane of the wave takes part.
__device__ void flush_stats(const SweepStats &S, double *out) {
  double v[28];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    v[k] = S.n[k];
    v[5 + k] = S.lg[k];
    v[10 + k] = S.mx[k];
    v[15 + k] = S.big[k];
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) v[20 + k] = S.kap[k];
  v[27] = S.cut;
#pragma unroll
  for (int k = 0; k < 28; ++k) {
    const bool is_max = k >= 10 && k < 15;
    for (int o = 32; o > 0; o >>= 1) {
      const double u = __shfl_down(v[k], o, 64);
      v[k] = is_max ? fmax(v[k], u) : v[k] + u;
    }
  }
  if (__lane_id() == 0)
    for (int k = 0; k < 28; ++k) {
      if (k >= 10 && k < 15) {
        // non-negative doubles order as their bit patterns
        atomicMax((unsigned long long *)&out[k], (unsigned long long)__double_as_longlong(v[k]));
      } else if (v[k] != 0.0) {
        unsafeAtomicAdd(&out[k], v[k]);
      }
    }
}

// Reverse sweep of the tape column `col` from entry `from` down to the
// window's first entry (0), l and mu the adjoints of the state and of the visit
// start.  LEAN: the hand-written reverse step only; at an iteration it does
// not cover the sweep stops and returns that entry's index (l and mu then
// hold the state to resume from).  Otherwise rare iterations use
// dual-number Jacobians.  Returns -1 when the window is done.
template <int NDIM, bool LEAN, bool STATS>
__device__ int64_t sweep(const double *__restrict__ cells, const double *__restrict__ tan,
                         const float *__restrict__ dd3, int64_t ncell, const WalkParams &W,
                         int64_t col, int64_t from, const uint64_t *__restrict__ tape,
                         double l[4], double mu[4],
                         double *g_al, double *g_at, double *g_dm, double &gs,
                         const SweepOpts &O, SweepStats &S) {
  using D = Dual<4>;
  double g[4] = {0.0, 0.0, 0.0, 0.0};
  int cur = 0;
  const bool any = l[0] != 0.0 || l[1] != 0.0 || l[2] != 0.0 || l[3] != 0.0 || mu[0] != 0.0 ||
                   mu[1] != 0.0 || mu[2] != 0.0 || mu[3] != 0.0;
  int64_t stop = -1;
  {
    for (int64_t k = from; any && k >= 0; --k) {
      const TapeEntry E = tape_get(tape, col, k);
      if (E.cell != cur) {
        flush(cur, g, g_al, g_at, g_dm, gs);
        g[0] = g[1] = g[2] = g[3] = 0.0;
        cur = E.cell;
      }
      const int64_t c0 = (int64_t)(E.cell - 1);
      const double *crec = cells + c0 * CELL_DOUBLES, *ctan = tan + c0 * kTan * 4;
      double ln[4] = {0.0, 0.0, 0.0, 0.0}, lo[4] = {l[0], l[1], l[2], l[3]};
      int restart = 0;
      double kappa = 1.0;
      if (reverse_step<NDIM>(crec, ctan, dd3 + c0 * 108, W, E, l, ln, g, O, kappa)) {
        if constexpr (STATS) {
          const bool exited = (E.flags >> 8) & 1, clipped = (E.flags >> 10) & 1;
          S.add(exited ? (kappa < 1e-3 ? 3 : 2) : (clipped ? 1 : 0), lo, ln);
          if (exited) S.exit_kappa(kappa, O.graze_cut);
        }
      } else {
        if constexpr (LEAN) {
          stop = k;
          break;
        } else {
        // a restart, rebound or unmatched exit: dual-number Jacobians
        restart = 0;
        const int rs = (E.flags >> 1) & 15, ef = (E.flags >> 5) & 7;
        // pass 1: directions x y z t, the record held fixed
        VisitG<D> V;
        V.x = dconst<4>(E.x); V.x.d[0] = 1.0;
        V.y = dconst<4>(E.y); V.y.d[1] = 1.0;
        V.z = dconst<4>(E.z); V.z.d[2] = 1.0;
        V.t = dconst<4>(E.t); V.t.d[3] = 1.0;
        V.ix = E.ix; V.iy = E.iy; V.iz = E.iz; V.t0 = E.t0;
        V.restarts = rs;
        V.exitFace = ef;
        Stream rng{E.ra, E.rb};
        bool zero_dt = false;
        const int out = iterate_g<D, RecD<4, 4>, NDIM>(RecD<4, 4>{crec, ctan}, W,
                                                      V, rng, zero_dt, restart);
        // the face coordinate after an exit is the constant face value
        const int ax = (out == IT_EXIT && V.exitFace > 0) ? (V.exitFace - 1) / 2 : -1;
        lo[0] = ax == 0 ? 0.0 : l[0];
        lo[1] = ax == 1 ? 0.0 : l[1];
        lo[2] = ax == 2 ? 0.0 : l[2];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          ln[j] = lo[0] * V.x.d[j] + lo[1] * V.y.d[j] + lo[2] * V.z.d[j] + lo[3] * V.t.d[j];
        // pass 2: directions aLH aTH Dm porosity scale, the state held fixed
        {
          VisitG<D> U;
          U.x = dconst<4>(E.x); U.y = dconst<4>(E.y); U.z = dconst<4>(E.z);
          U.t = dconst<4>(E.t);
          U.ix = E.ix; U.iy = E.iy; U.iz = E.iz; U.t0 = E.t0;
          U.restarts = rs;
          U.exitFace = ef;
          Stream r2{E.ra, E.rb};
          bool z2 = false;
          int rs2 = 0;
          (void)iterate_g<D, RecD<4, 0>, NDIM>(RecD<4, 0>{crec, ctan}, W, U, r2,
                                              z2, rs2);
#pragma unroll
          for (int j = 0; j < 4; ++j)
            g[j] += lo[0] * U.x.d[j] + lo[1] * U.y.d[j] + lo[2] * U.z.d[j] + lo[3] * U.t.d[j];
        }
        if constexpr (STATS) S.add(4, l, ln);
        }
      }
      if (restart == 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) mu[q] += lo[q];
      } else if (restart == 2) {
#pragma unroll
        for (int q = 0; q < 3; ++q) mu[q] += lo[q];
      }
      if (E.flags & 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          ln[q] += mu[q];
          mu[q] = 0.0;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) l[q] = ln[q];
    }
    flush(cur, g, g_al, g_at, g_dm, gs);
  }
  return stop;
}

// Reverse sweep of one window: lam (4 per listed particle: x y z t) holds
// the adjoint of the state after the window on entry and of the state
// before it on return.  The lean pass leaves the particles that reach a
// rare iteration in `deferred` (as thread indices j) with their resume
// entry and visit-start adjoint (resume[j], mus[4j..]); the full pass
// finishes them.  Thread j reverses listed particle ord[j].
template <int NDIM, bool STATS>
__global__ void __launch_bounds__(kBlk) adj_backward(
    const double *__restrict__ cells, const double *__restrict__ tan,
    const float *__restrict__ dd3, int64_t ncell,
    int64_t nsel, const int32_t *__restrict__ ord, const WalkParams W,
    const int64_t *__restrict__ cnt, const int64_t *__restrict__ wstart,
    const uint64_t *__restrict__ tape,
    double *__restrict__ lam, double *__restrict__ mus, int64_t *__restrict__ resume,
    int32_t *__restrict__ deferred, unsigned long long *__restrict__ ndeferred, double *g_al,
    double *g_at, double *g_dm, double *g_s, const SweepOpts O, double *stats) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double gs = 0.0;
  SweepStats S;
  if (j < nsel) {
    const int64_t i = ord[j];
    double l[4] = {lam[4 * i], lam[4 * i + 1], lam[4 * i + 2], lam[4 * i + 3]};
    double mu[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t k =
        sweep<NDIM, true, STATS>(cells, tan, dd3, ncell, W, tape_col(wstart, j), cnt[i] - 1,
                                 tape, l, mu, g_al, g_at, g_dm, gs, O, S);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      lam[4 * i + q] = l[q];
      mus[4 * j + q] = mu[q];
    }
    resume[j] = k;
    if (k >= 0) deferred[atomicAdd(ndeferred, 1ull)] = (int32_t)j;
  }
  for (int o = 32; o > 0; o >>= 1) gs += __shfl_down(gs, o, 64);
  if (__lane_id() == 0 && gs != 0.0) unsafeAtomicAdd(g_s, gs);
  if constexpr (STATS) flush_stats(S, stats);
}

template <int NDIM, bool STATS>
__global__ void __launch_bounds__(kBlk) adj_backward_full(
    const double *__restrict__ cells, const double *__restrict__ tan,
    const float *__restrict__ dd3, int64_t ncell,
    const WalkParams W, const int32_t *__restrict__ ord, const int64_t *__restrict__ wstart,
    const uint64_t *__restrict__ tape, double *__restrict__ lam, const double *__restrict__ mus,
    const int64_t *__restrict__ resume, const int32_t *__restrict__ deferred,
    const unsigned long long *__restrict__ ndef, double *g_al, double *g_at, double *g_dm,
    double *g_s, const SweepOpts O, double *stats) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double gs = 0.0;
  SweepStats S;
  if (t < (int64_t)*ndef) {
    const int64_t j = deferred[t], i = ord[j];
    double l[4] = {lam[4 * i], lam[4 * i + 1], lam[4 * i + 2], lam[4 * i + 3]};
    double mu[4] = {mus[4 * j], mus[4 * j + 1], mus[4 * j + 2], mus[4 * j + 3]};
    (void)sweep<NDIM, false, STATS>(cells, tan, dd3, ncell, W, tape_col(wstart, j), resume[j],
                                    tape, l, mu, g_al, g_at, g_dm, gs, O, S);
#pragma unroll
    for (int q = 0; q < 4; ++q) lam[4 * i + q] = l[q];
  }
  for (int o = 32; o > 0; o >>= 1) gs += __shfl_down(gs, o, 64);
  if (__lane_id() == 0 && gs != 0.0) unsafeAtomicAdd(g_s, gs);
  if constexpr (STATS) flush_stats(S, stats);
}

// Whole-horizon adjoint growth of the listed particles: nrm holds the L1
// norm of each particle's adjoint at the final time; lam the adjoint at
// the first window's start.  out[28] particles, out[29] largest log10
// growth (max with 0), out[30 .. 32] particles growing beyond 1e3, 1e6, 1e9.
__global__ void adj_growth(int64_t nsel, const double *__restrict__ nrm,
                           const double *__restrict__ lam, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < nsel) {
    const double b = fabs(lam[4 * i]) + fabs(lam[4 * i + 1]) + fabs(lam[4 * i + 2]) +
                     fabs(lam[4 * i + 3]);
    v[0] = 1.0;
    if (nrm[i] > 0.0 && b > nrm[i]) {
      const double g = log10(b / nrm[i]);
      v[1] = g;
      v[2] = g > 3.0 ? 1.0 : 0.0;
      v[3] = g > 6.0 ? 1.0 : 0.0;
      v[4] = g > 9.0 ? 1.0 : 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k)
    for (int o = 32; o > 0; o >>= 1) {
      const double u = __shfl_down(v[k], o, 64);
      v[k] = k == 1 ? fmax(v[k], u) : v[k] + u;
    }
  if (__lane_id() == 0) {
    if (v[0] != 0.0) unsafeAtomicAdd(&out[28], v[0]);
    atomicMax((unsigned long long *)&out[29], (unsigned long long)__double_as_longlong(v[1]));
    for (int k = 2; k < 5; ++k)
      if (v[k] != 0.0) unsafeAtomicAdd(&out[28 + k], v[k]);
  }
}

// L1 norms of the listed particles' adjoints.
__global__ void adj_norms(int64_t nsel, const double *__restrict__ lam, double *__restrict__ nrm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsel)
    nrm[i] = fabs(lam[4 * i]) + fabs(lam[4 * i + 1]) + fabs(lam[4 * i + 2]) +
             fabs(lam[4 * i + 3]);
}

// Concentrations at the observation points from the active particles'
// global coordinates: c_o += m_p prod_d N((X_d - x_od) / h_d) / h_d over the
// dimensions with h_d > 0.
__global__ void adj_conc(int64_t n, const int32_t *__restrict__ status,
                         const double *__restrict__ gx, const double *__restrict__ gy,
                         const double *__restrict__ gz, const double *__restrict__ mass,
                         int nobs, const double *__restrict__ obs, double h0, double h1,
                         double h2, double *__restrict__ conc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = p < n && status[p] == 1;
  const double X[3] = {on ? gx[p] : 0.0, on ? gy[p] : 0.0, on ? gz[p] : 0.0};
  const double m = on ? (mass ? mass[p] : 1.0) : 0.0;
  const double h[3] = {h0, h1, h2};
  const double c = 0.3989422804014327;  // 1 / sqrt(2 pi)
  for (int o = 0; o < nobs; ++o) {
    double k = m;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (h[d] > 0.0) {
        const double u = (X[d] - obs[3 * o + d]) / h[d];
        k = k * (c * exp(-0.5 * u * u) / h[d]);
      }
    for (int s = 32; s > 0; s >>= 1) k += __shfl_down(k, s, 64);
    if (__lane_id() == 0 && k != 0.0) unsafeAtomicAdd(&conc[o], k);
  }
}

// Adjoint of the listed particles' local coordinates from the misfit:
// dJ/dX_d = sum_o 2 (c_o - target_o) m K_o (-(X_d - x_od) / h_d^2), times
// dX/dlx = right - left of the cell (ConvertToModelXYZ), and 0 for t.
// Right - left, back - front and top - bottom of 0-based cell c0 as
// global_xyz_kernel forms them.
__device__ __forceinline__ void cell_widths(int c0, int ncol, int nrow,
                                            const double *__restrict__ delr,
                                            const double *__restrict__ delc,
                                            const double *__restrict__ top,
                                            const double *__restrict__ bot, double w[3]) {
  const int lc = nrow * ncol, rr = c0 % lc, ii = rr / ncol, jj = rr % ncol;
  w[0] = (delr[jj] / 2.0) + (delr[jj] / 2.0);
  w[1] = (delc[ii] / 2.0) + (delc[ii] / 2.0);
  w[2] = top[c0] - bot[c0];
}

__global__ void adj_seed(const int32_t *__restrict__ sel, int64_t nsel,
                         const int32_t *__restrict__ cellv, const double *__restrict__ gx,
                         const double *__restrict__ gy, const double *__restrict__ gz,
                         const double *__restrict__ mass, int nobs,
                         const double *__restrict__ obs, const double *__restrict__ resid,
                         double h0, double h1, double h2, int ncol, int nrow,
                         const double *__restrict__ delr, const double *__restrict__ delc,
                         const double *__restrict__ top, const double *__restrict__ bot,
                         double *__restrict__ lam) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsel) return;
  const int p = sel[i];
  const double X[3] = {gx[p], gy[p], gz[p]};
  const double m = mass ? mass[p] : 1.0;
  const double h[3] = {h0, h1, h2};
  const double c = 0.3989422804014327;
  double gX[3] = {0.0, 0.0, 0.0};
  for (int o = 0; o < nobs; ++o) {
    double k = m, u[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (h[d] > 0.0) {
        u[d] = (X[d] - obs[3 * o + d]) / h[d];
        k = k * (c * exp(-0.5 * u[d] * u[d]) / h[d]);
      }
    const double r = 2.0 * resid[o] * k;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (h[d] > 0.0) gX[d] += r * (-u[d] / h[d]);
  }
  double w[3];
  cell_widths(cellv[p] - 1, ncol, nrow, delr, delc, top, bot, w);
  lam[4 * i + 0] = gX[0] * w[0];
  lam[4 * i + 1] = gX[1] * w[1];
  lam[4 * i + 2] = gX[2] * w[2];
  lam[4 * i + 3] = 0.0;
}

// Straight-through restart of the adjoint at a window start (the option
// tangent_windows): the final-time adjoint of the particle's global
// position (lamT, local to its final cell) carried to the local
// coordinates of the cell it starts the window in, the transport from there
// to the final time taken as the identity; time adjoint 0.
__global__ void adj_restart(const int32_t *__restrict__ sel, int64_t nsel,
                            const int32_t *__restrict__ cell_w, const int32_t *__restrict__ cell_T,
                            const double *__restrict__ lamT, int ncol, int nrow,
                            const double *__restrict__ delr, const double *__restrict__ delc,
                            const double *__restrict__ top, const double *__restrict__ bot,
                            double *__restrict__ lam) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nsel) return;
  const int p = sel[i];
  double wT[3], ww[3];
  cell_widths(cell_T[p] - 1, ncol, nrow, delr, delc, top, bot, wT);
  cell_widths(cell_w[p] - 1, ncol, nrow, delr, delc, top, bot, ww);
#pragma unroll
  for (int a = 0; a < 3; ++a) lam[4 * i + a] = lamT[4 * i + a] / wT[a] * ww[a];
  lam[4 * i + 3] = 0.0;
}

// The drift terms' tangents for aLH, aTH and Dm in the reverse step's
// layout: dd3[cell][term j][corner pair q][param k] from the full table,
// in FP32.
__global__ void adj_dd3(int64_t ncell, const double *__restrict__ tan,
                        float *__restrict__ dd3) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncell * 108) return;
  const int64_t c = i / 108;
  const int r = (int)(i % 108), j = r / 12, q = (r % 12) / 3, k = r % 3;
  dd3[i] = (float)tan[(c * kTan + F_DD + 4 * j + q) * 4 + k];
}

__global__ void copy_ck(int64_t n, const int32_t *__restrict__ cell,
                        const int32_t *__restrict__ status, const double *__restrict__ lx,
                        const double *__restrict__ ly, const double *__restrict__ lz,
                        const double *__restrict__ t, const uint64_t *__restrict__ rng,
                        const int64_t *__restrict__ substeps, CkPtrs out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  out.cell[p] = cell[p];
  out.status[p] = status[p];
  out.lx[p] = lx[p];
  out.ly[p] = ly[p];
  out.lz[p] = lz[p];
  out.t[p] = t[p];
  out.rng[2 * p] = rng[2 * p];
  out.rng[2 * p + 1] = rng[2 * p + 1];
  out.substeps[p] = substeps[p];
}

}  // namespace
}  // namespace rwpt

using namespace rwpt;

namespace {

int adj_fail(rwpt_engine *e, int code, const char *msg) {
  e->err = msg;
  return code;
}

int adj_hip(rwpt_engine *e, hipError_t err, const char *where) {
  char buf[256];
  snprintf(buf, sizeof buf, "adjoint %s: %s", where, hipGetErrorString(err));
  e->err = buf;
  return err == hipErrorOutOfMemory ? RWPT_E_NOMEM : RWPT_E_DEVICE;
}

// Device buffers of one rwpt_adjoint call, freed on every return.
struct AdjBuffers {
  std::vector<void *> ptrs;
  template <class T> hipError_t alloc(T **p, size_t count) {
    *p = nullptr;
    hipError_t err = hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T));
    if (err == hipSuccess) ptrs.push_back(*p);
    return err;
  }
  ~AdjBuffers() {
    for (void *q : ptrs) (void)hipFree(q);
  }
};

}  // namespace

extern "C" int rwpt_adjoint_options(rwpt_engine *e, const rwpt_adjoint_opts *o) {
  if (!e || !o) return RWPT_E_ARG;
  if (!(o->graze_cut >= 0.0 && o->graze_cut <= 1.0))
    return adj_fail(e, RWPT_E_ARG, "graze_cut must lie in [0, 1]");
  if (o->reverse_windows < 0) return adj_fail(e, RWPT_E_ARG, "reverse_windows must be >= 0");
  if (o->tangent_windows < 0) return adj_fail(e, RWPT_E_ARG, "tangent_windows must be >= 0");
  if ((o->frozen_noise != 0 && o->frozen_noise != 1) ||
      (o->diagnostics != 0 && o->diagnostics != 1))
    return adj_fail(e, RWPT_E_ARG, "frozen_noise and diagnostics must be 0 or 1");
  e->adj_opt = *o;
  return RWPT_OK;
}

extern "C" int rwpt_adjoint_stats(rwpt_engine *e, double *out) {
  if (!e || !out) return RWPT_E_ARG;
  std::memcpy(out, e->adj_stats, sizeof e->adj_stats);
  return RWPT_OK;
}

extern "C" int rwpt_adjoint(rwpt_engine *e, int32_t n_windows, const double *t_windows,
                            int32_t n_obs, const double *obs_xyz, const double *target,
                            const double *bandwidth, const double *mass, double *misfit,
                            double *conc, double *grad_al, double *grad_at, double *grad_dm,
                            double *grad_porosity_scale, double *ms) {
  if (!e) return RWPT_E_ARG;
  if (e->usg) return adj_fail(e, RWPT_E_UNSUPPORTED, "the adjoint needs a structured grid");
  if (n_windows < 1 || !t_windows) return adj_fail(e, RWPT_E_ARG, "n_windows >= 1 times");
  for (int w = 0; w < n_windows; ++w)
    if (!std::isfinite(t_windows[w])) return adj_fail(e, RWPT_E_ARG, "window time not finite");
  if (n_obs < 1 || !obs_xyz || !target || !bandwidth)
    return adj_fail(e, RWPT_E_ARG, "observation points, targets and bandwidths required");
  for (int d = 0; d < 3; ++d)
    if (!(bandwidth[d] >= 0.0) || !std::isfinite(bandwidth[d]))
      return adj_fail(e, RWPT_E_ARG, "bandwidths must be finite and >= 0");
  if (!(bandwidth[0] > 0.0 || bandwidth[1] > 0.0 || bandwidth[2] > 0.0))
    return adj_fail(e, RWPT_E_ARG, "at least one bandwidth must be > 0");
  const bool want_grad = grad_al || grad_at || grad_dm || grad_porosity_scale;
  if (want_grad) {
    if (e->opt.precision != 0)
      return adj_fail(e, RWPT_E_UNSUPPORTED, "the adjoint needs an FP64 engine");
    if (e->opt.advection_kind != 2)
      return adj_fail(e, RWPT_E_UNSUPPORTED, "the adjoint needs EULERIAN advection");
    if (e->prob.disp_model != 0)
      return adj_fail(e, RWPT_E_UNSUPPORTED, "the adjoint needs the linear dispersion model");
  }
  int ndim = 0;
  for (int d = 0; d < 3; ++d) ndim += e->opt.dimension_mask[d] ? 1 : 0;
  if (want_grad && ndim < 2)
    return adj_fail(e, RWPT_E_UNSUPPORTED, "the adjoint needs a 2-D or 3-D walk");
  Particles &P = e->part;
  DeviceProblem &G = e->prob;
  const int64_t n = P.n;
  hipStream_t st = e->stream;
  AdjBuffers B;
  hipError_t err = hipSuccess;
  const int blk = 256;
  const dim3 pgrid((unsigned)((std::max<int64_t>(n, 1) + blk - 1) / blk));
  const auto t_begin = std::chrono::steady_clock::now();

  // checkpoints at every window start
  std::vector<CkPtrs> ck(want_grad ? n_windows : 0);
  std::vector<WalkParams> Ws(n_windows);
  {
    // one allocation for all checkpoints: per window the 8-byte fields
    // (lx ly lz t, rng, substeps), then cell and status, 256-B aligned
    const size_t stride = ((size_t)std::max<int64_t>(n, 1) * 64 + 255) / 256 * 256;
    char *ckmem = nullptr;
    if (!ck.empty() && (err = B.alloc(&ckmem, stride * ck.size())) != hipSuccess)
      return adj_hip(e, err, "checkpoint memory");
    for (size_t w = 0; w < ck.size(); ++w) {
      char *q = ckmem + w * stride;
      CkPtrs &C = ck[w];
      C.lx = reinterpret_cast<double *>(q);
      C.ly = C.lx + n;
      C.lz = C.ly + n;
      C.t = C.lz + n;
      C.rng = reinterpret_cast<uint64_t *>(C.t + n);
      C.substeps = reinterpret_cast<int64_t *>(C.rng + 2 * n);
      C.cell = reinterpret_cast<int32_t *>(C.substeps + n);
      C.status = C.cell + n;
    }
  }
  // forward: the engine's own advances
  for (int w = 0; w < n_windows; ++w) {
    if (want_grad && n > 0) {
      hipLaunchKernelGGL(copy_ck, pgrid, dim3(blk), 0, st, n, P.cell, P.status, P.lx, P.ly,
                         P.lz, P.t, P.rng, P.substeps, ck[w]);
      if ((err = hipGetLastError()) != hipSuccess) return adj_hip(e, err, "checkpoint");
    }
    Ws[w] = walk_params(e, t_windows[w]);
    const int rc = rwpt_advance(e, t_windows[w], nullptr, nullptr);
    if (rc != RWPT_OK) return rc;
  }
  // misfit
  double *gx, *gy, *gz, *dobs, *dconc, *dmass = nullptr;
  if ((err = B.alloc(&gx, n)) != hipSuccess || (err = B.alloc(&gy, n)) != hipSuccess ||
      (err = B.alloc(&gz, n)) != hipSuccess || (err = B.alloc(&dobs, 3 * (size_t)n_obs)) != hipSuccess ||
      (err = B.alloc(&dconc, n_obs)) != hipSuccess)
    return adj_hip(e, err, "misfit memory");
  if (mass) {
    if ((err = B.alloc(&dmass, n)) != hipSuccess) return adj_hip(e, err, "misfit memory");
    if ((err = hipMemcpyAsync(dmass, mass, n * 8, hipMemcpyHostToDevice, st)) != hipSuccess)
      return adj_hip(e, err, "masses");
  }
  if ((err = hipMemcpyAsync(dobs, obs_xyz, 24 * (size_t)n_obs, hipMemcpyHostToDevice, st)) !=
          hipSuccess ||
      (err = hipMemsetAsync(dconc, 0, 8 * (size_t)n_obs, st)) != hipSuccess)
    return adj_hip(e, err, "observations");
  if (n > 0) {
    if ((err = launch_global_xyz(e, gx, gy, gz)) != hipSuccess)
      return adj_hip(e, err, "global coordinates");
    hipLaunchKernelGGL(adj_conc, pgrid, dim3(blk), 0, st, n, (const int32_t *)P.status,
                       (const double *)gx, (const double *)gy, (const double *)gz,
                       (const double *)dmass, (int)n_obs, (const double *)dobs, bandwidth[0],
                       bandwidth[1], bandwidth[2], dconc);
    if ((err = hipGetLastError()) != hipSuccess) return adj_hip(e, err, "concentrations");
  }
  std::vector<double> hc(n_obs), resid(n_obs);
  if ((err = hipMemcpyAsync(hc.data(), dconc, 8 * (size_t)n_obs, hipMemcpyDeviceToHost, st)) !=
          hipSuccess ||
      (err = hipStreamSynchronize(st)) != hipSuccess)
    return adj_hip(e, err, "concentrations");
  double J = 0.0;
  for (int o = 0; o < n_obs; ++o) {
    resid[o] = hc[o] - target[o];
    J += resid[o] * resid[o];
  }
  if (misfit) *misfit = J;
  if (conc) std::memcpy(conc, hc.data(), 8 * (size_t)n_obs);
  if (!want_grad) {
    if (ms)
      *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                      t_begin).count();
    return RWPT_OK;
  }

  // reverse: the particles active at the final time carry the adjoint
  int32_t *sel;
  unsigned long long *nsel_d, nsel_h = 0;
  if ((err = B.alloc(&sel, n)) != hipSuccess || (err = B.alloc(&nsel_d, 1)) != hipSuccess ||
      (err = hipMemsetAsync(nsel_d, 0, 8, st)) != hipSuccess)
    return adj_hip(e, err, "adjoint memory");
  if (n > 0) {
    hipLaunchKernelGGL(adj_select, pgrid, dim3(blk), 0, st, n, (const int32_t *)P.status, sel,
                       nsel_d);
    if ((err = hipGetLastError()) != hipSuccess) return adj_hip(e, err, "statuses");
  }
  if ((err = hipMemcpyAsync(&nsel_h, nsel_d, 8, hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (err = hipStreamSynchronize(st)) != hipSuccess)
    return adj_hip(e, err, "statuses");
  const int64_t nsel = (int64_t)nsel_h;
  double *lam, *tan, *resid_d, *g_al, *g_at, *g_dm, *g_s;
  int64_t *wlen, *wstart, *resume;
  const int64_t nwaves = (nsel + kBlk - 1) / kBlk;
  double *mus;
  int32_t *deferred;
  const int64_t ncell = G.ncell;
  if ((err = B.alloc(&lam, 4 * nsel)) != hipSuccess ||
      (err = B.alloc(&tan, 4 * (size_t)ncell * kTan)) != hipSuccess ||
      (err = B.alloc(&resid_d, n_obs)) != hipSuccess || (err = B.alloc(&g_al, ncell)) != hipSuccess ||
      (err = B.alloc(&g_at, ncell)) != hipSuccess || (err = B.alloc(&g_dm, ncell)) != hipSuccess ||
      (err = B.alloc(&g_s, 1)) != hipSuccess ||
      (err = B.alloc(&wlen, nwaves + 1)) != hipSuccess ||
      (err = B.alloc(&wstart, nwaves + 1)) != hipSuccess ||
      (err = B.alloc(&resume, nsel)) != hipSuccess || (err = B.alloc(&mus, 4 * nsel)) != hipSuccess ||
      (err = B.alloc(&deferred, nsel)) != hipSuccess)
    return adj_hip(e, err, "adjoint memory");
  if ((err = hipMemcpyAsync(resid_d, resid.data(), 8 * (size_t)n_obs, hipMemcpyHostToDevice,
                            st)) != hipSuccess ||
      (err = hipMemsetAsync(g_al, 0, 8 * ncell, st)) != hipSuccess ||
      (err = hipMemsetAsync(g_at, 0, 8 * ncell, st)) != hipSuccess ||
      (err = hipMemsetAsync(g_dm, 0, 8 * ncell, st)) != hipSuccess ||
      (err = hipMemsetAsync(g_s, 0, 8, st)) != hipSuccess)
    return adj_hip(e, err, "adjoint set-up");
  if ((err = build_record_tangents(e, tan)) != hipSuccess)
    return adj_hip(e, err, "record tangents");
  float *dd3;
  if ((err = B.alloc(&dd3, (size_t)ncell * 108)) != hipSuccess)
    return adj_hip(e, err, "adjoint memory");
  hipLaunchKernelGGL(adj_dd3, dim3((unsigned)((ncell * 108 + blk - 1) / blk)), dim3(blk), 0, st,
                     ncell, (const double *)tan, dd3);
  if ((err = hipGetLastError()) != hipSuccess) return adj_hip(e, err, "record tangents");
  const dim3 sgrid((unsigned)((std::max<int64_t>(nsel, 1) + kBlk - 1) / kBlk));
  if (nsel > 0) {
    hipLaunchKernelGGL(adj_seed, sgrid, dim3(kBlk), 0, st, (const int32_t *)sel, nsel,
                       (const int32_t *)P.cell, (const double *)gx, (const double *)gy,
                       (const double *)gz, (const double *)dmass, (int)n_obs,
                       (const double *)dobs, (const double *)resid_d, bandwidth[0],
                       bandwidth[1], bandwidth[2], G.ncol, G.nrow, (const double *)G.delr,
                       (const double *)G.delc, (const double *)G.top, (const double *)G.bot,
                       lam);
    if ((err = hipGetLastError()) != hipSuccess) return adj_hip(e, err, "seed");
  }
  int64_t nlive = 0;
  if (nsel > 0) {
    int32_t *sel2;
    double *lam2;
    unsigned long long *nk;
    unsigned long long hk = 0;
    if ((err = B.alloc(&sel2, nsel)) != hipSuccess ||
        (err = B.alloc(&lam2, 4 * nsel)) != hipSuccess || (err = B.alloc(&nk, 1)) != hipSuccess ||
        (err = hipMemsetAsync(nk, 0, 8, st)) != hipSuccess)
      return adj_hip(e, err, "adjoint memory");
    hipLaunchKernelGGL(adj_keep, sgrid, dim3(kBlk), 0, st, (const int32_t *)sel,
                       (const double *)lam, nsel, sel2, lam2, nk);
    if ((err = hipGetLastError()) != hipSuccess ||
        (err = hipMemcpyAsync(&hk, nk, 8, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (err = hipStreamSynchronize(st)) != hipSuccess)
      return adj_hip(e, err, "seed");
    nlive = (int64_t)hk;
    sel = sel2;
    lam = lam2;
  }
  double *lamT = nullptr;  // the final-time adjoints, for adj_restart
  if (e->adj_opt.tangent_windows > 0 &&
      ((err = B.alloc(&lamT, 4 * std::max<int64_t>(nlive, 1))) != hipSuccess ||
       (nlive > 0 && (err = hipMemcpyAsync(lamT, lam, 32 * nlive, hipMemcpyDeviceToDevice,
                                           st)) != hipSuccess)))
    return adj_hip(e, err, "adjoint memory");
  const int64_t nwl = (nlive + kBlk - 1) / kBlk;
  const dim3 lgrid((unsigned)((std::max<int64_t>(nlive, 1) + kBlk - 1) / kBlk));
  // Per window, before the reverse pass: the listed particles' iteration
  // counts, their order (iteration count, then starting cell) and the tape
  // blocks of the wavefronts.  The windows' replays and sweeps are then
  // queued back to back with no host round trip: one tape of the largest
  // window's size, replay mismatches counted per window and checked at the
  // end, and the full pass reads its particle count on the device.
  int cell_bits = 1;
  while (cell_bits < 24 && (int64_t(1) << cell_bits) <= ncell) ++cell_bits;
  std::vector<int32_t *> ords(n_windows, nullptr);
  std::vector<int64_t *> cnts(n_windows, nullptr), wstarts(n_windows, nullptr);
  int32_t *okey, *okey2, *oidx;
  int64_t *totals;
  unsigned long long *bads, *ndefs;
  char *sort_tmp = nullptr, *scan_tmp = nullptr;
  size_t sort_bytes = 0, scan_bytes = 0;
  if ((err = B.alloc(&okey, nlive)) != hipSuccess || (err = B.alloc(&okey2, nlive)) != hipSuccess ||
      (err = B.alloc(&oidx, nlive)) != hipSuccess ||
      (err = B.alloc(&totals, n_windows)) != hipSuccess ||
      (err = B.alloc(&bads, n_windows)) != hipSuccess ||
      (err = B.alloc(&ndefs, n_windows)) != hipSuccess ||
      (err = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, sort_bytes, okey, okey2, oidx,
                                                          oidx, (int)nlive, 0, 31, st)) !=
          hipSuccess ||
      (err = B.alloc(&sort_tmp, sort_bytes)) != hipSuccess ||
      (err = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, wlen, wstart, (int)(nwl + 1),
                                              st)) != hipSuccess ||
      (err = B.alloc(&scan_tmp, scan_bytes)) != hipSuccess ||
      (err = hipMemsetAsync(totals, 0, 8 * (size_t)n_windows, st)) != hipSuccess ||
      (err = hipMemsetAsync(bads, 0, 8 * (size_t)n_windows, st)) != hipSuccess ||
      (err = hipMemsetAsync(ndefs, 0, 8 * (size_t)n_windows, st)) != hipSuccess)
    return adj_hip(e, err, "order memory");
  const CkPtrs now{P.cell, P.status, P.lx, P.ly, P.lz, P.t, P.rng, P.substeps};
  int32_t *ord_mem;
  int64_t *cnt_mem, *ws_mem;
  if ((err = B.alloc(&ord_mem, (size_t)n_windows * nlive)) != hipSuccess ||
      (err = B.alloc(&cnt_mem, (size_t)n_windows * (nlive + 1))) != hipSuccess ||
      (err = B.alloc(&ws_mem, (size_t)n_windows * (nwl + 1))) != hipSuccess)
    return adj_hip(e, err, "order memory");
  // the sweep covers the windows from w_first on (all with reverse_windows 0)
  const int w_first =
      e->adj_opt.reverse_windows > 0 ? std::max(0, n_windows - e->adj_opt.reverse_windows)
                                    : 0;
  for (int w = w_first; w < n_windows && nlive > 0; ++w) {
    const CkPtrs &in = ck[w];
    const CkPtrs &want = w + 1 < n_windows ? ck[w + 1] : now;
    ords[w] = ord_mem + (size_t)w * nlive;
    cnts[w] = cnt_mem + (size_t)w * (nlive + 1);
    wstarts[w] = ws_mem + (size_t)w * (nwl + 1);
    hipLaunchKernelGGL(adj_counts, dim3((unsigned)((nlive + 1 + blk - 1) / blk)), dim3(blk), 0,
                       st, (const int32_t *)sel, nlive, (const int64_t *)in.substeps,
                       (const int64_t *)want.substeps, cnts[w]);
    hipLaunchKernelGGL(adj_order_keys, lgrid, dim3(kBlk), 0, st, (const int32_t *)sel, nlive,
                       (const int32_t *)in.cell, (const int64_t *)cnts[w], cell_bits, okey,
                       oidx);
    size_t sb = sort_bytes, cb = scan_bytes;
    if ((err = hipGetLastError()) != hipSuccess ||
        (err = hipcub::DeviceRadixSort::SortPairsDescending(sort_tmp, sb, okey, okey2, oidx,
                                                            ords[w], (int)nlive, 0, 31, st)) !=
            hipSuccess)
      return adj_hip(e, err, "particle order");
    hipLaunchKernelGGL(adj_wave_len, dim3((unsigned)((nwl + 1 + blk - 1) / blk)), dim3(blk), 0,
                       st, (const int32_t *)ords[w], nlive, (const int64_t *)cnts[w], wlen);
    if ((err = hipGetLastError()) != hipSuccess ||
        (err = hipcub::DeviceScan::ExclusiveSum(scan_tmp, cb, wlen, wstarts[w], (int)(nwl + 1),
                                                st)) != hipSuccess ||
        (err = hipMemcpyAsync(totals + w, wstarts[w] + nwl, 8, hipMemcpyDeviceToDevice, st)) !=
            hipSuccess)
      return adj_hip(e, err, "tape offsets");
  }
  std::vector<int64_t> htotal(n_windows, 0);
  if ((err = hipMemcpyAsync(htotal.data(), totals, 8 * (size_t)n_windows,
                            hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (err = hipStreamSynchronize(st)) != hipSuccess)
    return adj_hip(e, err, "tape size");
  const int64_t most = *std::max_element(htotal.begin(), htotal.end());
  uint64_t *tape = nullptr;
  if ((err = B.alloc(&tape, (size_t)most * kTapeWords * kBlk)) != hipSuccess)
    return adj_hip(e, err, "tape memory");
  double *ga = grad_al ? g_al : nullptr, *gt = grad_at ? g_at : nullptr,
         *gd = grad_dm ? g_dm : nullptr;
  // diagnostics: the sweep's tallies and the adjoint norms at the final time
  double *dstats = nullptr, *nrm0 = nullptr;
  if (e->adj_opt.diagnostics) {
    if ((err = B.alloc(&dstats, RWPT_ADJ_NSTATS)) != hipSuccess ||
        (err = B.alloc(&nrm0, std::max<int64_t>(nlive, 1))) != hipSuccess ||
        (err = hipMemsetAsync(dstats, 0, 8 * RWPT_ADJ_NSTATS, st)) != hipSuccess)
      return adj_hip(e, err, "diagnostics memory");
    if (nlive > 0) hipLaunchKernelGGL(adj_norms, lgrid, dim3(kBlk), 0, st, nlive,
                                      (const double *)lam, nrm0);
  }
  for (int w = n_windows - 1; w >= w_first && nlive > 0; --w) {
    const CkPtrs &in = ck[w];
    const CkPtrs &want = w + 1 < n_windows ? ck[w + 1] : now;
    const WalkParams &W = Ws[w];
    const int32_t *ord = ords[w];
    const int64_t *cw = cnts[w], *ws = wstarts[w];
    if (ndim == 3)
      hipLaunchKernelGGL(adj_replay<3>, lgrid, dim3(kBlk), 0, st, (const double *)G.cells,
                         (const int32_t *)sel, nlive, ord, in, want, W, cw, ws, tape, bads + w);
    else
      hipLaunchKernelGGL(adj_replay<2>, lgrid, dim3(kBlk), 0, st, (const double *)G.cells,
                         (const int32_t *)sel, nlive, ord, in, want, W, cw, ws, tape, bads + w);
    const bool diag = e->adj_opt.diagnostics != 0;
    const SweepOpts sopt{e->adj_opt.graze_cut, e->adj_opt.frozen_noise};
#define RWPT_ADJ_SWEEP(ND, ST)                                                                   \
  hipLaunchKernelGGL((adj_backward<ND, ST>), lgrid, dim3(kBlk), 0, st, (const double *)G.cells,  \
                     (const double *)tan, (const float *)dd3, ncell, nlive, ord, W, cw, ws,     \
                     (const uint64_t *)tape, lam, mus, resume, deferred, ndefs + w, ga, gt, gd,  \
                     g_s, sopt, dstats);                                                         \
  hipLaunchKernelGGL((adj_backward_full<ND, ST>), lgrid, dim3(kBlk), 0, st,                      \
                     (const double *)G.cells, (const double *)tan, (const float *)dd3, ncell,   \
                     W, ord, ws, (const uint64_t *)tape, lam, (const double *)mus,               \
                     (const int64_t *)resume, (const int32_t *)deferred,                         \
                     (const unsigned long long *)(ndefs + w), ga, gt, gd, g_s, sopt, dstats)
    if (ndim == 3 && diag) {
      RWPT_ADJ_SWEEP(3, true);
    } else if (ndim == 3) {
      RWPT_ADJ_SWEEP(3, false);
    } else if (diag) {
      RWPT_ADJ_SWEEP(2, true);
    } else {
      RWPT_ADJ_SWEEP(2, false);
    }
#undef RWPT_ADJ_SWEEP
    const int tw = e->adj_opt.tangent_windows;
    if (tw > 0 && w > w_first && (n_windows - w) % tw == 0)
      hipLaunchKernelGGL(adj_restart, lgrid, dim3(kBlk), 0, st, (const int32_t *)sel, nlive,
                         (const int32_t *)in.cell, (const int32_t *)now.cell,
                         (const double *)lamT, G.ncol, G.nrow, (const double *)G.delr,
                         (const double *)G.delc, (const double *)G.top, (const double *)G.bot,
                         lam);
    if ((err = hipGetLastError()) != hipSuccess) return adj_hip(e, err, "reverse sweep");
  }
  if (e->adj_opt.diagnostics) {
    if (nlive > 0) hipLaunchKernelGGL(adj_growth, lgrid, dim3(kBlk), 0, st, nlive,
                                      (const double *)nrm0, (const double *)lam, dstats);
    if ((err = hipGetLastError()) != hipSuccess ||
        (err = hipMemcpyAsync(e->adj_stats, dstats, 8 * RWPT_ADJ_NSTATS, hipMemcpyDeviceToHost,
                              st)) != hipSuccess)
      return adj_hip(e, err, "diagnostics");
  }
  std::vector<unsigned long long> hbad(n_windows, 0);
  if ((err = hipMemcpyAsync(hbad.data(), bads, 8 * (size_t)n_windows, hipMemcpyDeviceToHost,
                            st)) != hipSuccess ||
      (err = hipStreamSynchronize(st)) != hipSuccess)
    return adj_hip(e, err, "replay");
  for (int w = 0; w < n_windows; ++w)
    if (hbad[w] != 0) {
      char buf[160];
      snprintf(buf, sizeof buf,
               "adjoint: the replay of window %d differs from the forward walk for %llu "
               "particles",
               w, hbad[w]);
      e->err = buf;
      return RWPT_E_STATE;
    }
  if ((grad_al && (err = hipMemcpyAsync(grad_al, g_al, 8 * ncell, hipMemcpyDeviceToHost, st)) !=
                      hipSuccess) ||
      (grad_at && (err = hipMemcpyAsync(grad_at, g_at, 8 * ncell, hipMemcpyDeviceToHost, st)) !=
                      hipSuccess) ||
      (grad_dm && (err = hipMemcpyAsync(grad_dm, g_dm, 8 * ncell, hipMemcpyDeviceToHost, st)) !=
                      hipSuccess) ||
      (grad_porosity_scale &&
       (err = hipMemcpyAsync(grad_porosity_scale, g_s, 8, hipMemcpyDeviceToHost, st)) !=
           hipSuccess) ||
      (err = hipStreamSynchronize(st)) != hipSuccess)
    return adj_hip(e, err, "gradients");
  if (ms)
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_begin)
              .count();
  return RWPT_OK;
}
