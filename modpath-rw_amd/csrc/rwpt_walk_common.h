// Pieces of the walk shared by the advance kernel (rwpt_walk.hip) and the
// adjoint (rwpt_adjoint.hip): the rng streams and normal deviates, the
// record accessors, the cell-entry checks and the packed loop counters.
#pragma once

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

// One xoroshiro128+ step on 32-bit halves: out = a + b; with s2 = b ^ a,
// a' = rotl(a, 24) ^ s2 ^ (s2 << 16) and b' = rotl(s2, 37).  Each 64-bit
// rotation or shift is two funnel shifts (v_alignbit_b32) and each
// three-way xor one v_bitop3_b32, the same bits as the 64-bit expressions.
__device__ __forceinline__ uint64_t draw_bits(Stream &s) {
  const uint32_t alo = (uint32_t)s.a, ahi = (uint32_t)(s.a >> 32);
  const uint32_t xlo = (uint32_t)s.b ^ alo, xhi = (uint32_t)(s.b >> 32) ^ ahi;
  const uint64_t out = s.a + s.b;
  const uint32_t rhi = __builtin_amdgcn_alignbit(ahi, alo, 8);   // rotl(a, 24)
  const uint32_t rlo = __builtin_amdgcn_alignbit(alo, ahi, 8);
  const uint32_t shi = __builtin_amdgcn_alignbit(xhi, xlo, 16);  // s2 << 16
  const uint32_t slo = xlo << 16;
  const uint32_t ahi2 = __builtin_amdgcn_bitop3_b32(rhi, xhi, shi, 0x96);  // x ^ y ^ z
  const uint32_t alo2 = __builtin_amdgcn_bitop3_b32(rlo, xlo, slo, 0x96);
  const uint32_t bhi2 = __builtin_amdgcn_alignbit(xlo, xhi, 27);  // rotl(s2, 37)
  const uint32_t blo2 = __builtin_amdgcn_alignbit(xhi, xlo, 27);
  s.a = ((uint64_t)ahi2 << 32) | alo2;
  s.b = ((uint64_t)bhi2 << 32) | blo2;
  return out;
}

__device__ __forceinline__ double draw_uniform(Stream &s) {
  const uint64_t out = draw_bits(s);
  return __longlong_as_double(
             (long long)(0x3FF0000000000000ull | (out >> 12))) - 1.0;
}

// Normal deviates by arithmetic type and draw policy.
//  NRM 0 (FP64): pr_GenerateStandardNormalRandomZig, 12 uniforms summed in
//        order, minus 6 -- the reference's numbers bit for bit.
//  NRM 1 (FP32): the same 12 draws; the top 24 bits of each summed exactly
//        as integers, one rounding to float: within 5e-7 of the FP64 normal,
//        so an FP32 walk follows the FP64 walk's random path.
//  NRM 2 (FP32): Irwin-Hall of 12 uniforms of 16 bits, three from bits
//        16..63 of each of 4 draws; (sum + 6)/65536 - 6 is exact up to the
//        final rounding and has mean 0 and variance 1 - 2^-32.
template <typename R, int NRM> struct Normal;
template <> struct Normal<double, 0> {
  static __device__ __forceinline__ double draw(Stream &s) {
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < 12; ++m) acc = acc + draw_uniform(s);
    return acc - 6.0;
  }
};
template <> struct Normal<float, 1> {
  static __device__ __forceinline__ float draw(Stream &s) {
    uint32_t acc = 0;
#pragma unroll
    for (int m = 0; m < 12; ++m) acc += (uint32_t)(draw_bits(s) >> 40);
    return (float)acc * 0x1p-24f - 6.0f;
  }
};
template <> struct Normal<float, 2> {
  static __device__ __forceinline__ float draw(Stream &s) {
    uint32_t acc = 0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const uint64_t w = draw_bits(s);
      const uint32_t hi = (uint32_t)(w >> 32), mid = (uint32_t)w >> 16;
      acc += (hi & 0xFFFFu) + (hi >> 16) + mid;
    }
    return ((float)acc + 6.0f) * 0x1p-16f - 6.0f;
  }
};

__device__ __forceinline__ double f32lit(float v) { return (double)v; }

template <typename R>
__device__ __forceinline__ const int32_t *ints(const R *rec) {
  return reinterpret_cast<const int32_t *>(rec);
}

// Guards against walks the reference never ends (not in the reference):
// sub-steps per particle and launch, and consecutive cell exits made with a
// zero time step (a particle on a face whose drift points into the face from
// both sides, with no random component across it, ping-pongs between the two
// cells forever in pr_RWPTrackPath).  Either reports status 9.
constexpr int64_t kMaxSteps = 20000000;
constexpr int kMaxStalls = 1000;

template <typename R>
__device__ __forceinline__ bool outside(R x, R y, R z) {
  return (x > R(1)) || (x < R(0)) || (y > R(1)) || (y < R(0)) || (z > R(1)) ||
         (z < R(0));
}

// outcome of one time-loop iteration of a cell visit
enum : int { IT_CONTINUE = 0, IT_EXIT = 2, IT_MAXTIME = 3, IT_UNDEFINED = 4 };

// Checks TrackCell makes before tracking (TrackCell.f90:375-510) and the
// early returns of TrackSubCell (:870-945).  Returns 0 to track, else the
// particle status to report at the cell-entry location.
template <typename R>
__device__ __forceinline__ int entry_checks(const R *__restrict__ rec,
                                            const WalkParams &W, R x, R y, R z) {
  if (outside(x, y, z)) return 9;  // InvalidLocation -> Undefined
  const int32_t *fl = ints(rec) + Lay<R>::FLAGS;
  if (fl[0] == 0) return 7;
  if (W.stop_zone > 0 && ints(rec)[Lay<R>::ZONE] == W.stop_zone) return 6;
  const int32_t b = fl[1];
  const bool src = (b & CF_SOURCE) != 0, snk = (b & CF_SINK) != 0;
  bool stopIfNoExit = true;
  if (W.backward) {
    if (W.stop_weak_sources) {
      if (src) return 4;
      if (!W.steady_state) stopIfNoExit = false;
    }
  } else if (W.stop_weak_sinks) {
    if (snk) return 3;
    if (!W.steady_state) stopIfNoExit = false;
  }
  if (!(b & CF_CELL_EXIT)) {
    if (W.steady_state || (W.backward && src) || (!W.backward && snk)) return 5;
  }
  if (stopIfNoExit && !(b & CF_SUBCELL_EXIT)) return 5;
  if (b & CF_DRY) return 7;
  if (b & CF_DT_NONPOS) return 9;
  return 0;
}

// Packed small per-particle loop counters.
struct Aux {
  int phase, restarts, exitFace, entryFace, stalls;
};
__device__ __forceinline__ Aux unpack_aux(int32_t a) {
  return Aux{a & 1, (a >> 1) & 15, (a >> 5) & 7, (a >> 8) & 7, (a >> 11) & 2047};
}
__device__ __forceinline__ int32_t pack_aux(const Aux &a) {
  return a.phase | (a.restarts << 1) | (a.exitFace << 5) | (a.entryFace << 8) |
         (a.stalls << 11);
}

// The last pathline point pr_RWPTrackPath records before an Undefined
// result, given the entry location (x, y, z, t) of the current cell: the
// advance's start location itself, or after a transfer the exit point in
// the previous cell (the entry point seen from the other side).
template <typename R>
__device__ __forceinline__ void last_point(const R *__restrict__ cells,
                                           int entryFace, int &c, R &x, R &y,
                                           R &z) {
  if (entryFace == 0) return;
  const R *rec = cells + (int64_t)(c - 1) * Lay<R>::STRIDE;
  c = ints(rec)[Lay<R>::CONN + entryFace - 1];
  switch (entryFace) {
    case 1: x = R(1); break;
    case 2: x = R(0); break;
    case 3: y = R(1); break;
    case 4: y = R(0); break;
    case 5: z = R(1); break;
    default: z = R(0); break;
  }
}

}  // namespace
}  // namespace rwpt
