/* Dual numbers for forward-mode tangents through the oracle (test
 * infrastructure, no part of the product): a value and ORC_ND directional
 * derivatives.  Every operation computes its value exactly as the double
 * operation it replaces (same operands, same order; the oracle is built
 * with -ffp-contract=off), so the branches of a dual run are those of the
 * FP64 oracle; comparisons look at values only. */
#pragma once
#include <cmath>

#ifndef ORC_ND
#define ORC_ND 4
#endif

struct orc_dual {
  double v;
  double d[ORC_ND];
  orc_dual() : v(0.0) {
    for (int k = 0; k < ORC_ND; ++k) d[k] = 0.0;
  }
  orc_dual(double x) : v(x) {
    for (int k = 0; k < ORC_ND; ++k) d[k] = 0.0;
  }
  explicit operator int() const { return (int)v; }
  explicit operator long() const { return (long)v; }
  explicit operator float() const { return (float)v; }
  orc_dual &operator+=(const orc_dual &b) { return *this = *this + b; }
  orc_dual &operator-=(const orc_dual &b) { return *this = *this - b; }
  orc_dual &operator*=(const orc_dual &b) { return *this = *this * b; }
  orc_dual &operator/=(const orc_dual &b) { return *this = *this / b; }

  friend orc_dual operator+(const orc_dual &a, const orc_dual &b) {
    orc_dual r(a.v + b.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
  }
  friend orc_dual operator-(const orc_dual &a, const orc_dual &b) {
    orc_dual r(a.v - b.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
  }
  friend orc_dual operator-(const orc_dual &a) {
    orc_dual r(-a.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = -a.d[k];
    return r;
  }
  friend orc_dual operator+(const orc_dual &a) { return a; }
  friend orc_dual operator*(const orc_dual &a, const orc_dual &b) {
    orc_dual r(a.v * b.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
  }
  friend orc_dual operator/(const orc_dual &a, const orc_dual &b) {
    orc_dual r(a.v / b.v);
    for (int k = 0; k < ORC_ND; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
  }
  friend bool operator<(const orc_dual &a, const orc_dual &b) { return a.v < b.v; }
  friend bool operator>(const orc_dual &a, const orc_dual &b) { return a.v > b.v; }
  friend bool operator<=(const orc_dual &a, const orc_dual &b) { return a.v <= b.v; }
  friend bool operator>=(const orc_dual &a, const orc_dual &b) { return a.v >= b.v; }
  friend bool operator==(const orc_dual &a, const orc_dual &b) { return a.v == b.v; }
  friend bool operator!=(const orc_dual &a, const orc_dual &b) { return a.v != b.v; }
};

/* f(a) with f'(a) = g: tangents scaled by g, left zero where they are zero
 * (so that a value where f' is infinite, sqrt at 0, stays finite when the
 * argument does not depend on the parameters) */
static inline orc_dual orc_chain(const orc_dual &a, double fv, double g) {
  orc_dual r(fv);
  for (int k = 0; k < ORC_ND; ++k) r.d[k] = a.d[k] == 0.0 ? 0.0 : g * a.d[k];
  return r;
}
static inline orc_dual sqrt(const orc_dual &a) {
  const double s = std::sqrt(a.v);
  return orc_chain(a, s, 0.5 / s);
}
static inline orc_dual exp(const orc_dual &a) {
  const double e = std::exp(a.v);
  return orc_chain(a, e, e);
}
static inline orc_dual log(const orc_dual &a) { return orc_chain(a, std::log(a.v), 1.0 / a.v); }
static inline orc_dual log10(const orc_dual &a) {
  return orc_chain(a, std::log10(a.v), 1.0 / (a.v * std::log(10.0)));
}
static inline orc_dual fabs(const orc_dual &a) { return a.v < 0.0 ? -a : a; }
static inline orc_dual fmin(const orc_dual &a, const orc_dual &b) { return b.v < a.v ? b : a; }
static inline orc_dual fmax(const orc_dual &a, const orc_dual &b) { return b.v > a.v ? b : a; }
static inline orc_dual floor(const orc_dual &a) { return orc_dual(std::floor(a.v)); }
static inline orc_dual ceil(const orc_dual &a) { return orc_dual(std::ceil(a.v)); }
static inline orc_dual erf(const orc_dual &a) {
  return orc_chain(a, std::erf(a.v), 2.0 / std::sqrt(M_PI) * std::exp(-a.v * a.v));
}
static inline orc_dual pow(const orc_dual &a, const orc_dual &b) {
  const double p = std::pow(a.v, b.v);
  orc_dual r(p);
  for (int k = 0; k < ORC_ND; ++k) {
    double t = 0.0;
    if (a.d[k] != 0.0) t += b.v * std::pow(a.v, b.v - 1.0) * a.d[k];
    if (b.d[k] != 0.0) t += p * std::log(a.v) * b.d[k];
    r.d[k] = t;
  }
  return r;
}
static inline bool isnan(const orc_dual &a) { return std::isnan(a.v); }
static inline bool isfinite(const orc_dual &a) { return std::isfinite(a.v); }
