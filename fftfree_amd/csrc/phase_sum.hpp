// phase_sum.hpp -- the double-precision running sum phi_f = fl(phi_{f-1} + d), phi_0 = 0,
// for any f in a number of steps that grows with the number of binades phi crosses, not
// with f. This is the linear phase accumulator of phase inference (phase_infer.hpp:47-61:
// phase[k] += dphi[k] once per frame, in double).
//
// Why it is exact: while phi stays in one binade [2^(e-1), 2^e), every representable value
// there is P * u with u = 2^(e-53) and P an integer, so fl(phi + d) = u * rne(P + d / u).
// If d / u = q is not half-way between integers, every addition in the binade adds
// round(q). If it is (q = n + 1/2), round-half-even picks the even one of P + n and
// P + n + 1: after one literal addition P is even, and from then on every addition adds
// the even one of n and n + 1. Either way m additions that stay in the binade add
// m * S * u, a multiple of u below 2^e, which double arithmetic forms exactly. The one
// addition that leaves the binade is made as fl(phi + d) itself.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define FFTFREE_HD __host__ __device__
#else
#define FFTFREE_HD
#endif

namespace fftfree {

FFTFREE_HD inline double running_sum(double d, uint64_t f) {
  if (f == 0 || d == 0.0) return 0.0;
  const bool negative = d < 0.0;            // round-to-nearest-even is symmetric
  d = fabs(d);
  double phi = d;                           // fl(0 + d)
  uint64_t k = 1;
  while (k < f) {
    if (!(phi <= 1.7976931348623157e308)) break;  // overflowed: inf + d stays inf
    int e = 0;
    (void)frexp(phi, &e);                          // phi in [2^(e-1), 2^e)
    const int se = 53 - e < 1074 ? 53 - e : 1074;  // u = 2^-se (subnormal floor)
    const double q = ldexp(d, se);                 // d / u, exact
    const double n = floor(q);
    const bool half = q - n == 0.5;
    if (half) {                                    // one literal addition makes P even
      phi = phi + d;
      ++k;
      if (k == f) break;
      (void)frexp(phi, &e);
      if (53 - e < 1074 ? 53 - e != se : se != 1074) continue;  // left the binade: start over
    }
    double S = q - n < 0.5 ? n : n + 1.0;
    if (half) S = fmod(n, 2.0) == 0.0 ? n : n + 1.0;
    if (S == 0.0) break;                           // phi no longer moves
    // room below 2^e in units of u (2^1024 itself is not a double: halve both there)
    const double P = e < 1024 ? ldexp(ldexp(1.0, e) - phi, se) : ldexp(ldexp(1.0, e - 1) - 0.5 * phi, se + 1);
    // m = floor((P - 1) / S): the rounded quotient can only land on the next integer up
    uint64_t m = (uint64_t)floor((P - 1.0) / S);
    if (m * (uint64_t)S > (uint64_t)P - 1) --m;
    const uint64_t take = m < f - k ? m : f - k;
    phi += (double)take * ldexp(S, -se);
    k += take;
    if (k < f) {                                   // the addition that reaches the next binade
      phi = phi + d;
      ++k;
    }
  }
  return negative ? -phi : phi;
}

// The additions from phi on that all add the same exact amount: after this call, the next
// `run` additions of d each add `step` (a multiple of the binade's spacing), and the one
// after them must be made literally as fl(phi + d). For d >= 0 and finite phi >= 0 (the
// linear phase); run == 0 means the next addition is literal. With S == 0, phi never moves
// again and run is UINT64_MAX.
FFTFREE_HD inline void running_segment(double phi, double d, double* step, uint64_t* run) {
  *step = 0.0;
  *run = 0;
  if (d == 0.0) {
    *run = UINT64_MAX;
    return;
  }
  if (phi == 0.0) return;  // fl(0 + d) = d: literal
  int e = 0;
  (void)frexp(phi, &e);
  const int se = 53 - e < 1074 ? 53 - e : 1074;
  const double q = ldexp(d, se);
  const double n = floor(q);
  double S = q - n < 0.5 ? n : n + 1.0;
  if (q - n == 0.5) {
    // half-way: round-half-even adds the even one of n, n + 1 once P = phi / u is even
    if (fmod(ldexp(phi, se), 2.0) != 0.0) return;
    S = fmod(n, 2.0) == 0.0 ? n : n + 1.0;
  }
  if (S == 0.0) {
    *run = UINT64_MAX;
    return;
  }
  const double P = e < 1024 ? ldexp(ldexp(1.0, e) - phi, se) : ldexp(ldexp(1.0, e - 1) - 0.5 * phi, se + 1);
  uint64_t m = (uint64_t)floor((P - 1.0) / S);
  if (m * (uint64_t)S > (uint64_t)P - 1) --m;
  *step = ldexp(S, -se);
  *run = m;
}

}  // namespace fftfree
